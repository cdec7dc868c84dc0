"""iactrace_amd -- MI355X (gfx950) implementation of iactrace's ray-tracing path.

Drop-in surface of the reference (iactrace/__init__.py:1-19): ``Telescope``,
``Mirror``, ``Integrator``, ``MCIntegrator``, ``SquareSensor``,
``HexagonalSensor``, ``load_telescope``; plus the renderer entry points
``render``, ``render_debug``, ``render_response_matrix`` (render.py) and the
batched / differentiable extensions.  Every render runs in the HIP kernels of
libiactrace_amd.so; there is no other backend.
"""
from . import filtering
from .autotune import autotune
from .filtering import make_filter, partition, show_structure, trainable_params
from .batched import AlignmentFit, pack_sources, render_batched
from .geometry import (AsphericSurface, DiskAperture, PolygonAperture, euler_to_matrix,
                       look_at_euler, look_at_rotation)
from .integrators import Integrator, MCIntegrator, split_key
from .mirrors import (AsphericDiskMirrorGroup, AsphericPolygonMirrorGroup, Mirror, MirrorGroup,
                      group_mirrors)
from .obstructions import (Box, BoxGroup, Cylinder, CylinderGroup, OrientedBox,
                           OrientedBoxGroup, Sphere, SphereGroup, Triangle, TriangleGroup,
                           group_obstructions)
from .render import LaunchConfig, render, render_debug, render_response_matrix
from .sensors import (DifferentiableHexagonalSensor, DifferentiableSquareSensor,
                      HexagonalSensor, SquareSensor)
from .telescope import Telescope
from .yaml_loader import build_telescope, load_telescope

__version__ = "0.4.0"

__all__ = [
    "Telescope", "Mirror", "Integrator", "MCIntegrator", "SquareSensor", "HexagonalSensor",
    "DifferentiableSquareSensor", "DifferentiableHexagonalSensor", "load_telescope",
    "build_telescope", "render", "render_debug", "render_response_matrix", "LaunchConfig",
    "AsphericSurface", "DiskAperture", "PolygonAperture", "MirrorGroup",
    "AsphericDiskMirrorGroup", "AsphericPolygonMirrorGroup", "group_mirrors", "Cylinder",
    "CylinderGroup", "Box", "BoxGroup", "Sphere", "SphereGroup", "OrientedBox",
    "OrientedBoxGroup", "Triangle", "TriangleGroup", "group_obstructions", "euler_to_matrix",
    "look_at_euler", "look_at_rotation", "split_key", "render_batched", "pack_sources",
    "AlignmentFit", "autotune", "make_filter", "partition", "trainable_params",
    "show_structure", "filtering",
]
