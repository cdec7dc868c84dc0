"""YAML telescope loader (iactrace/io/yaml_loader.py:29-172).

Same schema as the reference: ``telescope.name``, ``mirror_templates``,
``mirrors`` (template, position, orientation, aperture, optional stage and
offset), ``obstructions`` and ``sensors``.  Stage-0 groups are sampled on the
HIP device with keys split in sequence, one per sampled group.
"""
from __future__ import annotations

import torch
import yaml

from .geometry import AsphericSurface, DiskAperture, PolygonAperture
from .integrators import as_key, split_key
from .mirrors import Mirror, group_mirrors
from .obstructions import Box, Cylinder, OrientedBox, Sphere, Triangle, group_obstructions
from .sensors import HexagonalSensor, SquareSensor
from .telescope import Telescope

__all__ = ["load_telescope", "build_telescope"]


def load_telescope(filename, integrator, key=None, device=None):
    """Load a telescope from a YAML file and sample its primary facets.

    ``key`` is an integer seed (default 0).  ``device`` defaults to the first
    HIP device; pass ``integrator=None`` to parse without sampling.
    """
    with open(filename, "r") as f:
        config = yaml.safe_load(f)
    return build_telescope(config, integrator, key, device)


def build_telescope(config, integrator, key=None, device=None):
    """Telescope from a parsed config dict (yaml_loader.py:52-92)."""
    if device is None:
        device = "cuda" if integrator is not None else "cpu"
    name = config.get("telescope", {}).get("name", "telescope")
    templates = config.get("mirror_templates", {})
    mirrors = [_parse_mirror(m, templates) for m in config.get("mirrors", [])]
    groups = [g.to(device) for g in group_mirrors(mirrors)]
    key = as_key(key)
    sampled = []
    for g in groups:
        if g.optical_stage == 0 and integrator is not None:
            key, sub = split_key(key, 2)
            sampled.append(integrator.sample_group(g, sub))
        else:
            sampled.append(g)
    obstructions = group_obstructions([_parse_obstruction(o)
                                       for o in config.get("obstructions", [])])
    sensors = [_parse_sensor(s) for s in config.get("sensors", [])]
    return Telescope(sampled, obstructions, sensors, name)


def _parse_mirror(m, templates):
    aperture = _parse_aperture(m["aperture"])
    surface = AsphericSurface.from_template(templates[m["template"]])
    return Mirror(position=m["position"], rotation=m["orientation"], surface=surface,
                  aperture=aperture, optical_stage=m.get("stage", 0),
                  offset=m.get("offset", [0.0, 0.0]))


def _parse_aperture(cfg):
    atype = cfg["type"]
    if atype == "circular":
        return DiskAperture(cfg["radius"])
    if atype == "polygon":
        return PolygonAperture(cfg["vertices"])
    raise ValueError(f"Unknown aperture type: {atype}")


def _parse_obstruction(cfg):
    otype = cfg["type"]
    if otype == "cylinder":
        return Cylinder(cfg["p1"], cfg["p2"], cfg["r"])
    if otype == "box":
        return Box(cfg["p1"], cfg["p2"])
    if otype == "sphere":
        return Sphere(cfg["center"], cfg["r"])
    if otype == "oriented_box":
        return OrientedBox(cfg["center"], cfg["half_extents"], cfg["rotation"])
    if otype == "triangle":
        return Triangle(cfg["v0"], cfg["v1"], cfg["v2"])
    raise ValueError(f"Unknown obstruction type: {otype}")


def _parse_sensor(cfg):
    stype = cfg["type"]
    edge = cfg.get("edge_width", 0.0)
    if stype == "square":
        return SquareSensor(cfg["position"], cfg["orientation"], cfg["width"], cfg["height"],
                            tuple(cfg["bounds"]), edge)
    if stype == "hexagonal":
        centers = torch.tensor([cfg["centers_x"], cfg["centers_y"]], dtype=torch.float32).T
        return HexagonalSensor(cfg["position"], cfg["orientation"], centers.numpy(), edge)
    raise ValueError(f"Unknown sensor type: {stype}")
