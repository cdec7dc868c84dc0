"""Mirrors and mirror groups (iactrace/telescope/mirrors.py).

A group is a structure of arrays of facets that share surface parameters and
aperture type.  All per-facet arrays are torch tensors (fp32) so they can
live on the GPU and carry ``requires_grad`` (rotations, positions,
aspheric coefficients).  Groups are treated as immutable: the operations
module returns new groups via ``replace``.
"""
from __future__ import annotations

import dataclasses
from collections import defaultdict
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .geometry import AsphericSurface, DiskAperture, PolygonAperture

__all__ = ["Mirror", "MirrorGroup", "AsphericDiskMirrorGroup", "AsphericPolygonMirrorGroup",
           "group_mirrors"]


def _t(x, dtype=torch.float32):
    if isinstance(x, torch.Tensor):
        return x if x.dtype == dtype else x.to(dtype)
    return torch.as_tensor(np.asarray(x), dtype=dtype)


@dataclass(eq=False)
class Mirror:
    """One mirror: pose, surface, aperture, offset on the parent surface
    (mirrors.py:10-43).  ``rotation`` is Euler degrees (tip, tilt, rotation)."""
    position: object
    rotation: object
    surface: AsphericSurface
    aperture: object
    points: object = None
    normals: object = None
    weights: object = None
    optical_stage: int = 0
    offset: object = None

    def __post_init__(self):
        self.position = _t(self.position).reshape(3)
        self.rotation = _t(self.rotation).reshape(3)
        self.offset = torch.zeros(2) if self.offset is None else _t(self.offset).reshape(2)
        self.optical_stage = int(self.optical_stage)


@dataclass(eq=False)
class MirrorGroup:
    """Facets sharing one surface and aperture type (mirrors.py:46-98)."""
    positions: torch.Tensor                  # (N, 3)
    rotations: torch.Tensor                  # (N, 3) degrees
    curvature: float
    conic: float
    aspheric: torch.Tensor                   # (K,)
    offsets: torch.Tensor                    # (N, 2)
    optical_stage: int = 0
    points: torch.Tensor = None              # (N, M, 3)
    normals: torch.Tensor = None             # (N, M, 3)
    weights: torch.Tensor = None             # (N, M, 1)
    perturbation_delta: torch.Tensor = None  # (N, M, 3)
    perturbation_scale: torch.Tensor = None  # (N,)

    def __post_init__(self):
        n = self.positions.shape[0]
        dev = self.positions.device
        if self.points is None:
            self.points = torch.zeros((n, 0, 3), device=dev)
            self.normals = torch.zeros((n, 0, 3), device=dev)
            self.weights = torch.zeros((n, 0, 1), device=dev)
            self.perturbation_delta = torch.zeros((n, 0, 3), device=dev)
        if self.perturbation_scale is None:
            self.perturbation_scale = torch.zeros(n, device=dev)

    def __len__(self):
        return self.positions.shape[0]

    @property
    def n_samples(self):
        return self.points.shape[1]

    def replace(self, **kw):
        return dataclasses.replace(self, **kw)

    def get_surface(self):
        return AsphericSurface(self.curvature, self.conic, self.aspheric)

    def to(self, device):
        """Copy of the group with its per-facet arrays on ``device``.  Surface
        coefficients stay on the host: they are launch constants."""
        kw = {}
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            if isinstance(v, torch.Tensor) and f.name != "aspheric":
                kw[f.name] = v.to(device)
        return self.replace(**kw)

    def surface_abi(self):
        s = _lib.IactSurface()
        s.curvature = self.curvature
        s.conic = self.conic
        a = self.aspheric.detach().float().cpu().reshape(-1)
        if a.numel() > _lib.IACT_MAX_ASPHERIC:
            raise ValueError(f"{a.numel()} aspheric terms (max {_lib.IACT_MAX_ASPHERIC})")
        s.n_aspheric = a.numel()
        for i, v in enumerate(a.tolist()):
            s.aspheric[i] = v
        return s

    def transform_to_world(self):
        """World points, perturbed unit normals and weights of every sample,
        computed by the HIP world-transform kernel (mirrors.py:64-79)."""
        import ctypes
        from .render import facets_abi
        if self.positions.device.type != "cuda":
            raise RuntimeError("transform_to_world runs on the HIP device; call .to('cuda')")
        n, m = len(self), self.n_samples
        keep = []
        F = facets_abi([self], keep)
        out = torch.empty((7, n * m), dtype=torch.float32, device=self.positions.device)
        _lib.check(_lib.lib().iact_facet_world(ctypes.byref(F), _lib.ptr(out),
                                               _lib.current_stream()), "iact_facet_world")
        p = out[0:3].T.reshape(n, m, 3)
        nn = out[3:6].T.reshape(n, m, 3)
        w = out[6].reshape(n, m, 1)
        return p, nn, w


@dataclass(eq=False)
class AsphericDiskMirrorGroup(MirrorGroup):
    """Group with circular apertures (mirrors.py:101-158)."""
    radii: torch.Tensor = None               # (N,)

    def check_aperture(self, x, y, mirror_idx):
        return x ** 2 + y ** 2 <= self.radii[mirror_idx] ** 2

    def get_sampling_params(self):
        return {"type": "disk", "radii": self.radii, "offsets": self.offsets,
                "surface": self.get_surface()}

    def areas(self):
        return np.pi * self.radii.detach().double().cpu().numpy() ** 2


@dataclass(eq=False)
class AsphericPolygonMirrorGroup(MirrorGroup):
    """Group with convex polygon apertures sharing a vertex count
    (mirrors.py:161-229).  Inside tests need CCW vertices."""
    vertices: torch.Tensor = None            # (N, K, 2)

    @property
    def n_vertices(self):
        return self.vertices.shape[1]

    def check_aperture(self, x, y, mirror_idx):
        v = self.vertices[mirror_idx]
        n = v.shape[0]
        ok = torch.ones_like(x, dtype=torch.bool)
        for i in range(n):
            a, b = v[i], v[(i + 1) % n]
            ok &= (b[0] - a[0]) * (y - a[1]) - (b[1] - a[1]) * (x - a[0]) >= 0
        return ok

    def get_sampling_params(self):
        return {"type": "polygon", "vertices": self.vertices, "offsets": self.offsets,
                "surface": self.get_surface()}

    def areas(self):
        v = self.vertices.detach().double().cpu().numpy()
        x, y = v[..., 0], v[..., 1]
        return 0.5 * np.abs(np.sum(x * np.roll(y, -1, -1) - np.roll(x, -1, -1) * y, -1))


def _group_by_surface(mirrors):
    out = defaultdict(list)
    for m in mirrors:
        out[m.surface.key()].append(m)
    return out


def group_mirrors(mirrors):
    """Group by stage, aperture type (and vertex count), then identical surface
    parameters (mirrors.py:232-336)."""
    groups = []
    by_stage = defaultdict(list)
    for m in mirrors:
        by_stage[m.optical_stage].append(m)
    for stage, ms in sorted(by_stage.items()):
        disk = [m for m in ms if isinstance(m.aperture, DiskAperture)]
        for _, lst in _group_by_surface(disk).items():
            s = lst[0].surface
            groups.append(AsphericDiskMirrorGroup(
                positions=torch.stack([m.position for m in lst]),
                rotations=torch.stack([m.rotation for m in lst]),
                curvature=s.curvature, conic=s.conic, aspheric=s.aspheric.clone(),
                offsets=torch.stack([m.offset for m in lst]), optical_stage=stage,
                radii=torch.tensor([m.aperture.radius for m in lst], dtype=torch.float32)))
        poly = [m for m in ms if isinstance(m.aperture, PolygonAperture)]
        by_n = defaultdict(list)
        for m in poly:
            by_n[m.aperture.n_vertices].append(m)
        for _, lst0 in by_n.items():
            for _, lst in _group_by_surface(lst0).items():
                s = lst[0].surface
                groups.append(AsphericPolygonMirrorGroup(
                    positions=torch.stack([m.position for m in lst]),
                    rotations=torch.stack([m.rotation for m in lst]),
                    curvature=s.curvature, conic=s.conic, aspheric=s.aspheric.clone(),
                    offsets=torch.stack([m.offset for m in lst]), optical_stage=stage,
                    vertices=torch.as_tensor(np.stack([m.aperture.vertices for m in lst]),
                                             dtype=torch.float32)))
    return groups
