"""Pupil-sampling integrators (iactrace/core/integrators.py).

``MCIntegrator.sample_group`` fills a group's local-frame samples (points,
normals, unit roughness deltas, weights cos_z * M / area) with the HIP
sampling kernel.  Random streams are this build's counter-based Philox4x32-10
keyed by the integer ``key``; the reference's jax.random streams are not
reproduced (JAX is not part of this stack), so identical seeds give
statistically equivalent -- not identical -- samples.  Externally produced
samples can be injected with ``MirrorGroup.replace(points=..., ...)``.
"""
from __future__ import annotations

import ctypes
from abc import ABC, abstractmethod

import torch

from . import _lib
from .mirrors import AsphericDiskMirrorGroup, AsphericPolygonMirrorGroup, MirrorGroup

__all__ = ["Integrator", "MCIntegrator", "split_key", "as_key"]

_M64 = (1 << 64) - 1


def as_key(key):
    """Normalise a key (None, int, numpy/torch integer) to a uint64 int."""
    if key is None:
        return 0
    if isinstance(key, torch.Tensor):
        key = int(key.item())
    return int(key) & _M64


def _splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & _M64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def split_key(key, num=2):
    """Derive ``num`` independent keys from one (the analogue of jax.random.split)."""
    k = as_key(key)
    return [_splitmix64(k ^ _splitmix64(i + 1)) for i in range(num)]


class Integrator(ABC):
    """Abstract sampler of mirror groups (integrators.py:18-55)."""

    @abstractmethod
    def sample_group(self, group: MirrorGroup, key) -> MirrorGroup:
        ...

    def sample_mirror_groups(self, mirror_groups, key):
        if not mirror_groups:
            return []
        keys = split_key(key, len(mirror_groups) + 1)
        return [self.sample_group(g, k) for g, k in zip(mirror_groups, keys[:-1])]


class MCIntegrator(Integrator):
    """Monte Carlo integrator with ``n_samples`` samples per facet
    (integrators.py:58-95)."""

    def __init__(self, n_samples: int = 128):
        self.n_samples = int(n_samples)

    def sample_group(self, group: MirrorGroup, key) -> MirrorGroup:
        dev = group.positions.device
        if dev.type != "cuda":
            raise RuntimeError("MCIntegrator samples on the HIP device; move the group to "
                               "'cuda' first (Telescope.to('cuda'))")
        n, m = len(group), self.n_samples
        D = _lib.IactSampleDesc()
        D.seed = as_key(key)
        D.group = 0
        D.n_facets = n
        D.n_samples = m
        D.surface = group.surface_abi()
        offsets = group.offsets.detach().float().contiguous()
        keep = [offsets]
        D.offsets = offsets.data_ptr()
        if isinstance(group, AsphericDiskMirrorGroup):
            D.aperture = _lib.APERTURE_DISK
            radii = group.radii.detach().float().contiguous()
            keep.append(radii)
            D.radii = radii.data_ptr()
        elif isinstance(group, AsphericPolygonMirrorGroup):
            D.aperture = _lib.APERTURE_POLYGON
            verts = group.vertices.detach().float().contiguous()
            keep.append(verts)
            D.vertices = verts.data_ptr()
            D.n_vertices = verts.shape[1]
        else:
            raise TypeError(f"Unknown MirrorGroup type: {type(group).__name__}")
        pts = torch.empty((n, m, 3), dtype=torch.float32, device=dev)
        nrm = torch.empty((n, m, 3), dtype=torch.float32, device=dev)
        dlt = torch.empty((n, m, 3), dtype=torch.float32, device=dev)
        w = torch.empty((n, m, 1), dtype=torch.float32, device=dev)
        _lib.check(_lib.lib().iact_sample_facets(ctypes.byref(D), _lib.ptr(pts), _lib.ptr(nrm),
                                                 _lib.ptr(dlt), _lib.ptr(w),
                                                 _lib.current_stream()), "iact_sample_facets")
        return group.replace(points=pts, normals=nrm, perturbation_delta=dlt, weights=w)
