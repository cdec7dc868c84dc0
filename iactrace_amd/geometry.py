"""Host-side geometry: rotations, surfaces, apertures (set-up time only).

These compute the launch-constant numbers the kernels read (sensor frames,
stage>=1 mirror frames) in fp32 with the reference's operation order, and
the look-at helpers used by the telescope generators.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

__all__ = ["euler_to_matrix", "look_at_rotation", "look_at_euler", "AsphericSurface",
           "DiskAperture", "PolygonAperture"]


def euler_to_matrix(tip_tilt_rotation):
    """R = Rz @ Ry @ Rx from Euler degrees, fp32 (transforms.py:72-107)."""
    a = np.asarray(tip_tilt_rotation, np.float32).reshape(3)
    rx, ry, rz = a * np.float32(math.pi / 180.0)
    cx, sx = np.cos(rx), np.sin(rx)
    cy, sy = np.cos(ry), np.sin(ry)
    cz, sz = np.cos(rz), np.sin(rz)
    one, zero = np.float32(1), np.float32(0)
    Rx = np.array([[one, zero, zero], [zero, cx, -sx], [zero, sx, cx]], np.float32)
    Ry = np.array([[cy, zero, sy], [zero, one, zero], [-sy, zero, cy]], np.float32)
    Rz = np.array([[cz, -sz, zero], [sz, cz, zero], [zero, zero, one]], np.float32)
    return (Rz @ Ry @ Rx).astype(np.float32)


def look_at_rotation(mirror_pos, target_pos=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)):
    """Rotation whose columns are (right, up, forward) (transforms.py:4-24)."""
    f = np.asarray(target_pos, np.float64) - np.asarray(mirror_pos, np.float64)
    f = f / np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    r = r / np.linalg.norm(r)
    u = np.cross(r, f)
    return np.column_stack([r, u, f])


def look_at_euler(mirror_pos, target_pos=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)):
    """Euler degrees turning local +z toward target (transforms.py:27-69)."""
    f = np.asarray(target_pos, np.float64) - np.asarray(mirror_pos, np.float64)
    f = f / np.linalg.norm(f)
    r = np.cross(np.asarray(up, np.float64), f)
    r = r / np.linalg.norm(r)
    u = np.cross(f, r)
    R = np.column_stack([r, u, f])
    sy = -R[2, 0]
    cy = math.sqrt(max(1.0 - sy * sy, 0.0))
    if cy > 1e-6:
        tip, rot = math.atan2(R[2, 1], R[2, 2]), math.atan2(R[1, 0], R[0, 0])
    else:
        tip, rot = 0.0, math.atan2(-R[0, 1], R[1, 1])
    return np.rad2deg(np.array([tip, math.asin(sy), rot]))


@dataclass(eq=False)
class AsphericSurface:
    """Conic + even polynomial surface (surfaces.py:8-39).

    ``aspheric[i]`` multiplies (x^2+y^2)^(2(i+1)) -- r^4, r^8, r^12, ... -- as
    the reference's ``jnp.arange(2, 2+2K, 2)`` powers of r^2 do.  The
    coefficients may be a tensor with ``requires_grad``.
    """
    curvature: float
    conic: float
    aspheric: torch.Tensor = field(default_factory=lambda: torch.zeros(0))

    def __post_init__(self):
        self.curvature = float(self.curvature)
        self.conic = float(self.conic)
        if not isinstance(self.aspheric, torch.Tensor):
            self.aspheric = torch.as_tensor(np.asarray(self.aspheric, np.float32).reshape(-1))

    @staticmethod
    def from_template(tmpl):
        """Surface from a YAML template dict (surfaces.py:15-23)."""
        s = tmpl["surface"]
        return AsphericSurface(float(s["curvature"]), float(s["conic"]),
                               torch.tensor(s.get("aspheric", []) or [], dtype=torch.float32))

    def key(self):
        return (self.curvature, self.conic, tuple(float(a) for a in self.aspheric.detach().cpu()))

    def sag_raw(self, x, y):
        """Sag z(x, y) without offset (numpy float64, for set-up and checks)."""
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        r2 = x * x + y * y
        c, k = self.curvature, self.conic
        z = r2 * c / (1 + np.sqrt(1 - (1 + k) * c * c * r2))
        for i, a in enumerate(self.aspheric.detach().cpu().double().numpy()):
            z = z + a * r2 ** (2 * (i + 1))
        return z


@dataclass(eq=False)
class DiskAperture:
    """Circular aperture (apertures.py:22-34)."""
    radius: float = 1.0

    def __post_init__(self):
        self.radius = float(self.radius)

    def area(self):
        return math.pi * self.radius ** 2

    def check_aperture(self, x, y):
        return x ** 2 + y ** 2 <= self.radius ** 2


@dataclass(eq=False)
class PolygonAperture:
    """Convex polygon aperture, vertices (n, 2) (apertures.py:37-63)."""
    vertices: np.ndarray

    def __post_init__(self):
        self.vertices = np.asarray(self.vertices, np.float32).reshape(-1, 2)

    @property
    def n_vertices(self):
        return len(self.vertices)

    def area(self):
        x, y = self.vertices[:, 0], self.vertices[:, 1]
        return float(0.5 * abs(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)))

    def check_aperture(self, x, y):
        """Inside test; requires CCW vertices exactly like the reference."""
        v = self.vertices
        n = len(v)
        ok = np.ones(np.shape(x), bool)
        for i in range(n):
            a, b = v[i], v[(i + 1) % n]
            ok &= (b[0] - a[0]) * (np.asarray(y) - a[1]) - (b[1] - a[1]) * (np.asarray(x) - a[0]) >= 0
        return ok
