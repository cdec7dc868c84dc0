"""Obstructions (iactrace/core/obstructions.py:16-277).

Single primitives and per-type groups.  Groups pack themselves into the fp32
tables the kernels read (row layouts in include/iactrace_amd.h); per-primitive
constants (cylinder unit axis and height, box min/max) are hoisted here in
fp32 with the reference's operation order.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ["Obstruction", "ObstructionGroup", "Cylinder", "CylinderGroup", "Box", "BoxGroup",
           "Sphere", "SphereGroup", "OrientedBox", "OrientedBoxGroup", "Triangle",
           "TriangleGroup", "group_obstructions"]

F32 = np.float32


class Obstruction:
    """Base class of single obstructions."""


class ObstructionGroup:
    """Base class of grouped obstructions: ``table()`` gives the packed rows."""
    kind = None

    def __len__(self):
        raise NotImplementedError

    def table(self) -> np.ndarray:
        raise NotImplementedError


def _v3(x):
    return np.asarray(x, F32).reshape(3)


@dataclass(eq=False)
class Cylinder(Obstruction):
    """Capped cylinder from p1 to p2 (obstructions.py:38-51)."""
    p1: np.ndarray
    p2: np.ndarray
    radius: float

    def __post_init__(self):
        self.p1, self.p2, self.radius = _v3(self.p1), _v3(self.p2), float(self.radius)


class CylinderGroup(ObstructionGroup):
    """obstructions.py:54-79.  Row: p1[3], unit axis[3], height, radius."""
    kind = "cylinder"

    def __init__(self, cylinders=None, p1=None, p2=None, r=None):
        if cylinders is not None:
            self.p1 = np.stack([c.p1 for c in cylinders]).astype(F32)
            self.p2 = np.stack([c.p2 for c in cylinders]).astype(F32)
            self.r = np.array([c.radius for c in cylinders], F32)
        else:
            self.p1 = np.asarray(p1, F32).reshape(-1, 3)
            self.p2 = np.asarray(p2, F32).reshape(-1, 3)
            self.r = np.asarray(r, F32).reshape(-1)

    def __len__(self):
        return self.p1.shape[0]

    def table(self):
        axis = self.p2 - self.p1
        h = np.sqrt(axis[:, 0] * axis[:, 0] + axis[:, 1] * axis[:, 1] + axis[:, 2] * axis[:, 2])
        unit = axis / h[:, None]
        return np.concatenate([self.p1, unit, h[:, None], self.r[:, None]], 1).astype(F32)


@dataclass(eq=False)
class Box(Obstruction):
    """Axis-aligned box between corners p1, p2 (obstructions.py:84-95)."""
    p1: np.ndarray
    p2: np.ndarray

    def __post_init__(self):
        self.p1, self.p2 = _v3(self.p1), _v3(self.p2)


class BoxGroup(ObstructionGroup):
    """obstructions.py:98-120.  Row: min[3], max[3]."""
    kind = "box"

    def __init__(self, boxes=None, p1=None, p2=None):
        if boxes is not None:
            self.p1 = np.stack([b.p1 for b in boxes]).astype(F32)
            self.p2 = np.stack([b.p2 for b in boxes]).astype(F32)
        else:
            self.p1 = np.asarray(p1, F32).reshape(-1, 3)
            self.p2 = np.asarray(p2, F32).reshape(-1, 3)

    def __len__(self):
        return self.p1.shape[0]

    def table(self):
        return np.concatenate([np.minimum(self.p1, self.p2), np.maximum(self.p1, self.p2)],
                              1).astype(F32)


@dataclass(eq=False)
class Sphere(Obstruction):
    """Sphere (obstructions.py:125-136)."""
    center: np.ndarray
    radius: float

    def __post_init__(self):
        self.center, self.radius = _v3(self.center), float(self.radius)


class SphereGroup(ObstructionGroup):
    """obstructions.py:139-161.  Row: center[3], radius."""
    kind = "sphere"

    def __init__(self, spheres=None, centers=None, radii=None):
        if spheres is not None:
            self.centers = np.stack([s.center for s in spheres]).astype(F32)
            self.radii = np.array([s.radius for s in spheres], F32)
        else:
            self.centers = np.asarray(centers, F32).reshape(-1, 3)
            self.radii = np.asarray(radii, F32).reshape(-1)

    def __len__(self):
        return self.centers.shape[0]

    def table(self):
        return np.concatenate([self.centers, self.radii[:, None]], 1).astype(F32)


@dataclass(eq=False)
class OrientedBox(Obstruction):
    """Oriented box: centre, half extents, local->world rotation (obstructions.py:166-181)."""
    center: np.ndarray
    half_extents: np.ndarray
    rotation: np.ndarray

    def __post_init__(self):
        self.center, self.half_extents = _v3(self.center), _v3(self.half_extents)
        self.rotation = np.asarray(self.rotation, F32).reshape(3, 3)


class OrientedBoxGroup(ObstructionGroup):
    """obstructions.py:184-209.  Row: center[3], half[3], rotation[9] row-major."""
    kind = "oriented_box"

    def __init__(self, boxes=None, centers=None, half_extents=None, rotations=None):
        if boxes is not None:
            self.centers = np.stack([b.center for b in boxes]).astype(F32)
            self.half_extents = np.stack([b.half_extents for b in boxes]).astype(F32)
            self.rotations = np.stack([b.rotation for b in boxes]).astype(F32)
        else:
            self.centers = np.asarray(centers, F32).reshape(-1, 3)
            self.half_extents = np.asarray(half_extents, F32).reshape(-1, 3)
            self.rotations = np.asarray(rotations, F32).reshape(-1, 3, 3)

    def __len__(self):
        return self.centers.shape[0]

    def table(self):
        return np.concatenate([self.centers, self.half_extents,
                               self.rotations.reshape(-1, 9)], 1).astype(F32)


@dataclass(eq=False)
class Triangle(Obstruction):
    """Triangle (obstructions.py:214-227)."""
    v0: np.ndarray
    v1: np.ndarray
    v2: np.ndarray

    def __post_init__(self):
        self.v0, self.v1, self.v2 = _v3(self.v0), _v3(self.v1), _v3(self.v2)


class TriangleGroup(ObstructionGroup):
    """obstructions.py:230-255.  Row: v0[3], v1[3], v2[3]."""
    kind = "triangle"

    def __init__(self, triangles=None, v0=None, v1=None, v2=None):
        if triangles is not None:
            self.v0 = np.stack([t.v0 for t in triangles]).astype(F32)
            self.v1 = np.stack([t.v1 for t in triangles]).astype(F32)
            self.v2 = np.stack([t.v2 for t in triangles]).astype(F32)
        else:
            self.v0 = np.asarray(v0, F32).reshape(-1, 3)
            self.v1 = np.asarray(v1, F32).reshape(-1, 3)
            self.v2 = np.asarray(v2, F32).reshape(-1, 3)

    def __len__(self):
        return self.v0.shape[0]

    def table(self):
        return np.concatenate([self.v0, self.v1, self.v2], 1).astype(F32)


_ORDER = [(Cylinder, CylinderGroup), (Box, BoxGroup), (Sphere, SphereGroup),
          (OrientedBox, OrientedBoxGroup), (Triangle, TriangleGroup)]


def group_obstructions(obstructions):
    """One group per primitive type, in the reference order (obstructions.py:258-277)."""
    groups = []
    for single, grp in _ORDER:
        items = [o for o in obstructions if isinstance(o, single)]
        if items:
            groups.append(grp(items))
    return groups


def pack_obstruction_tables(groups):
    """Concatenate the groups of each type into one table per type."""
    out = {k: [] for k in ("cylinder", "box", "sphere", "oriented_box", "triangle")}
    for g in groups or []:
        if g.kind not in out:
            raise ValueError(f"unknown obstruction group {type(g).__name__}")
        if len(g):
            out[g.kind].append(g.table())
    width = {"cylinder": 8, "box": 6, "sphere": 4, "oriented_box": 15, "triangle": 9}
    return {k: (np.concatenate(v, 0) if v else np.zeros((0, width[k]), F32)) for k, v in out.items()}
