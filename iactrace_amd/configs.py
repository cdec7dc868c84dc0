"""Telescope generators and the benchmark workloads.

``TelescopeConfigBuilder`` writes the YAML schema ``load_telescope`` reads
(configs/builder.py:5-125).  The generators below build the five benchmark
telescopes programmatically (Davies-Cotton and parabolic segmented dishes
with hexagonal facets aimed by look-at angles as CreateHESSConfig.ipynb does,
a two-mirror Cassegrain, synthetic hexagonal cameras) and the matching
synthetic source fields.  All randomness is numpy PCG64 with fixed seeds.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch
import yaml

from .geometry import AsphericSurface, DiskAperture, PolygonAperture, look_at_euler
from .mirrors import Mirror, group_mirrors
from .obstructions import Box, Cylinder, group_obstructions
from .sensors import HexagonalSensor, SquareSensor
from .telescope import Telescope

__all__ = ["TelescopeConfigBuilder", "hex_lattice", "hex_facet_vertices", "davies_cotton_dish",
           "parabolic_dish", "hex_camera", "cassegrain", "point_source_field", "Workload",
           "workload", "BatchWorkload", "batch_workload", "parallel_grid", "support_truss"]


class TelescopeConfigBuilder:
    """Programmatic writer of telescope YAML files (configs/builder.py:5-125)."""

    def __init__(self, name, units="m"):
        self.config = {"telescope": {"name": name, "units": units}, "mirror_templates": {},
                       "mirrors": [], "obstructions": [], "sensors": []}

    def add_mirror_template(self, name, curvature, conic, aspheric_coeffs=()):
        self.config["mirror_templates"][name] = {"surface": {
            "curvature": float(curvature), "conic": float(conic),
            "aspheric": [float(a) for a in aspheric_coeffs]}}
        return self

    def add_mirror_circular(self, mirror_id, template, position, orientation, radius, stage=0,
                            offset=None):
        m = {"id": mirror_id, "template": template,
             "position": [float(x) for x in position],
             "orientation": [float(x) for x in orientation],
             "aperture": {"type": "circular", "radius": float(radius)}}
        if stage:
            m["stage"] = int(stage)
        if offset is not None:
            m["offset"] = [float(x) for x in offset]
        self.config["mirrors"].append(m)
        return self

    def add_mirror_polygon(self, mirror_id, template, position, orientation, vertices, stage=0,
                           offset=None):
        m = {"id": mirror_id, "template": template,
             "position": [float(x) for x in position],
             "orientation": [float(x) for x in orientation],
             "aperture": {"type": "polygon",
                          "vertices": [[float(x), float(y)] for x, y in vertices]}}
        if stage:
            m["stage"] = int(stage)
        if offset is not None:
            m["offset"] = [float(x) for x in offset]
        self.config["mirrors"].append(m)
        return self

    def add_obstruction_box(self, obs_id, p1, p2):
        self.config["obstructions"].append({"id": obs_id, "type": "box",
                                            "p1": [float(x) for x in p1],
                                            "p2": [float(x) for x in p2]})
        return self

    def add_obstruction_cylinder(self, obs_id, p1, p2, radius):
        self.config["obstructions"].append({"id": obs_id, "type": "cylinder",
                                            "p1": [float(x) for x in p1],
                                            "p2": [float(x) for x in p2], "r": float(radius)})
        return self

    def add_obstruction_sphere(self, obs_id, center, radius):
        self.config["obstructions"].append({"id": obs_id, "type": "sphere",
                                            "center": [float(x) for x in center],
                                            "r": float(radius)})
        return self

    def add_square_sensor_array(self, sensor_id, position, orientation, width, height, bounds,
                                edge_width=0.0):
        self.config["sensors"].append({"id": sensor_id, "type": "square",
                                       "position": [float(x) for x in position],
                                       "orientation": [float(x) for x in orientation],
                                       "width": int(width), "height": int(height),
                                       "bounds": [float(x) for x in bounds],
                                       "edge_width": float(edge_width)})
        return self

    def add_hexagon_sensor_array(self, sensor_id, position, orientation, pixel_x, pixel_y,
                                 edge_width=0.0):
        self.config["sensors"].append({"id": sensor_id, "type": "hexagonal",
                                       "position": [float(x) for x in position],
                                       "orientation": [float(x) for x in orientation],
                                       "centers_x": [float(x) for x in pixel_x],
                                       "centers_y": [float(y) for y in pixel_y],
                                       "edge_width": float(edge_width)})
        return self

    def dumps(self):
        return yaml.safe_dump(self.config, default_flow_style=False, sort_keys=False)

    def save(self, filename):
        with open(filename, "w") as f:
            f.write(self.dumps())


# --------------------------------------------------------------------- lattices

def hex_lattice(n, pitch):
    """The n points of a triangular lattice (spacing ``pitch``, one axis along x)
    nearest the origin, ordered by radius then azimuth."""
    k = int(math.ceil(math.sqrt(n))) + 3
    pts = []
    for j in range(-k, k + 1):
        for i in range(-k, k + 1):
            pts.append((pitch * (i + 0.5 * j), pitch * (math.sqrt(3) / 2) * j))
    pts = np.array(pts)
    r = np.round(np.hypot(pts[:, 0], pts[:, 1]) / pitch, 9)
    a = np.round(np.mod(np.arctan2(pts[:, 1], pts[:, 0]), 2 * np.pi), 9)
    order = np.lexsort((a, r))
    return pts[order[:n]]


def hex_facet_vertices(flat_to_flat):
    """Hexagon vertices (R sin a, R cos a), a = 0, 60, ... deg, R = D/sqrt(3),
    as CreateHESSConfig.ipynb builds the CT5 facets (clockwise order)."""
    R = flat_to_flat / math.sqrt(3)
    ang = np.deg2rad(np.arange(0, 360, 60))
    return np.stack([R * np.sin(ang), R * np.cos(ang)], 1)


def _dish(centres_xy, z_of_r, aim_z_of_r, flat_to_flat, curvature, name):
    verts = hex_facet_vertices(flat_to_flat)
    surf = AsphericSurface(curvature, 0.0, [])
    mirrors = []
    for x, y in centres_xy:
        r2 = x * x + y * y
        pos = np.array([x, y, z_of_r(r2)])
        rot = look_at_euler(pos, np.array([0.0, 0.0, aim_z_of_r(r2)]))
        rot[2] = 0.0
        mirrors.append(Mirror(pos, rot, surf, PolygonAperture(verts)))
    return mirrors


def davies_cotton_dish(n_facets=382, focal=15.0, pitch=0.61, flat_to_flat=0.60,
                       curvature=1.0 / 30.0):
    """Davies-Cotton dish: centres on the sphere z = f - sqrt(f^2 - r^2), each
    facet aimed at (0, 0, 2f)."""
    xy = hex_lattice(n_facets, pitch)
    return _dish(xy, lambda r2: focal - math.sqrt(focal * focal - r2), lambda r2: 2 * focal,
                 flat_to_flat, curvature, "dc")


def parabolic_dish(n_facets=875, focal=36.0, pitch=0.91, flat_to_flat=0.90, curvature=0.0137):
    """Parabolic dish: centres on z = r^2/(4f), each facet aimed at
    (0, 0, 2f + r^2/(4f)) (CreateHESSConfig.ipynb's CT5 layout)."""
    xy = hex_lattice(n_facets, pitch)
    return _dish(xy, lambda r2: r2 / (4 * focal), lambda r2: 2 * focal + r2 / (4 * focal),
                 flat_to_flat, curvature, "parabolic")


def hex_camera(n_pixels=2048, pitch=0.04193673807057477):
    """Centres of a round hexagonal-pixel camera (pointy-top lattice)."""
    return hex_lattice(n_pixels, pitch).astype(np.float32)


CASSEGRAIN_SECONDARY_CONIC = -16.0
"""Conic of the c = -0.05 secondary whose foci are the primary focus (z = 10 m)
and z = -2/3 m: for vertex curvature -1/20 at z = 6 the exact hyperboloid has
K = -(5.3333/1.3333)^2 = -16."""


def cassegrain(n_samples_hint=None, seed=5, sensor_z=-0.45, sensor_half=0.3, sensor_px=64):
    """Two-mirror telescope: aspheric disk primary (r = 2 m, c = 0.05, K = -1.05,
    four aspheric terms ~ N(0, 1e-9 r_max^-4(i+1))), hyperbolic secondary at
    z = 6 m rotated 180 deg, square sensor behind the primary."""
    rng = np.random.default_rng(seed)
    rmax = 2.0
    asph = np.array([rng.normal(0.0, 1e-9 * rmax ** (-4 * (i + 1))) for i in range(4)],
                    np.float32)
    prim = Mirror([0, 0, 0], [0, 0, 0], AsphericSurface(0.05, -1.05, asph), DiskAperture(rmax))
    sec = Mirror([0, 0, 6.0], [180.0, 0, 0],
                 AsphericSurface(-0.05, CASSEGRAIN_SECONDARY_CONIC, []), DiskAperture(0.5),
                 optical_stage=1)
    sensor = SquareSensor([0, 0, sensor_z], [0, 0, 0], sensor_px, sensor_px,
                          (-sensor_half, sensor_half, -sensor_half, sensor_half))
    return [prim, sec], sensor


# --------------------------------------------------------------------- sources

def point_source_field(n, radius_deg, seed, z=1e10):
    """n point sources uniform in a disk of ``radius_deg`` on the sky at
    distance z (angle = radius sqrt(U), azimuth 2 pi U), lognormal(0,1) values."""
    rng = np.random.default_rng(seed)
    ang = np.deg2rad(radius_deg) * np.sqrt(rng.random(n))
    az = 2 * np.pi * rng.random(n)
    t = np.tan(ang)
    src = np.stack([z * t * np.cos(az), z * t * np.sin(az), np.full(n, z)], 1)
    vals = np.exp(rng.normal(0.0, 1.0, n)).astype(np.float32)
    return src, vals


def parallel_grid(nx, ny, spacing_deg):
    """Unit directions normalize(x, y, -1) on a regular grid of angles
    (ResponseMatrix.ipynb cell 9), x fastest."""
    sx = np.deg2rad(spacing_deg) * (np.arange(nx) - (nx - 1) / 2)
    sy = np.deg2rad(spacing_deg) * (np.arange(ny) - (ny - 1) / 2)
    X, Y = np.meshgrid(sx, sy, indexing="xy")
    d = np.stack([X.ravel(), Y.ravel(), -np.ones(X.size)], 1)
    return d / np.linalg.norm(d, axis=1, keepdims=True)


# --------------------------------------------------------------------- workloads

@dataclass
class Workload:
    """A benchmark configuration: telescope parts plus sources."""
    name: str
    mirrors: list
    obstructions: list
    sensor: object
    n_samples: int
    sources: np.ndarray
    values: np.ndarray
    source_type: str
    tilt_sigma_deg: float = 0.0
    tilt_seed: int = 0

    @property
    def n_facets(self):
        return sum(1 for m in self.mirrors if m.optical_stage == 0)

    @property
    def n_rays(self):
        return self.n_facets * self.n_samples * len(self.sources)

    def telescope(self, device="cuda", key=0, integrator=None):
        """Build and sample the telescope (tilt errors applied to stage 0)."""
        from .integrators import MCIntegrator
        groups = [g.to(device) for g in group_mirrors(self.mirrors)]
        integ = integrator or MCIntegrator(self.n_samples)
        from .integrators import split_key
        out = []
        k = key
        for g in groups:
            if g.optical_stage == 0:
                k, sub = split_key(k, 2)
                g = integ.sample_group(g, sub)
                if self.tilt_sigma_deg > 0:
                    rng = np.random.default_rng(self.tilt_seed)
                    d = rng.normal(0.0, self.tilt_sigma_deg, (len(g), 2))
                    rot = g.rotations.detach().clone()
                    rot[:, 0:2] += torch.as_tensor(d, dtype=torch.float32, device=rot.device)
                    g = g.replace(rotations=rot)
            out.append(g)
        return Telescope(out, group_obstructions(self.obstructions), [self.sensor], self.name)


def _c3_obstructions(focal=36.0, n_facets=875, pitch=0.91):
    box = Box([-1.0, -1.0, 36.25], [1.0, 1.0, 38.75])
    rim = np.max(np.hypot(*hex_lattice(n_facets, pitch).T))
    cyl = []
    for k in range(4):
        a = np.deg2rad(45.0 + 90.0 * k)
        p1 = [rim * np.cos(a), rim * np.sin(a), rim * rim / (4 * focal)]
        p2 = [np.sign(np.cos(a)) * 1.0, np.sign(np.sin(a)) * 1.0, 36.25]
        cyl.append(Cylinder(p1, p2, 0.1))
    return cyl + [box]


def support_truss(focal=36.0, n_facets=875, pitch=0.91, seed=7):
    """A camera support structure of CT5's size and kind (270 cylinders and
    the camera body): 4 masts from the dish rim to the camera, each in 16
    pieces; cross bracing between neighbouring masts at 13 levels (a
    horizontal and two diagonals per side and level, 156 members); 44 short
    camera-mount struts and 6 camera supports around the camera."""
    rng = np.random.default_rng(seed)
    rim = np.max(np.hypot(*hex_lattice(n_facets, pitch).T))
    z_cam = 36.25
    obs = [Box([-1.5, -1.45, z_cam], [1.5, 1.45, z_cam + 1.1])]
    base = []
    top = []
    for k in range(4):
        a = np.deg2rad(45.0 + 90.0 * k)
        base.append(np.array([rim * np.cos(a), rim * np.sin(a), rim * rim / (4 * focal)]))
        top.append(np.array([1.6 * np.cos(a), 1.6 * np.sin(a), z_cam - 0.5]))
    # masts in pieces
    for k in range(4):
        pts = [base[k] + (top[k] - base[k]) * t for t in np.linspace(0, 1, 17)]
        for i in range(16):
            obs.append(Cylinder(pts[i], pts[i + 1], 0.162 - 0.005 * i))
    # bracing between neighbouring masts
    for lev in range(13):
        t0, t1 = (lev + 1) / 15.0, (lev + 2) / 15.0
        for k in range(4):
            k2 = (k + 1) % 4
            a0 = base[k] + (top[k] - base[k]) * t0
            b0 = base[k2] + (top[k2] - base[k2]) * t0
            a1 = base[k] + (top[k] - base[k]) * t1
            b1 = base[k2] + (top[k2] - base[k2]) * t1
            r = 0.03 + 0.03 * rng.random()
            obs += [Cylinder(a0, b0, r), Cylinder(a0, b1, r), Cylinder(b0, a1, r)]
    # camera mount: a ring of short struts below the camera, spokes to it
    ring = [np.array([2.2 * np.cos(a), 2.2 * np.sin(a), z_cam - 0.8])
            for a in np.linspace(0, 2 * np.pi, 23)[:-1]]
    for i in range(22):
        obs.append(Cylinder(ring[i], ring[(i + 1) % 22], 0.05))
        corner = np.array([np.sign(ring[i][0]) * 1.4, np.sign(ring[i][1]) * 1.4, z_cam])
        obs.append(Cylinder(ring[i], corner, 0.04))
    # camera supports
    for k in range(6):
        a = 2 * np.pi * k / 6
        obs.append(Cylinder([1.0 * np.cos(a), 1.0 * np.sin(a), z_cam - 1.5],
                            [1.0 * np.cos(a), 1.0 * np.sin(a), z_cam], 0.08))
    return obs


def workload(name, n_sources=None, n_samples=None):
    """The benchmark configurations C1..C4 (C5 is C3 batched; see
    iactrace_amd.batched).  Sizes may be reduced for tests."""
    mrad_deg = np.rad2deg(0.5e-3)
    p = 15.0 * math.tan(math.radians(0.16))
    if name in ("C1", "C2"):
        sensor = SquareSensor([0, 0, 15.0], [0, 0, 0], 32, 30, (-16 * p, 16 * p, -15 * p, 15 * p))
        if name == "C1":
            src = np.array([[0.0, 0.0, 1e10]])
            vals = np.ones(1, np.float32)
            m = n_samples or 1000
        else:
            src, vals = point_source_field(n_sources or 4096, 2.5, seed=2)
            m = n_samples or 4096
        if n_sources is not None and name == "C1":
            src, vals = point_source_field(n_sources, 2.5, seed=2)
        return Workload(name, davies_cotton_dish(), [], sensor, m, src, vals, "point")
    if name == "C3":
        sensor = HexagonalSensor([0, 0, 36.066], [0, 0, 0], hex_camera(2048))
        src, vals = point_source_field(n_sources or 16384, 2.5, seed=4)
        return Workload(name, parabolic_dish(), _c3_obstructions(), sensor, n_samples or 1024,
                        src, vals, "point", tilt_sigma_deg=mrad_deg, tilt_seed=3)
    if name == "C3T":
        # C3 with a CT5-sized support structure (271 obstructions)
        sensor = HexagonalSensor([0, 0, 36.066], [0, 0, 0], hex_camera(2048))
        src, vals = point_source_field(n_sources or 16384, 2.5, seed=4)
        return Workload(name, parabolic_dish(), support_truss(), sensor, n_samples or 1024,
                        src, vals, "point", tilt_sigma_deg=mrad_deg, tilt_seed=3)
    if name == "C4":
        mirrors, sensor = cassegrain()
        src = parallel_grid(128, 64, 0.01)
        if n_sources is not None:
            src = src[:n_sources]
        vals = np.ones(len(src), np.float32)
        return Workload(name, mirrors, [], sensor, n_samples or 2048, src, vals, "parallel")
    raise KeyError(name)


@dataclass
class BatchWorkload:
    """C5: instances of the C3 dish, each with its own 0.5 mrad tilt draw
    (numpy PCG64 seeds tilt_seed0 + i) and its own point-source field of
    N_i ~ U{n_min..n_max} sources (sizes from seed 6, field of instance i
    from seed source_seed0 + i).  The fit target of an instance is the
    render of the untilted dish with that instance's sources."""
    base: Workload
    n_instances: int
    sources: np.ndarray      # (total, 3) packed
    values: np.ndarray       # (total,)
    offsets: np.ndarray      # (n_instances + 1,) int64
    tilt_seed0: int = 100

    @property
    def n_rays(self):
        return self.base.n_facets * self.base.n_samples * int(self.offsets[-1])

    def telescope(self, device="cuda", key=0):
        """The untilted dish (sampled with ``key``)."""
        return self.base.telescope(device, key)

    def rotations(self, tel):
        """(n_instances, n_facets, 3) stage-0 rotations with the tilt draws."""
        base = tel.mirror_groups[0].rotations.detach().float()
        out = base.unsqueeze(0).repeat(self.n_instances, 1, 1)
        sig = np.rad2deg(0.5e-3)
        for i in range(self.n_instances):
            d = np.random.default_rng(self.tilt_seed0 + i).normal(0.0, sig, (base.shape[0], 2))
            out[i, :, 0:2] += torch.as_tensor(d, dtype=torch.float32, device=base.device)
        return out.contiguous()


def batch_workload(n_instances=64, n_min=256, n_max=16384, n_samples=None, source_seed0=200):
    """The C5 configuration (sizes may be reduced for tests)."""
    base = workload("C3", n_samples=n_samples)
    base.tilt_sigma_deg = 0.0
    sizes = np.random.default_rng(6).integers(n_min, n_max + 1, n_instances)
    src, vals = [], []
    for i, n in enumerate(sizes):
        a, b = point_source_field(int(n), 2.5, seed=source_seed0 + i)
        src.append(a)
        vals.append(b)
    off = np.zeros(n_instances + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    return BatchWorkload(base, n_instances, np.concatenate(src, 0), np.concatenate(vals, 0), off)
