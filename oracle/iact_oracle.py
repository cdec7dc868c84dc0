"""CPU oracle for the iactrace ray-tracing hot path -- TEST INFRASTRUCTURE ONLY.

This module restates, in numpy, the algorithm of the reference's JAX renderer
(iactrace/core/render.py and everything it calls).  It is the checker that the
HIP kernels are compared against; nothing in the product (iactrace_amd/) may
import it.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg use it.

Precision modes (argument ``prec``):
  "fp64"  -- every step in float64 (the accuracy reference).
  "ref32" -- the reference's own fp32 arithmetic (JAX x64 disabled), except for
             the two deviations the MI355X build sanctions: point-source
             directions are computed in float64 and rounded to fp32, and the
             stage>=1 conic initial guess + Newton refinement run in float64.
Python-float constants follow JAX weak typing: inside fp32 computations they
are rounded to fp32 (numpy 2's NEP-50 promotion gives the same behaviour).

Pinning (each function is checked in tests/test_oracle.py against something
other than its own formula): closed forms (rotation orthonormality and special
angles, on-axis sphere/paraboloid roots, analytic sag slopes vs. central finite
differences), brute force (dense ray marching for every obstruction primitive,
nearest-centre search for hexagonal pixels), and whole-path known answers
(flat facet flux == facet area, perfect paraboloid focuses in one pixel).
The exact jax.random streams of the reference cannot be reproduced without
importing JAX (absent here); the sampler below is this build's own
counter-based Philox generator, checked for integer bit-exactness against the
published Philox4x32-10 known-answer vectors and for distributional
properties.  Parity unpinned: the jax.random streams themselves, XLA's fp32
fusion/rounding order, XLA's float->int conversion of NaN coordinates.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32
F64 = np.float64

# ---------------------------------------------------------------------------
# precision helpers
# ---------------------------------------------------------------------------


def _dt(prec):
    if prec == "fp64":
        return F64
    if prec == "ref32":
        return F32
    raise ValueError(prec)


def _c(x, dt):
    """A Python constant as seen inside a computation of dtype dt (weak typing)."""
    return dt(x)


# ---------------------------------------------------------------------------
# transforms.py
# ---------------------------------------------------------------------------


def euler_to_matrix(angles_deg, prec="ref32"):
    """R = Rz @ Ry @ Rx from (tip, tilt, rotation) in degrees.

    Follows iactrace/core/transforms.py:72-107 (jnp.radians then the three
    elementary rotations, multiplied Rz @ Ry @ Rx).  Accepts (..., 3).
    """
    dt = _dt(prec)
    a = np.asarray(angles_deg, dtype=dt)
    rad = a * _c(math.pi / 180.0, dt)
    rx, ry, rz = rad[..., 0], rad[..., 1], rad[..., 2]
    cx, sx = np.cos(rx), np.sin(rx)
    cy, sy = np.cos(ry), np.sin(ry)
    cz, sz = np.cos(rz), np.sin(rz)
    one = np.ones_like(cx)
    zero = np.zeros_like(cx)
    Rx = np.stack([np.stack([one, zero, zero], -1),
                   np.stack([zero, cx, -sx], -1),
                   np.stack([zero, sx, cx], -1)], -2)
    Ry = np.stack([np.stack([cy, zero, sy], -1),
                   np.stack([zero, one, zero], -1),
                   np.stack([-sy, zero, cy], -1)], -2)
    Rz = np.stack([np.stack([cz, -sz, zero], -1),
                   np.stack([sz, cz, zero], -1),
                   np.stack([zero, zero, one], -1)], -2)
    return (Rz @ Ry @ Rx).astype(dt)


def look_at_euler(mirror_pos, target_pos, up=(0.0, 1.0, 0.0)):
    """Euler angles (deg) aiming local +z from mirror_pos at target_pos.

    Follows iactrace/core/transforms.py:27-69 (float64 here: it only feeds
    geometry generators).
    """
    mirror_pos = np.asarray(mirror_pos, F64)
    target_pos = np.asarray(target_pos, F64)
    up = np.asarray(up, F64)
    fwd = target_pos - mirror_pos
    fwd = fwd / np.linalg.norm(fwd)
    right = np.cross(up, fwd)
    right = right / np.linalg.norm(right)
    upc = np.cross(fwd, right)
    R = np.column_stack([right, upc, fwd])
    sy = -R[2, 0]
    cy = math.sqrt(max(1.0 - sy * sy, 0.0))
    if cy > 1e-6:
        tip = math.atan2(R[2, 1], R[2, 2])
        rot = math.atan2(R[1, 0], R[0, 0])
    else:
        tip = 0.0
        rot = math.atan2(-R[0, 1], R[1, 1])
    tilt = math.asin(sy)
    return np.rad2deg(np.array([tip, tilt, rot]))


# ---------------------------------------------------------------------------
# reflection.py
# ---------------------------------------------------------------------------


def _dot(a, b):
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def reflect(d, n):
    """(r, cos) with c = d.n, r = d - 2 c n, cos = -c.  reflection.py:5-19."""
    c = _dot(d, n)
    r = d - (2.0 * c)[..., None] * n
    return r, -c


def perturbation_delta(normals, theta1, theta2):
    """Tangent-plane roughness offset from two N(0,1) draws.

    Follows iactrace/core/reflection.py:22-50: reference axis x if |n.z|>0.9
    else z; t1 = normalize(n x ref); t2 = n x t1; delta = th1 t1 + th2 t2.
    The draws are passed in (the build's own RNG produces them).
    """
    dt = normals.dtype.type
    dot_z = np.abs(normals[..., 2])
    ref = np.where((dot_z > _c(0.9, dt))[..., None],
                   np.array([1.0, 0.0, 0.0], dt), np.array([0.0, 0.0, 1.0], dt))
    t1 = np.cross(normals, ref)
    t1 = t1 / np.sqrt(_dot(t1, t1))[..., None]
    t2 = np.cross(normals, t1)
    return theta1[..., None] * t1 + theta2[..., None] * t2


# ---------------------------------------------------------------------------
# intersections.py
# ---------------------------------------------------------------------------


def intersect_plane(o, d, center, R):
    """Sensor-plane hit in the sensor frame, sentinel 1e10 on miss.

    intersections.py:6-41.  R columns are (u1, u2, normal).
    """
    dt = o.dtype.type
    u1, u2, pn = R[:, 0], R[:, 1], R[:, 2]
    ndotd = _dot(d, pn)
    ndoto = _dot(o, pn)
    ndotp = np.sum(pn * center)
    parallel = np.abs(ndotd) < _c(1e-10, dt)
    safe = np.where(parallel, _c(1.0, dt), ndotd)
    t = (ndotp - ndoto) / safe
    inter = o + t[..., None] * d
    op = inter - center
    x = _dot(op, u1)
    y = _dot(op, u2)
    invalid = parallel | (t <= 0)
    x = np.where(invalid, _c(1e10, dt), x)
    y = np.where(invalid, _c(1e10, dt), y)
    return x, y


def cylinder_consts(p1, p2, r, dt):
    """Per-primitive hoisted values: axis unit vector and height (fp of dt)."""
    p1 = np.asarray(p1, dt)
    p2 = np.asarray(p2, dt)
    axis = p2 - p1
    height = np.sqrt(_dot(axis, axis))
    return p1, axis / height, height, dt(r)


def intersect_cylinder(o, d, p1, axis, height, radius):
    """Finite capped cylinder, min positive t.  intersections.py:44-87."""
    dt = o.dtype.type
    eps = _c(1e-8, dt)
    oc = o - p1
    oc_axial = _dot(oc, axis)
    rd_axial = _dot(d, axis)
    oc_perp = oc - oc_axial[..., None] * axis
    rd_perp = d - rd_axial[..., None] * axis
    a = _dot(rd_perp, rd_perp)
    b = 2 * _dot(oc_perp, rd_perp)
    c = _dot(oc_perp, oc_perp) - radius * radius
    disc = b * b - 4 * a * c
    sq = np.sqrt(np.maximum(disc, _c(0.0, dt)))
    t1 = (-b - sq) / (2 * a + eps)
    t2 = (-b + sq) / (2 * a + eps)
    y1 = oc_axial + t1 * rd_axial
    y2 = oc_axial + t2 * rd_axial
    inf = dt(np.inf)
    t1 = np.where((t1 > eps) & (y1 >= 0) & (y1 <= height) & (disc >= 0), t1, inf)
    t2 = np.where((t2 > eps) & (y2 >= 0) & (y2 <= height) & (disc >= 0), t2, inf)
    with np.errstate(divide="ignore", invalid="ignore"):
        t_bot = -oc_axial / (rd_axial + eps)
        t_top = (height - oc_axial) / (rd_axial + eps)
    pb = oc_perp + t_bot[..., None] * rd_perp
    pt = oc_perp + t_top[..., None] * rd_perp
    r2 = radius ** 2
    t_bot = np.where((t_bot > eps) & (_dot(pb, pb) <= r2), t_bot, inf)
    t_top = np.where((t_top > eps) & (_dot(pt, pt) <= r2), t_top, inf)
    return np.minimum(np.minimum(np.minimum(t1, t2), t_bot), t_top)


def intersect_box(o, d, bmin, bmax):
    """Slab AABB with the reference's +1e-8 direction bias.  intersections.py:90-110."""
    dt = o.dtype.type
    eps = _c(1e-8, dt)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / (d + eps)
        t1 = (bmin - o) * inv
        t2 = (bmax - o) * inv
    tn = np.minimum(t1, t2)
    tf = np.maximum(t1, t2)
    tmin = np.max(tn, axis=-1)
    tmax = np.min(tf, axis=-1)
    hit = (tmax >= tmin) & (tmax > eps)
    tres = np.where(tmin > eps, tmin, tmax)
    return np.where(hit, tres, dt(np.inf))


def intersect_sphere(o, d, center, radius):
    """Ray-sphere nearest positive root.  intersections.py:195-226."""
    dt = o.dtype.type
    eps = _c(1e-8, dt)
    oc = o - center
    a = _dot(d, d)
    b = 2.0 * _dot(oc, d)
    c = _dot(oc, oc) - radius * radius
    disc = b * b - 4.0 * a * c
    sq = np.sqrt(np.maximum(disc, _c(0.0, dt)))
    t1 = (-b - sq) / (2.0 * a + eps)
    t2 = (-b + sq) / (2.0 * a + eps)
    inf = dt(np.inf)
    t1 = np.where((t1 > eps) & (disc >= 0), t1, inf)
    t2 = np.where((t2 > eps) & (disc >= 0), t2, inf)
    return np.minimum(t1, t2)


def intersect_oriented_box(o, d, center, half, rot):
    """OBB via local-frame slab test, sign-aware eps.  intersections.py:113-149."""
    dt = o.dtype.type
    eps = _c(1e-8, dt)
    lo = (o - center) @ rot          # rot.T @ v  ==  v @ rot
    ld = d @ rot
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / (ld + eps * np.sign(ld + eps))
        t1 = (-half - lo) * inv
        t2 = (half - lo) * inv
    tn = np.minimum(t1, t2)
    tf = np.maximum(t1, t2)
    tmin = np.max(tn, axis=-1)
    tmax = np.min(tf, axis=-1)
    hit = (tmax >= tmin) & (tmax > eps)
    tres = np.where(tmin > eps, tmin, tmax)
    return np.where(hit & (tres > eps), tres, dt(np.inf))


def intersect_triangle(o, d, v0, v1, v2):
    """Moller-Trumbore with sign-aware eps.  intersections.py:152-192."""
    dt = o.dtype.type
    eps = _c(1e-8, dt)
    e1 = v1 - v0
    e2 = v2 - v0
    h = np.cross(d, e2)
    a = _dot(np.broadcast_to(e1, h.shape), h)
    parallel = np.abs(a) < eps
    with np.errstate(divide="ignore", invalid="ignore"):
        f = 1.0 / (a + eps * np.sign(a + eps))
    s = o - v0
    u = f * _dot(s, h)
    q = np.cross(s, e1)
    v = f * _dot(d, q)
    t = f * _dot(np.broadcast_to(e2, q.shape), q)
    valid = (~parallel) & (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1) & (t > eps)
    return np.where(valid, t, dt(np.inf))


def intersect_conic(o, d, c, k):
    """Closed-form conic root, smallest t > 1e-8.  intersections.py:229-285.

    o, d: (..., 3) arrays of the working dtype; c, k scalars of that dtype.
    """
    dt = o.dtype.type
    ox, oy, oz = o[..., 0], o[..., 1], o[..., 2]
    dx, dy, dz = d[..., 0], d[..., 1], d[..., 2]
    one_k = 1 + k
    A = c * (dx * dx + dy * dy + one_k * dz * dz)
    B = 2 * (c * (ox * dx + oy * dy + one_k * oz * dz) - dz)
    C = c * (ox * ox + oy * oy + one_k * oz * oz) - 2 * oz
    is_plane = np.abs(c) < _c(1e-12, dt)
    with np.errstate(divide="ignore", invalid="ignore"):
        t_plane = np.where(np.abs(dz) > _c(1e-10, dt), -oz / dz, dt(np.inf))
    disc = B * B - 4 * A * C
    sq = np.sqrt(np.maximum(disc, _c(0.0, dt)))
    with np.errstate(divide="ignore", invalid="ignore"):
        t1 = (-B - sq) / (2 * A + _c(1e-30, dt))
        t2 = (-B + sq) / (2 * A + _c(1e-30, dt))
    v1 = t1 > _c(1e-8, dt)
    v2 = t2 > _c(1e-8, dt)
    inf = dt(np.inf)
    tc = np.where(v1 & v2, np.minimum(t1, t2), np.where(v1, t1, np.where(v2, t2, inf)))
    tc = np.where(disc < 0, inf, tc)
    return np.where(is_plane, t_plane, tc)


# ---------------------------------------------------------------------------
# surfaces.py
# ---------------------------------------------------------------------------


@dataclass
class Surface:
    """Conic + even polynomial surface (surfaces.py:8-13).

    aspheric[i] multiplies (r^2)^(2(i+1)) -- i.e. r^4, r^8, r^12, ... exactly
    as jnp.arange(2, 2+2K, 2) is used as the power of r2 (surfaces.py:35-37).
    """
    curvature: float
    conic: float
    aspheric: np.ndarray = field(default_factory=lambda: np.zeros(0))

    def cast(self, dt):
        return dt(self.curvature), dt(self.conic), np.asarray(self.aspheric, dt)


def sag_raw(X, Y, surf: Surface, dt):
    """z(X,Y) = c r2 / (1+sqrt(1-(1+k)c^2 r2)) + sum a_i r2^(2(i+1)).  surfaces.py:25-39."""
    c, k, a = surf.cast(dt)
    r2 = X * X + Y * Y
    with np.errstate(invalid="ignore"):
        denom = 1 + np.sqrt(1 - (1 + k) * c * c * r2)
    z = r2 * c / denom
    for i in range(a.shape[0]):
        z = z + a[i] * r2 ** (2 * (i + 1))
    return z


def sag_slope(X, Y, surf: Surface, dt):
    """(dz/dX, dz/dY) of sag_raw, written as the chain rule jax.grad applies.

    surfaces.py:51-58 differentiates _sag_raw by autodiff; with u = 1-K r2,
    s = sqrt(u), den = 1+s:  dz/dr2 = c/den + (r2 c) K / (2 s den^2)
    + sum a_i p_i r2^(p_i-1), and dr2/dX = 2X.
    """
    c, k, a = surf.cast(dt)
    r2 = X * X + Y * Y
    K = (1 + k) * c * c
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.sqrt(1 - K * r2)
        den = 1 + s
        g = c / den + (r2 * c) * K / (2 * s * den * den)
    for i in range(a.shape[0]):
        p = 2 * (i + 1)
        g = g + a[i] * p * r2 ** (p - 1)
    return 2 * X * g, 2 * Y * g


def surface_point_normal(x, y, surf: Surface, offset, dt):
    """Local point (x, y, sag(x,y)) and unit normal (-zx, -zy, 1)/|.|.

    surfaces.py:41-65 (offset form: sag_raw(x+x0, y+y0) - sag_raw(x0, y0)).
    """
    x0, y0 = dt(offset[0]), dt(offset[1])
    z0 = sag_raw(x0, y0, surf, dt)
    z = sag_raw(x + x0, y + y0, surf, dt) - z0
    zx, zy = sag_slope(x + x0, y + y0, surf, dt)
    n = np.stack([-zx, -zy, np.ones_like(zx)], -1)
    n = n / np.sqrt(_dot(n, n))[..., None]
    return np.stack([x, y, z], -1), n


def newton_intersect(o, d, surf: Surface, offset, t_init, dt, max_iter=10, tol=1e-8):
    """Fixed-count Newton on g(t) = o_z + t d_z - sag(o_xy + t d_xy).

    intersections.py:290-367: exactly max_iter steps; a lane freezes using
    the *previous* iteration's converged flag; g' is replaced by +1e-12 when
    |g'| <= 1e-12; result valid iff t > 1e-8 and |g(t)| < 100 tol.
    """
    x0, y0 = dt(offset[0]), dt(offset[1])
    z0 = sag_raw(x0, y0, surf, dt)
    ox, oy, oz = o[..., 0], o[..., 1], o[..., 2]
    dx, dy, dz = d[..., 0], d[..., 1], d[..., 2]

    def g(t):
        return (oz + t * dz) - (sag_raw(ox + t * dx + x0, oy + t * dy + y0, surf, dt) - z0)

    def gp(t):
        sx, sy = sag_slope(ox + t * dx + x0, oy + t * dy + y0, surf, dt)
        return dz - (sx * dx + sy * dy)

    t = t_init.astype(dt)
    conv = np.zeros(t.shape, bool)
    tol_ = dt(tol)
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        for _ in range(max_iter):
            gv = g(t)
            gd = gp(t)
            gd = np.where(np.abs(gd) > dt(1e-12), gd, dt(1e-12))
            tn = t - gv / gd
            newc = conv | (np.abs(gv) < tol_)
            t = np.where(conv, t, tn)
            conv = newc
        res = np.abs(g(t))
    valid = (t > dt(1e-8)) & (res < tol_ * 100)
    return np.where(valid, t, dt(np.inf)), ox + t * dx, oy + t * dy


def surface_intersect(o, d, surf: Surface, offset, prec):
    """Stage>=1 mirror hit in the mirror's local frame.  surfaces.py:67-108.

    Returns (t, point(3), normal(3)) in the working dtype.  The conic initial
    guess and the Newton refinement always run in float64 on the working-dtype
    inputs (the sanctioned deviation for ref32); t and the hit (x, y) are then
    rounded back to the working dtype, and z = sag(x, y) and the normal are
    evaluated there (Q7: z is the sag, not o_z + t d_z).
    """
    dt = _dt(prec)
    nd = F64
    x0, y0 = dt(offset[0]), dt(offset[1])
    c, k, a = surf.cast(dt)
    s64 = Surface(float(c), float(k), a.astype(nd))
    o64 = o.astype(nd)
    d64 = d.astype(nd)
    z0 = sag_raw(nd(x0), nd(y0), s64, nd)
    o_raw = np.stack([o64[..., 0] + nd(x0), o64[..., 1] + nd(y0), o64[..., 2] + z0], -1)
    t_init = intersect_conic(o_raw, d64, nd(c), nd(k))
    t, xh, yh = newton_intersect(o64, d64, s64, (nd(x0), nd(y0)), t_init, nd)
    p, n = surface_point_normal(xh.astype(dt), yh.astype(dt), surf, offset, dt)
    return t.astype(dt), p, n


# ---------------------------------------------------------------------------
# sensors
# ---------------------------------------------------------------------------

SQRT3 = 1.7320508075688772
SQRT3_2 = 0.8660254037844386
SQRT3_3 = 0.5773502691896257


@dataclass
class SquareSensorSpec:
    """SquareSensor (sensors/square.py:28-64) as plain numbers."""
    position: tuple
    rotation: tuple
    width: int
    height: int
    bounds: tuple
    edge_width: float = 0.0
    splat: bool = False          # DifferentiableSquareSensor (square.py:94-173)
    sigma: float = 0.1
    kernel_size: int = 2

    @property
    def n_pixels(self):
        return self.width * self.height

    def params(self):
        xmin, xmax, ymin, ymax = self.bounds
        return (float(xmin), float(ymin), float((xmax - xmin) / self.width),
                float((ymax - ymin) / self.height))


def square_bin(x, y, s: SquareSensorSpec):
    """Pixel id (-1 = dropped) per ray.  square.py:66-91 (floor, range, edge).

    Out-of-range coordinates are tested in float before any integer cast, so
    the 1e10 miss sentinel can never wrap (Q5).
    """
    dt = x.dtype.type
    x0, y0, dx, dy = s.params()
    xc = (x - dt(x0)) / dt(dx)
    yc = (y - dt(y0)) / dt(dy)
    xf = np.floor(xc)
    yf = np.floor(yc)
    valid = (xf >= 0) & (xf < s.width) & (yf >= 0) & (yf < s.height)
    xfr = xc - xf
    yfr = yc - yf
    dist = np.minimum(np.minimum(xfr, 1 - xfr) * dt(dx), np.minimum(yfr, 1 - yfr) * dt(dy))
    valid &= ~(dist < dt(s.edge_width))
    xi = np.where(valid, xf, 0).astype(np.int64)
    yi = np.where(valid, yf, 0).astype(np.int64)
    return np.where(valid, yi * s.width + xi, -1), xc, yc


def square_splat(x, y, v, s: SquareSensorSpec):
    """Gaussian splat over a (2k+1)^2 neighbourhood.  square.py:144-173.

    Returns list of (pixel_id (-1 if dropped), weighted value) arrays per tap.
    """
    dt = x.dtype.type
    x0, y0, dx, dy = s.params()
    xp = (x - dt(x0)) / dt(dx)
    yp = (y - dt(y0)) / dt(dy)
    xb = np.floor(xp)
    yb = np.floor(yp)
    xfr = xp - xb
    yfr = yp - yb
    ks = s.kernel_size
    offs = [(ox, oy) for oy in range(-ks, ks + 1) for ox in range(-ks, ks + 1)]
    ws = []
    for ox, oy in offs:
        ddx = xfr - dt(ox)
        ddy = yfr - dt(oy)
        ws.append(np.exp(dt(-0.5) * (ddx ** 2 + ddy ** 2) / dt(s.sigma) ** 2))
    tot = ws[0]
    for w in ws[1:]:
        tot = tot + w
    out = []
    for (ox, oy), w in zip(offs, ws):
        xi = xb + ox
        yi = yb + oy
        valid = (xi >= 0) & (xi < s.width) & (yi >= 0) & (yi < s.height)
        pid = np.where(valid, np.where(valid, yi, 0) * s.width + np.where(valid, xi, 0), -1)
        out.append((pid.astype(np.int64), v * (w / tot) * valid.astype(dt)))
    return out


@dataclass
class HexSensorSpec:
    """HexagonalSensor (sensors/hexagonal.py:104-144) with its derived grid."""
    position: tuple
    rotation: tuple
    centers: np.ndarray
    edge_width: float = 0.0
    splat: bool = False          # DifferentiableHexagonalSensor (hexagonal.py:197-315)
    sigma: float = 0.5
    kernel_size: int = 1
    hex_size: float = 0.0
    inradius: float = 0.0
    grid_rotation: float = 0.0
    grid_offset: tuple = (0.0, 0.0)
    table: np.ndarray = None
    q_min: int = 0
    r_min: int = 0

    @property
    def n_pixels(self):
        return len(self.centers)

    def __post_init__(self):
        c = np.asarray(self.centers, F32)
        self.centers = c
        self.hex_size, self.grid_rotation, off = detect_hex_grid(c)
        self.inradius = float(F32(self.hex_size) * F32(SQRT3_2))
        self.grid_offset = (float(off[0]), float(off[1]))
        self.table, self.q_min, self.r_min = build_hex_table(c, self.hex_size, self.grid_rotation, off)


def detect_hex_grid(c):
    """(size, rotation, offset) of a hex pixel lattice.  hexagonal.py:61-80 (fp32)."""
    c = np.asarray(c, F32)
    n = len(c)
    diff = c[:, None, :] - c[None, :, :]
    d2 = diff[..., 0] ** 2 + diff[..., 1] ** 2
    d2[np.arange(n), np.arange(n)] = np.inf
    idx = int(np.argmin(d2))
    min_dist = np.sqrt(d2.reshape(-1)[idx])
    vec = diff[idx // n, idx % n]
    ang = np.mod(np.arctan2(vec[1], vec[0]), F32(np.pi / 3))
    off = c[int(np.argmin(c[:, 0] ** 2 + c[:, 1] ** 2))]
    return float(min_dist / F32(SQRT3)), float(ang), off


def _rot2(x, y, ang, dt):
    cs = dt(np.cos(dt(ang)))
    sn = dt(np.sin(dt(ang)))
    return cs * x - sn * y, sn * x + cs * y


def _to_axial(x, y, size, dt):
    return (dt(SQRT3_3) * x - y / 3) / dt(size), (2 * y / 3) / dt(size)


def _from_axial(q, r, size, dt):
    return dt(size) * dt(SQRT3) * (q + r / 2), dt(size) * dt(1.5) * r


def axial_round(q, r):
    """Cube rounding with half-to-even np.round (== jnp.round).  hexagonal.py:32-39."""
    s = -q - r
    qi, ri, si = np.round(q), np.round(r), np.round(s)
    dq, dr, ds = np.abs(qi - q), np.abs(ri - r), np.abs(si - s)
    qi2 = np.where((dq > dr) & (dq > ds), -ri - si, qi)
    ri2 = np.where((dr > dq) & (dr > ds), -qi2 - si, ri)
    return qi2, ri2


def hex_norm(x, y, inradius, dt):
    """max(|x|, |x|/2 + sqrt3/2 |y|)/inradius.  hexagonal.py:42-47."""
    return np.maximum(np.abs(x), dt(0.5) * np.abs(x) + dt(SQRT3_2) * np.abs(y)) / dt(inradius)


def build_hex_table(c, size, rot, off):
    """Axial lookup table, -1 = empty.  hexagonal.py:83-101."""
    dt = F32
    x = c[:, 0] - off[0]
    y = c[:, 1] - off[1]
    xr, yr = _rot2(x, y, -rot, dt)
    q, r = _to_axial(xr, yr, size, dt)
    qi = np.round(q).astype(np.int32)
    ri = np.round(r).astype(np.int32)
    qmin, qmax, rmin, rmax = int(qi.min()), int(qi.max()), int(ri.min()), int(ri.max())
    tab = np.full((qmax - qmin + 1, rmax - rmin + 1), -1, np.int32)
    tab[qi - qmin, ri - rmin] = np.arange(len(c), dtype=np.int32)
    return tab, qmin, rmin


def hex_grid_coords(x, y, s: HexSensorSpec):
    dt = x.dtype.type
    return _rot2(x - dt(s.grid_offset[0]), y - dt(s.grid_offset[1]), -s.grid_rotation, dt)


def _hex_lookup(qi, ri, s):
    qidx = qi - s.q_min
    ridx = ri - s.r_min
    nq, nr = s.table.shape
    inb = (qidx >= 0) & (qidx < nq) & (ridx >= 0) & (ridx < nr)
    pid = s.table[np.clip(qidx, 0, nq - 1).astype(np.int64), np.clip(ridx, 0, nr - 1).astype(np.int64)]
    valid = inb & (pid >= 0)
    return np.where(valid, pid, -1)


def hex_bin(x, y, s: HexSensorSpec):
    """Pixel id (-1 = dropped) per ray.  hexagonal.py:174-194."""
    dt = x.dtype.type
    xg, yg = hex_grid_coords(x, y, s)
    q, r = _to_axial(xg, yg, s.hex_size, dt)
    finite = np.isfinite(q) & np.isfinite(r) & (np.abs(q) < 1e6) & (np.abs(r) < 1e6)
    q = np.where(finite, q, 0)
    r = np.where(finite, r, 0)
    qi, ri = axial_round(q, r)
    pid = _hex_lookup(qi.astype(np.int64), ri.astype(np.int64), s)
    pid = np.where(finite, pid, -1)
    cx, cy = _from_axial(qi, ri, s.hex_size, dt)
    hd = hex_norm(xg - cx, yg - cy, s.inradius, dt)
    thr = dt(1.0 - s.edge_width / s.inradius)
    pid = np.where(hd > thr, -1, pid)
    return pid, q, r


def hex_splat(x, y, v, s: HexSensorSpec):
    """Gaussian splat over hex rings.  hexagonal.py:264-315."""
    dt = x.dtype.type
    xg, yg = hex_grid_coords(x, y, s)
    q, r = _to_axial(xg, yg, s.hex_size, dt)
    qb, rb = axial_round(q, r)
    bx, by = _from_axial(qb, rb, s.hex_size, dt)
    ddx = xg - bx
    ddy = yg - by
    ks = s.kernel_size
    offs = [(oq, orr) for oq in range(-ks, ks + 1) for orr in range(-ks, ks + 1)
            if max(abs(oq), abs(orr), abs(-oq - orr)) <= ks]
    ws = []
    for oq, orr in offs:
        ox, oy = _from_axial(dt(oq), dt(orr), s.hex_size, dt)
        hd = hex_norm(ddx - ox, ddy - oy, s.inradius, dt)
        ws.append(np.exp(dt(-0.5) * (hd / dt(s.sigma)) ** 2))
    tot = ws[0]
    for w in ws[1:]:
        tot = tot + w
    out = []
    for (oq, orr), w in zip(offs, ws):
        pid = _hex_lookup(qb.astype(np.int64) + oq, rb.astype(np.int64) + orr, s)
        out.append((pid, v * (w / tot) * (pid >= 0).astype(dt)))
    return out


def sensor_hits(x, y, v, sensor):
    """[(pixel_id, value)] contributions of every ray (hard: one tap)."""
    if isinstance(sensor, SquareSensorSpec):
        if sensor.splat:
            return square_splat(x, y, v, sensor)
        pid, _, _ = square_bin(x, y, sensor)
    else:
        if sensor.splat:
            return hex_splat(x, y, v, sensor)
        pid, _, _ = hex_bin(x, y, sensor)
    return [(pid, np.where(pid >= 0, v, v.dtype.type(0)))]


def accumulate(x, y, v, sensor, acc_dtype=F64):
    """segment_sum of valid rays into the sensor's pixels (flattened)."""
    img = np.zeros(sensor.n_pixels, acc_dtype)
    for pid, val in sensor_hits(x, y, v, sensor):
        m = pid >= 0
        np.add.at(img, pid[m], val[m].astype(acc_dtype))
    return img


# ---------------------------------------------------------------------------
# scene description + renderer (render.py)
# ---------------------------------------------------------------------------


@dataclass
class Obstructions:
    """Obstruction groups (obstructions.py:38-277) as arrays, grouped by type."""
    cylinders: list = field(default_factory=list)   # (p1, p2, r)
    boxes: list = field(default_factory=list)       # (p1, p2)
    spheres: list = field(default_factory=list)     # (center, r)
    oriented_boxes: list = field(default_factory=list)  # (center, half, rot3x3)
    triangles: list = field(default_factory=list)   # (v0, v1, v2)


@dataclass
class MirrorSpec:
    """A stage>=1 mirror (render.py:82-115): pose, surface, aperture."""
    position: tuple
    rotation: tuple
    surface: Surface
    offset: tuple = (0.0, 0.0)
    radius: float = None          # disk aperture
    vertices: np.ndarray = None   # polygon aperture (CCW)
    stage: int = 1
    group: int = 0


@dataclass
class Scene:
    """Everything render() reads.  Stage-0 samples are local-frame arrays."""
    points: np.ndarray       # (Nf, M, 3)
    normals: np.ndarray      # (Nf, M, 3)
    delta: np.ndarray        # (Nf, M, 3)
    weights: np.ndarray      # (Nf, M)
    positions: np.ndarray    # (Nf, 3)
    rotations: np.ndarray    # (Nf, 3) degrees
    scale: np.ndarray        # (Nf,)
    sensor: object
    obstructions: Obstructions = field(default_factory=Obstructions)
    mirrors: list = field(default_factory=list)   # stage >= 1 MirrorSpec


def transform_to_world(sc: Scene, prec):
    """World points, perturbed normals, weights.  mirrors.py:64-79."""
    dt = _dt(prec)
    R = euler_to_matrix(sc.rotations, prec)                      # (Nf,3,3)
    p = np.einsum("fij,fmj->fmi", R, sc.points.astype(dt)) + sc.positions.astype(dt)[:, None, :]
    n = np.einsum("fij,fmj->fmi", R, sc.normals.astype(dt))
    dl = np.einsum("fij,fmj->fmi", R, sc.delta.astype(dt))
    pert = n + sc.scale.astype(dt)[:, None, None] * dl
    pert = pert / np.sqrt(_dot(pert, pert))[..., None]
    return p.astype(dt), pert.astype(dt), sc.weights.astype(dt)


def check_occlusions(o, d, obs: Obstructions, dt):
    """1 if no obstruction has t < 1e10 along the unbounded ray.  render.py:21-41."""
    shadow = np.ones(o.shape[:-1], dt)
    big = dt(1e10)
    groups = []
    if obs.cylinders:
        groups.append([intersect_cylinder(o, d, *cylinder_consts(p1, p2, r, dt))
                       for p1, p2, r in obs.cylinders])
    if obs.boxes:
        groups.append([intersect_box(o, d, np.minimum(np.asarray(a, dt), np.asarray(b, dt)),
                                     np.maximum(np.asarray(a, dt), np.asarray(b, dt)))
                       for a, b in obs.boxes])
    if obs.spheres:
        groups.append([intersect_sphere(o, d, np.asarray(c, dt), dt(r)) for c, r in obs.spheres])
    if obs.oriented_boxes:
        groups.append([intersect_oriented_box(o, d, np.asarray(c, dt), np.asarray(h, dt),
                                              np.asarray(R, dt)) for c, h, R in obs.oriented_boxes])
    if obs.triangles:
        groups.append([intersect_triangle(o, d, np.asarray(a, dt), np.asarray(b, dt),
                                          np.asarray(c, dt)) for a, b, c in obs.triangles])
    for ts in groups:
        tmin = ts[0]
        for t in ts[1:]:
            tmin = np.minimum(tmin, t)
        shadow = shadow * np.where(tmin < big, dt(0.0), dt(1.0))
    return shadow


def _aperture_ok(m: MirrorSpec, x, y, dt):
    if m.vertices is None:
        return x ** 2 + y ** 2 <= dt(m.radius) ** 2           # mirrors.py:147-149
    v = np.asarray(m.vertices, dt)                             # mirrors.py:209-220 (CCW)
    ok = np.ones(x.shape, bool)
    n = len(v)
    for i in range(n):
        a, b = v[i], v[(i + 1) % n]
        cr = (b[0] - a[0]) * (y - a[1]) - (b[1] - a[1]) * (x - a[0])
        ok &= cr >= 0
    return ok


def reflect_at_stage(o, d, v, mirrors, obs, prec):
    """Closest stage mirror hit, reflect, value *= hit*shadow*|cos|.  render.py:44-115."""
    dt = _dt(prec)
    best_t = np.full(o.shape[:-1], np.inf, dt)
    best_p = np.zeros(o.shape, dt)
    best_n = np.zeros(o.shape, dt)
    groups = sorted({m.group for m in mirrors})
    for g in groups:
        gm = [m for m in mirrors if m.group == g]
        gt = np.full(o.shape[:-1], np.inf, dt)
        gp = np.zeros(o.shape, dt)
        gn = np.zeros(o.shape, dt)
        for m in gm:                                   # argmin keeps the first minimum
            R = euler_to_matrix(np.asarray(m.rotation), prec)
            pos = np.asarray(m.position, dt)
            ol = (o - pos) @ R                         # R^T (o - pos)
            dl = d @ R
            t, pl, nl = surface_intersect(ol, dl, m.surface, m.offset, prec)
            t = np.where(_aperture_ok(m, pl[..., 0], pl[..., 1], dt), t, dt(np.inf))
            pw = pl @ R.T + pos
            nw = nl @ R.T
            closer = t < gt
            gt = np.where(closer, t, gt)
            gp = np.where(closer[..., None], pw, gp)
            gn = np.where(closer[..., None], nw, gn)
        closer = gt < best_t                           # render.py:68 strict: earlier group wins ties
        best_t = np.where(closer, gt, best_t)
        best_p = np.where(closer[..., None], gp, best_p)
        best_n = np.where(closer[..., None], gn, best_n)
    r, cos = reflect(d, best_n)
    hit = (best_t < dt(1e10)).astype(dt)
    shadow = check_occlusions(o, d, obs, dt)
    return best_p, r, v * hit * shadow * np.abs(cos)


def source_dirs(p, sources, source_type, prec):
    """Per (source, sample) direction.  render.py:128-133.

    'point': normalize(p - src) -- computed in float64 in both modes and rounded
    to the working dtype (the sanctioned FP64 deviation).  Anything else is
    'parallel': src used as given, unnormalised (Q15/Q16).
    """
    dt = _dt(prec)
    if source_type == "point":
        d = p.astype(F64)[None, :, :] - np.asarray(sources, F64)[:, None, :]
        d = d / np.sqrt(_dot(d, d))[..., None]
        return d.astype(dt)
    src = np.asarray(sources).astype(dt)
    return np.broadcast_to(src[:, None, :], (src.shape[0], p.shape[0], 3))


def trace_facet(f, world, sc: Scene, sources, values, source_type, prec):
    """One primary facet: dirs, shadow, reflect, value, stages, sensor plane.

    render.py:118-157.  Returns (x, y, vals) each (Ns, M).
    """
    dt = _dt(prec)
    tp, tn, tw = world
    p, n, w = tp[f], tn[f], tw[f]
    d = source_dirs(p, sources, source_type, prec)
    o = np.broadcast_to(p[None], d.shape)
    nn = np.broadcast_to(n[None], d.shape)
    shadow = check_occlusions(o, -d, sc.obstructions, dt)
    r, cos = reflect(d, nn)
    vals = np.asarray(values, dt)[:, None] * cos / w[None, :] * shadow
    stages = sorted({m.stage for m in sc.mirrors})
    oc, rc = o, r
    for st in stages:
        oc, rc, vals = reflect_at_stage(oc, rc, vals, [m for m in sc.mirrors if m.stage == st],
                                        sc.obstructions, prec)
    sp = sc.sensor.position
    R = euler_to_matrix(np.asarray(sc.sensor.rotation), prec)
    x, y = intersect_plane(oc, rc, np.asarray(sp, dt), R)
    return x, y, vals


def render(sc: Scene, sources, values, source_type="point", prec="ref32", mode="image"):
    """render / render_response_matrix / render_debug (render.py:174-324).

    image:   (P,) float64 accumulation of the working-precision ray values
    response:(Ns, P)
    debug:   (x, y, vals, pixel_ids) of shape (Nf*Ns*M,), facet-major
    """
    world = transform_to_world(sc, prec)
    nf = sc.points.shape[0]
    ns = np.asarray(sources).shape[0]
    P = sc.sensor.n_pixels
    if mode == "image":
        img = np.zeros(P, F64)
    elif mode == "response":
        img = np.zeros((ns, P), F64)
    else:
        xs, ys, vs, ps = [], [], [], []
    for f in range(nf):
        x, y, v = trace_facet(f, world, sc, sources, values, source_type, prec)
        if mode == "image":
            img += accumulate(x.reshape(-1), y.reshape(-1), v.reshape(-1), sc.sensor)
        elif mode == "response":
            for s in range(ns):
                img[s] += accumulate(x[s], y[s], v[s], sc.sensor)
        else:
            xs.append(x.reshape(-1))
            ys.append(y.reshape(-1))
            vs.append(v.reshape(-1))
            if isinstance(sc.sensor, SquareSensorSpec):
                pid = square_bin(x.reshape(-1), y.reshape(-1), sc.sensor)[0]
            else:
                pid = hex_bin(x.reshape(-1), y.reshape(-1), sc.sensor)[0]
            ps.append(pid)
    if mode in ("image", "response"):
        return img
    if not xs:
        return np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0, np.int64)
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(vs), np.concatenate(ps)


def edge_margin(x, y, sensor):
    """Distance (in pixel units) of each ray's sensor coordinate from the
    nearest binning decision boundary -- used to define the edge band.
    """
    if isinstance(sensor, SquareSensorSpec):
        x0, y0, dx, dy = sensor.params()
        xc = (x.astype(F64) - x0) / dx
        yc = (y.astype(F64) - y0) / dy
        mx = np.abs(xc - np.round(xc))
        my = np.abs(yc - np.round(yc))
        m = np.minimum(mx, my)
        if sensor.edge_width > 0:
            fx = xc - np.floor(xc)
            fy = yc - np.floor(yc)
            dist = np.minimum(np.minimum(fx, 1 - fx) * dx, np.minimum(fy, 1 - fy) * dy)
            m = np.minimum(m, np.abs(dist - sensor.edge_width) / min(dx, dy))
        return m
    xg, yg = hex_grid_coords(x.astype(F64), y.astype(F64), sensor)
    q, r = _to_axial(xg, yg, sensor.hex_size, F64)
    qi, ri = axial_round(q, r)
    cx, cy = _from_axial(qi, ri, sensor.hex_size, F64)
    hd = hex_norm(xg - cx, yg - cy, sensor.inradius, F64)
    thr = 1.0 - sensor.edge_width / sensor.inradius
    # hex_norm == 1 on the cell boundary; the margin is the distance of the
    # normalised hex radius from the threshold (and from the rounding tie).
    return np.abs(thr - hd) if sensor.edge_width > 0 else np.abs(1.0 - hd)


# ---------------------------------------------------------------------------
# sampler: Philox4x32-10 counter RNG (Salmon et al., SC'11) + sample maps
# ---------------------------------------------------------------------------

_PHILOX_M0 = 0xD2511F53
_PHILOX_M1 = 0xCD9E8D57
_PHILOX_W0 = 0x9E3779B9
_PHILOX_W1 = 0xBB67AE85
_M32 = 0xFFFFFFFF


def philox4x32(ctr, key, rounds=10):
    """Philox4x32 on uint64-held uint32 arrays.  ctr: 4 arrays, key: 2 arrays."""
    c0, c1, c2, c3 = [np.asarray(c, np.uint64) & _M32 for c in ctr]
    k0, k1 = [np.asarray(k, np.uint64) & _M32 for k in key]
    for _ in range(rounds):
        p0 = np.uint64(_PHILOX_M0) * c0
        p1 = np.uint64(_PHILOX_M1) * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(_M32)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(_M32)
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
        k0 = (k0 + np.uint64(_PHILOX_W0)) & np.uint64(_M32)
        k1 = (k1 + np.uint64(_PHILOX_W1)) & np.uint64(_M32)
    return [c.astype(np.uint32) for c in (c0, c1, c2, c3)]


def u01(u):
    """[0,1) float32 from the top 24 bits (exact)."""
    return ((np.asarray(u, np.uint32) >> np.uint32(8)).astype(F32) * F32(2.0 ** -24))


def u01_open(u):
    """(0,1] float32 from the top 24 bits (exact)."""
    return (((np.asarray(u, np.uint32) >> np.uint32(8)) + np.uint32(1)).astype(F32) * F32(2.0 ** -24))


def sample_streams(seed, group, n_facets, n_samples):
    """Raw 32-bit words for every (facet, sample) of one mirror group.

    counter = (sample, facet, group, stream), key = (seed lo, seed hi);
    stream 0 drives the aperture position, stream 1 the roughness draws.
    """
    m = np.arange(n_samples, dtype=np.uint64)[None, :]
    f = np.arange(n_facets, dtype=np.uint64)[:, None]
    m, f = np.broadcast_arrays(m, f)
    k = (np.uint64(seed & _M32), np.uint64((seed >> 32) & _M32))
    g = np.uint64(group)
    a = philox4x32((m, f, np.full_like(m, g), np.zeros_like(m)), k)
    b = philox4x32((m, f, np.full_like(m, g), np.ones_like(m)), k)
    return a, b


def box_muller(b0, b1):
    """Two N(0,1) draws from two words (fp32)."""
    ua = u01_open(b0)
    ub = u01(b1)
    rad = np.sqrt(F32(-2.0) * np.log(ua))
    ang = F32(2.0 * np.pi) * ub
    return rad * np.cos(ang), rad * np.sin(ang)


def sample_disk_xy(a0, a1, radius):
    """sampling.py:8-27: r = sqrt(U1), theta = U2 * 2 * pi; scaled by radius."""
    r = np.sqrt(u01(a0))
    th = u01(a1) * F32(2) * F32(np.pi)
    return np.stack([r * np.cos(th), r * np.sin(th)], -1) * F32(radius)


def sample_polygon_xy(a0, a1, a2, verts):
    """sampling.py:30-67: fan triangles from v0, area-weighted choice, sqrt(U)
    barycentrics (1-u, u(1-v), uv).  Triangle index = #cumulative-area
    boundaries <= U0 * total (the build's definition of the choice)."""
    v = np.asarray(verts, F32)
    nt = len(v) - 2
    tris = np.stack([np.stack([v[0], v[i], v[i + 1]]) for i in range(1, nt + 1)])
    t0, t1, t2 = tris[:, 0], tris[:, 1], tris[:, 2]
    areas = np.abs((t1[:, 0] - t0[:, 0]) * (t2[:, 1] - t0[:, 1])
                   - (t2[:, 0] - t0[:, 0]) * (t1[:, 1] - t0[:, 1])) / F32(2)
    cum = np.cumsum(areas, dtype=F32)
    target = u01(a0) * cum[-1]
    tri = np.zeros(np.shape(a0), np.int64)
    for i in range(nt - 1):
        tri += (target >= cum[i]).astype(np.int64)
    u = np.sqrt(u01(a1))
    vv = u01(a2)
    A = F32(1) - u
    B = u * (F32(1) - vv)
    C = u * vv
    T = tris[tri]
    pts = A[..., None] * T[..., 0, :] + B[..., None] * T[..., 1, :] + C[..., None] * T[..., 2, :]
    return pts, tri


def polygon_area(verts):
    """Shoelace area (integrators.py:169-175)."""
    v = np.asarray(verts, F32)
    x, y = v[:, 0], v[:, 1]
    return F32(0.5) * np.abs(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y, dtype=F32))


def sample_group(seed, group, surf: Surface, offsets, n_samples, radii=None, vertices=None):
    """MCIntegrator._sample_disk_group / _sample_polygon_group (integrators.py:97-188)
    with the build's Philox streams.  Returns points, normals, delta, weights
    (fp32), weights shaped (Nf, M).
    """
    nf = len(offsets)
    a, b = sample_streams(seed, group, nf, n_samples)
    pts, nrm = [], []
    areas = []
    for i in range(nf):
        if vertices is None:
            xy = sample_disk_xy(a[0][i], a[1][i], radii[i])
            areas.append(F32(np.pi) * F32(radii[i]) ** 2)
        else:
            xy, _ = sample_polygon_xy(a[0][i], a[1][i], a[2][i], vertices[i])
            areas.append(polygon_area(vertices[i]))
        p, n = surface_point_normal(xy[:, 0], xy[:, 1], surf, offsets[i], F32)
        pts.append(p)
        nrm.append(n)
    points = np.stack(pts).astype(F32)
    normals = np.stack(nrm).astype(F32)
    th1, th2 = box_muller(b[0], b[1])
    delta = perturbation_delta(normals, th1, th2).astype(F32)
    area = np.asarray(areas, F32)[:, None]
    weights = normals[..., 2] / area * F32(n_samples)
    return points, normals, delta, weights.astype(F32)
