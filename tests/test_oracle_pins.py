"""Pins of the oracle's splat sensors, polygon apertures and polygon areas
(CPU, no GPU).

Each oracle function is checked against what the geometry fixes, not against
its own formula:
  * square splat (reference square.py:144-173): the taps of a ray sit on the
    (2k+1)^2 lattice points around floor(x) and weigh exp(-d^2 / (2 sigma^2))
    in the distance d (pixel units) from the ray to the lattice point --
    at a lattice point the neighbour / centre ratio is exp(-0.5/sigma^2),
    the diagonal one exp(-1/sigma^2); the heaviest tap is the lattice point
    nearest the ray; an interior ray's taps sum to its value;
  * hex splat (hexagonal.py:264-315): for a ray at a pixel centre the ring-1
    taps are exactly the six pixels whose centres lie 2 inradii away (found
    by brute force in the centre list), each weighing exp(-2/sigma^2) of the
    centre tap; on the segment between two neighbouring centres, at fraction
    t, the ratio of the two taps is exp(-2 (1 - 2t) / sigma^2); taps sum to
    the value;
  * polygon aperture (mirrors.py:209-220): agrees with scipy's Delaunay
    point location for counter-clockwise vertices, and (quirk Q12: the test
    requires every edge cross product >= 0) rejects every interior point of
    a clockwise polygon;
  * polygon area (integrators.py:169-175): the regular hexagon's
    sqrt(3)/2 D^2 and scipy's convex-hull area; and the whole path: a flat
    hexagonal facet lit normally sends exactly its area to the sensor.
Planted mistakes (a flipped tap offset, sigma for sigma^2, no weight
normalisation, a reversed edge test, a missing 1/2 in the shoelace) are each
caught by one of these pins.
"""
import math

import numpy as np
import pytest
from scipy.spatial import ConvexHull, Delaunay

import oracle.iact_oracle as O
from iactrace_amd.configs import hex_camera, hex_facet_vertices

F64 = np.float64


# ------------------------------------------------------------------ square splat
def _sq(sigma=0.7, k=2, w=8, h=6):
    return O.SquareSensorSpec((0, 0, 0), (0, 0, 0), w, h, (0.0, 0.25 * w, 0.0, 0.25 * h),
                              splat=True, sigma=sigma, kernel_size=k)


def _taps(out):
    """{pixel: weight} of one ray from the oracle's per-tap lists."""
    d = {}
    for pid, val in out:
        p = int(pid[0])
        if p >= 0:
            d[p] = d.get(p, 0.0) + float(val[0])
    return d


def test_square_splat_lattice_point_ratios():
    s = _sq()
    # (1.0, 0.75) is lattice point (4, 3) exactly (dx = dy = 0.25)
    t = _taps(O.square_splat(np.array([1.0]), np.array([0.75]), np.array([1.0]), s))
    c = t[3 * 8 + 4]
    sg2 = s.sigma ** 2
    for (ox, oy), d2 in [((1, 0), 1), ((-1, 0), 1), ((0, 1), 1), ((0, -1), 1), ((1, 1), 2),
                         ((-1, 1), 2), ((2, 0), 4), ((2, 1), 5), ((-2, -2), 8)]:
        got = t[(3 + oy) * 8 + 4 + ox] / c
        assert got == pytest.approx(math.exp(-0.5 * d2 / sg2), rel=1e-12), (ox, oy)
    assert len(t) == 25


def test_square_splat_nearest_tap_is_heaviest():
    s = _sq()
    rng = np.random.default_rng(5)
    x = rng.uniform(0.6, 1.4, 200)
    y = rng.uniform(0.6, 0.9, 200)
    outs = O.square_splat(x, y, np.ones_like(x), s)
    pids = np.stack([p for p, _ in outs])
    vals = np.stack([v for _, v in outs])
    best = pids[np.argmax(vals, 0), np.arange(len(x))]
    nx = np.floor(x / 0.25 + 0.5).astype(int)
    ny = np.floor(y / 0.25 + 0.5).astype(int)
    assert np.array_equal(best, ny * 8 + nx)


def test_square_splat_taps_sum_to_value():
    s = _sq(sigma=0.5, k=2, w=12, h=12)
    rng = np.random.default_rng(6)
    x = rng.uniform(0.5 + 0.01, 2.5 - 0.01, 300)   # >= 2 px from every border
    y = rng.uniform(0.5 + 0.01, 2.5 - 0.01, 300)
    v = rng.uniform(0.5, 2.0, 300)
    outs = O.square_splat(x, y, v, s)
    tot = np.sum([val for _, val in outs], 0)
    assert np.allclose(tot, v, rtol=1e-12)
    # near a border the taps outside the sensor are dropped: less than v
    outs = O.square_splat(np.array([0.01]), np.array([1.5]), np.array([1.0]), s)
    assert 0.2 < np.sum([val for _, val in outs]) < 0.99


# ------------------------------------------------------------------ hex splat
def _hex(rot_deg=0.0, sigma=0.6, k=1):
    c = hex_camera(300, 0.05).astype(F64)
    a = math.radians(rot_deg)
    R = np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])
    c = (c @ R.T).astype(np.float32)
    return O.HexSensorSpec((0, 0, 0), (0, 0, 0), c, splat=True, sigma=sigma, kernel_size=k)


@pytest.mark.parametrize("rot", [0.0, 17.0])
def test_hex_splat_ring_one_at_a_centre(rot):
    s = _hex(rot)
    c = s.centers.astype(F64)
    two_r = 2.0 * s.inradius
    checked = 0
    for p in [0, 7, 40, 123]:
        d = np.hypot(c[:, 0] - c[p, 0], c[:, 1] - c[p, 1])
        ring = set(np.nonzero(np.abs(d - two_r) < 1e-4 * two_r)[0].tolist())
        if len(ring) != 6:
            continue  # a pixel on the camera's border
        t = _taps(O.hex_splat(np.array([c[p, 0]]), np.array([c[p, 1]]), np.array([1.0]), s))
        assert set(t) == ring | {p}
        for j in ring:
            assert t[j] / t[p] == pytest.approx(math.exp(-2.0 / s.sigma ** 2), rel=1e-4)
        assert sum(t.values()) == pytest.approx(1.0, rel=1e-12)
        checked += 1
    assert checked >= 3


def test_hex_splat_between_two_centres():
    s = _hex(0.0)
    c = s.centers.astype(F64)
    p = 40
    d = np.hypot(c[:, 0] - c[p, 0], c[:, 1] - c[p, 1])
    q = int(np.nonzero(np.abs(d - 2 * s.inradius) < 1e-4)[0][0])
    for tfrac in (0.1, 0.3, 0.45, 0.6, 0.85):
        x, y = c[p] + tfrac * (c[q] - c[p])
        t = _taps(O.hex_splat(np.array([x]), np.array([y]), np.array([1.0]), s))
        want = math.exp(-2.0 * (1 - 2 * tfrac) / s.sigma ** 2)
        assert t[q] / t[p] == pytest.approx(want, rel=1e-4), tfrac


def test_hex_splat_taps_sum_to_value():
    s = _hex(5.0, sigma=0.4, k=2)
    rng = np.random.default_rng(8)
    c = s.centers.astype(F64)
    idx = rng.integers(0, 40, 100)             # rays near the camera's middle
    x = c[idx, 0] + rng.uniform(-0.02, 0.02, 100)
    y = c[idx, 1] + rng.uniform(-0.02, 0.02, 100)
    v = rng.uniform(0.5, 2.0, 100)
    ok = np.hypot(x, y) < 0.3
    outs = O.hex_splat(x[ok], y[ok], v[ok], s)
    assert np.allclose(np.sum([val for _, val in outs], 0), v[ok], rtol=1e-12)


# ------------------------------------------------------------------ polygon aperture
def _ap(verts):
    return O.MirrorSpec((0, 0, 0), (0, 0, 0), O.Surface(0.0, 0.0), (0, 0), vertices=verts)


@pytest.mark.parametrize("which", ["hexagon", "irregular"])
def test_polygon_aperture_vs_delaunay(which):
    if which == "hexagon":
        ccw = hex_facet_vertices(0.9)[::-1].astype(F64)   # builder order is clockwise
    else:
        ccw = np.array([[0, 0], [2, 0], [2.5, 1], [1, 2], [-0.2, 1.0]], F64)
    cr = lambda a, b: a[0] * b[1] - a[1] * b[0]  # noqa: E731
    assert cr(ccw[1] - ccw[0], ccw[2] - ccw[0]) > 0
    rng = np.random.default_rng(9)
    lo, hi = ccw.min(0) - 0.3, ccw.max(0) + 0.3
    pts = rng.uniform(lo, hi, (4000, 2))
    inside = Delaunay(ccw).find_simplex(pts) >= 0
    # distance to the boundary, to leave out points within rounding of it
    edge = np.full(len(pts), np.inf)
    for i in range(len(ccw)):
        a, b = ccw[i], ccw[(i + 1) % len(ccw)]
        ab = b - a
        tt = np.clip(((pts - a) @ ab) / (ab @ ab), 0, 1)
        edge = np.minimum(edge, np.linalg.norm(pts - (a + tt[:, None] * ab), axis=1))
    far = edge > 1e-9
    got = O._aperture_ok(_ap(ccw), pts[:, 0], pts[:, 1], F64)
    assert inside.sum() > 500 and (~inside).sum() > 500
    assert np.array_equal(got[far], inside[far])
    # clockwise vertices: no interior point passes (every cross product < 0)
    got_cw = O._aperture_ok(_ap(ccw[::-1].copy()), pts[:, 0], pts[:, 1], F64)
    assert not got_cw[far & inside].any()


# ------------------------------------------------------------------ polygon area
def test_polygon_area_closed_forms():
    for D in (0.6, 0.9, 1.3):
        v = hex_facet_vertices(D)
        assert float(O.polygon_area(v)) == pytest.approx(math.sqrt(3) / 2 * D * D, rel=1e-6)
    rng = np.random.default_rng(10)
    for _ in range(5):
        pts = rng.uniform(-1, 1, (40, 2))
        hull = pts[ConvexHull(pts).vertices]
        assert float(O.polygon_area(hull)) == pytest.approx(ConvexHull(pts).volume, rel=1e-5)
        assert float(O.polygon_area(hull[::-1])) == pytest.approx(ConvexHull(pts).volume, rel=1e-5)


def test_flat_hexagonal_facet_flux_equals_area():
    """Whole path: a flat hexagonal facet lit along its normal by a parallel
    source sends exactly its area to a large sensor (each ray carries
    area / M, whatever the samples)."""
    D = 0.9
    v = hex_facet_vertices(D)
    P, N, Dl, W = O.sample_group(12, 0, O.Surface(0.0, 0.0), np.zeros((1, 2)), 3000,
                                 vertices=v[None])
    sensor = O.SquareSensorSpec((0, 0, 5), (0, 0, 0), 8, 8, (-4, 4, -4, 4))
    sc = O.Scene(P, N, Dl, W, np.zeros((1, 3), np.float32), np.zeros((1, 3), np.float32),
                 np.zeros(1, np.float32), sensor, O.Obstructions(), [])
    for prec in ("fp64", "ref32"):
        img = O.render(sc, np.array([[0, 0, -1.0]]), np.ones(1, np.float32), "parallel", prec)
        assert img.sum() == pytest.approx(math.sqrt(3) / 2 * D * D, rel=1e-5), prec


# ------------------------------------------------------------------ planted mistakes
def test_planted_square_tap_sign_is_caught(monkeypatch):
    orig = O.square_splat

    def bad(x, y, v, s):  # the ray mirrored about its pixel corner: offsets flip
        x0, y0, dx, dy = s.params()
        xb = x0 + dx * np.floor((x - x0) / dx)
        yb = y0 + dy * np.floor((y - y0) / dy)
        return orig(2 * xb - x + dx * 1e-9, 2 * yb - y + dy * 1e-9, v, s)
    monkeypatch.setattr(O, "square_splat", bad)
    with pytest.raises(AssertionError):
        test_square_splat_nearest_tap_is_heaviest()


def test_planted_sigma_power_is_caught(monkeypatch):
    orig = O.square_splat

    def bad(x, y, v, s):  # sigma where sigma^2 belongs
        import copy
        s2 = copy.copy(s)
        s2.sigma = math.sqrt(s.sigma)
        return orig(x, y, v, s2)
    monkeypatch.setattr(O, "square_splat", bad)
    with pytest.raises(AssertionError):
        test_square_splat_lattice_point_ratios()


def test_planted_unnormalised_hex_weights_are_caught(monkeypatch):
    orig = O.hex_splat

    def bad(x, y, v, s):
        out = orig(x, y, v, s)
        return [(p, val * 1.25) for p, val in out]
    monkeypatch.setattr(O, "hex_splat", bad)
    with pytest.raises(AssertionError):
        test_hex_splat_taps_sum_to_value()


def test_planted_reversed_edge_test_is_caught(monkeypatch):
    orig = O._aperture_ok

    def bad(m, x, y, dt):
        if m.vertices is None:
            return orig(m, x, y, dt)
        return orig(O.MirrorSpec(m.position, m.rotation, m.surface, m.offset,
                                 vertices=np.asarray(m.vertices)[::-1].copy()), x, y, dt)
    monkeypatch.setattr(O, "_aperture_ok", bad)
    with pytest.raises(AssertionError):
        test_polygon_aperture_vs_delaunay("irregular")


def test_planted_shoelace_half_is_caught(monkeypatch):
    orig = O.polygon_area
    monkeypatch.setattr(O, "polygon_area", lambda v: 2 * orig(v))
    with pytest.raises(AssertionError):
        test_polygon_area_closed_forms()
