"""CPU restatements of numerical shortcuts the HIP kernels take, checked to
be conservative against the oracle (the kernels themselves are checked on the
GPU by the parity tests)."""
import numpy as np

import oracle.iact_oracle as O


def _far(o, d, p1, ax, r):
    """The cylinder test's early out (iact_device.h, hit_cylinder) in fp32:
    the ray's line is farther than ~5e-5 m outside the surface, with a bound
    on the rounding of the discriminant."""
    f = np.float32
    oc = (o - p1).astype(f)
    oca = oc @ ax
    rda = d @ ax
    ocp = oc - oca[:, None] * ax
    rdp = d - rda[:, None] * ax
    a = np.sum(rdp * rdp, 1)
    b = f(2) * np.sum(ocp * rdp, 1)
    cc = np.sum(ocp * ocp, 1) - r * r
    disc = b * b - f(4) * a * cc
    margin = f(4) * a * (f(1e-4) * r) + f(1e-6) * (f(4) * a * np.abs(cc) + b * b)
    return -disc > margin


def test_cylinder_early_out_never_skips_a_hit():
    """Rays aimed at random points close to a long thin cylinder (the
    telescope-support geometry): a ray the early out skips is a miss of the
    reference's fp32 test, and of the fp64 test with the radius grown by
    2e-5 m (so it is outside the shadow band)."""
    rng = np.random.default_rng(0)
    f = np.float32
    p1 = np.array([10.0, 10.0, 1.4])
    p2 = np.array([1.0, 1.0, 36.25])
    r = 0.1
    args32 = O.cylinder_consts(p1, p2, r, f)
    args64g = O.cylinder_consts(p1, p2, r + 2e-5, np.float64)
    n = 200000
    # origins on a dish-like disc, targets near the cylinder axis
    o = np.stack([rng.uniform(-12, 12, n), rng.uniform(-12, 12, n), rng.uniform(0, 3, n)], 1)
    s = rng.uniform(0, 1, n)[:, None]
    axis_pt = p1 + s * (p2 - p1)
    jitter = rng.normal(0, 1, (n, 3)) * rng.choice([0.05, 0.1, 0.15, 0.3], n)[:, None]
    tgt = axis_pt + jitter
    d = tgt - o
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    o32, d32 = o.astype(f), d.astype(f)
    far = _far(o32, d32, args32[0], args32[1], args32[3])
    hit32 = np.isfinite(O.intersect_cylinder(o32, d32, *args32))
    hit64g = np.isfinite(O.intersect_cylinder(o, d, *args64g))
    print(f"rays {n}: far {far.mean():.3f}, hits {hit32.mean():.3f}")
    assert 0.2 < far.mean() < 0.95 and hit32.mean() > 0.02
    assert not np.any(far & hit32)
    assert not np.any(far & hit64g)


def _hex_cell_staged(q, r, tq, tr):
    """hex_cell_staged (iact_device.h) in fp32: the centre from the floors of
    the three cube-coordinate differences, as the staged table's row and
    column (cell + 1) clamped into [0, tq + 1] x [0, tr + 1] before the
    conversion."""
    f = np.float32
    q, r = q.astype(f), r.astype(f)
    A = np.floor(f(2) * q + r)
    B = np.floor(-q - f(2) * r)
    C = np.floor(r - q)
    third = f(1.0 / 3.0)
    row = np.clip((A - C) * third + f(1.5), f(0), f(tq + 1))
    col = np.clip((C - B) * third + f(1.5), f(0), f(tr + 1))
    return np.floor(row).astype(np.int64), np.floor(col).astype(np.int64)


def test_hex_cell_staged_is_cube_rounding():
    """The kernels' nearest-hexagon formula against cube rounding (the
    oracle's axial_round, hexagonal.py:32-39) on random axial points: the
    same cell (row - 1, column - 1) for every point farther than 1e-4 from a
    cell boundary (every boundary lies on an integer line of 2q + r, -q - 2r
    or r - q); on a small table, exactly the cells off it land on the staged
    border (row or column 0 or tq + 1 / tr + 1); and a planted off-by-one in
    the division is caught."""
    rng = np.random.default_rng(3)
    q = rng.uniform(-300.0, 300.0, 400_000)
    r = rng.uniform(-300.0, 300.0, 400_000)
    qi, ri = (np.asarray(a) for a in O.axial_round(q, r))
    # distance (in the pairwise differences) from the nearest integer line
    d = np.stack([2 * q + r, -q - 2 * r, r - q])
    margin = np.abs(d - np.round(d)).min(0)
    clear = margin > 1e-4
    # (a table of 1000 x 1000 cells holding the points shifted by 400 cells)
    big = 1000
    fq, fr = _hex_cell_staged(q + 400.0, r + 400.0, big, big)
    assert clear.mean() > 0.99
    assert np.array_equal(fq[clear] - 401, qi[clear])
    assert np.array_equal(fr[clear] - 401, ri[clear])
    # a 40 x 25 table: on it, the cell; off it, a border row or column
    tq, tr = 40, 25
    sq, sr = _hex_cell_staged(q, r, tq, tr)
    on = (qi >= 0) & (qi < tq) & (ri >= 0) & (ri < tr)
    border = (sq == 0) | (sq == tq + 1) | (sr == 0) | (sr == tr + 1)
    assert on[clear].any() and (~on[clear]).any()
    assert np.array_equal(border[clear], ~on[clear])
    assert np.array_equal(sq[clear & on] - 1, qi[clear & on])
    assert np.array_equal(sr[clear & on] - 1, ri[clear & on])
    assert sq.min() >= 0 and sq.max() <= tq + 1 and sr.min() >= 0 and sr.max() <= tr + 1
    # integer centres map to themselves, also at the table's far corner
    c = np.arange(0, 1000, dtype=np.float64)
    cq, cr = _hex_cell_staged(c, c[::-1].copy(), big, big)
    assert np.array_equal(cq - 1, c.astype(np.int64))
    assert np.array_equal(cr - 1, c[::-1].astype(np.int64))
    # planted mistake: (A - C) div 3 without the +1 misses the cells
    A = np.floor(2 * q + r)
    C = np.floor(r - q)
    bad = np.floor((A - C) / 3.0).astype(np.int64)
    assert not np.array_equal(bad[clear], qi[clear])
