"""The C++/OpenMP restatement (CPU baseline, FP64 run, adjoint) agrees with the numpy oracle and the twin."""
import numpy as np
import pytest

import oracle.iact_oracle as O
from oracle import cpu_ref
from tests.test_oracle import _single_facet_scene


def _scenes():
    from iactrace_amd import configs
    from tests.helpers import oracle_obstructions, oracle_sensor
    from iactrace_amd.obstructions import group_obstructions
    out = []
    # C3-like: parabolic polygon facets, hex camera, cylinders + box
    wl = configs.workload("C3", n_sources=3, n_samples=8)
    ms = [m for m in wl.mirrors][:120]
    verts = ms[0].aperture.vertices
    P, N, D, W = O.sample_group(4, 0, O.Surface(0.0137, 0.0), np.zeros((len(ms), 2)), 8,
                                vertices=np.stack([verts] * len(ms)))
    sc = O.Scene(P, N, D, W, np.stack([m.position.numpy() for m in ms]),
                 np.stack([m.rotation.numpy() for m in ms]), np.zeros(len(ms), np.float32),
                 oracle_sensor(wl.sensor), oracle_obstructions(group_obstructions(wl.obstructions)))
    src, vals = configs.point_source_field(3, 1.0, seed=4)
    out.append(("C3", sc, src, vals, "point"))
    # Cassegrain with the asphere and hyperboloid, square sensor
    sec = O.MirrorSpec((0, 0, 6.0), (180.0, 0, 0), O.Surface(-0.05, -16.0), (0, 0), radius=0.5)
    sensor = O.SquareSensorSpec((0, 0, -0.45), (0, 0, 0), 64, 64, (-0.3, 0.3, -0.3, 0.3))
    sc = _single_facet_scene(O.Surface(0.05, -1.05, np.array([2e-4, -1e-6, 0, 0])), 2.0, sensor,
                             m=3000, mirrors=[sec],
                             obstructions=O.Obstructions(spheres=[(np.array([0, 0.5, 3.0]), 0.3)]))
    src = configs.parallel_grid(4, 3, 0.2)
    out.append(("C4", sc, src, np.ones(len(src), np.float32), "parallel"))
    return out


@pytest.mark.parametrize("case", [0, 1])
def test_cpu_ref_matches_oracle(case):
    name, sc, src, vals, st = _scenes()[case]
    keep = []
    s = cpu_ref.scene_struct(sc, keep)
    img = cpu_ref.render(s, src, vals, st, sc.sensor.n_pixels, nthreads=4)
    ref = O.render(sc, src, vals, st, "ref32", "image")
    tot = np.abs(ref).sum()
    assert tot > 0
    assert np.abs(img - ref).sum() <= 1e-4 * tot, (name, np.abs(img - ref).sum() / tot)


@pytest.mark.parametrize("case", [0, 1])
def test_cpu_ref_fp64_run_matches_oracle_fp64(case):
    """The FP64 instantiation against the numpy oracle's fp64 path.  Both run
    in double, but the library's scene descriptor holds fp32 numbers -- the
    stage mirrors' rotation matrices among them (sin 180 deg rounds to
    -8.7e-8 in fp32) -- so rays near a pixel edge may differ: every pixel
    agrees within its edge-band mass, and the image to 1e-3 in L1."""
    import copy
    name, sc, src, vals, st = _scenes()[case]
    # the descriptors hold surface parameters in fp32: give the oracle the
    # same numbers
    f32 = lambda x: float(np.float32(x))  # noqa: E731
    sc.mirrors = [copy.copy(m) for m in sc.mirrors]
    for m in sc.mirrors:
        m.surface = O.Surface(f32(m.surface.curvature), f32(m.surface.conic),
                              np.asarray(m.surface.aspheric, np.float32).astype(np.float64))
    keep = []
    s = cpu_ref.scene_struct(sc, keep)
    img, band, _ = cpu_ref.render_band(s, src, vals, st, sc.sensor.n_pixels, 4, 64, 1e-3)
    ref = O.render(sc, src, vals, st, "fp64", "image")
    tot = np.abs(ref).sum()
    assert tot > 0
    assert np.all(np.abs(img - ref) <= band + 1e-9 * np.abs(ref).max()), name
    assert np.abs(img - ref).sum() <= 1e-3 * tot, (name, np.abs(img - ref).sum() / tot)
    if case == 0:  # no stage mirrors: the same numbers in, the same image out
        assert np.abs(img - ref).sum() <= 1e-9 * tot


@pytest.mark.parametrize("case", [0, 1])
def test_band_image_bounds_fp32_vs_fp64(case):
    """Every pixel's fp32-vs-fp64 difference lies within its band mass: the
    rays that bin differently are all within the band of a boundary."""
    name, sc, src, vals, st = _scenes()[case]
    keep = []
    s = cpu_ref.scene_struct(sc, keep)
    P = sc.sensor.n_pixels
    i64, b64, stats = cpu_ref.render_band(s, src, vals, st, P, 4, 64, 1e-3)
    i32 = cpu_ref.render(s, src, vals, st, P, 4, 32)
    assert stats["traced"] == len(src) * sc.points.shape[0] * sc.points.shape[1]
    assert 0 < stats["binned"] <= stats["traced"]
    slack = 1e-5 * np.abs(i64) + 1e-7 * np.abs(i64).max()
    assert np.all(np.abs(i32 - i64) <= b64 + slack), name
    # a band of zero width explains nothing; a planted offset of the fp32
    # image is caught
    assert not np.all(np.abs(i32 * (1 + 1e-3) - i64) <= b64 + slack)


@pytest.mark.parametrize("case", ["hex-point-obstructed", "square-parallel"])
def test_cpu_ref_grad_matches_twin(case):
    """The CPU port's adjoint (rotation and value gradients of a weighted
    hard-binned image) against float64 torch autograd of the oracle twin."""
    import torch
    import oracle.torch_twin as TW
    if case == "hex-point-obstructed":
        name, sc, src, vals, st = _scenes()[0]
        sc.points, sc.normals, sc.delta, sc.weights = (a[:40] for a in (sc.points, sc.normals,
                                                                         sc.delta, sc.weights))
        sc.positions, sc.rotations, sc.scale = sc.positions[:40], sc.rotations[:40], sc.scale[:40]
        rng = np.random.default_rng(3)
        sc.scale = rng.uniform(0, 2e-3, 40).astype(np.float32)
    else:
        sensor = O.SquareSensorSpec((0, 0, 5.0), (0, 0, 0), 16, 16, (-0.5, 0.5, -0.5, 0.5))
        sc = _single_facet_scene(O.Surface(0.1, -1.0), 1.0, sensor, m=2000,
                                 rotation=(0.4, -0.3, 10.0))
        src = np.array([[0.01, 0.0, -1.0], [-0.02, 0.01, -1.0]])
        vals = np.array([1.0, 0.6], np.float32)
        st = "parallel"
    P = sc.sensor.n_pixels
    W = 1.0 + 0.5 * np.sin(np.arange(P) * 0.37)
    keep = []
    s = cpu_ref.scene_struct(sc, keep)
    g_rot, g_val = cpu_ref.grad(s, src, vals, st, W, sc.points.shape[0], nthreads=4)
    rot = torch.tensor(sc.rotations, dtype=torch.float64, requires_grad=True)
    v = torch.tensor(vals, dtype=torch.float64, requires_grad=True)
    img = TW.render_image(sc, src, v, st, rot)
    (img * torch.as_tensor(W)).sum().backward()
    for label, a, b in (("rotations", g_rot, rot.grad.numpy()), ("values", g_val, v.grad.numpy())):
        err = np.abs(a - b).max() / np.abs(b).max()
        print(f"[cpu grad {case}] d/d{label}: max rel err {err:.2e}")
        assert np.abs(b).max() > 0 and err < 1e-3, (label, err)
