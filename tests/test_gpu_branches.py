"""Per-ray parity on the branches of the path that the configuration tests do
not reach (GPU against the oracle, reduced sizes).

Every case goes through tests.test_gpu_forward._trace_and_compare: the same
injected samples through the kernels (image in the requested layout, debug
trace with per-ray pixel ids) and the oracle; pixel ids identical outside
the edge band, values to 1e-5, the image to 1e-5 per pixel.
  * every obstruction kind -- sphere, oriented box, triangle, box, cylinder
    (reference intersections.py:44-226, obstructions.py:125-255) -- over a
    Davies-Cotton dish, both image layouts, culling on and off; each kind
    blocks rays on its own (counted with the oracle);
  * a Cassegrain with obstructions on both legs (render.py:76: the stage-leg
    shadow, tested from the primary hit toward the secondary), with parallel
    and point sources (at 1e4 m and 1e10 m) through the secondary, and its
    response matrix;
  * two stage-1 groups (render.py:62-71, 108-113): a group of disks with
    two identical mirrors (a tie: the first index wins) and a polygon group
    (mirrors.py:209-220) at the same pose (a tie across groups: the earlier
    group wins) or lower (the nearer hit wins); the aspheric gradients of
    the two groups' surfaces against the float64 twin show which group took
    each tied ray; a clockwise polygon, which the reference's edge test
    rejects everywhere (quirk Q12);
  * sensors with an edge mask (square.py:79, hexagonal.py:188-189) through
    the sample-major image trace, the source-major trace, the debug trace,
    the response matrix and the adjoint;
  * the reference's own telescopes, CT3 (380 disk facets, 33 masts, the
    hexagonal science camera with edge_width 0.0004, the 1536 x 1024 lid
    camera whose image does not fit in LDS) and CT5 (876 hexagonal facets,
    271 obstructions; every 6th facet, so that the oracle's facet loop stays
    short), from tests/golden/hess_ct*.npz.
"""
import math

import numpy as np
import pytest
import torch

import iactrace_amd as ia
import oracle.iact_oracle as O
import oracle.torch_twin as TW
from iactrace_amd import configs
from iactrace_amd.obstructions import Box, Cylinder, OrientedBox, Sphere, Triangle
from tests.helpers import oracle_obstructions, oracle_scene
from tests.test_gpu_forward import _trace_and_compare

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _telescope(mirrors, obstructions, sensor, m, key=0):
    w = configs.Workload("case", mirrors, obstructions, sensor, m, np.zeros((1, 3)),
                         np.ones(1, np.float32), "point")
    return w.telescope(DEV, key=key)


def _rot(z_deg, x_deg):
    a, b = math.radians(z_deg), math.radians(x_deg)
    Rz = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
    Rx = np.array([[1, 0, 0], [0, math.cos(b), -math.sin(b)], [0, math.sin(b), math.cos(b)]])
    return Rz @ Rx


def _every_kind():
    return [Sphere([0.6, 0.3, 8.0], 0.35),
            OrientedBox([-0.8, 0.6, 6.0], [0.3, 0.12, 0.2], _rot(30.0, 20.0)),
            Triangle([0.2, -1.2, 5.0], [1.3, -0.6, 5.5], [0.5, -0.2, 4.5]),
            Triangle([-1.5, -1.0, 4.0], [-0.9, -1.4, 4.2], [-1.2, -1.2, 6.0]),
            Box([-1.0, -1.2, 9.0], [-0.5, -0.7, 9.5]),
            Cylinder([-2.0, 0.0, 7.0], [2.0, 0.2, 7.5], 0.05)]


@pytest.mark.parametrize("layout,cull", [(1, 0), (2, 0), (2, -1), (1, -1)])
def test_every_obstruction_kind(layout, cull):
    wl = configs.workload("C2", n_sources=150, n_samples=16)
    tel = _telescope(configs.davies_cotton_dish(n_facets=37), _every_kind(), wl.sensor, 16,
                     key=21)
    src, vals = configs.point_source_field(150, 1.5, seed=21)
    if layout == 1 and cull == 0:
        # each kind shadows rays on its own (oracle, one kind at a time)
        for ob in _every_kind():
            sc1 = oracle_scene(tel)
            sc1.obstructions = oracle_obstructions(ia.group_obstructions([ob]))
            _, _, v, _ = O.render(sc1, src[:20], vals[:20], "point", "fp64", "debug")
            n1 = int((v == 0).sum())  # the facets face the sources: 0 = shadowed
            print(f"[kinds] {type(ob).__name__}: {n1} of {v.size} rays shadowed alone",
                  flush=True)
            assert n1 > 0, type(ob).__name__
    _trace_and_compare(tel, src, vals, "point", f"every kind layout {layout} cull {cull}",
                       launch=ia.LaunchConfig(layout=layout, culling=cull))


def _cassegrain_with(obstructions, stage_mirrors=None, secondary_aspheric=None):
    mirrors, sensor = configs.cassegrain()
    if stage_mirrors is not None:
        mirrors = [mirrors[0]] + stage_mirrors
    return _telescope(mirrors, obstructions, sensor, 256, key=4)


def test_cassegrain_stage_leg_shadow_parallel_and_point_sources():
    ob = [Sphere([0.8, 0.0, 3.0], 0.1), Cylinder([-2.2, -0.05, 4.5], [2.2, 0.05, 4.5], 0.02),
          Box([0.9, 0.9, 1.0], [1.2, 1.3, 1.4])]
    tel = _cassegrain_with(ob)
    sc = oracle_scene(tel)
    par = configs.parallel_grid(4, 3, 0.05)
    # rays lost on the stage leg only: shadowed with the obstructions, lit
    # without them, and not shadowed toward the source
    sc0 = oracle_scene(tel)
    sc0.obstructions = O.Obstructions()
    _, _, v, _ = O.render(sc, par, np.ones(len(par), np.float32), "parallel", "fp64", "debug")
    _, _, v0, _ = O.render(sc0, par, np.ones(len(par), np.float32), "parallel", "fp64", "debug")
    wx = O.transform_to_world(sc, "fp64")[0].reshape(-1, 3)
    d = np.asarray(par, np.float64)
    ns, M = len(par), sc.points.shape[1]
    o = np.repeat(wx[None], ns, 0).reshape(-1, 3)
    dd = np.repeat(d[:, None], M, 1).reshape(-1, 3)
    toward = O.check_occlusions(o, -dd, sc.obstructions, np.float64)
    leg = (v == 0) & (v0 != 0) & (toward != 0)
    print(f"[cassegrain] rays shadowed on the stage leg only: {int(leg.sum())}", flush=True)
    assert leg.sum() > 0
    _trace_and_compare(tel, par, np.linspace(0.5, 1.5, len(par)).astype(np.float32), "parallel",
                       "cassegrain obstructed, parallel")
    for z in (1e4, 1e10):
        src, vals = configs.point_source_field(10, 0.05, seed=3, z=z)
        _trace_and_compare(tel, src, vals, "point", f"cassegrain obstructed, point z={z:g}")
    # response matrix: rows are the per-source images of the debug trace
    src, vals = configs.point_source_field(10, 0.05, seed=3, z=1e4)
    rm = ia.render_response_matrix(tel, torch.as_tensor(src), torch.as_tensor(vals),
                                   "point").cpu().numpy()
    _, vk, pk = ia.render_debug(tel, torch.as_tensor(src), torch.as_tensor(vals), "point",
                                return_pixels=True)
    vk, pk = vk.cpu().numpy(), pk.cpu().numpy()
    sidx = (np.arange(len(vk)) // M) % len(src)
    ref = np.zeros_like(rm, dtype=np.float64)
    m = pk >= 0
    np.add.at(ref, (sidx[m], pk[m]), vk[m].astype(np.float64))
    assert np.abs(ref).max() > 0
    assert np.allclose(rm, ref, rtol=1e-5, atol=1e-7 * np.abs(ref).max())


def _stage_one(kind, aspheric=(0.0, 0.0)):
    """Stage-1 mirrors: group A = three disks (two identical, radius 0.3, at
    the secondary's pose; one off to the side), group B = a hexagon
    (circumradius 0.5) at the same pose ('tie'), 5 cm lower ('lower') or with
    clockwise vertices ('clockwise')."""
    sa = ia.AsphericSurface(-0.05, configs.CASSEGRAIN_SECONDARY_CONIC, list(aspheric))
    sb = ia.AsphericSurface(-0.05, configs.CASSEGRAIN_SECONDARY_CONIC, list(aspheric))
    pose = ([0.0, 0.0, 6.0], [180.0, 0.0, 0.0])
    a = [ia.Mirror(*pose, sa, ia.DiskAperture(0.3), optical_stage=1),
         ia.Mirror(*pose, sa, ia.DiskAperture(0.3), optical_stage=1),
         ia.Mirror([0.55, 0.1, 6.0], [180.0, 0.0, 0.0], sa, ia.DiskAperture(0.12),
                   optical_stage=1)]
    hexv = configs.hex_facet_vertices(0.5 * math.sqrt(3))[::-1].copy()   # CCW
    if kind == "clockwise":
        hexv = hexv[::-1].copy()
    zb = 5.95 if kind == "lower" else 6.0
    b = [ia.Mirror([0.0, 0.0, zb], [180.0, 0.0, 0.0], sb, ia.PolygonAperture(hexv),
                   optical_stage=1)]
    return a + b


@pytest.mark.parametrize("kind", ["tie", "lower", "clockwise"])
def test_two_stage_one_groups(kind):
    tel = _cassegrain_with([], _stage_one(kind))
    assert [g.optical_stage for g in tel.mirror_groups] == [0, 1, 1]
    src = configs.parallel_grid(3, 3, 0.05)
    vals = np.linspace(0.5, 1.5, len(src)).astype(np.float32)
    img, _ = _trace_and_compare(tel, src, vals, "parallel", f"two stage-1 groups ({kind})")
    if kind == "clockwise":
        # only rays inside group A's disks reach the sensor
        sc = oracle_scene(tel)
        x, y, v, _ = O.render(sc, src, vals, "parallel", "fp64", "debug")
        assert (v != 0).sum() > 0
    assert img.sum() > 0


def test_two_stage_one_groups_aspheric_gradients_follow_the_tie():
    """Group A (disks) and group B (hexagon, same pose) tie wherever both
    apertures hold the hit: the earlier group, A, takes the ray, so B's
    surface only gets the gradient of rays outside A's disks.  Kernel and
    twin agree on both surfaces' aspheric gradients."""
    tel = _cassegrain_with([], _stage_one("tie", aspheric=(1e-3, -2e-5)))
    src = configs.parallel_grid(3, 2, 0.05)
    vals = np.array([1.0, 0.7, 1.3, 0.9, 1.1, 0.8], np.float32)
    sensor = tel.sensors[0]
    yy, xx = np.mgrid[0:sensor.height, 0:sensor.width]
    W = (1.0 + 0.3 * np.sin(xx / sensor.width * 3.0) + 0.2 * yy / sensor.height).reshape(-1)
    g0, ga, gb = tel.mirror_groups
    aa = ga.aspheric.detach().clone().requires_grad_(True)
    ab = gb.aspheric.detach().clone().requires_grad_(True)
    t2 = ia.Telescope([g0, ga.replace(aspheric=aa), gb.replace(aspheric=ab)], [], tel.sensors)
    img = ia.render(t2, torch.as_tensor(src), torch.as_tensor(vals), "parallel")
    (img.reshape(-1).double() * torch.as_tensor(W, device=DEV)).sum().backward()
    sc = oracle_scene(tel)
    ta = torch.tensor(ga.aspheric.double().numpy(), requires_grad=True)
    tb = torch.tensor(gb.aspheric.double().numpy(), requires_grad=True)
    # the scene's stage mirrors are A0, A1, A2, B
    timg = TW.render_image(sc, src, vals, "parallel", aspheric_stage={0: ta, 1: ta, 2: ta, 3: tb})
    (timg * torch.as_tensor(W)).sum().backward()
    for name, k, t in (("A", aa, ta), ("B", ab, tb)):
        a, b = k.grad.double().cpu().numpy(), t.grad.numpy()
        err = np.abs(a - b).max() / np.abs(b).max()
        print(f"[tie gradients] group {name}: kernel {a} twin {b} rel err {err:.2e}", flush=True)
        assert np.abs(b).max() > 0
        assert err < 1e-3
    # with A's disks removed every ray goes to B: B's gradient grows
    tel_b = _cassegrain_with([], _stage_one("tie", aspheric=(1e-3, -2e-5))[3:])
    gb2 = tel_b.mirror_groups[1]
    ab2 = gb2.aspheric.detach().clone().requires_grad_(True)
    t3 = ia.Telescope([tel_b.mirror_groups[0], gb2.replace(aspheric=ab2)], [], tel_b.sensors)
    img = ia.render(t3, torch.as_tensor(src), torch.as_tensor(vals), "parallel")
    (img.reshape(-1).double() * torch.as_tensor(W, device=DEV)).sum().backward()
    g_all, g_tie = ab2.grad.double().cpu().numpy(), ab.grad.double().cpu().numpy()
    tb2 = torch.tensor(gb2.aspheric.double().numpy(), requires_grad=True)
    timg = TW.render_image(oracle_scene(tel_b), src, vals, "parallel", aspheric_stage={0: tb2})
    (timg * torch.as_tensor(W)).sum().backward()
    t_all = tb2.grad.numpy()
    print(f"[tie gradients] group B alone: kernel {g_all} twin {t_all}", flush=True)
    assert np.abs(g_all - t_all).max() < 1e-3 * np.abs(t_all).max()
    # the tied rays moved part of B's gradient to A (by far more than the
    # kernel-twin difference)
    assert np.abs(g_all - g_tie).max() > 1e-2 * np.abs(g_all).max()


def _edge_tel(kind):
    if kind == "square":
        wl = configs.workload("C2", n_sources=150, n_samples=16)
        s = wl.sensor
        p = (float(s.bounds[1]) - float(s.bounds[0])) / s.width
        sensor = ia.SquareSensor(s.position, s.rotation, s.width, s.height, s.bounds,
                                 edge_width=0.15 * p)
        tel = _telescope(configs.davies_cotton_dish(n_facets=61), [], sensor, 16, key=8)
    else:
        wl = configs.workload("C3", n_sources=150, n_samples=8)
        s = wl.sensor
        sensor = ia.HexagonalSensor(s.position, s.rotation, s.hex_centers, edge_width=0.003)
        tel = _telescope(configs.parabolic_dish(n_facets=91), [], sensor, 8, key=8)
    src, vals = configs.point_source_field(150, 1.5, seed=8)
    return tel, src, vals


@pytest.mark.parametrize("kind", ["square", "hex"])
def test_edge_mask_through_every_kernel(kind):
    tel, src, vals = _edge_tel(kind)
    assert tel.sensors[0].edge_width > 0
    # sample-major and source-major images, debug pixel ids
    for layout in (1, 2):
        _trace_and_compare(tel, src, vals, "point", f"edge {kind} layout {layout}",
                           launch=ia.LaunchConfig(layout=layout))
    _, vk, pk = ia.render_debug(tel, torch.as_tensor(src), torch.as_tensor(vals), "point",
                                return_pixels=True)
    vk, pk = vk.cpu().numpy().astype(np.float64), pk.cpu().numpy()
    sc = oracle_scene(tel)
    x64, y64, v64, p64 = O.render(sc, src, vals, "point", "fp64", "debug")
    import copy
    s0 = copy.copy(sc.sensor)
    s0.edge_width = 0.0
    if kind == "square":
        p0 = O.square_bin(x64, y64, s0)[0]
    else:
        s0.__post_init__()
        p0 = O.hex_bin(x64, y64, s0)[0]
    dropped = (v64 != 0) & (p64 < 0) & (p0 >= 0)
    print(f"[edge {kind}] rays dropped by the edge mask (fp64): {int(dropped.sum())} of "
          f"{int(((v64 != 0) & (p0 >= 0)).sum())} on pixels", flush=True)
    assert dropped.sum() > 0
    M = sc.points.shape[1]
    ns = len(src)
    sidx = (np.arange(len(vk)) // M) % ns
    # response matrix rows = the debug trace's per-source images
    rm = ia.render_response_matrix(tel, torch.as_tensor(src), torch.as_tensor(vals),
                                   "point").cpu().numpy()
    ref = np.zeros_like(rm, dtype=np.float64)
    m = pk >= 0
    np.add.at(ref, (sidx[m], pk[m]), vk[m])
    assert np.allclose(rm, ref, rtol=1e-5, atol=1e-7 * np.abs(ref).max())
    # adjoint wrt values: d sum(W img) / d v_s = sum over the source's rays
    # of W[pixel] * value / v_s, at the debug trace's pixel ids
    W = np.linspace(0.5, 1.5, tel.sensors[0].n_pixels)
    v = torch.tensor(vals, device=DEV, requires_grad=True)
    img = ia.render(tel, torch.as_tensor(src), v, "point")
    (img.reshape(-1).double() * torch.as_tensor(W, device=DEV)).sum().backward()
    want = np.zeros(ns)
    np.add.at(want, sidx[m], W[pk[m]] * vk[m] / vals[sidx[m]])
    got = v.grad.double().cpu().numpy()
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"[edge {kind}] d/dvalues adjoint vs debug-trace sums: max rel err {err:.2e}",
          flush=True)
    assert np.abs(want).max() > 0 and err < 1e-5


def _facet_subset(tel, idx):
    """The telescope with only the stage-0 facets idx (obstructions and
    sensors kept): the oracle's facet loop then stays short."""
    g = tel.mirror_groups[0]
    i = torch.as_tensor(idx, device=g.positions.device)
    kw = {}
    for name in ("positions", "rotations", "points", "normals", "weights", "perturbation_delta",
                 "perturbation_scale", "offsets", "radii", "vertices"):
        t = getattr(g, name, None)
        if isinstance(t, torch.Tensor):
            kw[name] = t[i].detach().clone()
    return ia.Telescope([g.replace(**kw)] + list(tel.mirror_groups[1:]), tel.obstruction_groups,
                        tel.sensors)


@pytest.fixture(scope="module")
def hess():
    from iactrace_amd.yaml_loader import build_telescope
    from tests.golden import hess as H
    out = {}
    for name, m in (("CT3", 8), ("CT5", 4)):
        out[name] = build_telescope(H.config(name), ia.MCIntegrator(m), key=5, device=DEV)
    # every 6th of CT5's 876 facets (all 271 obstructions)
    out["CT5"] = _facet_subset(out["CT5"], np.arange(0, 876, 6))
    return out


@pytest.mark.parametrize("name,sensor,layout", [("CT3", 0, 1), ("CT3", 0, 2), ("CT3", 1, 2),
                                                ("CT5", 1, 2), ("CT5", 0, 1)])
def test_reference_telescopes(hess, name, sensor, layout):
    tel = hess[name]
    if sensor:
        tel = tel.replace_sensor(tel.sensors[sensor], 0)
    s = tel.sensors[0]
    print(f"\n[{name}] facets {sum(len(g) for g in tel.mirror_groups)}, obstructions "
          f"{sum(len(g) for g in tel.obstruction_groups)}, sensor {type(s).__name__} "
          f"{s.n_pixels} px, edge width {s.edge_width}", flush=True)
    src, vals = configs.point_source_field(6 if name == "CT3" else 3, 1.5, seed=17)
    _trace_and_compare(tel, src, vals, "point", f"{name} sensor {sensor} layout {layout}",
                       launch=ia.LaunchConfig(layout=layout))
