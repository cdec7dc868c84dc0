"""Forward parity of the HIP kernels against the oracle (GPU).

Each case traces the same injected MC samples through the kernels and the
oracle (reference fp32 arithmetic with fp64 source directions and Newton):
  * per-ray pixel ids must be identical except for rays in the edge band
    (within 2.5e-5 m of a binning boundary; a shadow/aperture decision that
    flips when obstructions and apertures move by 2e-5 m in the fp64 oracle;
    or a cylinder/sphere discriminant within its fp32 rounding bound --
    tests.helpers.band_mask);
  * per-ray values within 1e-5 relative;
  * the image within 1e-5 relative per pixel of the oracle's values summed
    at the kernel's pixel ids (band rays with the kernel's own value, so each
    is counted once, consistently),
    and the edge-band count and the error against the fp64 oracle are printed.
"""
import numpy as np
import pytest
import torch

import iactrace_amd as ia
import oracle.iact_oracle as O
from iactrace_amd import configs
from tests.helpers import band_mask, oracle_scene

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _trace_and_compare(tel, src, vals, stype, label, rtol=1e-5, launch=None):
    print(f"\n[{label}] facets={sum(len(g) for g in tel.mirror_groups if g.optical_stage == 0)} "
          f"samples={tel.mirror_groups[0].n_samples} sources={len(src)} launch={launch}",
          flush=True)
    img = ia.render(tel, torch.as_tensor(src), torch.as_tensor(vals), stype,
                    launch=launch).cpu().numpy()
    xy, vk, pk = ia.render_debug(tel, torch.as_tensor(src), torch.as_tensor(vals), stype,
                                 return_pixels=True)
    xy, vk, pk = xy.cpu().numpy(), vk.cpu().numpy(), pk.cpu().numpy()
    sc = oracle_scene(tel)
    xo, yo, vo, po = O.render(sc, src, vals, stype, "ref32", "debug")
    band, x64, y64, v64 = band_mask(sc, src, vals, stype)
    n = len(vo)
    assert vk.shape == (n,) and pk.shape == (n,)
    bad_pix = (pk != po) & ~band
    bad_val = (np.abs(vk - vo) > rtol * np.abs(vo) + 1e-30) & ~band
    print(f"[{label}] rays={n} band={int(band.sum())} pix_mismatch_in_band="
          f"{int(((pk != po) & band).sum())} pix_mismatch_outside={int(bad_pix.sum())} "
          f"val_mismatch_outside={int(bad_val.sum())}", flush=True)
    # edge-band evidence: how far the kernel's sensor coordinate is from the
    # fp64 one (pixel pitch / hex inradius units), and the fp64 distance of
    # every ray binned differently to its nearest binning boundary
    unit = (sc.sensor.params()[2] if isinstance(sc.sensor, O.SquareSensorSpec)
            else sc.sensor.inradius)
    ok = (np.abs(xy[:, 0]) < 1e9) & (np.abs(x64) < 1e9) & (vo != 0) & ((po >= 0) | (pk >= 0))
    cerr = (np.hypot(xy[ok, 0] - x64[ok], xy[ok, 1] - y64[ok]) / unit).max() if ok.any() else 0.0
    marg = O.edge_margin(x64, y64, sc.sensor)
    mm = (pk != po) & (np.abs(x64) < 1e9)
    print(f"[{label}] kernel vs fp64 sensor coordinate: max error {cerr:.2e} px; rays binned "
          f"differently from ref32: {int((pk != po).sum())}, their max fp64 edge distance "
          f"{marg[mm].max() if mm.any() else 0:.2e} px", flush=True)
    assert bad_pix.sum() == 0, np.nonzero(bad_pix)[0][:10]
    assert bad_val.sum() == 0, np.nonzero(bad_val)[0][:10]
    # image: oracle ray values accumulated at the kernel's pixel ids (the
    # kernel's own value for a band ray, whose shadow decision either side
    # may take)
    P = sc.sensor.n_pixels
    ref = np.zeros(P)
    m = pk >= 0
    vsel = np.where(band, vk, vo).astype(np.float64)
    np.add.at(ref, pk[m], vsel[m])
    got = img.reshape(-1).astype(np.float64)
    err = np.abs(got - ref)
    rel = err / np.maximum(np.abs(ref), 1e-30)
    sig = np.abs(ref) > 1e-6 * np.abs(ref).max()
    print(f"[{label}] image sum kernel={got.sum():.6g} oracle={ref.sum():.6g} "
          f"max_rel_err(sig px)={rel[sig].max() if sig.any() else 0:.3g}", flush=True)
    assert np.all(err[sig] <= rtol * np.abs(ref[sig]))
    assert np.all(err[~sig] <= 1e-6 * np.abs(ref).max() + 1e-30)
    # stated: error against the all-fp64 oracle image
    img64 = O.render(sc, src, vals, stype, "fp64", "image")
    print(f"[{label}] vs fp64 oracle: max|diff|/max = "
          f"{np.abs(got - img64).max() / max(np.abs(img64).max(), 1e-30):.3g}", flush=True)
    return img, band


def _tel(name, ns=None, m=None, key=0):
    wl = configs.workload(name, n_sources=ns, n_samples=m)
    return wl.telescope(DEV, key=key), wl


def test_c1_point_source_square():
    tel, wl = _tel("C1", m=64)
    img, _ = _trace_and_compare(tel, wl.sources, wl.values, "point", "C1")
    assert img.shape == (30, 32)
    # known answer: total flux ~ sum(facet area) * <cos> for an on-axis source
    assert img.sum() > 0


def test_c2_point_field_square():
    tel, wl = _tel("C2", ns=12, m=32)
    _trace_and_compare(tel, wl.sources, wl.values, "point", "C2")


def test_c3_parabolic_hex_obstructions():
    tel, wl = _tel("C3", ns=6, m=8)
    img, _ = _trace_and_compare(tel, wl.sources, wl.values, "point", "C3")
    assert img.shape == (2048,)


@pytest.mark.parametrize("layout", [1, 2])
def test_layouts_c2_many_sources(layout):
    """Both image layouts (lanes = samples of a facet / lanes = sources) on
    more sources than a wave holds, against the oracle."""
    tel, wl = _tel("C2", ns=80, m=16, key=4)
    _trace_and_compare(tel, wl.sources, wl.values, "point", f"C2 layout {layout}",
                       launch=ia.LaunchConfig(layout=layout))


@pytest.mark.parametrize("layout", [1, 2])
def test_layouts_c3_hex_obstructions(layout):
    tel, wl = _tel("C3", ns=70, m=8, key=5)
    _trace_and_compare(tel, wl.sources, wl.values, "point", f"C3 layout {layout}",
                       launch=ia.LaunchConfig(layout=layout, sample_group=5))


@pytest.mark.parametrize("layout", [1, 2])
def test_non_square_pixels(layout):
    """Pixels 0.8 times as tall as wide: the source-major trace takes its
    generic loop (the aspect ratio enters the hit's row coordinate)."""
    tel, wl = _tel("C2", ns=80, m=16, key=9)
    s = tel.sensors[0]
    b = [float(v) for v in s.bounds]
    tel = tel.replace_sensor(ia.SquareSensor(s.position, s.rotation, s.width, s.height,
                                             (b[0], b[1], 0.8 * b[2], 0.8 * b[3])), 0)
    img, _ = _trace_and_compare(tel, wl.sources, wl.values, "point", f"aspect 0.8 {layout}",
                                launch=ia.LaunchConfig(layout=layout))
    assert img.sum() > 0


@pytest.mark.parametrize("layout", [1, 2])
def test_facet_in_the_sensor_plane(layout):
    """A flat, untilted facet lying in the sensor plane (its samples are at
    distance 0 from it): the source-major trace then keeps the parallel-ray
    test for every sample, which it otherwise leaves to the binning; the
    image still matches the oracle."""
    from iactrace_amd.configs import AsphericSurface, Mirror, PolygonAperture, hex_facet_vertices
    wl = configs.workload("C2", n_sources=80, n_samples=16)
    mirrors = configs.davies_cotton_dish(n_facets=19) + [
        Mirror(np.array([5.0, 0.0, 15.0]), np.zeros(3), AsphericSurface(0.0, 0.0, []),
               PolygonAperture(hex_facet_vertices(0.6)))]
    w2 = configs.Workload("plane", mirrors, [], wl.sensor, 16, wl.sources, wl.values, "point")
    tel = w2.telescope(DEV, key=3)
    img, _ = _trace_and_compare(tel, w2.sources, w2.values, "point", f"in-plane facet {layout}",
                                launch=ia.LaunchConfig(layout=layout))
    assert img.sum() > 0


@pytest.mark.parametrize("kind", ["square", "hex"])
def test_layouts_agree_on_splat_and_parallel(kind):
    """The two layouts trace identical rays; only the order of the float
    additions differs."""
    if kind == "square":
        tel, wl = _tel("C2", ns=8, m=64, key=6)
        s = tel.sensors[0]
        tel = tel.replace_sensor(ia.DifferentiableSquareSensor(s.position, s.rotation, s.width,
                                                               s.height, s.bounds, 0.5, 2), 0)
        src = configs.parallel_grid(10, 13, 0.2)
        vals = np.linspace(0.5, 2.0, len(src)).astype(np.float32)
        stype = "parallel"
    else:
        tel, wl = _tel("C3", ns=8, m=32, key=7)
        s = tel.sensors[0]
        tel = tel.replace_sensor(ia.DifferentiableHexagonalSensor(s.position, s.rotation,
                                                                  s.hex_centers, 0.6, 1), 0)
        src, vals = configs.point_source_field(150, 2.0, seed=11)
        stype = "point"
    a = ia.render(tel, torch.as_tensor(src), torch.as_tensor(vals), stype,
                  launch=ia.LaunchConfig(layout=1)).reshape(-1)
    b = ia.render(tel, torch.as_tensor(src), torch.as_tensor(vals), stype,
                  launch=ia.LaunchConfig(layout=2, sample_group=33)).reshape(-1)
    err = ((a - b).abs().max() / a.abs().max()).item()
    print(f"[layouts {kind} splat] sum {a.sum().item():.6g} vs {b.sum().item():.6g}, "
          f"max rel err {err:.2e}", flush=True)
    assert a.abs().max().item() > 0
    assert err < 2e-6


def test_source_major_layout_rejects_stages_and_deterministic():
    tel, wl = _tel("C4", ns=70, m=16)
    with pytest.raises(ia._lib.IactError):
        ia.render(tel, torch.as_tensor(wl.sources), torch.as_tensor(wl.values), "parallel",
                  launch=ia.LaunchConfig(layout=2))
    tel, wl = _tel("C2", ns=70, m=16)
    with pytest.raises(ia._lib.IactError):
        ia.render(tel, torch.as_tensor(wl.sources), torch.as_tensor(wl.values), "point",
                  launch=ia.LaunchConfig(layout=2, deterministic=True))


def test_c4_cassegrain_response_matrix():
    tel, wl = _tel("C4", ns=24, m=256)
    src, vals = wl.sources, wl.values
    _trace_and_compare(tel, src, vals, "parallel", "C4")
    rm = ia.render_response_matrix(tel, torch.as_tensor(src), torch.as_tensor(vals),
                                   "parallel").cpu().numpy()
    assert rm.shape == (len(src), 4096)
    _, vk, pk = ia.render_debug(tel, torch.as_tensor(src), torch.as_tensor(vals), "parallel",
                                return_pixels=True)
    vk, pk = vk.cpu().numpy(), pk.cpu().numpy()
    ns, M = len(src), tel.mirror_groups[0].n_samples
    sidx = (np.arange(len(vk)) // M) % ns
    ref = np.zeros_like(rm, dtype=np.float64)
    m = pk >= 0
    np.add.at(ref, (sidx[m], pk[m]), vk[m].astype(np.float64))
    assert np.allclose(rm, ref, rtol=1e-5, atol=1e-7 * np.abs(ref).max())
    img = ia.render(tel, torch.as_tensor(src), torch.as_tensor(vals), "parallel").cpu().numpy()
    assert np.allclose(rm.sum(0), img.reshape(-1), rtol=1e-4, atol=1e-6 * img.max())


def test_response_matrix_hex_and_point():
    tel, wl = _tel("C3", ns=5, m=8)
    rm = ia.render_response_matrix(tel, torch.as_tensor(wl.sources), torch.as_tensor(wl.values),
                                   "point").cpu().numpy()
    sc = oracle_scene(tel)
    ref = O.render(sc, wl.sources, wl.values, "point", "ref32", "response")
    tot = np.abs(ref).sum()
    assert np.abs(rm - ref).sum() <= 1e-3 * tot


@pytest.mark.parametrize("kind", ["square", "hex"])
def test_splat_sensors(kind):
    if kind == "square":
        tel, wl = _tel("C2", ns=6, m=32)
        s = tel.sensors[0]
        sp = ia.DifferentiableSquareSensor(s.position, s.rotation, s.width, s.height, s.bounds,
                                           sigma=0.5, kernel_size=2)
    else:
        tel, wl = _tel("C3", ns=4, m=8)
        s = tel.sensors[0]
        sp = ia.DifferentiableHexagonalSensor(s.position, s.rotation, s.hex_centers, sigma=0.5,
                                              kernel_size=1)
    tel = tel.replace_sensor(sp, 0)
    img = ia.render(tel, torch.as_tensor(wl.sources), torch.as_tensor(wl.values),
                    "point").cpu().numpy().reshape(-1)
    sc = oracle_scene(tel)
    ref = O.render(sc, wl.sources, wl.values, "point", "ref32", "image")
    tot = np.abs(ref).sum()
    print(f"\n[splat-{kind}] sum kernel={img.sum():.6g} oracle={ref.sum():.6g} "
          f"L1 rel={np.abs(img - ref).sum() / tot:.3g}", flush=True)
    assert abs(img.sum() - ref.sum()) <= 1e-4 * tot
    assert np.abs(img - ref).sum() <= 2e-3 * tot


@pytest.mark.parametrize("aperture", ["disk", "polygon"])
def test_sampler_matches_oracle(aperture):
    rng = np.random.default_rng(7)
    if aperture == "disk":
        mirrors, _ = configs.cassegrain()
        m0 = mirrors[0]
        ms = [ia.Mirror([0, 0, 0], [0, 0, 0], m0.surface, ia.DiskAperture(r), offset=off)
              for r, off in [(2.0, (0.0, 0.0)), (0.7, (1.5, -0.4))]]
    else:
        surf = ia.AsphericSurface(0.0137, -0.3, [1e-4, -2e-6])
        v = configs.hex_facet_vertices(0.9)
        ms = [ia.Mirror([0, 0, 0], [0, 0, 0], surf, ia.PolygonAperture(v), offset=off)
              for off in [(0.0, 0.0), (3.0, 1.0), (-5.0, 2.0)]]
    g = ia.group_mirrors(ms)[0].to(DEV)
    seed = int(rng.integers(1 << 62))
    gs = ia.MCIntegrator(500).sample_group(g, seed)
    surf = O.Surface(g.curvature, g.conic, g.aspheric.double().numpy())
    offs = g.offsets.cpu().numpy()
    if aperture == "disk":
        ref = O.sample_group(seed, 0, surf, offs, 500, radii=g.radii.cpu().numpy())
    else:
        ref = O.sample_group(seed, 0, surf, offs, 500, vertices=g.vertices.cpu().numpy())
    got = [gs.points, gs.normals, gs.perturbation_delta, gs.weights[..., 0]]
    for name, a, b in zip(["points", "normals", "delta", "weights"], got, ref):
        a = a.cpu().numpy()
        assert np.allclose(a, b, rtol=2e-5, atol=2e-6 * max(1.0, np.abs(b).max())), name


def test_sensor_accumulate_pixel_ids():
    rng = np.random.default_rng(11)
    for s in [configs.workload("C1").sensor, configs.workload("C3").sensor,
              ia.SquareSensor([0, 0, 0], [0, 0, 0], 17, 9, (-1, 1, -0.5, 0.5), edge_width=0.01),
              ia.HexagonalSensor([0, 0, 0], [0, 0, 0], configs.hex_camera(300, 0.05),
                                 edge_width=0.003)]:
        n = 20000
        ext = 0.8 if isinstance(s, ia.SquareSensor) else 1.1
        x = rng.uniform(-ext, ext, n).astype(np.float32)
        y = rng.uniform(-ext, ext, n).astype(np.float32)
        v = rng.random(n).astype(np.float32)
        pid = s.pixel_ids(torch.as_tensor(x, device=DEV), torch.as_tensor(y, device=DEV))
        pid = pid.cpu().numpy()
        from tests.helpers import oracle_sensor
        spec = oracle_sensor(s)
        if isinstance(s, ia.SquareSensor):
            ref = O.square_bin(x, y, spec)[0]
        else:
            ref = O.hex_bin(x, y, spec)[0]
        band = O.edge_margin(x, y, spec) < 1e-4
        assert np.all((pid == ref) | band), np.nonzero((pid != ref) & ~band)[0][:5]
        img = s.accumulate(torch.as_tensor(x, device=DEV), torch.as_tensor(y, device=DEV),
                           torch.as_tensor(v, device=DEV)).cpu().numpy().reshape(-1)
        r = np.zeros(s.n_pixels)
        np.add.at(r, pid[pid >= 0], v[pid >= 0])
        assert np.allclose(img, r, rtol=1e-5, atol=1e-5)


def test_transform_to_world():
    tel, _ = _tel("C3", ns=1, m=16)
    g = tel.mirror_groups[0]
    p, n, w = g.transform_to_world()
    sc = oracle_scene(tel)
    rp, rn, rw = O.transform_to_world(sc, "ref32")
    assert np.allclose(p.cpu().numpy(), rp, atol=2e-5)
    assert np.allclose(n.cpu().numpy(), rn, atol=2e-6)
    assert np.array_equal(w.cpu().numpy()[..., 0], rw)


def test_empty_cases():
    tel, wl = _tel("C1", m=16)
    z = ia.render(tel, torch.zeros((0, 3), dtype=torch.float64), torch.zeros(0), "point")
    assert z.shape == (30, 32) and float(z.abs().sum()) == 0.0
    stage1_only = ia.Telescope([], [], tel.sensors)
    z = ia.render(stage1_only, torch.as_tensor(wl.sources), torch.as_tensor(wl.values), "point")
    assert float(z.abs().sum()) == 0.0
    rm = ia.render_response_matrix(tel, torch.zeros((0, 3), dtype=torch.float64), torch.zeros(0))
    assert rm.shape == (0, 960)


def test_grazing_upward_and_onaxis_parallel_sources():
    # (1,0,0): grazing light, reflected rays mostly miss the camera plane region;
    # (0,0,1): light travelling up hits the mirror backs, reflected rays leave
    # the sensor plane behind them (t <= 0, 1e10 sentinel); (0,0,-1): on-axis star.
    tel, _ = _tel("C1", m=32)
    src = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])
    _trace_and_compare(tel, src, np.ones(3, np.float32), "parallel", "C1-parallel")
    xy, v, pix = ia.render_debug(tel, torch.as_tensor(src), torch.ones(3), "parallel",
                                 return_pixels=True)
    M, nf = 32, 382
    up = (np.arange(nf * 3 * M) // M) % 3 == 1
    assert np.all(xy.cpu().numpy()[up] == 1e10) and np.all(pix.cpu().numpy()[up] == -1)


def test_dlpack_inputs():
    """Sources and values handed over as DLPack exporters (the way CuPy or
    JAX arrays arrive) give the image of the same tensors."""
    from tests.test_dlpack import Exporter
    tel, wl = _tel("C2", ns=70, m=16, key=2)
    s = torch.as_tensor(wl.sources, device=DEV)
    v = torch.as_tensor(wl.values, device=DEV)
    a = ia.render(tel, s, v, "point", launch=ia.LaunchConfig(deterministic=True))
    b = ia.render(tel, Exporter(s), Exporter(v), "point",
                  launch=ia.LaunchConfig(deterministic=True))
    assert a.abs().max().item() > 0 and torch.equal(a, b)


def test_deterministic_mode_bitwise():
    tel, wl = _tel("C2", ns=64, m=64)
    L = ia.LaunchConfig(deterministic=True)
    a = ia.render(tel, torch.as_tensor(wl.sources), torch.as_tensor(wl.values), "point", launch=L)
    b = ia.render(tel, torch.as_tensor(wl.sources), torch.as_tensor(wl.values), "point", launch=L)
    assert torch.equal(a, b)
    c = ia.render(tel, torch.as_tensor(wl.sources), torch.as_tensor(wl.values), "point")
    assert torch.allclose(a, c, rtol=1e-5, atol=1e-6 * float(a.abs().max()))


@pytest.mark.parametrize("threads,tile,bpc,acc", [(64, 1, 1, 0), (128, 7, 2, 1), (256, 64, 8, 2),
                                                  (256, 16, 4, 1), (192, 3, 3, 2)])
def test_launch_shapes_agree(threads, tile, bpc, acc):
    tel, wl = _tel("C3", ns=9, m=16)
    ref = ia.render(tel, torch.as_tensor(wl.sources), torch.as_tensor(wl.values), "point")
    L = ia.LaunchConfig(block_threads=threads, source_tile=tile, blocks_per_cu=bpc,
                        accumulate=acc)
    got = ia.render(tel, torch.as_tensor(wl.sources), torch.as_tensor(wl.values), "point",
                    launch=L)
    assert torch.allclose(got, ref, rtol=1e-5, atol=1e-6 * float(ref.abs().max()))
