"""Gradients of the HIP adjoint kernel (GPU).

* against torch float64 autograd of the oracle twin (oracle/torch_twin.py),
  which stands in for jax.grad of the reference: rotations, positions,
  sources, values and primary aspheric coefficients, for hard and
  Gaussian-splat sensors, point and parallel sources, with obstructions;
* against central finite differences of the GPU forward render
  (h = 1e-4 rad) on splat sensors and on the total flux of a hard sensor.
Losses weight the image with a smooth pixel map so that rays in the binning
edge band (assigned to a neighbouring pixel in fp32 vs fp64) barely matter.
"""
import math

import numpy as np
import pytest
import torch

import iactrace_amd as ia
import oracle.iact_oracle as O
import oracle.torch_twin as TW
from iactrace_amd import configs
from tests.helpers import band_mask, hex_kink_margin, oracle_scene

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _smooth_weights(sensor):
    if isinstance(sensor, ia.SquareSensor):
        yy, xx = np.mgrid[0:sensor.height, 0:sensor.width]
        return (1.0 + 0.3 * np.sin(xx / sensor.width * 3.0) + 0.2 * yy / sensor.height).reshape(-1)
    c = np.asarray(sensor.hex_centers, np.float64)
    return 1.0 + 0.3 * np.sin(c[:, 0] * 3.0) + 0.2 * c[:, 1]


def _kernel_grads(tel, src, vals, stype, W, asph=False):
    g = tel.mirror_groups[0]
    rot = g.rotations.detach().clone().requires_grad_(True)
    pos = g.positions.detach().clone().requires_grad_(True)
    groups = list(tel.mirror_groups)
    a0 = g.aspheric.detach().clone().requires_grad_(asph)
    groups[0] = g.replace(rotations=rot, positions=pos, aspheric=a0)
    t2 = ia.Telescope(groups, tel.obstruction_groups, tel.sensors)
    s = torch.tensor(src, dtype=torch.float64, device=DEV, requires_grad=True)
    v = torch.tensor(vals, dtype=torch.float32, device=DEV, requires_grad=True)
    img = ia.render(t2, s, v, stype)
    loss = (img.reshape(-1).double() * torch.as_tensor(W, device=DEV)).sum()
    loss.backward()
    out = {"rot": rot.grad.double().cpu().numpy(), "pos": pos.grad.double().cpu().numpy(),
           "src": s.grad.double().cpu().numpy(), "val": v.grad.double().cpu().numpy()}
    if asph:
        out["asph"] = a0.grad.double().cpu().numpy()
    return out, float(loss)


def _twin_grads(tel, src, vals, stype, W, asph=False):
    sc = oracle_scene(tel)
    g = tel.mirror_groups[0]
    rot = torch.tensor(sc.rotations, dtype=torch.float64, requires_grad=True)
    pos = torch.tensor(sc.positions, dtype=torch.float64, requires_grad=True)
    s = torch.tensor(src, dtype=torch.float64, requires_grad=True)
    v = torch.tensor(vals, dtype=torch.float64, requires_grad=True)
    kw = {}
    if asph:
        kw = dict(aspheric0=torch.tensor(g.aspheric.double().numpy(), requires_grad=True),
                  surface0=(g.curvature, g.conic), offsets0=g.offsets.cpu().numpy())
    img = TW.render_image(sc, s, v, stype, rot, pos, **kw)
    loss = (img * torch.as_tensor(W)).sum()
    loss.backward()
    z = lambda t: (t.grad if t.grad is not None else torch.zeros_like(t)).numpy()  # noqa: E731
    out = {"rot": z(rot), "pos": z(pos), "src": z(s), "val": z(v)}
    if asph:
        out["asph"] = z(kw["aspheric0"])
    return out, float(loss)


def _compare(label, gk, gt, keys, tol=1e-3, facet_mask=None, zero_keys=()):
    for k in keys:
        a, b = gk[k], gt[k]
        if facet_mask is not None and k in ("rot", "pos"):
            a, b = a[facet_mask], b[facet_mask]
        scale = np.abs(b).max()
        err = np.abs(a - b).max() / max(scale, 1e-300)
        print(f"[{label}] d/d{k}: max|kernel - twin| / max|twin| = {err:.2e} (scale {scale:.3g})",
              flush=True)
        if k in zero_keys:
            assert scale == 0 and np.abs(a).max() == 0, k
            continue
        assert scale > 0, k
        assert err < tol, (k, err)


def test_grad_hard_square_point_sources():
    wl = configs.workload("C2", n_sources=6, n_samples=16)
    tel = wl.telescope(DEV, key=3)
    W = _smooth_weights(tel.sensors[0])
    gk, lk = _kernel_grads(tel, wl.sources, wl.values, "point", W)
    gt, lt = _twin_grads(tel, wl.sources, wl.values, "point", W)
    assert abs(lk - lt) < 1e-4 * abs(lt)
    _compare("hard-square", gk, gt, ["rot", "pos", "src", "val"])


def test_grad_splat_hex_obstructions():
    wl = configs.workload("C3", n_sources=3, n_samples=8)
    tel = wl.telescope(DEV, key=5)
    s = tel.sensors[0]
    tel = tel.replace_sensor(ia.DifferentiableHexagonalSensor(s.position, s.rotation,
                                                              s.hex_centers, 0.5, 1), 0)
    src, vals = configs.point_source_field(3, 1.2, seed=9)
    W = _smooth_weights(tel.sensors[0])
    gk, _ = _kernel_grads(tel, src, vals, "point", W)
    gt, _ = _twin_grads(tel, src, vals, "point", W)
    # a facet with a ray whose nearest hexagon (which fixes the splat's tap
    # set) or shadow decision the kernels take differently from the float64
    # oracle (the twin's decisions), or whose hex-norm branch (a kink of the
    # tap weight) fp32 may take differently, carries an O(1/rays) difference;
    # those facets are left out of the pose comparison
    sc = oracle_scene(tel)
    x, y, v64, _ = O.render(sc, src, vals, "point", "fp64", "debug")
    xyk, vk = ia.render_debug(tel, torch.as_tensor(src), torch.as_tensor(vals), "point")
    xyk, vk = xyk.cpu().numpy(), vk.cpu().numpy()
    base = O.hex_grid_coords(np.asarray(x, np.float64), np.asarray(y, np.float64), sc.sensor)
    kern = O.hex_grid_coords(xyk[:, 0].astype(np.float64), xyk[:, 1].astype(np.float64),
                             sc.sensor)
    to_ax = lambda g: O.axial_round(*O._to_axial(g[0], g[1], sc.sensor.hex_size,  # noqa: E731
                                                  np.float64))
    (q64, r64), (qk, rk) = to_ax(base), to_ax(kern)
    differ = (q64 != qk) | (r64 != rk) | ((vk != 0) != (v64 != 0))
    differ |= hex_kink_margin(x, y, sc.sensor) < 2e-4
    nf = sc.points.shape[0]
    keep = ~differ.reshape(nf, len(src), -1).any((1, 2))
    print(f"[splat-hex] facets compared {int(keep.sum())} of {nf}", flush=True)
    _compare("splat-hex", gk, gt, ["rot", "pos", "src", "val"], facet_mask=keep)


def test_grad_splat_square_parallel():
    wl = configs.workload("C1", n_samples=16)
    tel = wl.telescope(DEV, key=2)
    s = tel.sensors[0]
    tel = tel.replace_sensor(ia.DifferentiableSquareSensor(s.position, s.rotation, s.width,
                                                           s.height, s.bounds, 0.4, 2), 0)
    src = configs.parallel_grid(3, 2, 0.5)
    vals = np.array([1.0, 0.5, 2.0, 1.5, 0.7, 1.1], np.float32)
    W = _smooth_weights(tel.sensors[0])
    gk, _ = _kernel_grads(tel, src, vals, "parallel", W)
    gt, _ = _twin_grads(tel, src, vals, "parallel", W)
    _compare("splat-square-parallel", gk, gt, ["rot", "pos", "src", "val"])


def _aspheric_primary_tel(m=512, splat=False):
    mirrors, _ = configs.cassegrain()
    prim = mirrors[0]
    prim.surface.aspheric = torch.tensor([3e-4, -2e-6, 1e-8, -5e-11])
    if splat:
        sensor = ia.DifferentiableSquareSensor([0, 0, 9.5], [0, 0, 0], 32, 32,
                                               (-0.4, 0.4, -0.4, 0.4), sigma=0.6, kernel_size=2)
    else:
        sensor = ia.SquareSensor([0, 0, 9.5], [0, 0, 0], 32, 32, (-0.4, 0.4, -0.4, 0.4))
    g = ia.group_mirrors([prim])[0].to(DEV)
    g = ia.MCIntegrator(m).sample_group(g, 11)
    return ia.Telescope([g], [], [sensor])


@pytest.mark.parametrize("splat", [False, True])
def test_grad_primary_aspheric(splat):
    tel = _aspheric_primary_tel(splat=splat)
    src = configs.parallel_grid(2, 2, 0.3)
    vals = np.ones(4, np.float32)
    W = _smooth_weights(tel.sensors[0])
    gk, _ = _kernel_grads(tel, src, vals, "parallel", W, asph=True)
    gt, _ = _twin_grads(tel, src, vals, "parallel", W, asph=True)
    # parallel rays on a hard sensor: moving a facet only moves its spot
    # across pixel boundaries, so the position gradient is exactly zero
    _compare(f"aspheric-{'splat' if splat else 'hard'}", gk, gt, ["asph", "rot", "pos"],
             tol=1e-3, zero_keys=() if splat else ("pos",))


def _facet(tel, f):
    """Telescope holding only facet f of the primary group (its rays are the
    only ones that depend on its pose)."""
    g = tel.mirror_groups[0]
    kw = {}
    for name in ("positions", "rotations", "points", "normals", "weights", "perturbation_delta",
                 "perturbation_scale", "offsets", "radii", "vertices"):
        t = getattr(g, name, None)
        if isinstance(t, torch.Tensor):
            kw[name] = t[f:f + 1].detach().clone()
    return ia.Telescope([g.replace(**kw)], tel.obstruction_groups, tel.sensors)


DET = ia.LaunchConfig(deterministic=True)


def _loss_fd(tel1, src, vals, stype, W, rot):
    t2 = tel1.set_mirror_rotations(0, rot)
    img = ia.render(t2, torch.as_tensor(src), torch.as_tensor(vals), stype, launch=DET)
    return float((img.reshape(-1).double() * torch.as_tensor(W, device=DEV)).sum())


def _fd_check(label, tel, src, vals, stype, W, facets, h_rad, rtol):
    """Richardson-extrapolated central differences of single-facet
    deterministic renders (error O(h^4)) against the adjoint of the full
    render."""
    gk, _ = _kernel_grads(tel, src, vals, stype, W)
    scale = np.abs(gk["rot"]).max()
    for f in facets:
        t1 = _facet(tel, int(f))
        rot0 = t1.mirror_groups[0].rotations.detach().clone()
        for k in (0, 1):
            def cd(h):
                rp, rm = rot0.clone(), rot0.clone()
                rp[0, k] += h
                rm[0, k] -= h
                return (_loss_fd(t1, src, vals, stype, W, rp) -
                        _loss_fd(t1, src, vals, stype, W, rm)) / (2 * h)
            h = math.degrees(h_rad)
            fd = (4 * cd(h / 2) - cd(h)) / 3
            ad = gk["rot"][f, k]
            print(f"[{label}] facet {f} angle {k}: adjoint {ad:.7g} fd {fd:.7g}", flush=True)
            assert abs(fd - ad) <= rtol * max(abs(fd), 0.05 * scale)


def test_grad_vs_central_differences_splat():
    """Square Gaussian splat (sigma 0.7 px) on the Davies-Cotton dish; no
    obstructions, whose shadow edges would make the difference quotient jump.
    The reference's splat renormalises the weights over a window of
    (2k+1)^2 taps around floor(x) (square.py:144-173): when a ray crosses a
    pixel edge the window moves and the image jumps by the weight of the tap
    that enters, ~1e-4 of a ray for k = 3, which biases a difference quotient
    by ~1e-3 at any step size (tests/test_twin.py pins this on the float64
    twin); with k = 4 (9 x 9 taps) the jump is ~1e-7 and the adjoint matches
    central differences to 1e-3."""
    wl = configs.workload("C2", n_sources=2, n_samples=256)
    tel = wl.telescope(DEV, key=7)
    sensor = ia.DifferentiableSquareSensor([0, 0, 15.0], [0, 0, 0], 64, 64,
                                           (-0.8, 0.8, -0.8, 0.8), sigma=0.7, kernel_size=4)
    tel = tel.replace_sensor(sensor, 0)
    src, vals = configs.point_source_field(2, 0.5, seed=1)
    yy, xx = np.mgrid[0:64, 0:64]
    W = (1.0 + 0.5 * np.sin(xx / 64 * 5.0) + 0.4 * np.cos(yy / 64 * 4.0)).reshape(-1)
    _fd_check("fd-splat", tel, src, vals, "point", W, [0, 41, 190, 300, 377], 1e-4, 1e-3)


def test_grad_vs_central_differences_total_flux_hard():
    """Total flux on a hard sensor that catches every ray: only the facet
    weight depends on the tilt, a relative change of ~1e-3 per mrad, so the
    step is large (1e-2 rad) and Richardson removes its O(h^2) error."""
    wl = configs.workload("C2", n_sources=4, n_samples=1024)
    tel = wl.telescope(DEV, key=8)
    big = ia.SquareSensor([0, 0, 15.0], [0, 0, 0], 16, 16, (-2.0, 2.0, -2.0, 2.0))
    tel = tel.replace_sensor(big, 0)
    _fd_check("fd-flux", tel, wl.sources, wl.values, "point", np.ones(big.n_pixels),
              [0, 17, 200, 381], 1e-2, 1e-3)


def _cassegrain_tel(splat, m=512):
    """C4's two-mirror telescope with aspheric terms on both mirrors."""
    mirrors, sensor = configs.cassegrain()
    mirrors[0].surface.aspheric = torch.tensor([3e-4, -2e-6, 1e-8, -5e-11])
    mirrors[1].surface.aspheric = torch.tensor([2e-3, -4e-5])
    if splat:
        sensor = ia.DifferentiableSquareSensor(sensor.position, sensor.rotation, sensor.width,
                                               sensor.height, sensor.bounds, sigma=0.7,
                                               kernel_size=2)
    groups = [g.to(DEV) for g in ia.group_mirrors(mirrors)]
    groups[0] = ia.MCIntegrator(m).sample_group(groups[0], 13)
    return ia.Telescope(groups, [], [sensor])


@pytest.mark.parametrize("splat", [False, True])
def test_grad_through_stage_two(splat):
    """Gradients through the secondary mirror of a Cassegrain (hit by the
    implicit function theorem on the Newton root, normal via the sag's
    Hessian): primary pose, both mirrors' aspheric coefficients, source
    directions and values, against the float64 twin (whose Newton iterations
    are differentiated step by step)."""
    tel = _cassegrain_tel(splat)
    src = configs.parallel_grid(3, 2, 0.05)
    vals = np.array([1.0, 0.7, 1.3, 0.9, 1.1, 0.8], np.float32)
    W = _smooth_weights(tel.sensors[0])
    # kernel
    g0, g1 = tel.mirror_groups
    rot = g0.rotations.detach().clone().requires_grad_(True)
    pos = g0.positions.detach().clone().requires_grad_(True)
    a0 = g0.aspheric.detach().clone().requires_grad_(True)
    a1 = g1.aspheric.detach().clone().requires_grad_(True)
    t2 = ia.Telescope([g0.replace(rotations=rot, positions=pos, aspheric=a0),
                       g1.replace(aspheric=a1)], [], tel.sensors)
    s = torch.tensor(src, dtype=torch.float64, device=DEV, requires_grad=True)
    v = torch.tensor(vals, dtype=torch.float32, device=DEV, requires_grad=True)
    img = ia.render(t2, s, v, "parallel")
    (img.reshape(-1).double() * torch.as_tensor(W, device=DEV)).sum().backward()
    gk = {"rot": rot.grad.double().cpu().numpy(), "pos": pos.grad.double().cpu().numpy(),
          "asph": a0.grad.double().cpu().numpy(), "asph2": a1.grad.double().cpu().numpy(),
          "src": s.grad.double().cpu().numpy(), "val": v.grad.double().cpu().numpy()}
    # twin
    sc = oracle_scene(tel)
    trot = torch.tensor(sc.rotations, dtype=torch.float64, requires_grad=True)
    tpos = torch.tensor(sc.positions, dtype=torch.float64, requires_grad=True)
    ta0 = torch.tensor(g0.aspheric.double().numpy(), requires_grad=True)
    ta1 = torch.tensor(g1.aspheric.double().numpy(), requires_grad=True)
    ts = torch.tensor(src, dtype=torch.float64, requires_grad=True)
    tv = torch.tensor(vals, dtype=torch.float64, requires_grad=True)
    timg = TW.render_image(sc, ts, tv, "parallel", trot, tpos, aspheric0=ta0,
                           aspheric_stage={0: ta1}, surface0=(g0.curvature, g0.conic),
                           offsets0=g0.offsets.cpu().numpy())
    (timg * torch.as_tensor(W)).sum().backward()
    z = lambda t: (t.grad if t.grad is not None else torch.zeros_like(t)).numpy()  # noqa: E731
    gt = {"rot": z(trot), "pos": z(tpos), "asph": z(ta0), "asph2": z(ta1), "src": z(ts),
          "val": z(tv)}
    keys = ["rot", "pos", "asph", "asph2", "src", "val"]
    _compare(f"stage2-{'splat' if splat else 'hard'}", gk, gt, keys, tol=1e-3)


def _grads_both_layouts(tel, src, vals, stype, W):
    """Rotation, position, source and value gradients of sum(W image) from
    the source-major adjoint (the default from 64 sources) and from the
    sample-major one (layout 1)."""
    out = []
    for layout in (0, 1):
        g = tel.mirror_groups[0]
        rot = g.rotations.detach().clone().requires_grad_(True)
        pos = g.positions.detach().clone().requires_grad_(True)
        t2 = ia.Telescope([g.replace(rotations=rot, positions=pos)], tel.obstruction_groups,
                          tel.sensors)
        s = torch.tensor(src, dtype=torch.float64, device=DEV, requires_grad=True)
        v = torch.tensor(vals, dtype=torch.float32, device=DEV, requires_grad=True)
        img = ia.render(t2, s, v, stype, launch=ia.LaunchConfig(layout=layout))
        (img.reshape(-1).double() * torch.as_tensor(W, device=DEV)).sum().backward()
        out.append({"rot": rot.grad.double().cpu().numpy(), "pos": pos.grad.double().cpu().numpy(),
                    "src": s.grad.double().cpu().numpy(), "val": v.grad.double().cpu().numpy()})
    return out


def _mixed_sources(n, seed):
    """n point sources, far (1e10 m) and near (2e4 m) alternating, so that
    every wave of the source-major kernels holds both kinds."""
    far, fv = configs.point_source_field(n // 2, 2.0, seed)
    near, nv = configs.point_source_field(n - n // 2, 2.0, seed + 1, z=2.0e4)
    src = np.empty((n, 3))
    vals = np.empty(n, np.float32)
    src[0::2], vals[0::2] = far, fv
    src[1::2], vals[1::2] = near, nv
    return src, vals


@pytest.mark.parametrize("case", ["hex-point-obstructed", "hex-point-mixed", "square-point",
                                  "square-parallel"])
def test_source_major_adjoint(case):
    """The source-major adjoint (lanes = sources over the sample records,
    per-facet torques) against the sample-major adjoint and, at a size the
    twin follows, against float64 autograd of the oracle twin.  The far
    sources (1e10 m) take the far-source loops, the mixed case (far and
    2e4 m sources in every wave) the per-lane directions."""
    if case.startswith("hex-point"):
        # C3's camera, box and masts over the inner 91 facets of its dish;
        # 64 samples, so that a ray decided differently by the kernels and
        # the twin moves its facet's gradient by well under 1e-3
        wl = configs.workload("C3", n_sources=150, n_samples=64)
        w2 = configs.Workload("c3-inner", configs.parabolic_dish(n_facets=91), wl.obstructions,
                              wl.sensor, 64, wl.sources, wl.values, "point",
                              tilt_sigma_deg=wl.tilt_sigma_deg, tilt_seed=wl.tilt_seed)
        tel = w2.telescope(DEV, key=5)
        src, vals, stype = wl.sources, wl.values, "point"
        if case == "hex-point-mixed":
            src, vals = _mixed_sources(150, 21)
    elif case == "square-point":
        wl = configs.workload("C2", n_sources=100, n_samples=16)
        tel = wl.telescope(DEV, key=6)
        src, vals, stype = wl.sources, wl.values, "point"
    else:
        wl = configs.workload("C1", n_samples=16)
        tel = wl.telescope(DEV, key=7)
        src = configs.parallel_grid(9, 8, 0.25)
        vals = np.linspace(0.5, 1.5, len(src)).astype(np.float32)
        stype = "parallel"
    W = _smooth_weights(tel.sensors[0])
    sm, smaj = _grads_both_layouts(tel, src, vals, stype, W)
    for k in ("rot", "pos", "src", "val"):
        a, b = sm[k], smaj[k]
        scale = np.abs(b).max()
        err = np.abs(a - b).max() / max(scale, 1e-300)
        print(f"[adjoint {case}] d/d{k}: source-major vs sample-major max rel err {err:.2e} "
              f"(scale {scale:.3g})", flush=True)
        if stype == "parallel" and k in ("pos", "src"):
            # parallel rays on a hard sensor: position gradients vanish
            # (the sample-major kernel's tiny residue aside)
            assert np.abs(a).max() <= 1e-6 * np.abs(sm["rot"]).max() + 1e-30 or err < 1e-4
            continue
        assert scale > 0 and err < 1e-4, (k, err)
    gt, _ = _twin_grads(tel, src, vals, stype, W)
    keys = ["rot", "val"] if stype == "parallel" else ["rot", "pos", "src", "val"]
    # the twin takes its pixel and shadow decisions from the float64 oracle:
    # a facet holding a ray that the kernels decide differently (an edge-band
    # ray, printed by the forward parity tests) differs from the twin by that
    # ray's weight jump, and is left out of the pose comparison
    sc = oracle_scene(tel)
    _, vk, pk = ia.render_debug(tel, torch.as_tensor(src), torch.as_tensor(vals), stype,
                                return_pixels=True)
    vk, pk = vk.cpu().numpy(), pk.cpu().numpy()
    _, _, v64, p64 = O.render(sc, src, vals, stype, "fp64", "debug")
    differ = (pk != p64) | ((vk != 0) != (v64 != 0))
    nf = sc.points.shape[0]
    keep = ~differ.reshape(nf, len(src), -1).any((1, 2))
    print(f"[source-major {case}] rays decided differently from float64: "
          f"{int(differ.sum())}, in {int((~keep).sum())} of {nf} facets", flush=True)
    # on a hard sensor d/dval_s = sum over s's rays of W[pixel] * value / val_s
    # exactly, so the differently decided rays shift the twin's value
    # gradient by a known amount
    def wsum(p, v):
        return np.where(p >= 0, W[np.clip(p, 0, None)], 0.0) * v
    shift = np.where(differ, wsum(pk, vk.astype(np.float64)) - wsum(p64, v64), 0.0)
    gt["val"] = gt["val"] + shift.reshape(nf, len(src), -1).sum((0, 2)) / np.asarray(vals,
                                                                                    np.float64)
    # every facet, at the gradient tolerance; and the pose gradients of the
    # facets without such a ray and the (shifted) value gradients, at the
    # kernels' own rounding
    _compare(f"source-major {case}", sm, gt, keys, tol=1e-3)
    _compare(f"source-major {case} (agreeing facets)", sm, gt,
             [k for k in keys if k != "src"], tol=1e-4, facet_mask=keep)


def test_stage_mirror_pose_gradients_refused():
    """A stage >= 1 mirror's pose that requires grad is refused with a clear
    error (the adjoint covers stage-0 poses, sources, values and aspheric
    coefficients), not silently left without a gradient."""
    wl = configs.workload("C4", n_samples=64)
    tel = wl.telescope(DEV, key=1)
    g0, g1 = tel.mirror_groups
    src = torch.as_tensor(wl.sources[:16])
    vals = torch.as_tensor(wl.values[:16])
    for field in ("rotations", "positions"):
        leaf = getattr(g1, field).detach().clone().requires_grad_(True)
        t2 = ia.Telescope([g0, g1.replace(**{field: leaf})], [], tel.sensors)
        with pytest.raises(NotImplementedError, match="stage-1"):
            ia.render(t2, src, vals, wl.source_type)
        with torch.no_grad():
            img = ia.render(t2, src, vals, wl.source_type)
        assert float(img.sum()) > 0
