"""Batched renders, their adjoint and the alignment-fit step (GPU).

The batched kernels are the plain trace/adjoint kernels over an extended
work list, so every instance must reproduce the plain render of a telescope
carrying that instance's rotations and sources (same pixel decisions; only
the float summation order of the atomics differs), and the per-instance
rotation gradients must match the plain adjoint's.  The fit step is checked
against torch: L2 loss/gradient, torch.optim.Adam, and graph replays against
eager calls of the same entry points.
"""
import numpy as np
import pytest
import torch

import iactrace_amd as ia
from iactrace_amd import configs
from iactrace_amd.batched import AlignmentFit, BatchedPlan

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _instances(tel, rotations, src, vals, off, stype):
    """Per-instance plain renders and rotation gradients (weights W)."""
    imgs = []
    for i in range(len(off) - 1):
        t = tel.set_mirror_rotations(0, rotations[i])
        s, v = src[off[i]:off[i + 1]], vals[off[i]:off[i + 1]]
        imgs.append(ia.render(t, torch.as_tensor(s), torch.as_tensor(v), stype).reshape(-1))
    return torch.stack(imgs)


def _small_c3(n_inst=5, sizes=(0, 7, 300, 1, 64)):
    wl = configs.workload("C3", n_sources=8, n_samples=96)
    tel = wl.telescope(DEV, key=1)
    base = tel.mirror_groups[0].rotations.detach()
    rng = np.random.default_rng(0)
    rot = base.unsqueeze(0).repeat(n_inst, 1, 1).clone()
    rot[:, :, :2] += torch.as_tensor(rng.normal(0, 0.03, (n_inst, base.shape[0], 2)),
                                     dtype=torch.float32, device=DEV)
    src, vals = [], []
    for i, n in enumerate(sizes):
        a, b = configs.point_source_field(n, 1.5, seed=50 + i)
        src.append(a)
        vals.append(b)
    s, v, off = ia.pack_sources(src, vals)
    return tel, rot.contiguous(), s, v, off


@pytest.mark.parametrize("tile,layout", [(0, 0), (0, 1), (16, 1)])
def test_batched_forward_matches_plain_hex_obstructions(tile, layout):
    """layout 0: instance by instance with the plain render's kernels;
    layout 1: the sample-major kernel over the whole batch's work list."""
    tel, rot, s, v, off = _small_c3()
    launch = ia.LaunchConfig(source_tile=tile, layout=layout)
    img_b = ia.render_batched(tel, rot, s, v, off, "point", launch=launch)
    img_p = _instances(tel, rot, s, v, off, "point")
    print(f"[batched-hex tile={tile} layout={layout}] instances {len(off) - 1}, "
          f"sources {np.diff(off).tolist()}", flush=True)
    assert img_b.shape == img_p.shape
    for i in range(len(off) - 1):
        a, b = img_b[i].cpu().numpy(), img_p[i].cpu().numpy()
        if off[i + 1] == off[i]:
            assert np.all(a == 0)
            continue
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
        print(f"  instance {i}: max rel err {err:.2e}, sum {b.sum():.6g}", flush=True)
        assert err < 1e-5


def test_batched_forward_square_parallel():
    wl = configs.workload("C2", n_sources=8, n_samples=128)
    tel = wl.telescope(DEV, key=2)
    base = tel.mirror_groups[0].rotations.detach()
    rot = torch.stack([base, base + 0.02, base - 0.01]).contiguous()
    d = configs.parallel_grid(4, 3, 0.3)
    src = [d[:5], d[5:], d[:2]]
    vals = [np.ones(5, np.float32), np.full(7, 0.5, np.float32), np.array([2.0, 3.0], np.float32)]
    s, v, off = ia.pack_sources(src, vals)
    img_b = ia.render_batched(tel, rot, s, v, off, "parallel")
    img_p = _instances(tel, rot, s, v, off, "parallel")
    err = (img_b - img_p).abs().max().item() / img_p.abs().max().item()
    print(f"[batched-square-parallel] max rel err {err:.2e}", flush=True)
    assert img_p.abs().max().item() > 0
    assert err < 1e-5


def test_batched_rotation_gradients_match_plain():
    tel, rot, s, v, off = _small_c3(3, (40, 5, 129))
    P = tel.sensors[0].n_pixels
    W = torch.linspace(0.5, 1.5, P, device=DEV)
    r = rot.clone().requires_grad_(True)
    img = ia.render_batched(tel, r, s, v, off, "point")
    (img * W).sum().backward()
    for i in range(3):
        g = tel.mirror_groups[0]
        ri = rot[i].clone().requires_grad_(True)
        t = ia.Telescope([g.replace(rotations=ri)], tel.obstruction_groups, tel.sensors)
        im = ia.render(t, torch.as_tensor(s[off[i]:off[i + 1]]),
                       torch.as_tensor(v[off[i]:off[i + 1]]), "point")
        (im.reshape(-1) * W).sum().backward()
        a, b = r.grad[i].double().cpu().numpy(), ri.grad.double().cpu().numpy()
        err = np.abs(a - b).max() / np.abs(b).max()
        print(f"[batched-grad] instance {i}: max rel err {err:.2e} (scale {np.abs(b).max():.3g})",
              flush=True)
        assert err < 1e-4


def test_l2_loss_grad_and_adam_match_torch():
    from iactrace_amd import _lib
    import ctypes
    lib = _lib.lib()
    torch.manual_seed(0)
    img = torch.rand(3, 1000, device=DEV)
    tgt = torch.rand(3, 1000, device=DEV)
    g = torch.empty_like(img)
    loss = torch.empty(3, dtype=torch.float64, device=DEV)
    _lib.check(lib.iact_l2_loss_grad(_lib.ptr(img), _lib.ptr(tgt), 1000, 3, _lib.ptr(g),
                                     _lib.ptr(loss), _lib.current_stream()), "l2")
    d = (img - tgt).double()
    torch.testing.assert_close(loss, (d * d).sum(1), rtol=1e-12, atol=0)
    torch.testing.assert_close(g, 2 * (img - tgt), rtol=0, atol=0)
    # Adam vs torch.optim.Adam over 5 steps
    p = torch.randn(300, device=DEV)
    p_ref = p.clone().requires_grad_(True)
    opt = torch.optim.Adam([p_ref], lr=1e-2)
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    step = torch.zeros(1, dtype=torch.int32, device=DEV)
    for k in range(5):
        grad = torch.randn(300, device=DEV, dtype=torch.float64)
        p_ref.grad = grad.float()
        opt.step()
        _lib.check(lib.iact_adam_step(_lib.ptr(p), _lib.ptr(grad), _lib.ptr(m), _lib.ptr(v),
                                      _lib.ptr(step), 300, ctypes.c_float(1e-2),
                                      ctypes.c_float(0.9), ctypes.c_float(0.999),
                                      ctypes.c_float(1e-8), None, _lib.current_stream()), "adam")
        torch.testing.assert_close(p, p_ref.detach(), rtol=1e-6, atol=1e-7)
    assert int(step.item()) == 5


def test_fit_graph_replays_match_eager_and_torch_adam():
    tel, rot, s, v, off = _small_c3(3, (20, 3, 70))
    with torch.no_grad():
        base = tel.mirror_groups[0].rotations.detach()
        targets = ia.render_batched(tel, base.unsqueeze(0).repeat(3, 1, 1).contiguous(), s, v,
                                    off, "point")
    lr = 2e-3
    fit_g = AlignmentFit(tel, rot, s, v, off, targets, lr=lr, use_graph=True)
    fit_e = AlignmentFit(tel, rot, s, v, off, targets, lr=lr, use_graph=False)
    # torch reference: autograd through the batched render + torch.optim.Adam
    r_ref = rot.clone().requires_grad_(True)
    opt = torch.optim.Adam([r_ref], lr=lr)
    losses = []
    for k in range(4):
        lg = fit_g.step().clone()
        torch.cuda.synchronize()
        le = fit_e.step().clone()
        torch.cuda.synchronize()
        opt.zero_grad()
        img = ia.render_batched(tel, r_ref, s, v, off, "point")
        lt = ((img - targets) ** 2).sum(1)
        lt.sum().backward()
        opt.step()
        torch.cuda.synchronize()
        # eager work between replays (a library reduction on the same stream)
        _ = fit_g.images.sum().item()
        rl = (lg - le).abs().max().item() / le.abs().max().item()
        # images are summed with float atomics in no fixed order, so the
        # gradients of two runs differ at the 1e-7 level; Adam's first steps
        # are ~lr * sign(g), so a parameter whose gradient is that close to
        # zero may move differently -- allow 1 in 1000 such parameters
        dr = (fit_g.rotations - fit_e.rotations).abs()
        dt = (fit_e.rotations - r_ref.detach()).abs()
        fr = (dr > 1e-3 * lr).float().mean().item()
        ft = (dt > 1e-3 * lr).float().mean().item()
        step_move = (fit_e.rotations - rot).abs().max().item()
        print(f"[fit] replay {k}: loss {le.sum().item():.6g} graph-vs-eager loss {rl:.2e}, "
              f"rotations off {fr:.1e} (max {dr.max().item():.2e}); eager-vs-torch off {ft:.1e} "
              f"(max {dt.max().item():.2e}); moved {step_move:.3g} deg", flush=True)
        losses.append(le.sum().item())
        assert rl < 1e-4
        assert fr <= 1e-3 and dr.max().item() <= 2.5 * lr * (k + 1)
        assert ft <= 1e-3 and dt.max().item() <= 2.5 * lr * (k + 1)
        assert int(fit_g.step_count.item()) == k + 1
    assert losses[-1] < losses[0]
    fit_g.close()


def test_batched_plan_validation():
    tel, rot, s, v, off = _small_c3(2, (3, 4))
    bp = BatchedPlan(tel, s, v, off)
    with pytest.raises(ValueError):
        bp.forward(rot[:1].contiguous())
    with pytest.raises(ia._lib.IactError):
        ia.render_batched(tel, rot[:2].contiguous(), s, v, off, launch=ia.LaunchConfig(
            deterministic=True))


def test_c5_reduced_instances_against_oracle():
    """A reduced C5 -- 4 instances with their own tilt draws, 4 samples, 64
    point sources each (the source-major kernels): every batched image
    equals the plain render of its instance's telescope, and instances 0
    and 3 meet the oracle per ray (pixel ids outside the edge band, values,
    image; tests.test_gpu_forward)."""
    from tests.test_gpu_forward import _trace_and_compare
    bw = configs.batch_workload(4, n_min=64, n_max=64, n_samples=4)
    tel = bw.telescope(DEV, key=3)
    rot = bw.rotations(tel)
    imgs = ia.render_batched(tel, rot, bw.sources, bw.values, bw.offsets, "point")
    for i in range(4):
        t = tel.set_mirror_rotations(0, rot[i])
        a, b = int(bw.offsets[i]), int(bw.offsets[i + 1])
        ref = ia.render(t, torch.as_tensor(bw.sources[a:b]), torch.as_tensor(bw.values[a:b]),
                        "point").reshape(-1).double()
        got = imgs[i].double()
        sig = ref > 1e-3 * ref.max()
        rel = ((got - ref).abs()[sig] / ref[sig]).max().item()
        print(f"[C5 reduced instance {i}] batched vs plain max rel {rel:.2e}", flush=True)
        assert rel < 1e-6
        assert (got - ref).abs()[~sig].max().item() <= 1e-6 * ref.max().item()
        if i in (0, 3):
            _trace_and_compare(t, bw.sources[a:b], bw.values[a:b], "point",
                               f"C5 reduced instance {i}")


def test_c5_full_size_instances_and_fit():
    """Full C5 (64 instances, 875 facets x 1024 samples, 548694 sources):
    batched instances 0, 17 and 63 equal plain renders of their telescopes,
    and the loss of the captured fit iteration falls over three steps."""
    bw = configs.batch_workload(64)
    tel = bw.telescope(DEV, key=0)
    rot = bw.rotations(tel)
    bp = BatchedPlan(tel, bw.sources, bw.values, bw.offsets, "point")
    with torch.no_grad():
        imgs = bp.forward(rot)
    src = torch.as_tensor(bw.sources, device=DEV)
    val = torch.as_tensor(bw.values, device=DEV)
    for i in (0, 17, 63):
        a, b = int(bw.offsets[i]), int(bw.offsets[i + 1])
        ref = ia.render(tel.set_mirror_rotations(0, rot[i]), src[a:b], val[a:b], "point")
        ref = ref.reshape(-1).double()
        got = imgs[i].double()
        sig = ref > 1e-3 * ref.max()
        rel = ((got - ref).abs()[sig] / ref[sig]).max().item()
        print(f"[C5 instance {i}] {b - a} sources, max rel vs plain render {rel:.2e}", flush=True)
        assert rel < 1e-6
    base = tel.mirror_groups[0].rotations.detach()
    with torch.no_grad():
        targets = bp.forward(base.unsqueeze(0).repeat(64, 1, 1).contiguous())
    fit = AlignmentFit(tel, rot, bw.sources, bw.values, bw.offsets, targets, lr=1e-4)
    losses = []
    for _ in range(4):
        fit.step()
        losses.append(float(fit.loss.sum()))
    fit.close()
    print(f"[C5 fit] losses {losses}", flush=True)
    assert losses[-1] < losses[0]
