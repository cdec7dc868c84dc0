"""iact_trace_fwd_noalloc (scratch in a static device array) against
iact_trace_fwd (caller workspace), through the C ABI (GPU).

Two calls in a row on one device, the second with a smaller input (the
scratch keeps the first call's contents): per-ray debug output bit for bit,
images to 1e-6 of each significant pixel (the two differ only in the order
of float additions).  Also: an image render whose default layout needs more
scratch than the array holds runs the sample-major layout and matches
iact_trace_fwd with that layout; a render that cannot fit is rejected before
any launch.
"""
import ctypes

import numpy as np
import pytest
import torch

from iactrace_amd import _lib, configs
from iactrace_amd.render import LaunchConfig, RenderPlan

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _inputs(name, ns, m, key=0):
    wl = configs.workload(name, n_sources=ns, n_samples=m)
    tel = wl.telescope(DEV, key=key)
    src = torch.as_tensor(wl.sources, dtype=torch.float64, device=DEV).contiguous()
    val = torch.as_tensor(wl.values, dtype=torch.float32, device=DEV).contiguous()
    st = _lib.SOURCE_POINT if wl.source_type == "point" else _lib.SOURCE_PARALLEL
    return RenderPlan(tel), src, val, st


def _outputs(plan, ns, mode):
    P = plan.sensor.n_pixels
    nf = sum(len(g) for g in plan.primary)
    m = plan.primary[0].n_samples
    if mode == _lib.MODE_IMAGE:
        return torch.full((P,), -7.0, device=DEV), None, None
    n = nf * ns * m
    return (torch.full((n,), -7.0, device=DEV), torch.full((n, 2), -7.0, device=DEV),
            torch.full((n,), -7, dtype=torch.int32, device=DEV))


def _fwd(plan, src, val, st, mode, launch, noalloc):
    keep = [src, val]
    sc = plan.scene(keep)
    L = launch.abi()
    ns = src.shape[0]
    out, xy, pix = _outputs(plan, ns, mode)
    lib = _lib.lib()
    if noalloc:
        rc = lib.iact_trace_fwd_noalloc(ctypes.byref(sc), _lib.ptr(src), _lib.ptr(val), ns, st,
                                        mode, ctypes.byref(L), _lib.ptr(out), _lib.ptr(xy),
                                        _lib.ptr(pix), _lib.current_stream())
        _lib.check(rc, "iact_trace_fwd_noalloc")
    else:
        nb = lib.iact_workspace_size(ctypes.byref(sc), ns, mode, ctypes.byref(L))
        ws = torch.empty(max(int(nb), 1), dtype=torch.uint8, device=DEV)
        rc = lib.iact_trace_fwd(ctypes.byref(sc), _lib.ptr(src), _lib.ptr(val), ns, st, mode,
                                ctypes.byref(L), _lib.ptr(ws), ws.numel(), _lib.ptr(out),
                                _lib.ptr(xy), _lib.ptr(pix), _lib.current_stream())
        _lib.check(rc, "iact_trace_fwd")
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (out, xy, pix) if t is not None]


def _close_images(a, b, label):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    sig = np.abs(b) > 1e-6 * np.abs(b).max()
    rel = (np.abs(a - b)[sig] / np.abs(b[sig])).max()
    print(f"[{label}] peak {np.abs(b).max():.4g}, max rel err {rel:.2e}", flush=True)
    assert np.abs(b).max() > 0
    assert rel <= 1e-6
    assert np.all(np.abs(a - b)[~sig] <= 1e-6 * np.abs(b).max())


def test_noalloc_matches_workspace_entry_two_calls():
    L = LaunchConfig()
    for step, (ns, m) in enumerate([(300, 64), (40, 16)]):  # the second call is smaller
        plan, src, val, st = _inputs("C3", ns, m, key=step)
        print(f"\n[noalloc call {step + 1}] C3 geometry, {ns} sources x {m} samples", flush=True)
        a = _fwd(plan, src, val, st, _lib.MODE_IMAGE, L, True)[0]
        b = _fwd(plan, src, val, st, _lib.MODE_IMAGE, L, False)[0]
        _close_images(a, b, f"image call {step + 1}")
        da = _fwd(plan, src, val, st, _lib.MODE_DEBUG, L, True)
        db = _fwd(plan, src, val, st, _lib.MODE_DEBUG, L, False)
        for x, y in zip(da, db):
            assert np.array_equal(x, y)
        assert (da[2] >= 0).sum() > 0


def test_noalloc_takes_the_sample_major_layout_when_records_do_not_fit():
    # 382 facets x 4096 samples: the source-major record table alone is
    # 100 MB, more than the scratch holds
    plan, src, val, st = _inputs("C2", 200, 4096)
    print("\n[noalloc layout fallback] C2 geometry, 200 sources x 4096 samples", flush=True)
    lib = _lib.lib()
    keep = [src, val]
    sc = plan.scene(keep)
    L0 = LaunchConfig().abi()
    assert lib.iact_workspace_size(ctypes.byref(sc), 200, _lib.MODE_IMAGE,
                                   ctypes.byref(L0)) > _lib.IACT_NOALLOC_SCRATCH_BYTES
    a = _fwd(plan, src, val, st, _lib.MODE_IMAGE, LaunchConfig(), True)[0]
    b = _fwd(plan, src, val, st, _lib.MODE_IMAGE, LaunchConfig(layout=1), False)[0]
    _close_images(a, b, "layout fallback")


def test_noalloc_rejects_what_does_not_fit():
    # response matrix of 382 x 4096 samples: 44 MB of world-space samples
    plan, src, val, st = _inputs("C2", 4, 4096)
    keep = [src, val]
    sc = plan.scene(keep)
    L = LaunchConfig().abi()
    out = torch.zeros((4, plan.sensor.n_pixels), device=DEV)
    lib = _lib.lib()
    rc = lib.iact_trace_fwd_noalloc(ctypes.byref(sc), _lib.ptr(src), _lib.ptr(val), 4, st,
                                    _lib.MODE_RESPONSE, ctypes.byref(L), _lib.ptr(out), None,
                                    None, _lib.current_stream())
    assert rc == _lib.IACT_ERR_UNSUPPORTED
    assert b"scratch" in lib.iact_last_error()
    assert float(out.abs().sum()) == 0.0
