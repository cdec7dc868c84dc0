"""Benchmark of the MI355X ray-tracing path (one JSON line on stdout).

Headline: configuration C5 -- 64 instances of the 28 m parabolic dish (875
hexagonal facets x 1024 MC samples, hexagonal 2048-pixel camera, camera box
and four masts), each with its own 0.5 mrad tilt draw and its own field of
U{256..16384} point sources at 1e10 m -- one alignment-fit iteration per
step: batched forward render, L2 loss against the untilted dish's images,
batched adjoint (rotation gradients), Adam on every facet's Euler angles,
replayed as one captured HIP graph.  `value` is traced rays per second of
the whole iteration (a ray = one (source, facet, sample) triple, traced
forward and backward).  Inputs are resident in HBM before timing starts.

Also reported:
  * roofline of the dominant kernel, the source-major adjoint k_adj_sm
    (per instance): its algorithmic FP32 flops per batched backward (rays x
    2 x the forward's FP32 flops per ray, the backward's share under the
    fwd+bwd = 3 x fwd convention, no credit for obstruction tests) over its
    HIP-event time, against the FP32 vector peak, with the VALU issue rate
    and HBM bytes of the committed rocprofv3 counter summary; the whole
    iteration's fraction beside it;
  * the same iteration done plainly (per-instance render + autograd);
  * image parity of batched instances against oracle/cpu_ref's FP64 run
    (and, with the C2 extra, of the full 6.4e9-ray C2 image);
  * the CPU baseline: oracle/cpu_ref (the reference's algorithm in C++ /
    OpenMP, fp32) forward + adjoint on a C5 subsample, median of 5;
  * extras: C2 plain forward (the previous headline), C3 forward and
    forward+backward, C3T obstruction culling, C4 response matrix.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-parity]
                       [--extras C2,C3,C3T,C4 | none] [--c5-instances 64]
                       [--only C2|C3|C3T|C4|copy]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32 = 157.3e12
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12
METRIC = "rays/s fwd and fwd+bwd, batched vs plain, 1 MI355X at ≤1e-5 image rel err"


def _sync():
    import torch
    torch.cuda.synchronize()


def _threads():
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)


def time_steps(fn, steps, warmup):
    """Wall time of `steps` calls and the HIP-event time of each call (events
    on the current stream, which every launch of the library uses)."""
    import torch
    for _ in range(warmup):
        fn()
    _sync()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record()
        fn()
        b.record()
    _sync()
    wall = time.perf_counter() - t0
    return wall, [a.elapsed_time(b) for a, b in evs]


def profile_value(name, key):
    """A number from a committed profile summary (`key = value` lines)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith(key + " ="):
                return float(line.split("=")[1])
    return None


def fp32_flops(wl, n_sources):
    """FP32 algorithmic flops per ray of a workload's forward, without the
    obstruction tests (culling skips most; the ones done are not credited)."""
    import iactrace_amd as ia
    from iactrace_amd.flop_model import split_per_ray
    f32, _ = split_per_ray(wl.source_type, 0, 0, 0,
                           sum(m.optical_stage > 0 for m in wl.mirrors), 0,
                           "square" if isinstance(wl.sensor, ia.SquareSensor) else "hex",
                           n_sources)
    return f32


def image_parity(sc, src, vals, source_type, gpu_image, label):
    """A GPU image against oracle/cpu_ref's FP64 run of the same samples
    (every step in double, image summed in double), on the host cores:
    relative L1 error, the largest relative error of a significant pixel
    (>= 1e-3 of the peak) and the same after taking away the pixel's
    edge-band mass (|value| of the FP64 rays within the edge band of a
    binning boundary -- 2.5e-5 m on the sensor, tests.helpers.band_units -- at
    their pixel and the one across the nearest edge: the rays an fp32
    coordinate may put on either side); and the fp32 CPU port's L1 error,
    the fp32 floor that any evaluation of the reference's arithmetic has."""
    import numpy as np
    from oracle import cpu_ref
    from tests.helpers import band_units
    band = band_units(sc.sensor)
    keep = []
    s = cpu_ref.scene_struct(sc, keep)
    P = sc.sensor.n_pixels
    t0 = time.perf_counter()
    i64, b64, st = cpu_ref.render_band(s, src, vals, source_type, P, _threads(), 64, band)
    dt = time.perf_counter() - t0
    i32 = cpu_ref.render(s, src, vals, source_type, P, _threads(), 32)
    g = np.asarray(gpu_image, np.float64).reshape(-1)
    sig = i64 > 1e-3 * i64.max()
    d = np.abs(g - i64)
    return {"case": label,
            "reference": f"oracle/cpu_ref.cc FP64 run of the same samples, {st['traced']:.4g} "
                         f"rays ({dt:.1f} s on {_threads()} host threads)",
            "l1_rel": float(d.sum() / np.abs(i64).sum()),
            "cpu_fp32_l1_rel": float(np.abs(i32 - i64).sum() / np.abs(i64).sum()),
            "significant_px": int(sig.sum()),
            "max_rel_significant_px": float((d[sig] / i64[sig]).max()),
            "max_rel_significant_px_band_resolved":
                float((np.maximum(d - b64, 0.0)[sig] / i64[sig]).max()),
            "band_units": band, "band_rays": st["band"], "traced_rays": st["traced"],
            "band_fraction": st["band"] / max(st["traced"], 1)}


# ------------------------------------------------------------------ C5 headline
def c5(ia, configs, n_instances, steps, warmup, want_parity, want_plain):
    import numpy as np
    import torch
    from iactrace_amd.batched import AlignmentFit, BatchedPlan
    dev = torch.device("cuda:0")
    bw = configs.batch_workload(n_instances)
    tel = bw.telescope(dev, key=0)
    rot = bw.rotations(tel)
    base = tel.mirror_groups[0].rotations.detach()
    bp = BatchedPlan(tel, bw.sources, bw.values, bw.offsets, "point")
    with torch.no_grad():
        targets = bp.forward(base.unsqueeze(0).repeat(n_instances, 1, 1).contiguous())
    n = bw.n_rays
    fit = AlignmentFit(tel, rot, bw.sources, bw.values, bw.offsets, targets, lr=1e-4)
    losses = []

    def step():
        fit.step()
        losses.append(fit.loss.sum())

    wall, ev = time_steps(step, steps, warmup)
    loss_curve = [float(x) for x in losses]
    fit.close()
    it_s = wall / steps
    # the two kernels of the iteration, timed eagerly with HIP events
    g_img = (2.0 * (fit.images - targets)).contiguous()
    _, ev_f = time_steps(lambda: bp.forward(rot), 3, 1)
    _, ev_b = time_steps(lambda: bp.backward(rot, g_img), 3, 1)
    f_ms, b_ms = float(np.mean(ev_f)), float(np.mean(ev_b))
    f32 = fp32_flops(bw.base, int(bw.offsets[-1]) / n_instances)
    bwd_tf = n * 2 * f32 / (b_ms * 1e-3) / 1e12
    fwd_tf = n * f32 / (f_ms * 1e-3) / 1e12
    it_tf = n * 3 * f32 / it_s / 1e12
    flops = {"forward": f32, "adjoint": 2 * f32,
             "convention": "fwd+bwd = 3 x fwd; FP32 only; no credit for obstruction tests; "
                           "the 12 FP64 flops of a point source's direction not counted"}
    peak_note = ("157.3 TFLOP/s assumes packed FP32 (two FMAs per lane per instruction); these "
                 "kernels are built without packed FP32.  frac counts algorithmic flops, which "
                 "include work the kernels do once per facet and source instead of per ray, "
                 "and the adjoint at twice the forward's flops (so its frac can exceed 1); "
                 "valu_issue (rocprofv3 counters: VALU instructions x 4 cycles / (SIMDs x "
                 "cycles)) is the hardware's measured utilisation")

    def roof(tf, ms, name, what, counters, key):
        return {"bound": "valu", "achieved": tf, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                "frac": tf * 1e12 / PEAK_FP32,
                "traffic": profile_value(counters, "hbm_bytes_per_launch"),
                "kernel": f"{name} ({what}, one launch per instance), HIP events around the "
                          "batched call on the launch stream, mean of 3",
                "kernel_ms": ms, "rocprof_kernel_ms": profile_value(counters, key),
                "fp32_flops_per_ray": flops, "peak_note": peak_note,
                "valu_issue": profile_value(counters, "valu_issue_frac")}

    fwd_roof = roof(fwd_tf, f_ms, "k_trace_sm", "source-major forward",
                    "k_trace_sm_C5_counters.txt", "kernel_ms_per_forward")
    adj_roof = roof(bwd_tf, b_ms, "k_adj_sm", "source-major adjoint",
                    "k_adj_sm_C5_counters.txt", "kernel_ms_per_backward")
    out = {
        "value": n / it_s, "ms_per_step": 1e3 * it_s, "rays_per_step": n,
        "event_ms_per_step": float(np.mean(ev)),
        "loss": {"first_step": loss_curve[0], "last_step": loss_curve[-1],
                 "steps": len(loss_curve), "falls": loss_curve[-1] < loss_curve[0]},
        "forward_kernel": fwd_roof,
        "adjoint_kernel": adj_roof,
        # the dominant kernel: the one of the two per-pass kernels that took
        # longer in this run (HIP events)
        "roofline": dict(fwd_roof if f_ms >= b_ms else adj_roof),
        "iteration_fp32_frac": it_tf * 1e12 / PEAK_FP32,
    }
    if want_plain:
        g = tel.mirror_groups[0]
        src = torch.as_tensor(bw.sources, device=dev)
        val = torch.as_tensor(bw.values, device=dev)

        def plain():
            for i in range(n_instances):
                a, b = int(bw.offsets[i]), int(bw.offsets[i + 1])
                r = rot[i].clone().requires_grad_(True)
                t = ia.Telescope([g.replace(rotations=r)], tel.obstruction_groups, tel.sensors)
                img = ia.render(t, src[a:b], val[a:b], "point").reshape(-1)
                ((img - targets[i]) ** 2).sum().backward()

        wp, _ = time_steps(plain, 1, 0)
        out["plain_fwd_bwd"] = {"s": wp, "rays_per_s": n / wp,
                                "how": f"{n_instances} x (render + autograd backward wrt tilts)",
                                "batched_speedup": wp / it_s}
    if want_parity:
        # batched images of 4 instances, each with the first 64 of its
        # sources, against the FP64 run (the same batched kernels)
        from tests.helpers import oracle_scene
        k = min(4, n_instances)
        idx = [int(bw.offsets[i]) for i in range(k)]
        s_sub = np.concatenate([bw.sources[a:a + 64] for a in idx])
        v_sub = np.concatenate([bw.values[a:a + 64] for a in idx])
        off = np.arange(k + 1, dtype=np.int64) * 64
        imgs = ia.render_batched(tel, rot[:k].contiguous(), s_sub, v_sub, off, "point")
        imgs = imgs.cpu().numpy()
        par = []
        for i in range(k):
            sc = oracle_scene(tel.set_mirror_rotations(0, rot[i]))
            par.append(image_parity(sc, s_sub[64 * i:64 * (i + 1)], v_sub[64 * i:64 * (i + 1)],
                                    "point", imgs[i], f"C5 instance {i}"))
        out["parity"] = {
            "case": f"C5 batched images, instances 0..{k - 1} with 64 sources each "
                    f"(875 facets x 1024 samples: {k} x 5.7e7 rays)",
            "l1_rel_max": max(p["l1_rel"] for p in par),
            "cpu_fp32_l1_rel_max": max(p["cpu_fp32_l1_rel"] for p in par),
            "max_rel_significant_px_band_resolved":
                max(p["max_rel_significant_px_band_resolved"] for p in par),
            "max_rel_significant_px": max(p["max_rel_significant_px"] for p in par),
            "band_fraction": max(p["band_fraction"] for p in par),
            "reference": par[0]["reference"]}
    out["workload"] = (f"C5: {n_instances} x C3 dish (875 hex facets x 1024 samples, hex 2048 "
                       f"px, box + 4 cylinders), independent 0.5 mrad tilt draws, "
                       f"{int(bw.offsets[-1])} point sources in total (U{{256..16384}} per "
                       f"instance); one step = batched fwd + L2 loss + batched adjoint + Adam, "
                       f"replayed as a HIP graph")
    out["_state"] = (bw, tel, rot, targets)
    return out


def cpu_fwd_bwd(bw, tel, rot, targets, reps=5):
    """oracle/cpu_ref forward image + adjoint (rotation and value gradients
    of the L2 loss) of C5 instance 0 with its first 64 sources: median of
    `reps` runs after one warm-up, on every host thread."""
    import numpy as np
    from oracle import cpu_ref
    from tests.helpers import oracle_scene
    sc = oracle_scene(tel.set_mirror_rotations(0, rot[0]))
    keep = []
    s = cpu_ref.scene_struct(sc, keep)
    ns = 64
    src, val = bw.sources[:ns], bw.values[:ns]
    P = sc.sensor.n_pixels
    nf, M = sc.points.shape[:2]
    tgt = targets[0].detach().cpu().numpy().astype(np.float64)
    nt = _threads()

    def once():
        t0 = time.perf_counter()
        img = cpu_ref.render(s, src, val, "point", P, nt, 32)
        t1 = time.perf_counter()
        cpu_ref.grad(s, src, val, "point", 2.0 * (img - tgt), nf, nt)
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1

    once()
    runs = [once() for _ in range(reps)]
    tf = statistics.median(r[0] for r in runs)
    tb = statistics.median(r[1] for r in runs)
    rays = nf * M * ns
    host = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            host = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")),
                        host)
    except OSError:
        pass
    return {"value": rays / (tf + tb), "unit": "rays/s (forward + backward)", "cores": nt,
            "kind": "port", "host_cpu": host, "nproc": os.cpu_count(),
            "extrapolated_full_iteration_s": bw.n_rays / (rays / (tf + tb)),
            "sample": f"C5 instance 0 with its first {ns} of {int(bw.offsets[1])} sources x "
                      f"{nf} facets x {M} samples ({rays:.3g} rays): oracle/cpu_ref.cc forward "
                      f"image ({1e3 * tf:.0f} ms) + adjoint of the L2 loss wrt rotations and "
                      f"values ({1e3 * tb:.0f} ms), fp32 with fp64 directions, g++ -O3 "
                      f"-march=x86-64-v3 -fopenmp, median of {reps}",
            "forward_rays_per_s": rays / tf, "backward_rays_per_s": rays / tb}


# ------------------------------------------------------------------ extras
def hbm_copy(steps, warmup, nbytes=1 << 30):
    """A plain device copy of `nbytes` (the calibration of the HBM byte
    counters: FETCH_SIZE and WRITE_SIZE of this kernel are nbytes each)."""
    import numpy as np
    import torch
    a = torch.ones(nbytes, dtype=torch.uint8, device="cuda:0")
    b = torch.empty_like(a)
    _, ev = time_steps(lambda: b.copy_(a), steps, warmup)
    ms = float(np.mean(ev))
    return {"workload": f"device copy of {nbytes} bytes", "ms": ms,
            "gb_per_s": 2 * nbytes / (ms * 1e-3) / 1e9}


def c2_forward(ia, configs, steps, warmup, want_parity):
    import numpy as np
    import torch
    dev = torch.device("cuda:0")
    wl = configs.workload("C2")
    tel = wl.telescope(dev, key=0)
    src = torch.as_tensor(wl.sources, dtype=torch.float64, device=dev)
    vals = torch.as_tensor(wl.values, dtype=torch.float32, device=dev)
    wall, ev = time_steps(lambda: ia.render(tel, src, vals, "point"), steps, warmup)
    f32 = fp32_flops(wl, len(wl.sources))
    k_ms = float(np.mean(ev))
    out = {"workload": "C2 plain forward: Davies-Cotton f=15 m, 382 hex facets x 4096 samples, "
                       "4096 point sources at 1e10 m (2.5 deg field, lognormal values), square "
                       "32x30 camera",
           "rays_per_step": wl.n_rays, "steps": steps, "fwd_rays_per_s": wl.n_rays * steps / wall,
           "fwd_ms": 1e3 * wall / steps,
           "roofline": {"kernel": "k_sample_records + k_trace_sm (HIP events around the render)",
                        "kernel_ms": k_ms, "fp32_flops_per_ray": f32,
                        "achieved_tflops": wl.n_rays * f32 / (k_ms * 1e-3) / 1e12,
                        "frac": wl.n_rays * f32 / (k_ms * 1e-3) / PEAK_FP32,
                        "traffic": profile_value("k_trace_sm_C2_traffic.txt",
                                                 "hbm_bytes_per_launch")}}
    if want_parity:
        from tests.helpers import oracle_scene
        img = ia.render(tel, src, vals, "point").detach().cpu().numpy()
        out["parity"] = image_parity(oracle_scene(tel), wl.sources, wl.values, "point", img,
                                     "C2 full: 6.4e9 rays")
    return out


def c3_fwd_bwd(ia, configs, steps, warmup):
    """C3: forward, and forward + backward of a weighted image sum wrt facet
    tilts, positions and source values."""
    import torch
    dev = torch.device("cuda:0")
    wl = configs.workload("C3")
    tel = wl.telescope(dev, key=0)
    src = torch.as_tensor(wl.sources, dtype=torch.float64, device=dev)
    vals = torch.as_tensor(wl.values, dtype=torch.float32, device=dev)
    W = torch.linspace(0.5, 1.5, tel.sensors[0].n_pixels, device=dev)
    g = tel.mirror_groups[0]
    rot = g.rotations.detach().clone().requires_grad_(True)
    pos = g.positions.detach().clone().requires_grad_(True)
    v = vals.clone().requires_grad_(True)
    tel_g = ia.Telescope([g.replace(rotations=rot, positions=pos)] + list(tel.mirror_groups[1:]),
                         tel.obstruction_groups, tel.sensors)

    def fwd_bwd():
        rot.grad = pos.grad = v.grad = None
        (ia.render(tel_g, src, v, wl.source_type) * W).sum().backward()

    wf, _ = time_steps(lambda: ia.render(tel, src, vals, wl.source_type), steps, warmup)
    wb, _ = time_steps(fwd_bwd, steps, warmup)
    return {"workload": "C3: parabolic f=36 m, 875 facets x 1024 samples x 16384 point sources, "
                        "hex 2048 px, box + 4 cylinders, tilt errors 0.5 mrad",
            "rays_per_step": wl.n_rays, "steps": steps,
            "fwd_rays_per_s": wl.n_rays * steps / wf, "fwd_ms": 1e3 * wf / steps,
            "fwd_bwd_rays_per_s": wl.n_rays * steps / wb, "fwd_bwd_ms": 1e3 * wb / steps,
            "bwd_grad": "d(sum(W*image))/d(facet tilts, facet positions, source values)"}


def c3t_culling(ia, configs, steps, warmup):
    """C3 with a CT5-sized support truss (271 obstructions): forward with
    obstruction culling (the default) and without."""
    import torch
    dev = torch.device("cuda:0")
    wl = configs.workload("C3T")
    tel = wl.telescope(dev, key=0)
    src = torch.as_tensor(wl.sources, dtype=torch.float64, device=dev)
    vals = torch.as_tensor(wl.values, dtype=torch.float32, device=dev)
    out = {"workload": f"C3T: C3's dish and camera, 875 facets x 1024 samples x 16384 point "
                       f"sources, {len(wl.obstructions)} obstructions (masts, bracing, mount)",
           "rays_per_step": wl.n_rays, "steps": steps}
    for name, cull in (("culled", 1), ("unculled", -1)):
        L = ia.LaunchConfig(culling=cull)
        w, _ = time_steps(lambda: ia.render(tel, src, vals, wl.source_type, launch=L), steps,
                          warmup)
        out[f"{name}_rays_per_s"] = wl.n_rays * steps / w
        out[f"{name}_ms"] = 1e3 * w / steps
    return out


def c4_response(ia, configs, steps, warmup):
    """C4: Cassegrain (aspheric primary, hyperbolic secondary), square 64x64:
    the 8192 x 4096 response matrix, the image, and forward + backward of a
    weighted image wrt both mirrors' aspheric coefficients."""
    import torch
    dev = torch.device("cuda:0")
    wl = configs.workload("C4")
    tel = wl.telescope(dev, key=0)
    src = torch.as_tensor(wl.sources, dtype=torch.float64, device=dev)
    vals = torch.as_tensor(wl.values, dtype=torch.float32, device=dev)
    g0, g1 = tel.mirror_groups
    a0 = g0.aspheric.detach().clone().requires_grad_(True)
    a1 = torch.zeros(2).requires_grad_(True)
    t2 = ia.Telescope([g0.replace(aspheric=a0), g1.replace(aspheric=a1)], [], tel.sensors)
    W = torch.linspace(0.5, 1.5, tel.sensors[0].n_pixels, device=dev)

    def fwd_bwd():
        a0.grad = a1.grad = None
        (ia.render(t2, src, vals, wl.source_type).reshape(-1) * W).sum().backward()

    wr, _ = time_steps(lambda: ia.render_response_matrix(tel, src, vals, wl.source_type), steps,
                       warmup)
    wi, evi = time_steps(lambda: ia.render(tel, src, vals, wl.source_type), steps, warmup)
    wb, _ = time_steps(fwd_bwd, steps, warmup)
    from iactrace_amd.flop_model import split_per_ray
    f32, f64 = split_per_ray("parallel", 0, 0, 0, 1, 0, "square", len(wl.sources))
    ki = float(sum(evi) / len(evi)) * 1e-3
    return {"workload": "C4: Cassegrain, 1 primary x 2048 samples x 8192 parallel sources "
                        "(128 x 64 grid, 0.01 deg), square 64 x 64 sensor",
            "rays_per_step": wl.n_rays, "steps": steps,
            "image_roofline": {"kernel_ms": ki * 1e3, "fp32_flops_per_ray": f32,
                               "fp64_flops_per_ray": f64,
                               "fp32_frac": wl.n_rays * f32 / ki / PEAK_FP32,
                               "fp64_frac": wl.n_rays * f64 / ki / PEAK_FP64,
                               "note": "the secondary's Newton solve is FP64 (280 flops per "
                                       "ray), weighed against the FP64 vector peak"},
            "response_matrix_rays_per_s": wl.n_rays * steps / wr,
            "response_matrix_ms": 1e3 * wr / steps,
            "image_rays_per_s": wl.n_rays * steps / wi, "image_ms": 1e3 * wi / steps,
            "fwd_bwd_aspheric_rays_per_s": wl.n_rays * steps / wb,
            "fwd_bwd_aspheric_ms": 1e3 * wb / steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed fit iterations")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-parity", action="store_true",
                    help="skip the image comparisons with the FP64 CPU run")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain (per-instance) forward+backward comparison")
    ap.add_argument("--extras", default="C2,C3,C3T,C4",
                    help="comma list of extra lines (C2, C3, C3T, C4) or 'none'")
    ap.add_argument("--c5-instances", type=int, default=64)
    ap.add_argument("--only", default=None,
                    help="run one extra (C2, C3, C3T, C4, or copy: a 1 GiB device copy) "
                         "alone and print its record instead of the headline line (for "
                         "profiling one workload)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the path would shard instances or sources across GPUs with one small "
              "all-reduce; only the one-GPU share is built and measured)", file=sys.stderr)
        sys.exit(2)
    want = set() if args.extras == "none" else {e.strip() for e in args.extras.split(",")
                                                if e.strip()}
    unknown = want - {"C2", "C3", "C3T", "C4"}
    if unknown:
        sys.exit(f"unknown extras: {sorted(unknown)}")

    import iactrace_amd as ia
    from iactrace_amd import configs

    if args.only:
        fns = {"copy": lambda: hbm_copy(args.steps, args.warmup),
               "C2": lambda: c2_forward(ia, configs, args.steps, args.warmup, False),
               "C3": lambda: c3_fwd_bwd(ia, configs, args.steps, args.warmup),
               "C3T": lambda: c3t_culling(ia, configs, args.steps, args.warmup),
               "C4": lambda: c4_response(ia, configs, args.steps, args.warmup)}
        if args.only not in fns:
            sys.exit(f"unknown workload for --only: {args.only}")
        print(json.dumps({"only": args.only, **fns[args.only]()}), flush=True)
        return
    r = c5(ia, configs, args.c5_instances, args.steps, args.warmup, not args.no_parity,
           not args.no_plain)
    bw, tel, rot, targets = r.pop("_state")
    out = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 source directions)",
        "data": "synthetic: generated dish geometry, Philox MC samples (key 0), numpy PCG64 "
                "tilt draws (seeds 100..) and source fields (seeds 200.., sizes seed 6)",
        "config": {"workload": r["workload"], "rays_per_step": r["rays_per_step"],
                   "instances": args.c5_instances, "facets": 875, "samples": 1024,
                   "sources": int(bw.offsets[-1]),
                   "mode": "batched forward + backward (alignment-fit iteration)"},
        "roofline": r["roofline"],
        "forward_kernel": r["forward_kernel"],
        "adjoint_kernel": r["adjoint_kernel"],
        "iteration_fp32_frac": r["iteration_fp32_frac"],
        "loss": r["loss"],
    }
    for k in ("plain_fwd_bwd", "parity"):
        if k in r:
            out[k] = r[k]
    ex = {}
    if "C2" in want:
        ex["C2_forward"] = c2_forward(ia, configs, 10, 2, not args.no_parity)
    if "C3" in want:
        ex["C3_fwd_bwd"] = c3_fwd_bwd(ia, configs, 3, 1)
    if "C3T" in want:
        ex["C3T_culling"] = c3t_culling(ia, configs, 2, 1)
    if "C4" in want:
        ex["C4_response"] = c4_response(ia, configs, 5, 1)
    if ex:
        out["extras"] = ex
    if not args.no_cpu:
        cb = cpu_fwd_bwd(bw, tel, rot, targets)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu_baseline"] = out["value"] / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
