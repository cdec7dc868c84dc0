"""Benchmark of the MI355X ray-tracing path (one JSON line on stdout).

Headline: configuration C2 -- Davies-Cotton dish (382 hexagonal facets,
f = 15 m) x 4096 MC samples per facet x 4096 point sources at 1e10 m, square
32x30 camera -- plain forward render, rays/s.  A step is one render of the
whole workload; sources, values and samples are resident in HBM before the
timed region starts.

Also reported: the roofline of the trace (algorithmic flops per step over the
step's HIP-event duration, against MI355X's fp32 vector peak), HBM traffic
from the committed rocprofv3 counter summary if present, the C3
forward+backward and C5 batched lines, and a CPU baseline: the oracle's
C/OpenMP port timed on the host cores on a bounded source subsample.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
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


def time_steps(fn, steps, warmup):
    """Wall time of `steps` calls and the HIP-event time of each call."""
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
    ev_ms = [a.elapsed_time(b) for a, b in evs]
    return wall, ev_ms


def traffic_from_profiles(name):
    """HBM bytes per step from the committed counter summary (FETCH_SIZE +
    WRITE_SIZE of the step's kernels, scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", f"k_trace_sm_{name}_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("hbm_bytes_per_launch"):
                return float(line.split("=")[1])
    return None


def c3_fwd_bwd(ia, configs, steps, warmup):
    """C3 (875 facets x 1024 samples x 16384 point sources, hex camera, box +
    4 cylinders): forward rays/s and forward+backward rays/s, the backward
    being the gradient of a weighted image sum wrt all facet tilts."""
    import numpy as np
    import torch
    dev = torch.device("cuda:0")
    wl = configs.workload("C3")
    tel = wl.telescope(dev, key=0)
    src = torch.as_tensor(wl.sources, dtype=torch.float64, device=dev)
    vals = torch.as_tensor(wl.values, dtype=torch.float32, device=dev)
    W = torch.linspace(0.5, 1.5, tel.sensors[0].n_pixels, device=dev)
    g = tel.mirror_groups[0]
    rot = g.rotations.detach().clone().requires_grad_(True)
    tel_g = ia.Telescope([g.replace(rotations=rot)] + list(tel.mirror_groups[1:]),
                         tel.obstruction_groups, tel.sensors)

    def fwd():
        return ia.render(tel, src, vals, wl.source_type)

    def fwd_bwd():
        rot.grad = None
        loss = (ia.render(tel_g, src, vals, wl.source_type) * W).sum()
        loss.backward()

    wf, _ = time_steps(fwd, steps, warmup)
    wb, _ = time_steps(fwd_bwd, steps, warmup)
    return {"workload": f"C3: parabolic f=36 m, {wl.n_facets} facets x {wl.n_samples} samples "
                        f"x {len(wl.sources)} point sources, hex 2048 px, box + 4 cylinders, "
                        "tilt errors 0.5 mrad",
            "rays_per_step": wl.n_rays, "steps": steps,
            "fwd_rays_per_s": wl.n_rays * steps / wf, "fwd_ms": 1e3 * wf / steps,
            "fwd_bwd_rays_per_s": wl.n_rays * steps / wb, "fwd_bwd_ms": 1e3 * wb / steps,
            "bwd_grad": "d(sum(W*image))/d(facet tilts)"}


def c3t_culling(ia, configs, steps, warmup):
    """C3 with a CT5-sized support truss (271 obstructions): forward rays/s
    with obstruction culling (the default) and without."""
    import torch
    dev = torch.device("cuda:0")
    wl = configs.workload("C3T")
    tel = wl.telescope(dev, key=0)
    src = torch.as_tensor(wl.sources, dtype=torch.float64, device=dev)
    vals = torch.as_tensor(wl.values, dtype=torch.float32, device=dev)
    out = {"workload": f"C3T: C3's dish and camera, {wl.n_facets} facets x {wl.n_samples} "
                       f"samples x {len(wl.sources)} point sources, "
                       f"{len(wl.obstructions)} obstructions (masts, bracing, camera mount)",
           "rays_per_step": wl.n_rays, "steps": steps}
    for name, cull in (("culled", 1), ("unculled", -1)):
        L = ia.LaunchConfig(culling=cull)
        w, _ = time_steps(lambda: ia.render(tel, src, vals, wl.source_type, launch=L), steps,
                          warmup)
        out[f"{name}_rays_per_s"] = wl.n_rays * steps / w
        out[f"{name}_ms"] = 1e3 * w / steps
    return out


def c4_response(ia, configs, steps, warmup):
    """C4: two-mirror Cassegrain (aspheric primary, K = -1.05, four even
    terms; hyperbolic secondary), square 64x64 sensor: the 8192 x 4096
    response matrix of a 128 x 64 grid of parallel sources 0.01 deg apart
    (2048 samples), and forward+backward of a weighted image wrt both
    mirrors' aspheric coefficients (the adjoint through the secondary)."""
    import torch
    dev = torch.device("cuda:0")
    wl = configs.workload("C4")
    tel = wl.telescope(dev, key=0)
    src = torch.as_tensor(wl.sources, dtype=torch.float64, device=dev)
    vals = torch.as_tensor(wl.values, dtype=torch.float32, device=dev)
    g0, g1 = tel.mirror_groups
    a0 = g0.aspheric.detach().clone().requires_grad_(True)
    a1 = torch.zeros(2).requires_grad_(True)  # secondary: two even terms, 0 here (host: launch constants)
    t2 = ia.Telescope([g0.replace(aspheric=a0), g1.replace(aspheric=a1)], [], tel.sensors)
    W = torch.linspace(0.5, 1.5, tel.sensors[0].n_pixels, device=dev)

    def resp():
        return ia.render_response_matrix(tel, src, vals, wl.source_type)

    def fwd_bwd():
        a0.grad = None
        a1.grad = None
        (ia.render(t2, src, vals, wl.source_type).reshape(-1) * W).sum().backward()

    wr, _ = time_steps(resp, steps, warmup)
    wb, _ = time_steps(fwd_bwd, steps, warmup)
    return {"workload": f"C4: Cassegrain, {wl.n_facets} primary x {wl.n_samples} samples x "
                        f"{len(wl.sources)} parallel sources (128 x 64 grid, 0.01 deg), "
                        "square 64 x 64 sensor",
            "rays_per_step": wl.n_rays, "steps": steps,
            "response_matrix_rays_per_s": wl.n_rays * steps / wr,
            "response_matrix_ms": 1e3 * wr / steps,
            "fwd_bwd_aspheric_rays_per_s": wl.n_rays * steps / wb,
            "fwd_bwd_aspheric_ms": 1e3 * wb / steps}


def cpu_baseline(wl, tel, target_s=15.0):
    """Time the C/OpenMP port on a source subsample of the same workload."""
    import numpy as np
    from oracle import cpu_ref
    from tests.helpers import oracle_scene
    sc = oracle_scene(tel)
    keep = []
    s = cpu_ref.scene_struct(sc, keep)
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    P = sc.sensor.n_pixels
    rays_per_src = sc.points.shape[0] * sc.points.shape[1]
    probe = min(16, len(wl.sources))
    t0 = time.perf_counter()
    cpu_ref.render(s, wl.sources[:probe], wl.values[:probe], wl.source_type, P, nthreads)
    dt = time.perf_counter() - t0
    rate = probe * rays_per_src / max(dt, 1e-9)
    n = int(max(probe, min(len(wl.sources), target_s * rate / rays_per_src)))
    t0 = time.perf_counter()
    img = cpu_ref.render(s, wl.sources[:n], wl.values[:n], wl.source_type, P, nthreads)
    dt = time.perf_counter() - t0
    return {"value": n * rays_per_src / dt, "unit": "rays/s", "cores": nthreads,
            "kind": "port",
            "sample": f"{wl.name}: first {n} of {len(wl.sources)} sources x all "
                      f"{sc.points.shape[0]} facets x {sc.points.shape[1]} samples "
                      f"({n * rays_per_src:.3g} rays, {dt:.2f} s), oracle/cpu_ref.cc fp32 "
                      f"with fp64 directions, gcc -O3 -march=x86-64-v3 -fopenmp"}, img, n


def image_parity(ia, wl, tel, src, vals):
    """The GPU image of the whole workload against oracle/cpu_ref's FP64 run
    of the same samples (every step in double, image summed in double), on
    the host cores: relative L1 error, the largest relative error of a
    significant pixel (>= 1e-3 of the peak) and the same after taking away
    the pixel's edge-band mass (|value| of the FP64 rays within the edge band
    of a binning boundary -- 1e-3 px or 2.5e-5 m, tests.helpers.band_units --
    at their pixel and the one across the nearest edge: the rays an fp32
    coordinate may put on either side)."""
    import numpy as np
    from oracle import cpu_ref
    from tests.helpers import band_units, oracle_scene
    sc = oracle_scene(tel)
    band = band_units(sc.sensor)
    keep = []
    s = cpu_ref.scene_struct(sc, keep)
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    t0 = time.perf_counter()
    i64, b64, st = cpu_ref.render_band(s, wl.sources, wl.values, wl.source_type,
                                       sc.sensor.n_pixels, nthreads, 64, band)
    dt = time.perf_counter() - t0
    g = ia.render(tel, src, vals, wl.source_type).detach().double().cpu().numpy().reshape(-1)
    sig = i64 > 1e-3 * i64.max()
    d = np.abs(g - i64)
    return {"reference": f"oracle/cpu_ref.cc FP64 run of the same samples, all {wl.n_rays:.4g} "
                         f"rays ({dt:.1f} s on {nthreads} host threads)",
            "l1_rel": float(d.sum() / np.abs(i64).sum()),
            "significant_px": int(sig.sum()),
            "max_rel_significant_px": float((d[sig] / i64[sig]).max()),
            "max_rel_significant_px_band_resolved":
                float((np.maximum(d - b64, 0.0)[sig] / i64[sig]).max()),
            "band_px": band, "band_rays": st["band"], "traced_rays": st["traced"],
            "band_fraction": st["band"] / max(st["traced"], 1)}


def c5_batched(ia, configs, n_instances):
    """C5: instances of the C3 dish with their own tilt draws and uneven
    source lists (U{256..16384}).  Times (a) the batched forward, (b) one
    alignment-fit iteration -- batched forward, L2 loss vs the untilted
    render, batched adjoint, Adam -- replayed as a HIP graph, and (c) the
    same forward+backward done plainly, one render + autograd per instance."""
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
    wf, _ = time_steps(lambda: bp.forward(rot), 1, 0)
    fit = AlignmentFit(tel, rot, bw.sources, bw.values, bw.offsets, targets, lr=1e-4)
    wg, _ = time_steps(fit.step, 1, 1)
    loss = float(fit.loss.sum())
    fit.close()
    g = tel.mirror_groups[0]
    src = torch.as_tensor(bw.sources, device=dev)
    val = torch.as_tensor(bw.values, device=dev)
    tgt = targets

    def plain():
        for i in range(n_instances):
            a, b = int(bw.offsets[i]), int(bw.offsets[i + 1])
            r = rot[i].clone().requires_grad_(True)
            t = ia.Telescope([g.replace(rotations=r)], tel.obstruction_groups, tel.sensors)
            img = ia.render(t, src[a:b], val[a:b], "point").reshape(-1)
            ((img - tgt[i]) ** 2).sum().backward()

    wp, _ = time_steps(plain, 1, 0)
    n = bw.n_rays
    # VALU roofline of the fit iteration, lower bound: the algorithmic flops
    # of a ray without any obstruction test (culling skips most of them, and
    # the tests done are not credited), the adjoint counted as 3x the forward
    # (recompute + 2x, the convention of the kernel table)
    from iactrace_amd.flop_model import split_per_ray
    f32, f64 = split_per_ray("point", 0, 0, 0, 0, 0, "hex", int(bw.offsets[-1]) // n_instances)
    fb_flops = 4.0 * (f32 + f64)
    achieved = n * fb_flops / wg / 1e12
    roof = {"flops_per_ray_fwd_without_obstructions": f32 + f64, "fwd_plus_bwd_factor": 4.0,
            "achieved_tflops": achieved, "peak_tflops": PEAK_FP32 / 1e12,
            "frac_lower_bound": achieved * 1e12 / PEAK_FP32}
    return {"workload": f"C5: {n_instances} x C3 dish (875 facets x 1024 samples, hex 2048 px, "
                        f"box + 4 cylinders), independent 0.5 mrad tilt draws, "
                        f"{int(bw.offsets[-1])} point sources in total (U{{256..16384}} per "
                        "instance)",
            "rays_per_iteration": n,
            "batched_fwd_rays_per_s": n / wf, "batched_fwd_s": wf,
            "fit_iteration_rays_per_s": n / wg, "fit_iteration_s": wg,
            "fit_iteration": "HIP graph: batched fwd + L2 loss grad + batched adjoint + Adam",
            "plain_fwd_bwd_rays_per_s": n / wp, "plain_fwd_bwd_s": wp,
            "plain": f"{n_instances} x (render + autograd backward wrt tilts)",
            "batched_vs_plain_fwd_bwd": wp / wg, "loss_at_start": loss,
            "fit_iteration_valu_roofline": roof}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="C2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-parity", action="store_true",
                    help="skip the image comparison with the FP64 CPU run")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the C3 forward+backward and C5 batched lines")
    ap.add_argument("--extras", default="C3,C3T,C4,C5",
                    help="comma list of the extra lines to run (C3 forward+backward, C3T "
                         "culling, C4 response matrix, C5 batched)")
    ap.add_argument("--c5-instances", type=int, default=64)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device copy (for counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the path would shard sources across GPUs with one image all-reduce)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import iactrace_amd as ia
    from iactrace_amd import configs
    from iactrace_amd.flop_model import split_per_ray

    dev = torch.device("cuda:0")
    wl = configs.workload(args.config)
    tel = wl.telescope(dev, key=0)
    src = torch.as_tensor(wl.sources, dtype=torch.float64, device=dev)
    vals = torch.as_tensor(wl.values, dtype=torch.float32, device=dev)
    _sync()

    def step():
        return ia.render(tel, src, vals, wl.source_type)

    wall, ev_ms = time_steps(step, args.steps, args.warmup)
    n_rays = wl.n_rays
    ms_per_step = 1e3 * wall / args.steps
    value = n_rays * args.steps / wall
    kern_ms = float(np.mean(ev_ms))
    ob = wl.obstructions
    n_cyl = sum(isinstance(o, ia.Cylinder) for o in ob)
    n_box = sum(isinstance(o, ia.Box) for o in ob)
    n_sph = sum(isinstance(o, ia.Sphere) for o in ob)
    f32, f64 = split_per_ray(wl.source_type, n_cyl, n_box, n_sph,
                             sum(m.optical_stage > 0 for m in wl.mirrors), 0,
                             "square" if isinstance(wl.sensor, ia.SquareSensor) else "hex",
                             len(wl.sources))
    flops = n_rays * (f32 + f64)
    achieved = flops / (kern_ms * 1e-3)
    traffic = traffic_from_profiles(args.config)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 source directions)",
        "data": "synthetic: generated dish geometry, Philox MC samples (key 0), "
                "numpy PCG64 source field (seed 2)",
        "config": {"workload": f"{wl.name} plain forward: Davies-Cotton f=15 m, "
                               f"{wl.n_facets} hex facets x {wl.n_samples} samples, "
                               f"{len(wl.sources)} point sources at 1e10 m (2.5 deg field, "
                               f"lognormal values), square 32x30 camera",
                   "rays_per_step": n_rays, "facets": wl.n_facets, "samples": wl.n_samples,
                   "sources": len(wl.sources), "mode": "image, plain (not batched), forward"},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
                     "traffic": traffic,
                     "kernel": "k_trace_sm, timed per step with HIP events around the "
                               "render call (k_fill_f32 + k_sample_records + k_trace_sm; "
                               "the rocprofv3 summary gives each)",
                     "kernel_ms": kern_ms,
                     "flops_per_ray": {"fp32": f32, "fp64": f64},
                     "fp64_frac_of_fp64_peak": n_rays * f64 / (kern_ms * 1e-3) / PEAK_FP64},
    }
    if args.copy_calibration:
        a = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        _sync()
        b.copy_(a)
        _sync()
        out["copy_calibration_bytes"] = {"fetch": a.numel() * 4, "write": b.numel() * 4}
    if not args.no_extras and args.config == "C2":
        want = {e.strip() for e in args.extras.split(",") if e.strip()}
        unknown = want - {"C3", "C3T", "C4", "C5"}
        if unknown:
            sys.exit(f"unknown extras: {sorted(unknown)}")
        ex = {}
        if "C3" in want:
            ex["C3_fwd_bwd"] = c3_fwd_bwd(ia, configs, 3, 1)
        if "C3T" in want:
            ex["C3T_culling"] = c3t_culling(ia, configs, 2, 1)
        if "C4" in want:
            ex["C4_response"] = c4_response(ia, configs, 5, 1)
        if "C5" in want:
            ex["C5_batched"] = c5_batched(ia, configs, args.c5_instances)
        out["extras"] = ex
    if not args.no_parity:
        out["parity"] = image_parity(ia, wl, tel, src, vals)
    if not args.no_cpu:
        cb, cimg, n = cpu_baseline(wl, tel)
        out["cpu_baseline"] = cb
        # sanity: the same subsample on the GPU agrees with the CPU port
        g = ia.render(tel, src[:n], vals[:n], wl.source_type).detach().cpu().numpy().reshape(-1)
        out["cpu_baseline"]["gpu_vs_cpu_image_l1_rel"] = float(
            np.abs(g - cimg).sum() / max(np.abs(cimg).sum(), 1e-30))
        out["speedup_vs_cpu_baseline"] = value / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
