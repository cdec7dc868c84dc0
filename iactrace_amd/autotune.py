"""Launch-shape autotuning (no counterpart in the reference, whose launch
shapes XLA chooses).

``autotune(tel, sources, values, ...)`` asks the library (``iact_autotune``)
to time candidate launch settings -- image layout (lanes = samples of one
facet, or lanes = sources), block size, blocks per CU, source tile or sample
group -- on the given inputs and returns the fastest as a ``LaunchConfig``;
with ``apply`` it becomes the telescope's default (``tel.launch``).  Results
are cached in a JSON file keyed by the GPU architecture and the workload's
shape (facets, samples, source count rounded to a power of two, sensor,
obstructions, stages, source type, mode), so later processes skip the search.
The cache file is ``$IACTRACE_AMD_TUNE_CACHE`` or
``~/.cache/iactrace_amd/launch.json``.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import tempfile

import torch

from . import _lib
from .render import LaunchConfig, RenderPlan

__all__ = ["autotune", "tuning_key", "cache_path"]

_MODES = {"image": _lib.MODE_IMAGE, "response": _lib.MODE_RESPONSE}


def cache_path():
    p = os.environ.get("IACTRACE_AMD_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "iactrace_amd", "launch.json")


def _arch():
    n_cu = ctypes.c_int32()
    lds = ctypes.c_int32()
    buf = ctypes.create_string_buffer(64)
    _lib.check(_lib.lib().iact_device_info(torch.cuda.current_device(), ctypes.byref(n_cu),
                                           ctypes.byref(lds), buf, 64), "iact_device_info")
    return f"{buf.value.decode()}:{n_cu.value}cu"


def tuning_key(tel, n_sources, source_type="point", sensor_idx=0, mode="image", arch=None):
    """Cache key of a workload shape (the source count is rounded up to a
    power of two: nearby counts share a setting)."""
    prim = [g for g in tel.mirror_groups if g.optical_stage == 0]
    s = tel.sensors[sensor_idx]
    obst = {}
    for g in tel.obstruction_groups or []:
        obst[g.kind] = obst.get(g.kind, 0) + len(g)
    key = {
        "arch": arch or _arch(),
        "facets": sum(len(g) for g in prim),
        "samples": prim[0].n_samples if prim else 0,
        "sources": 1 << max(0, math.ceil(math.log2(max(n_sources, 1)))),
        "sensor": type(s).__name__,
        "pixels": s.n_pixels,
        "obstructions": dict(sorted(obst.items())),
        "stages": sorted({g.optical_stage for g in tel.mirror_groups if g.optical_stage > 0}),
        "source_type": source_type,
        "mode": mode,
    }
    return json.dumps(key, sort_keys=True)


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, table):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=os.path.dirname(path), prefix=".launch.")
    with os.fdopen(fd, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _from_abi(L):
    return LaunchConfig(L.block_threads, L.source_tile, L.blocks_per_cu, bool(L.deterministic),
                        L.accumulate, L.chunk_block, L.block_order, L.layout, L.sample_group)


def autotune(tel, sources, values, source_type="point", sensor_idx=0, mode="image", reps=3,
             base=None, cache=True, apply=True, retune=False):
    """Fastest launch settings for rendering ``sources``/``values`` through
    ``tel`` (see the module docstring).  Returns a LaunchConfig."""
    if mode not in _MODES:
        raise ValueError(f"autotune covers modes {sorted(_MODES)}, not {mode!r}")
    plan = RenderPlan(tel, sensor_idx, base)
    dev = plan.device
    src = torch.as_tensor(sources).to(device=dev, dtype=torch.float64).reshape(-1, 3).contiguous()
    val = torch.as_tensor(values).to(device=dev, dtype=torch.float32).reshape(-1).contiguous()
    ns = src.shape[0]
    key = tuning_key(tel, ns, source_type, sensor_idx, mode)
    path = cache_path()
    if cache and not retune:
        hit = _load(path).get(key)
        if hit is not None:
            cfg = LaunchConfig(**hit["launch"])
            if base is not None:
                cfg.deterministic = base.deterministic
            if apply:
                tel.launch = cfg
            return cfg
    keep = [src, val]
    sc = plan.scene(keep)
    L0 = (base or LaunchConfig()).abi()
    lib = _lib.lib()
    # workspace for the largest candidate (the source-major layout's records)
    need = 0
    for lay in (1, 2):
        Lc = (base or LaunchConfig()).abi()
        Lc.layout = lay
        need = max(need, int(lib.iact_workspace_size(ctypes.byref(sc), ns, _MODES[mode],
                                                     ctypes.byref(Lc))))
    ws = torch.empty(max(need, 1), dtype=torch.uint8, device=dev)
    P = plan.sensor.n_pixels
    out = torch.empty((ns, P) if mode == "response" else P, dtype=torch.float32, device=dev)
    best = _lib.IactLaunch()
    ms = ctypes.c_float()
    st = _lib.SOURCE_POINT if source_type == "point" else _lib.SOURCE_PARALLEL
    _lib.check(lib.iact_autotune(ctypes.byref(sc), _lib.ptr(src), _lib.ptr(val), ns, st,
                                 _MODES[mode], ctypes.byref(L0), _lib.ptr(ws), ws.numel(),
                                 _lib.ptr(out), int(reps), ctypes.byref(best), ctypes.byref(ms),
                                 _lib.current_stream()), "iact_autotune")
    cfg = _from_abi(best)
    if cache:
        table = _load(path)
        table[key] = {"launch": vars(cfg), "ms": float(ms.value)}
        _store(path, table)
    if apply:
        tel.launch = cfg
    return cfg
