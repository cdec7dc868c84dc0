"""Renderer entry points (iactrace/core/render.py:174-324) on the HIP kernels.

``render``, ``render_debug`` and ``render_response_matrix`` keep the
reference's signatures.  Each packs the telescope into the C-ABI scene
descriptor (device pointers into the telescope's own tensors, plus cached
device tables for obstructions, stage>=1 mirrors and the hex lookup table)
and launches ``iact_trace_fwd``.  ``render`` is differentiable: a
torch.autograd.Function routes dL/dimage to ``iact_trace_bwd``, which
recomputes every ray and returns gradients for stage-0 rotations and
positions, sources, values and the aspheric coefficients of every surface.
"""
from __future__ import annotations

import ctypes
import math
import weakref
from collections import defaultdict

import numpy as np
import torch

from . import _lib
from .dlpack import as_tensor
from .geometry import euler_to_matrix
from .mirrors import AsphericDiskMirrorGroup, AsphericPolygonMirrorGroup
from .obstructions import pack_obstruction_tables

__all__ = ["render", "render_debug", "render_response_matrix", "RenderPlan", "facets_abi",
           "LaunchConfig"]


class LaunchConfig:
    """Launch shape handed to the kernels (see iactrace_amd.autotune)."""

    def __init__(self, block_threads=0, source_tile=0, blocks_per_cu=0, deterministic=False,
                 accumulate=0, chunk_block=0, block_order=0, layout=0, sample_group=0,
                 culling=0):
        self.block_threads = int(block_threads)
        self.source_tile = int(source_tile)
        self.blocks_per_cu = int(blocks_per_cu)
        self.deterministic = bool(deterministic)
        self.accumulate = int(accumulate)
        self.chunk_block = int(chunk_block)
        self.block_order = int(block_order)
        self.layout = int(layout)
        self.sample_group = int(sample_group)
        self.culling = int(culling)

    def abi(self):
        L = _lib.IactLaunch()
        L.block_threads = self.block_threads
        L.source_tile = self.source_tile
        L.blocks_per_cu = self.blocks_per_cu
        L.deterministic = int(self.deterministic)
        L.accumulate = self.accumulate
        L.chunk_block = self.chunk_block
        L.block_order = self.block_order
        L.layout = self.layout
        L.sample_group = self.sample_group
        L.culling = self.culling
        return L

    def __repr__(self):
        return (f"LaunchConfig(block_threads={self.block_threads}, source_tile="
                f"{self.source_tile}, blocks_per_cu={self.blocks_per_cu}, "
                f"deterministic={self.deterministic}, accumulate={self.accumulate}, "
                f"chunk_block={self.chunk_block}, block_order={self.block_order}, "
                f"layout={self.layout}, sample_group={self.sample_group}, "
                f"culling={self.culling})")


def _stages(groups):
    by = defaultdict(list)
    for g in groups:
        by[g.optical_stage].append(g)
    return dict(sorted(by.items()))


def _f32c(t, device):
    return t.detach().to(device=device, dtype=torch.float32).contiguous()


def _cat(ts, device):
    ts = [_f32c(t, device) for t in ts]
    return ts[0] if len(ts) == 1 else torch.cat(ts, 0).contiguous()


_RADII = {}


def sample_radius(points):
    """max |point| over a group's samples (facet frame), once per version of
    the points tensor (one device read; later calls are free)."""
    key = id(points)
    hit = _RADII.get(key)
    if hit is not None and hit[0]() is points and hit[1] == points._version:
        return hit[2]
    r = float(points.detach().float().norm(dim=-1).max()) if points.numel() else 0.0
    if len(_RADII) > 256:
        for k in [k for k, v in _RADII.items() if v[0]() is None]:
            del _RADII[k]
    _RADII[key] = (weakref.ref(points), points._version, r)
    return r


def facets_abi(groups, keep, rotations=None, positions=None):
    """IactFacets over the concatenation of stage-0 ``groups`` (render.py:201-205)."""
    dev = groups[0].positions.device
    F = _lib.IactFacets()
    F.n_facets = sum(len(g) for g in groups)
    F.n_samples = groups[0].n_samples
    for g in groups:
        if g.n_samples != F.n_samples:
            raise ValueError("stage-0 groups must share n_samples (render.py:203-205)")
    arrs = {
        "points": _cat([g.points for g in groups], dev),
        "normals": _cat([g.normals for g in groups], dev),
        "delta": _cat([g.perturbation_delta for g in groups], dev),
        "weights": _cat([g.weights for g in groups], dev),
        "positions": _f32c(positions, dev) if positions is not None else
        _cat([g.positions for g in groups], dev),
        "rotations": _f32c(rotations, dev) if rotations is not None else
        _cat([g.rotations for g in groups], dev),
        "perturbation_scale": _cat([g.perturbation_scale.reshape(-1) for g in groups], dev),
        "offsets": _cat([g.offsets for g in groups], dev),
    }
    # the far-source test of the kernels bounds every sample's offset from
    # its facet position by this radius
    F.max_sample_radius = max(sample_radius(g.points) for g in groups)
    sidx = torch.cat([torch.full((len(g),), i, dtype=torch.int32) for i, g in enumerate(groups)])
    arrs["surface_index"] = sidx.to(dev)
    for k, v in arrs.items():
        keep.append(v)
        setattr(F, k, v.data_ptr())
    return F


def _mirror_table(stage_groups, surface_base):
    """Rows of the stage>=1 mirror table, ordered by (stage, group, index)."""
    rows = []
    stages = []
    sidx = surface_base
    for stage, groups in stage_groups:
        stages.append(stage)
        for gi, g in enumerate(groups):
            pos = g.positions.detach().float().cpu().numpy()
            rot = g.rotations.detach().float().cpu().numpy()
            off = g.offsets.detach().float().cpu().numpy()
            for i in range(len(g)):
                r = np.zeros(_lib.IACT_MIRROR_STRIDE, np.float32)
                r[0:3] = pos[i]
                r[3:6] = rot[i]
                r[6:15] = euler_to_matrix(rot[i]).reshape(-1)
                r[15:17] = off[i]
                if isinstance(g, AsphericDiskMirrorGroup):
                    r[17] = 0
                    r[18] = float(g.radii[i])
                else:
                    v = g.vertices[i].detach().float().cpu().numpy()
                    if len(v) > _lib.IACT_MAX_POLY_VERTS:
                        raise ValueError(f"polygon with {len(v)} vertices "
                                         f"(max {_lib.IACT_MAX_POLY_VERTS})")
                    r[17] = 1
                    r[19] = len(v)
                    r[20:20 + 2 * len(v)] = v.reshape(-1)
                r[36] = stage
                r[37] = gi
                r[38] = sidx
                rows.append(r)
            sidx += 1
    return (np.stack(rows) if rows else np.zeros((0, _lib.IACT_MIRROR_STRIDE), np.float32)), stages


class RenderPlan:
    """Packed scene for one (telescope, sensor) pair on one device.

    Tables that do not change between calls (obstructions, stage>=1 mirrors,
    hex lookup table) are built once; facet arrays are referenced in place,
    so in-place updates of rotations/positions are seen by the next call.
    """

    def __init__(self, tel, sensor_idx=0, launch=None):
        if not (-len(tel.sensors) <= sensor_idx < len(tel.sensors)):
            raise IndexError(f"sensor index {sensor_idx} out of range ({len(tel.sensors)} "
                             "sensors)")
        self.tel = tel
        self.sensor = tel.sensors[sensor_idx]
        st = _stages(tel.mirror_groups)
        self.primary = st.get(0, [])
        self.secondary = [(k, v) for k, v in st.items() if k > 0]
        dev = None
        for g in tel.mirror_groups:
            dev = g.positions.device
            break
        self.device = dev if dev is not None else torch.device("cuda")
        if self.device.type != "cuda":
            raise RuntimeError("render runs on the HIP device; call telescope.to('cuda')")
        self.launch = launch or LaunchConfig()
        self.keep = []
        self._build_static()

    @property
    def empty(self):
        return not self.primary

    def _build_static(self):
        dev = self.device
        self.surface_groups = list(self.primary) + [g for _, gs in self.secondary for g in gs]
        if len(self.surface_groups) > _lib.IACT_MAX_SURFACES:
            raise ValueError(f"{len(self.surface_groups)} surfaces (max {_lib.IACT_MAX_SURFACES})")
        tab, stages = _mirror_table(self.secondary, len(self.primary))
        self.mirror_table = torch.as_tensor(tab).to(dev).contiguous()
        self.stages = stages
        ob = pack_obstruction_tables(self.tel.obstruction_groups or [])
        self.obstruction_tables = {k: torch.as_tensor(v).to(dev).contiguous() for k, v in ob.items()}
        self.sensor_abi = self.sensor.to_abi(dev)

    def scene(self, keep, rotations=None, positions=None):
        sc = _lib.IactScene()
        if self.primary:
            sc.facets = facets_abi(self.primary, keep, rotations, positions)
        sc.n_surfaces = len(self.surface_groups)
        for i, g in enumerate(self.surface_groups):
            sc.surfaces[i] = g.surface_abi()
        sc.n_mirrors = self.mirror_table.shape[0]
        sc.mirrors = self.mirror_table.data_ptr() if sc.n_mirrors else None
        sc.n_stages = len(self.stages)
        for i, s in enumerate(self.stages):
            sc.stages[i] = int(s)
        ob = self.obstruction_tables
        o = sc.obstructions
        for kind, cnt, field in (("cylinder", "n_cylinders", "cylinders"),
                                 ("box", "n_boxes", "boxes"),
                                 ("sphere", "n_spheres", "spheres"),
                                 ("oriented_box", "n_oriented_boxes", "oriented_boxes"),
                                 ("triangle", "n_triangles", "triangles")):
            t = ob[kind]
            setattr(o, cnt, t.shape[0])
            setattr(o, field, t.data_ptr() if t.shape[0] else None)
        sc.sensor = self.sensor_abi
        return sc

    # ------------------------------------------------------------ forward
    def forward(self, sources, values, source_type, mode, rotations=None, positions=None):
        dev = self.device
        P = self.sensor.n_pixels
        src = as_tensor(sources).to(device=dev, dtype=torch.float64).reshape(-1, 3)
        src = src.contiguous()
        val = as_tensor(values).to(device=dev, dtype=torch.float32).reshape(-1).contiguous()
        ns = src.shape[0]
        if val.shape[0] != ns:
            raise ValueError(f"{ns} sources but {val.shape[0]} values")
        st = _lib.SOURCE_POINT if source_type == "point" else _lib.SOURCE_PARALLEL
        if self.empty:  # render.py:198-199
            if mode == _lib.MODE_IMAGE:
                return torch.zeros(P, device=dev)
            if mode == _lib.MODE_RESPONSE:
                return torch.zeros((ns, P), device=dev)
            return (torch.zeros((0, 2), device=dev), torch.zeros(0, device=dev),
                    torch.zeros(0, dtype=torch.int32, device=dev))
        keep = [src, val]
        sc = self.scene(keep, rotations, positions)
        L = self.launch.abi()
        lib = _lib.lib()
        ws_bytes = lib.iact_workspace_size(ctypes.byref(sc), ns, mode, ctypes.byref(L))
        ws = torch.empty(max(int(ws_bytes), 1), dtype=torch.uint8, device=dev)
        nf, m = sc.facets.n_facets, sc.facets.n_samples
        dbg_xy = dbg_pix = None
        if mode == _lib.MODE_IMAGE:
            out = torch.empty(P, dtype=torch.float32, device=dev)
        elif mode == _lib.MODE_RESPONSE:
            out = torch.empty((ns, P), dtype=torch.float32, device=dev)
        else:
            n = nf * ns * m
            out = torch.empty(n, dtype=torch.float32, device=dev)
            dbg_xy = torch.empty((n, 2), dtype=torch.float32, device=dev)
            dbg_pix = torch.empty(n, dtype=torch.int32, device=dev)
        _lib.check(lib.iact_trace_fwd(ctypes.byref(sc), _lib.ptr(src), _lib.ptr(val), ns, st, mode,
                                      ctypes.byref(L), _lib.ptr(ws), ws.numel(), _lib.ptr(out),
                                      _lib.ptr(dbg_xy), _lib.ptr(dbg_pix), _lib.current_stream()),
                   "iact_trace_fwd")
        if mode == _lib.MODE_DEBUG:
            return dbg_xy, out, dbg_pix
        return out

    # ------------------------------------------------------------ backward
    def backward(self, grad_image, sources, values, source_type, rotations, positions, need):
        """need = (sources, values, rotations, positions, aspheric) booleans."""
        dev = self.device
        src = as_tensor(sources).to(device=dev, dtype=torch.float64).reshape(-1, 3)
        src = src.contiguous()
        val = as_tensor(values).to(device=dev, dtype=torch.float32).reshape(-1).contiguous()
        ns = src.shape[0]
        g_img = grad_image.detach().to(device=dev, dtype=torch.float32).reshape(-1).contiguous()
        keep = [src, val, g_img]
        sc = self.scene(keep, rotations, positions)
        nf = sc.facets.n_facets
        G = _lib.IactGrads()
        outs = {}
        if need[0]:
            outs["sources"] = torch.empty((ns, 3), dtype=torch.float64, device=dev)
        if need[1]:
            outs["values"] = torch.empty(ns, dtype=torch.float64, device=dev)
        if need[2]:
            outs["rotations"] = torch.empty((nf, 3), dtype=torch.float64, device=dev)
        if need[3]:
            outs["positions"] = torch.empty((nf, 3), dtype=torch.float64, device=dev)
        if need[4]:
            outs["aspheric"] = torch.empty((_lib.IACT_MAX_SURFACES, _lib.IACT_MAX_ASPHERIC),
                                           dtype=torch.float64, device=dev)
        for k, v in outs.items():
            setattr(G, k, v.data_ptr())
        if not outs:
            return outs
        st = _lib.SOURCE_POINT if source_type == "point" else _lib.SOURCE_PARALLEL
        L = self.launch.abi()
        lib = _lib.lib()
        ws_bytes = lib.iact_bwd_workspace_size(ctypes.byref(sc), ns, ctypes.byref(L))
        ws = torch.empty(max(int(ws_bytes), 1), dtype=torch.uint8, device=dev)
        _lib.check(lib.iact_trace_bwd(ctypes.byref(sc), _lib.ptr(src), _lib.ptr(val), ns, st,
                                      _lib.ptr(g_img), ctypes.byref(L), _lib.ptr(ws), ws.numel(),
                                      ctypes.byref(G), _lib.current_stream()), "iact_trace_bwd")
        return outs


class _TraceFunction(torch.autograd.Function):
    """Image = render(...); backward through the HIP adjoint kernel."""

    @staticmethod
    def forward(ctx, plan, source_type, sources, values, rotations, positions, *aspheric):
        img = plan.forward(sources, values, source_type, _lib.MODE_IMAGE, rotations, positions)
        ctx.plan = plan
        ctx.source_type = source_type
        ctx.n_asph = len(aspheric)
        ctx.save_for_backward(sources, values, rotations, positions)
        return img

    @staticmethod
    def backward(ctx, grad):
        sources, values, rotations, positions = ctx.saved_tensors
        nig = ctx.needs_input_grad
        need = (nig[2], nig[3], nig[4], nig[5], any(nig[6:]))
        outs = ctx.plan.backward(grad, sources, values, ctx.source_type, rotations, positions,
                                 need)
        def _like(g, ref):
            return g.to(device=ref.device, dtype=ref.dtype).reshape(ref.shape)
        g_src = _like(outs["sources"], sources) if nig[2] else None
        g_val = _like(outs["values"], values) if nig[3] else None
        g_rot = _like(outs["rotations"], rotations) if nig[4] else None
        g_pos = _like(outs["positions"], positions) if nig[5] else None
        g_asph = []
        for i in range(ctx.n_asph):
            if nig[6 + i]:
                a = ctx.plan.surface_groups[i].aspheric
                g = outs["aspheric"][i, :a.numel()].to(device=a.device, dtype=a.dtype)
                g_asph.append(g.reshape(a.shape))
            else:
                g_asph.append(None)
        return (None, None, g_src, g_val, g_rot, g_pos, *g_asph)


def _plan(tel, sensor_idx, launch=None):
    """Cached RenderPlan of (telescope, sensor, launch shape).  Telescopes are
    immutable by convention (operations return new ones), so the static
    tables stay valid for the telescope's lifetime."""
    launch = launch or getattr(tel, "launch", None)
    key = (sensor_idx, repr(launch))
    cache = tel.__dict__.setdefault("_plan_cache", {})
    plan = cache.get(key)
    if plan is None:
        plan = RenderPlan(tel, sensor_idx, launch)
        cache[key] = plan
    return plan


def render(tel, sources, values, source_type="point", sensor_idx=0, launch=None):
    """Render sources through the telescope onto a sensor (render.py:174-220).

    Returns the accumulated image with the sensor's shape ((H, W) square,
    (P,) hexagonal).  Differentiable w.r.t. sources, values, stage-0 facet
    rotations/positions and every surface's aspheric coefficients; raises
    NotImplementedError when a stage >= 1 mirror's position or rotation
    requires grad (the reference differentiates those through
    _reflect_at_stage, render.py:44-79).
    """
    plan = _plan(tel, sensor_idx, launch)
    shape = plan.sensor.get_accumulator_shape()
    sources = as_tensor(sources)
    values = as_tensor(values)
    if torch.is_grad_enabled():
        for stage, groups in plan.secondary:
            for g in groups:
                if g.positions.requires_grad or g.rotations.requires_grad:
                    raise NotImplementedError(
                        f"gradients with respect to the positions or rotations of stage-{stage} "
                        "mirrors are not implemented (the adjoint covers stage-0 poses, "
                        "sources, values and aspheric coefficients); detach them or render "
                        "under torch.no_grad()")
    if plan.empty:
        return torch.zeros(shape, device=plan.device)
    rot = torch.cat([g.rotations for g in plan.primary], 0) if len(plan.primary) > 1 \
        else plan.primary[0].rotations
    pos = torch.cat([g.positions for g in plan.primary], 0) if len(plan.primary) > 1 \
        else plan.primary[0].positions
    asph = [g.aspheric for g in plan.surface_groups]
    needs = torch.is_grad_enabled() and any(
        t.requires_grad for t in [sources, values, rot, pos, *asph])
    if needs:
        img = _TraceFunction.apply(plan, source_type, sources, values, rot, pos, *asph)
    else:
        img = plan.forward(sources, values, source_type, _lib.MODE_IMAGE, rot.detach(),
                           pos.detach())
    return img.reshape(shape)


def render_debug(tel, sources, values, source_type="point", sensor_idx=0, launch=None,
                 return_pixels=False):
    """Raw hits without accumulation (render.py:223-268): (points (R,2), values (R,))
    with R = n_facets * n_sources * n_samples in facet-major order; optionally
    also the hard pixel id of each hit (-1 = dropped)."""
    plan = _plan(tel, sensor_idx, launch)
    with torch.no_grad():
        xy, v, pix = plan.forward(sources, values, source_type, _lib.MODE_DEBUG)
    return (xy, v, pix) if return_pixels else (xy, v)


def render_response_matrix(tel, sources, values, source_type="point", sensor_idx=0,
                           launch=None):
    """Per-source images, shape (n_sources, n_pixels) (render.py:271-324)."""
    plan = _plan(tel, sensor_idx, launch)
    with torch.no_grad():
        return plan.forward(sources, values, source_type, _lib.MODE_RESPONSE)
