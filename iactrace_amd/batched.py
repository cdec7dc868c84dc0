"""Batched renders of many instances of one telescope, and the alignment fit.

The reference renders one telescope per jitted call (render.py:174-220) and
leaves fitting to the user (jax.grad + optax over the parameters picked by
utils/filtering.py:14-112).  Here many instances of one telescope -- same
facet samples, obstructions and sensor, each with its own stage-0 rotations
(e.g. independent tilt draws) and its own, unevenly sized source list -- are
traced by ``iact_trace_batched_fwd`` and differentiated by
``iact_trace_batched_bwd`` (per-instance rotation gradients).  For
primary-only scenes these run the plain source-major kernels instance by
instance on one stream (each instance already fills the GPU); the
sample-major layout and scenes with stage >= 1 mirrors use one persistent
launch over every (instance, facet chunk, source tile) work item.

``AlignmentFit`` runs the loop  image -> L2 loss vs a target -> rotation
gradient -> Adam  entirely on the device and, by default, replays it as one
captured HIP graph per iteration (``iact_fit_graph_create``).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .dlpack import as_tensor
from .render import LaunchConfig, RenderPlan

__all__ = ["pack_sources", "render_batched", "AlignmentFit", "BatchedPlan"]


def pack_sources(sources, values):
    """Concatenate per-instance source lists: returns (sources (N, 3) f64,
    values (N,) f32, offsets (n+1,) int64 on the host)."""
    if len(sources) != len(values):
        raise ValueError(f"{len(sources)} source lists but {len(values)} value lists")
    s = [np.asarray(x, np.float64).reshape(-1, 3) for x in sources]
    v = [np.asarray(x, np.float32).reshape(-1) for x in values]
    for i, (a, b) in enumerate(zip(s, v)):
        if a.shape[0] != b.shape[0]:
            raise ValueError(f"instance {i}: {a.shape[0]} sources but {b.shape[0]} values")
    off = np.zeros(len(s) + 1, np.int64)
    off[1:] = np.cumsum([a.shape[0] for a in s])
    src = np.concatenate(s, 0) if s else np.zeros((0, 3))
    val = np.concatenate(v, 0) if v else np.zeros(0, np.float32)
    return src, val, off


class BatchedPlan:
    """Scene + batch descriptors for one telescope and one packed source set.

    Every buffer the descriptors point to is owned here, so a graph captured
    from them stays valid while the plan lives.
    """

    def __init__(self, tel, sources, values, offsets, source_type="point", sensor_idx=0,
                 launch=None):
        self.plan = RenderPlan(tel, sensor_idx, launch)
        if self.plan.empty:
            raise ValueError("batched renders need stage-0 mirrors")
        dev = self.plan.device
        self.device = dev
        self.source_type = source_type
        self.st = _lib.SOURCE_POINT if source_type == "point" else _lib.SOURCE_PARALLEL
        self.offsets_host = np.ascontiguousarray(np.asarray(offsets, np.int64))
        self.n_inst = len(self.offsets_host) - 1
        if self.n_inst < 1:
            raise ValueError("need at least one instance")
        self.src = as_tensor(sources).to(device=dev, dtype=torch.float64)
        self.src = self.src.reshape(-1, 3).contiguous()
        self.val = as_tensor(values).to(device=dev, dtype=torch.float32)
        self.val = self.val.reshape(-1).contiguous()
        if self.src.shape[0] != self.offsets_host[-1] or self.val.shape[0] != self.src.shape[0]:
            raise ValueError("offsets do not match the packed sources/values")
        self.offsets_dev = torch.as_tensor(self.offsets_host).to(dev)
        self.keep = []
        self.scene = self.plan.scene(self.keep)
        self.nf = self.scene.facets.n_facets
        self.n_pixels = self.plan.sensor.n_pixels
        self.L = self.plan.launch.abi()
        self._host_off = self.offsets_host.ctypes.data_as(ctypes.c_void_p)

    def batch_abi(self, rotations):
        rot = rotations.detach()
        if rot.shape != (self.n_inst, self.nf, 3) or rot.dtype != torch.float32 or \
                not rot.is_contiguous() or rot.device != self.device:
            raise ValueError(f"rotations must be a contiguous float32 ({self.n_inst}, {self.nf}, 3)"
                             f" tensor on {self.device}")
        B = _lib.IactBatch()
        B.n_instances = self.n_inst
        B.rotations = rot.data_ptr()
        B.source_offsets = self.offsets_dev.data_ptr()
        B.host_source_offsets = self._host_off
        B.sources = self.src.data_ptr() if self.src.numel() else None
        B.values = self.val.data_ptr() if self.val.numel() else None
        return B

    def workspace(self, B):
        n = _lib.lib().iact_batch_workspace_size(ctypes.byref(self.scene), ctypes.byref(B),
                                                 ctypes.byref(self.L))
        return torch.empty(max(int(n), 1), dtype=torch.uint8, device=self.device)

    def forward(self, rotations):
        B = self.batch_abi(rotations)
        ws = self.workspace(B)
        out = torch.empty((self.n_inst, self.n_pixels), dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().iact_trace_batched_fwd(
            ctypes.byref(self.scene), ctypes.byref(B), self.st, ctypes.byref(self.L),
            _lib.ptr(ws), ws.numel(), _lib.ptr(out), _lib.current_stream()),
            "iact_trace_batched_fwd")
        return out

    def backward(self, rotations, grad_images):
        B = self.batch_abi(rotations)
        ws = self.workspace(B)
        g = grad_images.detach().to(device=self.device, dtype=torch.float32).contiguous()
        out = torch.empty((self.n_inst, self.nf, 3), dtype=torch.float64, device=self.device)
        _lib.check(_lib.lib().iact_trace_batched_bwd(
            ctypes.byref(self.scene), ctypes.byref(B), self.st, _lib.ptr(g), ctypes.byref(self.L),
            _lib.ptr(ws), ws.numel(), _lib.ptr(out), _lib.current_stream()),
            "iact_trace_batched_bwd")
        return out


class _BatchedTrace(torch.autograd.Function):
    @staticmethod
    def forward(ctx, bplan, rotations):
        ctx.bplan = bplan
        ctx.save_for_backward(rotations)
        return bplan.forward(rotations.detach().float().contiguous())

    @staticmethod
    def backward(ctx, grad):
        (rotations,) = ctx.saved_tensors
        g = ctx.bplan.backward(rotations.detach().float().contiguous(), grad)
        return None, g.to(dtype=rotations.dtype)


def render_batched(tel, rotations, sources, values, offsets, source_type="point", sensor_idx=0,
                   launch=None):
    """Images (n_instances, n_pixels) of ``tel`` with per-instance stage-0
    rotations (n_instances, n_facets, 3) and per-instance sources
    sources[offsets[i]:offsets[i+1]].  Differentiable w.r.t. ``rotations``."""
    bplan = BatchedPlan(tel, sources, values, offsets, source_type, sensor_idx, launch)
    rot = as_tensor(rotations)
    if torch.is_grad_enabled() and rot.requires_grad:
        return _BatchedTrace.apply(bplan, rot)
    return bplan.forward(rot.detach().to(device=bplan.device, dtype=torch.float32).contiguous())


class AlignmentFit:
    """Fit per-instance stage-0 rotations to target images with Adam.

    One iteration = batched render, loss sum((image - target)^2) per
    instance, batched adjoint, Adam (torch.optim.Adam semantics, betas
    (0.9, 0.999), eps 1e-8) on all three Euler angles of every facet of every
    instance.  ``use_graph`` replays the iteration as a captured HIP graph;
    otherwise the same entry points are called one by one.
    """

    def __init__(self, tel, rotations, sources, values, offsets, targets, lr=1e-3,
                 source_type="point", sensor_idx=0, launch=None, use_graph=True):
        self.bp = BatchedPlan(tel, sources, values, offsets, source_type, sensor_idx, launch)
        dev = self.bp.device
        n, nf, P = self.bp.n_inst, self.bp.nf, self.bp.n_pixels
        self.rotations = as_tensor(rotations).to(device=dev, dtype=torch.float32)
        self.rotations = self.rotations.reshape(n, nf, 3).clone().contiguous()
        self.targets = as_tensor(targets).to(device=dev, dtype=torch.float32)
        self.targets = self.targets.reshape(n, P).contiguous()
        self.lr = float(lr)
        self.images = torch.zeros((n, P), dtype=torch.float32, device=dev)
        self.grad_images = torch.zeros_like(self.images)
        self.loss = torch.zeros(n, dtype=torch.float64, device=dev)
        self.grad_rotations = torch.zeros((n, nf, 3), dtype=torch.float64, device=dev)
        self.exp_avg = torch.zeros_like(self.rotations)
        self.exp_avg_sq = torch.zeros_like(self.rotations)
        self.step_count = torch.zeros(1, dtype=torch.int32, device=dev)
        self.B = self.bp.batch_abi(self.rotations)
        self.ws = self.bp.workspace(self.B)
        self.graph = None
        if use_graph:
            h = ctypes.c_void_p()
            # the graph's kernels run on its own stream; order it after
            # whatever the current stream has queued for these buffers
            torch.cuda.current_stream().synchronize()
            _lib.check(_lib.lib().iact_fit_graph_create(
                ctypes.byref(self.bp.scene), ctypes.byref(self.B), self.bp.st,
                ctypes.byref(self.bp.L), _lib.ptr(self.ws), self.ws.numel(),
                _lib.ptr(self.images), _lib.ptr(self.targets), _lib.ptr(self.grad_images),
                _lib.ptr(self.loss), _lib.ptr(self.grad_rotations), _lib.ptr(self.rotations),
                _lib.ptr(self.exp_avg), _lib.ptr(self.exp_avg_sq), _lib.ptr(self.step_count),
                ctypes.c_float(self.lr), ctypes.byref(h)), "iact_fit_graph_create")
            self.graph = h

    def step(self):
        """One iteration on the current stream; returns the per-instance loss
        of the images rendered BEFORE the update (a device tensor)."""
        lib = _lib.lib()
        st = _lib.current_stream()
        if self.graph is not None:
            _lib.check(lib.iact_graph_launch(self.graph, st), "iact_graph_launch")
            return self.loss
        bp, n = self.bp, self.bp.n_inst
        _lib.check(lib.iact_trace_batched_fwd(
            ctypes.byref(bp.scene), ctypes.byref(self.B), bp.st, ctypes.byref(bp.L),
            _lib.ptr(self.ws), self.ws.numel(), _lib.ptr(self.images), st),
            "iact_trace_batched_fwd")
        _lib.check(lib.iact_l2_loss_grad(_lib.ptr(self.images), _lib.ptr(self.targets),
                                         bp.n_pixels, n, _lib.ptr(self.grad_images),
                                         _lib.ptr(self.loss), st), "iact_l2_loss_grad")
        _lib.check(lib.iact_trace_batched_bwd(
            ctypes.byref(bp.scene), ctypes.byref(self.B), bp.st, _lib.ptr(self.grad_images),
            ctypes.byref(bp.L), _lib.ptr(self.ws), self.ws.numel(),
            _lib.ptr(self.grad_rotations), st), "iact_trace_batched_bwd")
        _lib.check(lib.iact_adam_step(_lib.ptr(self.rotations), _lib.ptr(self.grad_rotations),
                                      _lib.ptr(self.exp_avg), _lib.ptr(self.exp_avg_sq),
                                      _lib.ptr(self.step_count), self.rotations.numel(),
                                      ctypes.c_float(self.lr), ctypes.c_float(0.9),
                                      ctypes.c_float(0.999), ctypes.c_float(1e-8), None, st),
                   "iact_adam_step")
        return self.loss

    def close(self):
        if self.graph is not None:
            _lib.lib().iact_graph_destroy(self.graph)
            self.graph = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
