"""ctypes binding of libiactrace_amd.so (declared in include/iactrace_amd.h).

The structures below mirror the C header field by field; ``check_abi()``
compares their sizes with the library's own ``iact_struct_size``.  The
library is mandatory: every render goes through it, and a missing or stale
build raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_char_p, c_double, c_float, c_int32, c_int64,
                    c_size_t, c_uint64, c_void_p)

IACT_ABI_VERSION = 1
IACT_MAX_ASPHERIC = 8
IACT_MAX_POLY_VERTS = 8
IACT_MAX_SURFACES = 16
IACT_MIRROR_STRIDE = 40
IACT_NOALLOC_SCRATCH_BYTES = 32 << 20

IACT_OK = 0
IACT_ERR_INVALID = -1
IACT_ERR_UNSUPPORTED = -2
IACT_ERR_HIP = -3
IACT_ERR_WORKSPACE = -4

SOURCE_POINT = 0
SOURCE_PARALLEL = 1
MODE_IMAGE = 0
MODE_RESPONSE = 1
MODE_DEBUG = 2
LAYOUT_AUTO = 0
LAYOUT_SAMPLES = 1
LAYOUT_SOURCES = 2
SENSOR_SQUARE = 0
SENSOR_HEX = 1
APERTURE_DISK = 0
APERTURE_POLYGON = 1

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libiactrace_amd.so")


class IactSurface(Structure):
    _fields_ = [("curvature", c_float), ("conic", c_float), ("n_aspheric", c_int32),
                ("aspheric", c_float * IACT_MAX_ASPHERIC)]


class IactFacets(Structure):
    _fields_ = [("n_facets", c_int32), ("n_samples", c_int32),
                ("points", c_void_p), ("normals", c_void_p), ("delta", c_void_p),
                ("weights", c_void_p), ("positions", c_void_p), ("rotations", c_void_p),
                ("perturbation_scale", c_void_p), ("surface_index", c_void_p),
                ("offsets", c_void_p), ("area_over_m", c_void_p),
                ("max_sample_radius", c_float)]


class IactObstructions(Structure):
    _fields_ = [("n_cylinders", c_int32), ("cylinders", c_void_p),
                ("n_boxes", c_int32), ("boxes", c_void_p),
                ("n_spheres", c_int32), ("spheres", c_void_p),
                ("n_oriented_boxes", c_int32), ("oriented_boxes", c_void_p),
                ("n_triangles", c_int32), ("triangles", c_void_p)]


class IactSensor(Structure):
    _fields_ = [("kind", c_int32), ("splat", c_int32), ("n_pixels", c_int32),
                ("position", c_float * 3), ("rotation", c_float * 9),
                ("width", c_int32), ("height", c_int32),
                ("x0", c_float), ("y0", c_float), ("dx", c_float), ("dy", c_float),
                ("edge_width", c_float),
                ("hex_size", c_float), ("hex_inradius", c_float), ("grid_rotation", c_float),
                ("grid_offset", c_float * 2), ("edge_threshold", c_float),
                ("q_min", c_int32), ("r_min", c_int32), ("table_q", c_int32),
                ("table_r", c_int32), ("lookup_table", c_void_p),
                ("sigma", c_float), ("kernel_size", c_int32)]


class IactScene(Structure):
    _fields_ = [("facets", IactFacets), ("n_surfaces", c_int32),
                ("surfaces", IactSurface * IACT_MAX_SURFACES),
                ("n_mirrors", c_int32), ("mirrors", c_void_p),
                ("n_stages", c_int32), ("stages", c_int32 * 8),
                ("obstructions", IactObstructions), ("sensor", IactSensor)]


class IactLaunch(Structure):
    _fields_ = [("block_threads", c_int32), ("source_tile", c_int32),
                ("blocks_per_cu", c_int32), ("deterministic", c_int32),
                ("accumulate", c_int32), ("chunk_block", c_int32), ("block_order", c_int32),
                ("layout", c_int32), ("sample_group", c_int32),
                ("culling", c_int32)]


class IactGrads(Structure):
    _fields_ = [("rotations", c_void_p), ("positions", c_void_p), ("sources", c_void_p),
                ("values", c_void_p), ("aspheric", c_void_p)]


class IactSampleDesc(Structure):
    _fields_ = [("seed", c_uint64), ("group", c_int32), ("n_facets", c_int32),
                ("n_samples", c_int32), ("aperture", c_int32), ("n_vertices", c_int32),
                ("radii", c_void_p), ("vertices", c_void_p), ("offsets", c_void_p),
                ("surface", IactSurface)]


class IactBatch(Structure):
    _fields_ = [("n_instances", c_int32), ("rotations", c_void_p),
                ("source_offsets", c_void_p), ("host_source_offsets", c_void_p),
                ("sources", c_void_p), ("values", c_void_p)]


STRUCTS = {0: IactSurface, 1: IactFacets, 2: IactObstructions, 3: IactSensor, 4: IactScene,
           5: IactLaunch, 6: IactGrads, 7: IactSampleDesc, 8: IactBatch}

# name -> (restype, argtypes)
_SIGNATURES = {
    "iact_abi_version": (c_int32, []),
    "iact_last_error": (c_char_p, []),
    "iact_struct_size": (c_int64, [c_int32]),
    "iact_device_info": (c_int32, [c_int32, POINTER(c_int32), POINTER(c_int32), c_char_p,
                                   c_size_t]),
    "iact_sample_facets": (c_int32, [POINTER(IactSampleDesc), c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p]),
    "iact_facet_world": (c_int32, [POINTER(IactFacets), c_void_p, c_void_p]),
    "iact_sensor_accumulate": (c_int32, [POINTER(IactSensor), c_void_p, c_void_p, c_void_p,
                                         c_int64, c_void_p, c_void_p, c_void_p]),
    "iact_workspace_size": (c_size_t, [POINTER(IactScene), c_int32, c_int32,
                                       POINTER(IactLaunch)]),
    "iact_trace_fwd": (c_int32, [POINTER(IactScene), c_void_p, c_void_p, c_int32, c_int32,
                                 c_int32, POINTER(IactLaunch), c_void_p, c_size_t, c_void_p,
                                 c_void_p, c_void_p, c_void_p]),
    "iact_trace_fwd_noalloc": (c_int32, [POINTER(IactScene), c_void_p, c_void_p, c_int32,
                                         c_int32, c_int32, POINTER(IactLaunch), c_void_p,
                                         c_void_p, c_void_p, c_void_p]),
    "iact_bwd_workspace_size": (c_size_t, [POINTER(IactScene), c_int32, POINTER(IactLaunch)]),
    "iact_trace_bwd": (c_int32, [POINTER(IactScene), c_void_p, c_void_p, c_int32, c_int32,
                                 c_void_p, POINTER(IactLaunch), c_void_p, c_size_t,
                                 POINTER(IactGrads), c_void_p]),
    "iact_batch_workspace_size": (c_size_t, [POINTER(IactScene), POINTER(IactBatch),
                                             POINTER(IactLaunch)]),
    "iact_trace_batched_fwd": (c_int32, [POINTER(IactScene), POINTER(IactBatch), c_int32,
                                         POINTER(IactLaunch), c_void_p, c_size_t, c_void_p,
                                         c_void_p]),
    "iact_trace_batched_bwd": (c_int32, [POINTER(IactScene), POINTER(IactBatch), c_int32,
                                         c_void_p, POINTER(IactLaunch), c_void_p, c_size_t,
                                         c_void_p, c_void_p]),
    "iact_l2_loss_grad": (c_int32, [c_void_p, c_void_p, c_int64, c_int32, c_void_p, c_void_p,
                                    c_void_p]),
    "iact_adam_step": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64,
                                 c_float, c_float, c_float, c_float, c_void_p, c_void_p]),
    "iact_fit_graph_create": (c_int32, [POINTER(IactScene), POINTER(IactBatch), c_int32,
                                        POINTER(IactLaunch), c_void_p, c_size_t, c_void_p,
                                        c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_void_p, c_void_p, c_void_p, c_float,
                                        POINTER(c_void_p)]),
    "iact_autotune": (c_int32, [POINTER(IactScene), c_void_p, c_void_p, c_int32, c_int32, c_int32,
                                POINTER(IactLaunch), c_void_p, c_size_t, c_void_p, c_int32,
                                POINTER(IactLaunch), POINTER(c_float), c_void_p]),
    "iact_graph_launch": (c_int32, [c_void_p, c_void_p]),
    "iact_graph_destroy": (c_int32, [c_void_p]),
}

_lib = None


class IactError(RuntimeError):
    """A non-zero status from libiactrace_amd (message from iact_last_error)."""

    def __init__(self, status, where, message):
        super().__init__(f"{where} failed with status {status}: {message}")
        self.status = status


def lib():
    """Load (once) and return the HIP library; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: the HIP kernels are not built. Run "
            "`python -c 'import __graft_entry__ as g; g.build()'` (or `make -C "
            "iactrace_amd/csrc`) first; there is no non-HIP fallback.")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.iact_abi_version() != IACT_ABI_VERSION:
        raise RuntimeError(f"libiactrace_amd ABI {L.iact_abi_version()} != {IACT_ABI_VERSION}")
    _lib = L
    check_abi()
    return L


def check_abi():
    """Compare every ctypes structure size with the compiled library's."""
    L = _lib if _lib is not None else lib()
    for which, cls in STRUCTS.items():
        n = L.iact_struct_size(which)
        if n != ctypes.sizeof(cls):
            raise RuntimeError(f"ABI mismatch for {cls.__name__}: C {n} vs ctypes "
                               f"{ctypes.sizeof(cls)}")


def check(status, where):
    if status != IACT_OK:
        msg = lib().iact_last_error()
        raise IactError(status, where, msg.decode() if msg else "")


def ptr(t):
    """Device (or host) address of a tensor, 0 for None."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def current_stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
