"""ctypes loader of oracle/libcpu_ref.so (test infrastructure: the CPU baseline,
the second restatement and the FP64 run of full-size images)."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libcpu_ref.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing; run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.cpu_ref_render.restype = ctypes.c_int
        L.cpu_ref_render.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                     ctypes.c_void_p]
        L.cpu_ref_grad.restype = ctypes.c_int
        L.cpu_ref_grad.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_int]
        L.cpu_ref_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def render(scene_struct, sources, values, source_type, n_pixels, nthreads=0, precision=32):
    """Image (n_pixels,) float64 of a scene descriptor whose pointers are HOST
    arrays.  precision 32: the reference's fp32 arithmetic (FP64 source
    directions and Newton); 64: FP64 throughout.  Accumulated in double."""
    return render_band(scene_struct, sources, values, source_type, n_pixels, nthreads,
                       precision, 0.0)[0]


def render_band(scene_struct, sources, values, source_type, n_pixels, nthreads=0,
                precision=64, band=1e-3):
    """(image, band_image, stats): band_image[p] sums |value| of the rays whose
    sensor-grid coordinate lies within `band` (pixels; hex-norm units) of a
    binning boundary, at their own pixel and at the pixel across the nearest
    cell edge; stats = {binned, band, traced} ray counts."""
    src = np.ascontiguousarray(sources, np.float64)
    val = np.ascontiguousarray(values, np.float32)
    out = np.zeros(n_pixels, np.float64)
    bimg = np.zeros(n_pixels, np.float64)
    st3 = np.zeros(3, np.int64)
    st = 0 if source_type == "point" else 1
    rc = lib().cpu_ref_render(ctypes.byref(scene_struct), src.ctypes.data, val.ctypes.data,
                              len(val), st, out.ctypes.data, int(nthreads), int(precision),
                              float(band), bimg.ctypes.data, st3.ctypes.data)
    if rc != 0:
        raise RuntimeError("cpu_ref_render failed")
    return out, bimg, {"binned": int(st3[0]), "band": int(st3[1]), "traced": int(st3[2])}


def grad(scene_struct, sources, values, source_type, grad_image, n_facets, nthreads=0):
    """(dL/drotations (n_facets, 3), dL/dvalues (n_sources,)) of
    L = sum(grad_image * image), hard binning, primary-only (fp32 arithmetic)."""
    src = np.ascontiguousarray(sources, np.float64)
    val = np.ascontiguousarray(values, np.float32)
    G = np.ascontiguousarray(grad_image, np.float64).reshape(-1)
    g_rot = np.zeros((n_facets, 3), np.float64)
    g_val = np.zeros(len(val), np.float64)
    st = 0 if source_type == "point" else 1
    rc = lib().cpu_ref_grad(ctypes.byref(scene_struct), src.ctypes.data, val.ctypes.data,
                            len(val), st, G.ctypes.data, g_rot.ctypes.data, g_val.ctypes.data,
                            int(nthreads))
    if rc != 0:
        raise RuntimeError(f"cpu_ref_grad failed ({rc})")
    return g_rot, g_val


def max_threads():
    return lib().cpu_ref_max_threads()


def _arr(a, dt, keep):
    a = np.ascontiguousarray(a, dt)
    keep.append(a)
    return a.ctypes.data


def scene_struct(sc, keep):
    """Host-pointer iact_scene of an oracle Scene (obstruction rows and mirror
    rows packed in the layout include/iactrace_amd.h documents)."""
    from iactrace_amd import _lib as L
    from oracle import iact_oracle as O
    s = L.IactScene()
    F = s.facets
    nf, M = sc.points.shape[:2]
    F.n_facets, F.n_samples = nf, M
    F.points = _arr(sc.points, np.float32, keep)
    F.normals = _arr(sc.normals, np.float32, keep)
    F.delta = _arr(sc.delta, np.float32, keep)
    F.weights = _arr(sc.weights, np.float32, keep)
    F.positions = _arr(sc.positions, np.float32, keep)
    F.rotations = _arr(sc.rotations, np.float32, keep)
    F.perturbation_scale = _arr(sc.scale, np.float32, keep)
    ob = sc.obstructions
    o = s.obstructions
    f32 = np.float32
    if ob.cylinders:
        rows = []
        for p1, p2, r in ob.cylinders:
            q1, ax, h, rr = O.cylinder_consts(p1, p2, r, f32)
            rows.append(np.concatenate([q1, ax, [h, rr]]))
        o.n_cylinders, o.cylinders = len(rows), _arr(np.stack(rows), f32, keep)
    if ob.boxes:
        rows = [np.concatenate([np.minimum(np.asarray(a, f32), np.asarray(b, f32)),
                                np.maximum(np.asarray(a, f32), np.asarray(b, f32))])
                for a, b in ob.boxes]
        o.n_boxes, o.boxes = len(rows), _arr(np.stack(rows), f32, keep)
    if ob.spheres:
        rows = [np.concatenate([np.asarray(c, f32), [r]]) for c, r in ob.spheres]
        o.n_spheres, o.spheres = len(rows), _arr(np.stack(rows), f32, keep)
    if ob.oriented_boxes:
        rows = [np.concatenate([np.asarray(c, f32), np.asarray(h, f32),
                                np.asarray(R, f32).reshape(-1)]) for c, h, R in ob.oriented_boxes]
        o.n_oriented_boxes, o.oriented_boxes = len(rows), _arr(np.stack(rows), f32, keep)
    if ob.triangles:
        rows = [np.concatenate([np.asarray(a, f32), np.asarray(b, f32), np.asarray(c, f32)])
                for a, b, c in ob.triangles]
        o.n_triangles, o.triangles = len(rows), _arr(np.stack(rows), f32, keep)
    # stage >= 1 mirrors, surfaces
    surfs = []
    rows = []
    for m in sorted(sc.mirrors, key=lambda m: (m.stage, m.group)):
        key = (m.surface.curvature, m.surface.conic, tuple(np.asarray(m.surface.aspheric)))
        if key not in surfs:
            surfs.append(key)
        r = np.zeros(L.IACT_MIRROR_STRIDE, f32)
        r[0:3] = m.position
        r[3:6] = m.rotation
        r[6:15] = O.euler_to_matrix(np.asarray(m.rotation), "ref32").reshape(-1)
        r[15:17] = m.offset
        if m.vertices is None:
            r[17], r[18] = 0, m.radius
        else:
            v = np.asarray(m.vertices, f32)
            r[17], r[19] = 1, len(v)
            r[20:20 + 2 * len(v)] = v.reshape(-1)
        r[36], r[37], r[38] = m.stage, m.group, surfs.index(key)
        rows.append(r)
    s.n_surfaces = len(surfs)
    for i, (c, k, a) in enumerate(surfs):
        s.surfaces[i].curvature, s.surfaces[i].conic = c, k
        s.surfaces[i].n_aspheric = len(a)
        for j, v in enumerate(a):
            s.surfaces[i].aspheric[j] = v
    if rows:
        s.n_mirrors, s.mirrors = len(rows), _arr(np.stack(rows), f32, keep)
        st = sorted({m.stage for m in sc.mirrors})
        s.n_stages = len(st)
        for i, v in enumerate(st):
            s.stages[i] = v
    # sensor
    S = s.sensor
    sen = sc.sensor
    S.position[:] = [float(v) for v in sen.position]
    S.rotation[:] = [float(v) for v in O.euler_to_matrix(np.asarray(sen.rotation), "ref32").reshape(-1)]
    S.n_pixels = sen.n_pixels
    if isinstance(sen, O.SquareSensorSpec):
        S.kind = L.SENSOR_SQUARE
        S.width, S.height = sen.width, sen.height
        S.x0, S.y0, S.dx, S.dy = sen.params()
        S.edge_width = sen.edge_width
    else:
        S.kind = L.SENSOR_HEX
        S.hex_size, S.hex_inradius, S.grid_rotation = sen.hex_size, sen.inradius, sen.grid_rotation
        S.grid_offset[:] = list(sen.grid_offset)
        S.edge_threshold = float(np.float32(1.0 - sen.edge_width / sen.inradius))
        S.q_min, S.r_min = sen.q_min, sen.r_min
        S.table_q, S.table_r = sen.table.shape
        S.lookup_table = _arr(sen.table, np.int32, keep)
    return s
