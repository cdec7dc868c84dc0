"""Differentiable float64 twin of the renderer -- TEST INFRASTRUCTURE ONLY.

Restates render.py:118-220 (and the sampling maps of integrators.py:97-188)
with torch float64 tensors so torch.autograd stands in for the reference's
jax.grad (JAX is not available here).  Decisions with zero derivative in the
reference -- shadow masks (render.py:40 `where`), hard pixel indices
(square.py:71-72 floor, hexagonal.py:178-182 round), aperture and hit masks
-- are taken from the float64 numpy oracle on detached values, exactly as
jax.grad treats them.  Stage>=1 mirrors are differentiated through the ten
unrolled Newton iterations with the freeze rule (intersections.py:332-365),
which is what jax.grad does.

The aspheric gradient of a primary group is the gradient of render composed
with re-sampling at fixed aperture positions (x, y) and fixed roughness
deltas: points z = sag(x+x0, y+y0) - sag(x0, y0), normals from the sag
slope, weights nz * M / area (the reference's render alone does not read the
primary's coefficients).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import iact_oracle as O

T = torch.float64


def euler(angles_deg):
    a = angles_deg * (math.pi / 180.0)
    cx, sx = torch.cos(a[..., 0]), torch.sin(a[..., 0])
    cy, sy = torch.cos(a[..., 1]), torch.sin(a[..., 1])
    cz, sz = torch.cos(a[..., 2]), torch.sin(a[..., 2])
    o, z = torch.ones_like(cx), torch.zeros_like(cx)
    Rx = torch.stack([torch.stack([o, z, z], -1), torch.stack([z, cx, -sx], -1),
                      torch.stack([z, sx, cx], -1)], -2)
    Ry = torch.stack([torch.stack([cy, z, sy], -1), torch.stack([z, o, z], -1),
                      torch.stack([-sy, z, cy], -1)], -2)
    Rz = torch.stack([torch.stack([cz, -sz, z], -1), torch.stack([sz, cz, z], -1),
                      torch.stack([z, z, o], -1)], -2)
    return Rz @ Ry @ Rx


def sag_raw(X, Y, c, k, asph):
    r2 = X * X + Y * Y
    z = r2 * c / (1 + torch.sqrt(1 - (1 + k) * c * c * r2))
    for i in range(asph.shape[0]):
        z = z + asph[i] * r2 ** (2 * (i + 1))
    return z


def normalize(v):
    return v / torch.sqrt((v * v).sum(-1, keepdim=True))


def sample_points_normals(xy, offsets, c, k, asph):
    """Re-derive local points and unit normals at fixed aperture positions."""
    x0 = offsets[:, None, 0]
    y0 = offsets[:, None, 1]
    X = xy[..., 0] + x0
    Y = xy[..., 1] + y0
    z = sag_raw(X, Y, c, k, asph) - sag_raw(x0, y0, c, k, asph)
    r2 = X * X + Y * Y
    K = (1 + k) * c * c
    s = torch.sqrt(1 - K * r2)
    den = 1 + s
    g = c / den + (r2 * c) * K / (2 * s * den * den)
    for i in range(asph.shape[0]):
        p = 2 * (i + 1)
        g = g + asph[i] * p * r2 ** (p - 1)
    n = normalize(torch.stack([-2 * X * g, -2 * Y * g, torch.ones_like(X)], -1))
    return torch.stack([xy[..., 0], xy[..., 1], z], -1), n


def _newton(o, d, c, k, asph, x0, y0):
    """Unrolled 10-step Newton of surfaces.py:67-108 (differentiable)."""
    z0 = sag_raw(x0, y0, c, k, asph)
    with torch.no_grad():
        t0 = torch.as_tensor(O.intersect_conic(
            np.stack([o[..., 0].detach().numpy() + x0, o[..., 1].detach().numpy() + y0,
                      o[..., 2].detach().numpy() + float(z0)], -1),
            d.detach().numpy(), np.float64(float(c)), np.float64(float(k))))
    t = t0.clone()
    conv = torch.zeros(t.shape, dtype=torch.bool)

    def g(tt):
        return (o[..., 2] + tt * d[..., 2]) - (sag_raw(o[..., 0] + tt * d[..., 0] + x0,
                                                       o[..., 1] + tt * d[..., 1] + y0,
                                                       c, k, asph) - z0)
    for _ in range(10):
        gv = g(t)
        # derivative of g w.r.t. t, differentiable in the remaining inputs
        X = o[..., 0] + t * d[..., 0] + x0
        Y = o[..., 1] + t * d[..., 1] + y0
        r2 = X * X + Y * Y
        K = (1 + k) * c * c
        s = torch.sqrt(1 - K * r2)
        den = 1 + s
        gg = c / den + (r2 * c) * K / (2 * s * den * den)
        for i in range(asph.shape[0]):
            p = 2 * (i + 1)
            gg = gg + asph[i] * p * r2 ** (p - 1)
        gpd = d[..., 2] - (2 * X * gg * d[..., 0] + 2 * Y * gg * d[..., 1])
        gpd = torch.where(gpd.abs() > 1e-12, gpd, torch.full_like(gpd, 1e-12))
        tn = t - gv / gpd
        newc = conv | (gv.detach().abs() < 1e-8)
        t = torch.where(conv, t, tn)
        conv = newc
    res = g(t).detach().abs()
    valid = (t.detach() > 1e-8) & (res < 1e-6)
    return t, valid


def render_image(sc, sources, values, source_type="point", rotations=None, positions=None,
                 aspheric0=None, aspheric_stage=None, surface0=None, offsets0=None):
    """Image (P,) in float64 with autograd.  Tensors passed for rotations
    (Nf,3), positions (Nf,3), sources (Ns,3), values (Ns,), aspheric0 (K,) of
    the (single) primary surface and aspheric_stage {mirror index: (K,)}
    may require grad.
    """
    nf, M = sc.points.shape[:2]
    rot = rotations if rotations is not None else torch.as_tensor(sc.rotations, dtype=T)
    pos = positions if positions is not None else torch.as_tensor(sc.positions, dtype=T)
    src = sources if isinstance(sources, torch.Tensor) else torch.as_tensor(sources, dtype=T)
    val = values if isinstance(values, torch.Tensor) else torch.as_tensor(values, dtype=T)
    src = src.to(T)
    val = val.to(T)
    pts = torch.as_tensor(sc.points, dtype=T)
    nrm = torch.as_tensor(sc.normals, dtype=T)
    wts = torch.as_tensor(sc.weights, dtype=T)
    if aspheric0 is not None:
        # surface0 = (curvature, conic) of the primary group, offsets0 (Nf, 2)
        offs = torch.as_tensor(np.asarray(offsets0, np.float64), dtype=T)
        pts, nrm = sample_points_normals(pts[..., :2], offs, torch.tensor(surface0[0], dtype=T),
                                         torch.tensor(surface0[1], dtype=T), aspheric0.to(T))
        area_over_m = torch.as_tensor(sc.normals[..., 2] / sc.weights, dtype=T)
        wts = nrm[..., 2] / area_over_m
    dl = torch.as_tensor(sc.delta, dtype=T)
    scale = torch.as_tensor(sc.scale, dtype=T)
    R = euler(rot.to(T))
    p = torch.einsum("fij,fmj->fmi", R, pts) + pos.to(T)[:, None, :]
    n = normalize(torch.einsum("fij,fmj->fmi", R, nrm + scale[:, None, None] * dl))
    ns = src.shape[0]
    # detached decisions from the float64 numpy oracle
    x64, y64, v64, _ = O.render(sc, src.detach().numpy(), val.detach().numpy(), source_type,
                                "fp64", "debug")
    if source_type == "point":
        d = normalize(p[:, None, :, :] - src[None, :, None, :])          # (Nf, Ns, M, 3)
    else:
        d = src[None, :, None, :].expand(nf, ns, M, 3)
    # shadow masks: recompute with the oracle on detached values
    with torch.no_grad():
        o_np = np.broadcast_to(p.detach().numpy()[:, None], (nf, ns, M, 3))
        shadow = O.check_occlusions(o_np.reshape(-1, 3), -d.detach().numpy().reshape(-1, 3),
                                    sc.obstructions, np.float64).reshape(nf, ns, M)
    nn = n[:, None].expand(nf, ns, M, 3)
    c = (d * nn).sum(-1)
    r = d - 2 * c[..., None] * nn
    v = val[None, :, None] * (-c) / wts[:, None, :] * torch.as_tensor(shadow)
    o = p[:, None].expand(nf, ns, M, 3)
    for st in sorted({m.stage for m in sc.mirrors}):
        o, r, v = _stage(sc, [i for i, m in enumerate(sc.mirrors) if m.stage == st], o, r, v,
                         aspheric_stage or {})
    S = sc.sensor
    Rs = euler(torch.as_tensor(np.asarray(S.rotation, np.float64), dtype=T))
    cpos = torch.as_tensor(np.asarray(S.position, np.float64), dtype=T)
    u1, u2, pn = Rs[:, 0], Rs[:, 1], Rs[:, 2]
    ndotd = (r * pn).sum(-1)
    t = (torch.dot(pn, cpos) - (o * pn).sum(-1)) / ndotd
    inter = o + t[..., None] * r
    x = ((inter - cpos) * u1).sum(-1)
    y = ((inter - cpos) * u2).sum(-1)
    ok = torch.as_tensor((x64 != 1e10).reshape(nf, ns, M))
    img = torch.zeros(S.n_pixels, dtype=T)
    xf, yf, vf, okf = x.reshape(-1), y.reshape(-1), v.reshape(-1), ok.reshape(-1)
    if not S.splat:
        if isinstance(S, O.SquareSensorSpec):
            pid = O.square_bin(x64, y64, S)[0]
        else:
            pid = O.hex_bin(x64, y64, S)[0]
        m = torch.as_tensor(pid >= 0) & okf
        idx = torch.as_tensor(np.where(pid >= 0, pid, 0))
        return img.index_add(0, idx[m], vf[m])
    return _splat(S, xf, yf, vf, okf, x64, y64, img)


def _stage(sc, idxs, o, d, v, asph_over):
    best_t = torch.full(o.shape[:-1], float("inf"), dtype=T)
    best_p = torch.zeros_like(o)
    best_n = torch.zeros_like(o)
    for i in idxs:
        m = sc.mirrors[i]
        R = euler(torch.as_tensor(np.asarray(m.rotation, np.float64), dtype=T))
        pos = torch.as_tensor(np.asarray(m.position, np.float64), dtype=T)
        ol = (o - pos) @ R
        dlc = d @ R
        c = torch.tensor(m.surface.curvature, dtype=T)
        k = torch.tensor(m.surface.conic, dtype=T)
        asph = asph_over.get(i, torch.as_tensor(np.asarray(m.surface.aspheric, np.float64),
                                                dtype=T)).to(T)
        x0, y0 = float(m.offset[0]), float(m.offset[1])
        t, valid = _newton(ol.reshape(-1, 3), dlc.reshape(-1, 3), c, k, asph, x0, y0)
        t = t.reshape(o.shape[:-1])
        valid = valid.reshape(o.shape[:-1])
        xh = ol[..., 0] + t * dlc[..., 0]
        yh = ol[..., 1] + t * dlc[..., 1]
        # the reference evaluates the hit at t_final (not the masked t)
        z = sag_raw(xh + x0, yh + y0, c, k, asph) - sag_raw(torch.tensor(x0, dtype=T),
                                                            torch.tensor(y0, dtype=T), c, k, asph)
        _, nl = sample_points_normals(torch.stack([xh, yh], -1).reshape(1, -1, 2),
                                      torch.tensor([[x0, y0]], dtype=T), c, k, asph)
        nl = nl.reshape(o.shape)
        pl = torch.stack([xh, yh, z], -1)
        with torch.no_grad():
            ap = O._aperture_ok(m, xh.detach().numpy(), yh.detach().numpy(), np.float64)
        tt = torch.where(torch.as_tensor(ap) & valid, t, torch.full_like(t, float("inf")))
        closer = tt.detach() < best_t.detach()
        best_t = torch.where(closer, tt, best_t)
        best_p = torch.where(closer[..., None], pl @ R.T + pos, best_p)
        best_n = torch.where(closer[..., None], nl @ R.T, best_n)
    c = (d * best_n).sum(-1)
    r = d - 2 * c[..., None] * best_n
    hit = (best_t.detach() < 1e10).to(T)
    with torch.no_grad():
        shadow = O.check_occlusions(o.detach().numpy().reshape(-1, 3),
                                    d.detach().numpy().reshape(-1, 3), sc.obstructions,
                                    np.float64).reshape(o.shape[:-1])
    return best_p, r, v * hit * torch.as_tensor(shadow) * c.abs()


def _splat(S, x, y, v, ok, x64, y64, img):
    if isinstance(S, O.SquareSensorSpec):
        x0, y0, dx, dy = S.params()
        xp = (x - x0) / dx
        yp = (y - y0) / dy
        xb = torch.floor(xp.detach())
        yb = torch.floor(yp.detach())
        xf = xp - xb
        yf = yp - yb
        ks = S.kernel_size
        offs = [(ox, oy) for oy in range(-ks, ks + 1) for ox in range(-ks, ks + 1)]
        ws = [torch.exp(-0.5 * ((xf - ox) ** 2 + (yf - oy) ** 2) / S.sigma ** 2) for ox, oy in offs]
        tot = sum(ws)
        for (ox, oy), w in zip(offs, ws):
            xi, yi = xb + ox, yb + oy
            valid = (xi >= 0) & (xi < S.width) & (yi >= 0) & (yi < S.height) & ok
            idx = (torch.where(valid, yi, torch.zeros_like(yi)) * S.width
                   + torch.where(valid, xi, torch.zeros_like(xi))).long()
            img = img.index_add(0, idx[valid], (v * w / tot)[valid])
        return img
    xg, yg = O.hex_grid_coords(x64, y64, S)
    # grid transform restated in torch for the gradient
    ang = -S.grid_rotation
    cs, sn = math.cos(ang), math.sin(ang)
    ux = x - S.grid_offset[0]
    uy = y - S.grid_offset[1]
    xg_t = cs * ux - sn * uy
    yg_t = sn * ux + cs * uy
    q = (O.SQRT3_3 * xg_t - yg_t / 3) / S.hex_size
    r = (2 * yg_t / 3) / S.hex_size
    qb_np, rb_np = O.axial_round(q.detach().numpy(), r.detach().numpy())
    qb, rb = torch.as_tensor(qb_np), torch.as_tensor(rb_np)
    dq, dr = q - qb, r - rb
    ks = S.kernel_size
    offs = [(a, b) for a in range(-ks, ks + 1) for b in range(-ks, ks + 1)
            if max(abs(a), abs(b), abs(a + b)) <= ks]

    def hd(a, b):
        e = (a + 0.5 * b).abs()
        return torch.maximum(2 * e, e + 1.5 * b.abs())
    ws = [torch.exp(-0.5 * (hd(dq - a, dr - b) / S.sigma) ** 2) for a, b in offs]
    tot = sum(ws)
    for (a, b), w in zip(offs, ws):
        pid = O._hex_lookup(qb_np.astype(np.int64) + a, rb_np.astype(np.int64) + b, S)
        valid = torch.as_tensor(pid >= 0) & ok
        idx = torch.as_tensor(np.where(pid >= 0, pid, 0)).long()
        img = img.index_add(0, idx[valid], (v * w / tot)[valid])
    return img
