"""Obstruction culling (GPU).

Culling only skips shadow tests of obstructions that no ray of a facet's
source-facing bundle can reach, so every per-ray result must be bitwise the
same with culling on and off: hit points, values and pixel ids of
render_debug, deterministic images, and (to float-atomic order) images,
response matrices and gradients.  The random scenes put obstructions of every
kind near the edges of the bundles, where a wrong bound would drop a
candidate that does block.
"""
import numpy as np
import pytest
import torch

import iactrace_amd as ia
from iactrace_amd import configs
from iactrace_amd.obstructions import Box, Cylinder, OrientedBox, Sphere, Triangle

pytestmark = pytest.mark.gpu
DEV = "cuda"

ON = ia.LaunchConfig(culling=1)
OFF = ia.LaunchConfig(culling=-1)


def _debug_pair(tel, src, vals, stype, label):
    s, v = torch.as_tensor(src), torch.as_tensor(vals)
    a = ia.render_debug(tel, s, v, stype, launch=ON, return_pixels=True)
    b = ia.render_debug(tel, s, v, stype, launch=OFF, return_pixels=True)
    xa, va, pa = (t.cpu().numpy() for t in a)
    xb, vb, pb = (t.cpu().numpy() for t in b)
    hit = (pb >= 0)
    shadowed = int(((vb == 0) & (np.abs(xb).sum(1) < 1e9)).sum())
    print(f"[{label}] rays={len(vb)} on-sensor={int(hit.sum())} zero-valued={shadowed} "
          f"value mismatches={int((va != vb).sum())} pixel mismatches={int((pa != pb).sum())}",
          flush=True)
    assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32))
    assert np.array_equal(va.view(np.uint32), vb.view(np.uint32))
    assert np.array_equal(pa, pb)
    return vb


@pytest.mark.parametrize("stype,ns", [("point", 12), ("parallel", 12), ("point", 160),
                                      ("parallel", 192)])
def test_truss_debug_bitwise(stype, ns):
    """(from 128 sources the masks are per source chunk)"""
    wl = configs.workload("C3T", n_sources=ns, n_samples=64 if ns < 100 else 16)
    tel = wl.telescope(DEV, key=5)
    if stype == "point":
        src, vals = wl.sources, wl.values
    else:
        src = configs.parallel_grid(4 if ns < 100 else 16, 3 if ns < 100 else 12, 0.3)
        vals = np.linspace(0.5, 2.0, len(src)).astype(np.float32)
    vb = _debug_pair(tel, src, vals, stype, f"truss-{stype}-{len(src)}")
    # the truss does cast shadows on this workload
    assert (vb == 0).mean() > 0.01


def _random_edge_scene(seed, n_obs=240, fov_deg=3.0, z_src=1e10):
    """A small dish with obstructions of every kind placed around the rays
    from random facets towards random sources, at lateral offsets of the
    order of the facet size (near the edges of the facet bundles)."""
    rng = np.random.default_rng(seed)
    mirrors = configs.parabolic_dish(n_facets=61)
    sensor = ia.HexagonalSensor([0, 0, 36.066], [0, 0, 0], configs.hex_camera(512))
    src, vals = configs.point_source_field(24, fov_deg, seed=seed + 1, z=z_src)
    centres = np.array([m.position for m in mirrors], np.float64)
    obs = []
    for k in range(n_obs):
        c = centres[rng.integers(len(centres))]
        s = src[rng.integers(len(src))]
        u = s - c
        u /= np.linalg.norm(u)
        # a unit vector across the ray
        w = np.cross(u, rng.normal(size=3))
        w /= np.linalg.norm(w)
        t = rng.uniform(-2.0, 40.0)
        lat = rng.uniform(0.0, 1.6)
        p = c + t * u + lat * w
        r = rng.uniform(0.01, 0.4)
        kind = k % 5
        if kind == 0:
            ax = rng.normal(size=3)
            ax /= np.linalg.norm(ax)
            h = rng.uniform(0.1, 6.0)
            obs.append(Cylinder(p - 0.5 * h * ax, p + 0.5 * h * ax, r))
        elif kind == 1:
            e = rng.uniform(0.02, 0.5, 3)
            obs.append(Box(p - e, p + e))
        elif kind == 2:
            obs.append(Sphere(p, r))
        elif kind == 3:
            q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
            obs.append(OrientedBox(p, rng.uniform(0.02, 0.5, 3), q))
        else:
            v = rng.normal(scale=0.5, size=(3, 3))
            obs.append(Triangle(p + v[0], p + v[1], p + v[2]))
    wl = configs.Workload("edge", mirrors, obs, sensor, 48, src, vals, "point")
    return wl


@pytest.mark.parametrize("seed,fov,z", [(0, 3.0, 1e10), (1, 20.0, 1e10), (2, 3.0, 2000.0),
                                        (3, 60.0, 50.0)])
def test_random_obstructions_near_bundle_edges(seed, fov, z):
    wl = _random_edge_scene(seed, fov_deg=fov, z_src=z)
    tel = wl.telescope(DEV, key=seed)
    vb = _debug_pair(tel, wl.sources, wl.values, "point", f"edges seed={seed} fov={fov} z={z}")
    assert (vb == 0).mean() > 0.001


def test_truss_images_response_and_gradients():
    wl = configs.workload("C3T", n_sources=96, n_samples=64)
    tel = wl.telescope(DEV, key=6)
    s, v = torch.as_tensor(wl.sources), torch.as_tensor(wl.values)
    # deterministic sample-major images: bitwise
    Ld_on = ia.LaunchConfig(deterministic=True, culling=1)
    Ld_off = ia.LaunchConfig(deterministic=True, culling=-1)
    a = ia.render(tel, s, v, "point", launch=Ld_on)
    b = ia.render(tel, s, v, "point", launch=Ld_off)
    assert torch.equal(a, b)
    # both layouts with float atomics
    for lay in (1, 2):
        a = ia.render(tel, s, v, "point", launch=ia.LaunchConfig(layout=lay, culling=1))
        err = (a - b).abs().max().item() / b.abs().max().item()
        print(f"[truss-image layout={lay}] max rel err vs unculled {err:.2e}", flush=True)
        assert err < 1e-5
    ra = ia.render_response_matrix(tel, s[:16], v[:16], "point", launch=ON)
    rb = ia.render_response_matrix(tel, s[:16], v[:16], "point", launch=OFF)
    err = (ra - rb).abs().max().item() / rb.abs().max().item()
    print(f"[truss-response] max rel err {err:.2e}", flush=True)
    assert err < 1e-5
    # gradients of a weighted image sum through the adjoint
    P = tel.sensors[0].n_pixels
    W = torch.linspace(0.5, 1.5, P, device=DEV)
    grads = []
    for L in (ON, OFF):
        g = tel.mirror_groups[0]
        rot = g.rotations.detach().clone().requires_grad_(True)
        t = ia.Telescope([g.replace(rotations=rot)], tel.obstruction_groups, tel.sensors)
        (ia.render(t, s, v, "point", launch=L).reshape(-1) * W).sum().backward()
        grads.append(rot.grad.double())
    err = (grads[0] - grads[1]).abs().max().item() / grads[1].abs().max().item()
    print(f"[truss-grad] max rel err {err:.2e}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("stype", ["point", "parallel"])
def test_chunk_culling_sorted_sources_and_source_gradients(stype):
    """From 128 sources the masks are per (source chunk, facet) and image and
    gradient traces sort the sources into direction cells: images, facet and
    per-source gradients (written back through the permutation) must match
    the unculled traces."""
    wl = configs.workload("C3T", n_sources=300, n_samples=32)
    tel = wl.telescope(DEV, key=8)
    W = torch.linspace(0.5, 1.5, tel.sensors[0].n_pixels, device=DEV)
    if stype == "point":
        src0, val0 = wl.sources, wl.values
    else:
        rng = np.random.default_rng(9)
        src0 = configs.parallel_grid(20, 15, 0.15)[rng.permutation(300)]
        val0 = rng.uniform(0.5, 2.0, 300).astype(np.float32)
    out = []
    for L in (ON, OFF):
        g = tel.mirror_groups[0]
        rot = g.rotations.detach().clone().requires_grad_(True)
        t = ia.Telescope([g.replace(rotations=rot)], tel.obstruction_groups, tel.sensors)
        s = torch.as_tensor(src0, device=DEV).clone().requires_grad_(True)
        v = torch.as_tensor(val0, device=DEV).clone().requires_grad_(True)
        img = ia.render(t, s, v, stype, launch=L)
        (img.reshape(-1) * W).sum().backward()
        out.append((img.detach(), rot.grad.double(), s.grad.double(), v.grad.double()))
    names = ["image", "rotations", "sources", "values"]
    for k in range(4):
        a, b = out[0][k], out[1][k]
        scale = b.abs().max().item()
        err = (a - b).abs().max().item() / max(scale, 1e-300)
        print(f"[chunk-culling] {names[k]}: max rel err {err:.2e} (scale {scale:.3g})", flush=True)
        assert scale > 0, names[k]
        assert err < 1e-5, names[k]
    # the truss shadows part of the image
    assert out[1][0].abs().sum().item() > 0


def test_culling_mode_validation():
    wl = configs.workload("C3T", n_sources=4, n_samples=16)
    tel = wl.telescope(DEV, key=0)
    with pytest.raises(ia._lib.IactError):
        ia.render(tel, torch.as_tensor(wl.sources), torch.as_tensor(wl.values), "point",
                  launch=ia.LaunchConfig(culling=2))


def test_batched_culling_matches_unculled():
    """One candidate table serves every instance of a batch (cones over all
    the batch's sources); images and rotation gradients must not change."""
    wl = configs.workload("C3T", n_sources=8, n_samples=64)
    tel = wl.telescope(DEV, key=7)
    base = tel.mirror_groups[0].rotations.detach()
    rng = np.random.default_rng(3)
    n_inst = 3
    rot = base.unsqueeze(0).repeat(n_inst, 1, 1).clone()
    rot[:, :, :2] += torch.as_tensor(rng.normal(0, 0.05, (n_inst, base.shape[0], 2)),
                                     dtype=torch.float32, device=DEV)
    src, vals = [], []
    # 270 sources in all: per-chunk masks over sources sorted per instance
    for i, n in enumerate((100, 0, 170)):
        a, b = configs.point_source_field(n, 2.0, seed=80 + i)
        src.append(a)
        vals.append(b)
    s, v, off = ia.pack_sources(src, vals)
    W = torch.linspace(0.5, 1.5, tel.sensors[0].n_pixels, device=DEV)
    out = []
    for L in (ON, OFF):
        r = rot.clone().requires_grad_(True)
        img = ia.render_batched(tel, r, s, v, off, "point", launch=L)
        (img * W).sum().backward()
        out.append((img.detach(), r.grad.double()))
    ei = (out[0][0] - out[1][0]).abs().max().item() / out[1][0].abs().max().item()
    eg = (out[0][1] - out[1][1]).abs().max().item() / out[1][1].abs().max().item()
    print(f"[batched-culling] image rel err {ei:.2e}, gradient rel err {eg:.2e}", flush=True)
    assert ei < 1e-5 and eg < 1e-5
