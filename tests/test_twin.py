"""Pins of the float64 autograd twin (CPU): its image equals the float64
oracle's, and its gradients equal central differences of itself."""
import numpy as np
import pytest
import torch

import oracle.iact_oracle as O
import oracle.torch_twin as TW
from tests.test_oracle import _single_facet_scene


def _scene(splat, mirrors=(), m=300, asph=np.array([3e-4, -2e-6])):
    if splat:
        sensor = O.SquareSensorSpec((0, 0, 9.5), (0, 0, 0), 24, 24, (-0.3, 0.3, -0.3, 0.3),
                                    splat=True, sigma=0.6, kernel_size=2)
    else:
        sensor = O.SquareSensorSpec((0, 0, 9.5), (0, 0, 0), 24, 24, (-0.3, 0.3, -0.3, 0.3))
    surf = O.Surface(0.05, -1.05, asph)
    sc = _single_facet_scene(surf, 1.2, sensor, m=m, position=(0.1, -0.2, 0.0),
                             rotation=(0.3, -0.2, 0.0), mirrors=list(mirrors))
    return sc, surf


SRC = np.array([[0.002, 0.001, -1.0], [-0.003, 0.0, -1.0]])
VAL = np.array([1.0, 0.7], np.float32)


@pytest.mark.parametrize("splat", [False, True])
def test_twin_forward_equals_oracle(splat):
    sc, _ = _scene(splat)
    img = TW.render_image(sc, SRC, VAL, "parallel").detach().numpy()
    ref = O.render(sc, SRC, VAL.astype(np.float32), "parallel", "fp64", "image")
    assert np.allclose(img, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def test_twin_forward_equals_oracle_cassegrain():
    sec = O.MirrorSpec((0, 0, 6.0), (180.0, 0, 0), O.Surface(-0.05, -16.0, np.array([1e-4])),
                       (0, 0), radius=0.5)
    sensor = O.SquareSensorSpec((0, 0, -0.45), (0, 0, 0), 16, 16, (-0.3, 0.3, -0.3, 0.3))
    sc = _single_facet_scene(O.Surface(0.05, -1.05, np.array([1e-6])), 2.0, sensor, m=300,
                             mirrors=[sec])
    img = TW.render_image(sc, SRC, VAL, "parallel").detach().numpy()
    ref = O.render(sc, SRC, VAL.astype(np.float32), "parallel", "fp64", "image")
    assert ref.sum() > 0
    assert np.allclose(img, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def _fd(f, x, h):
    g = np.zeros_like(x)
    for i in np.ndindex(x.shape):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        g[i] = (f(xp) - f(xm)) / (2 * h)
    return g


def test_twin_gradients_match_its_finite_differences():
    sc, surf = _scene(True)
    W = torch.linspace(0.5, 1.5, sc.sensor.n_pixels, dtype=torch.float64)

    def loss(rot=None, src=None, asph=None):
        r = torch.as_tensor(rot if rot is not None else sc.rotations, dtype=torch.float64)
        s = torch.as_tensor(src if src is not None else SRC, dtype=torch.float64)
        a = torch.as_tensor(asph if asph is not None else surf.aspheric, dtype=torch.float64)
        return float((TW.render_image(sc, s, VAL, "parallel", r, aspheric0=a,
                                      surface0=(surf.curvature, surf.conic),
                                      offsets0=np.zeros((1, 2))) * W).sum())
    rot = torch.tensor(sc.rotations, dtype=torch.float64, requires_grad=True)
    src = torch.tensor(SRC, requires_grad=True)
    a = torch.tensor(surf.aspheric, dtype=torch.float64, requires_grad=True)
    L = (TW.render_image(sc, src, VAL, "parallel", rot, aspheric0=a,
                         surface0=(surf.curvature, surf.conic), offsets0=np.zeros((1, 2))) * W).sum()
    L.backward()
    fr = _fd(lambda r: loss(rot=r), sc.rotations.astype(np.float64), 1e-5)
    assert np.allclose(rot.grad.numpy(), fr, rtol=1e-4, atol=1e-6 * np.abs(fr).max())
    fs = _fd(lambda s: loss(src=s), SRC.copy(), 1e-7)
    assert np.allclose(src.grad.numpy(), fs, rtol=1e-4, atol=1e-6 * np.abs(fs).max())
    fa = _fd(lambda x: loss(asph=x), np.asarray(surf.aspheric, np.float64), 1e-9)
    assert np.allclose(a.grad.numpy(), fa, rtol=1e-4, atol=1e-6 * np.abs(fa).max())


def test_splat_window_truncation_biases_differences():
    """The reference's square splat renormalises its weights over the
    (2k+1)^2 taps around floor(x) (square.py:144-173), so the image jumps by
    the entering tap's weight whenever a ray crosses a pixel edge.  A central
    difference of a tilt then carries a bias that does not shrink with the
    step: ~1e-3 for sigma 0.7 px and k = 3, ~1e-6 for k = 4, on the float64
    twin itself.  (Its autograd, like the kernels' adjoint, is the derivative
    of the smooth part.)"""
    import math
    import torch
    import iactrace_amd as ia
    from iactrace_amd import configs
    from tests.helpers import cpu_scene, oracle_sensor
    wl = configs.workload("C2", n_sources=2, n_samples=256)
    wl.mirrors = configs.davies_cotton_dish(n_facets=1)
    src, vals = configs.point_source_field(2, 0.5, seed=1)
    yy, xx = np.mgrid[0:64, 0:64]
    W = torch.as_tensor((1.0 + 0.5 * np.sin(xx / 64 * 5.0) + 0.4 * np.cos(yy / 64 * 4.0))
                        .reshape(-1))
    rel = {}
    for k in (3, 4):
        sc = cpu_scene(wl, key=7)
        sc.sensor = oracle_sensor(ia.DifferentiableSquareSensor(
            [0, 0, 15.0], [0, 0, 0], 64, 64, (-0.8, 0.8, -0.8, 0.8), sigma=0.7, kernel_size=k))
        rot0 = torch.tensor(sc.rotations, dtype=torch.float64)

        def loss(rot):
            return (TW.render_image(sc, src, vals, "point", rot) * W).sum()
        r = rot0.clone().requires_grad_(True)
        loss(r).backward()
        ad = r.grad[0, 0].item()

        def cd(h):
            rp, rm = rot0.clone(), rot0.clone()
            rp[0, 0] += h
            rm[0, 0] -= h
            return (loss(rp) - loss(rm)).item() / (2 * h)
        h = math.degrees(1e-4)
        fd = (4 * cd(h / 2) - cd(h)) / 3
        rel[k] = abs(fd - ad) / abs(fd)
    assert rel[3] > 2e-4 and rel[4] < 1e-5, rel
