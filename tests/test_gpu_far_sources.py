"""Far point sources on the GPU.

A point source farther from a facet than 2^30 times the samples' radius
bound takes the direction u = e (its offset from the facet position) for
every sample of the facet; any other takes u = q + e per ray (q the
sample's offset).  Cases: waves holding far (1e10 m) and near (2e4 m)
sources together -- the per-lane path of the source-major kernels --
against the oracle in both image layouts; and the far-source rays against
per-ray directions of the same samples (the radius bound withheld).
"""
import importlib

import numpy as np
import pytest
import torch

import iactrace_amd as ia
from iactrace_amd import configs
from tests.helpers import band_mask, oracle_scene
from tests.test_gpu_forward import _trace_and_compare
from tests.test_gpu_grad import _mixed_sources

pytestmark = pytest.mark.gpu
DEV = "cuda"
# the module (the package exports its render function under the same name)
R = importlib.import_module("iactrace_amd.render")


def _inner_dish(m=8):
    """C3's camera, box and masts over the inner 91 facets of its dish."""
    wl = configs.workload("C3", n_sources=128, n_samples=m)
    w2 = configs.Workload("c3-inner", configs.parabolic_dish(n_facets=91), wl.obstructions,
                          wl.sensor, m, wl.sources, wl.values, "point",
                          tilt_sigma_deg=wl.tilt_sigma_deg, tilt_seed=wl.tilt_seed)
    return w2.telescope(DEV, key=9), wl


def test_far_threshold_of_the_descriptor():
    """The scene carries the samples' radius bound; 1e10 m sources are far
    and 2e4 m sources near for the dish's facets."""
    tel, _ = _inner_dish()
    r = R.sample_radius(tel.mirror_groups[0].points)
    pts = tel.mirror_groups[0].points.detach().double()
    assert r == pytest.approx(float(pts.norm(dim=-1).max()), rel=1e-6)
    assert 0.1 < r < 5.0
    thresh = max(r, 1e-3) * 2.0 ** 30
    assert 2.0e4 < thresh < 1e10


@pytest.mark.parametrize("layout", [1, 2])
def test_mixed_far_and_near_sources(layout):
    tel, _ = _inner_dish()
    src, vals = _mixed_sources(128, 31)
    _trace_and_compare(tel, src, vals, "point", f"far+near layout {layout}",
                       launch=ia.LaunchConfig(layout=layout))


def test_far_directions_match_per_ray_directions(monkeypatch):
    """The same far sources traced with one direction per (facet, source)
    and, with the radius bound withheld (0: not given), per ray: sensor
    coordinates within 1e-5 m, the same pixel for every ray outside the
    edge band, and images that differ only by edge-band rays."""
    tel, wl = _inner_dish(m=16)
    src, vals = wl.sources, wl.values
    s, v = torch.as_tensor(src), torch.as_tensor(vals)
    xy_f, v_f, p_f = ia.render_debug(tel, s, v, "point", return_pixels=True)
    img_f = ia.render(tel, s, v, "point").cpu().numpy().astype(np.float64)
    monkeypatch.setattr(R, "sample_radius", lambda points: 0.0)
    xy_r, v_r, p_r = ia.render_debug(tel, s, v, "point", return_pixels=True)
    img_r = ia.render(tel, s, v, "point").cpu().numpy().astype(np.float64)
    xy_f, xy_r = xy_f.cpu().numpy(), xy_r.cpu().numpy()
    p_f, p_r = p_f.cpu().numpy(), p_r.cpu().numpy()
    v_f, v_r = v_f.cpu().numpy(), v_r.cpu().numpy()
    hit = (np.abs(xy_r[:, 0]) < 1e9) & (np.abs(xy_f[:, 0]) < 1e9)
    dxy = np.abs(xy_f[hit] - xy_r[hit]).max()
    band, _, _, _ = band_mask(oracle_scene(tel), src, vals, "point")
    differ = (p_f != p_r) | ((v_f != 0) != (v_r != 0))
    print(f"[far vs per-ray] rays {len(p_f)}, max |dxy| {dxy:.2e} m, pixel or shadow differs "
          f"for {int(differ.sum())} rays ({int((differ & ~band).sum())} outside the band)",
          flush=True)
    assert dxy < 1e-5
    assert not (differ & ~band).any()
    same = ~differ
    assert np.allclose(v_f[same], v_r[same], rtol=1e-6, atol=0)
    # image difference within the values of the rays binned differently
    moved = np.abs(v_f[differ]).sum() + np.abs(v_r[differ]).sum()
    assert np.abs(img_f - img_r).sum() <= moved * (1 + 1e-5) + 1e-6 * np.abs(img_r).sum()
