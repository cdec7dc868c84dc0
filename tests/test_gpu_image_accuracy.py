"""Image accuracy at large sizes: the kernels' image against the FP64 run (GPU).

The per-ray parity tests stop at ~5e5 rays per image.  Here whole dishes at
their full sample counts are traced for a source subsample (C2: 382 facets x
4096 samples x 64 point sources = 1.0e8 rays, square camera; C3: 875 facets
x 1024 samples x 64 point sources = 5.7e7 rays, hexagonal camera, box and four
cylinders) and the image is compared with oracle/cpu_ref's FP64 run of the
same samples (every step in double, image summed in double):
  * relative L1 error of the image at most twice that of the reference's
    own fp32 arithmetic (oracle/cpu_ref's fp32 run of the same rays) plus
    1e-6: rays within fp32 rounding of a pixel edge land on either side in
    any fp32 evaluation, and at C3's 36 m focal length and 2 cm hexagonal
    pixels that alone moves ~3e-5 of the image between neighbours -- the
    edge band, which the per-ray tests count;
  * on every significant pixel (>= 1e-3 of the peak), the error that the edge
    band does not explain <= 1e-5 of the pixel: |gpu - fp64| minus the band
    mass of that pixel (|value| of the FP64 rays within the edge band of a
    binning boundary -- 2.5e-5 m on the sensor, tests.helpers.band_units -- at
    their pixel and the one across the nearest edge; the kernels' sensor
    coordinates stay within 8.6e-6 m of the FP64 ones, see the per-ray tests'
    printed coordinate errors);
  * the raw per-pixel error and the band counts are printed.
"""
import numpy as np
import pytest
import torch

import iactrace_amd as ia
from iactrace_amd import configs
from oracle import cpu_ref
from tests.helpers import band_units, oracle_scene

pytestmark = pytest.mark.gpu
DEV = "cuda"


def image_vs_fp64(img, i64, b64):
    """(L1 rel, max rel on significant px, band-resolved max rel on them)."""
    g = np.asarray(img, np.float64).reshape(-1)
    sig = i64 > 1e-3 * i64.max()
    d = np.abs(g - i64)
    l1 = d.sum() / np.abs(i64).sum()
    raw = (d[sig] / i64[sig]).max()
    res = (np.maximum(d - b64, 0.0)[sig] / i64[sig]).max()
    return float(l1), float(raw), float(res), int(sig.sum())


@pytest.mark.parametrize("name", ["C2", "C3"])
def test_image_vs_fp64_run(name):
    wl = configs.workload(name, n_sources=64)
    print(f"\n[{name} x 64 sources] facets={wl.n_facets} samples={wl.n_samples} "
          f"rays={wl.n_rays:.3g}", flush=True)
    tel = wl.telescope(DEV, key=0)
    img = ia.render(tel, torch.as_tensor(wl.sources, device=DEV),
                    torch.as_tensor(wl.values, device=DEV), wl.source_type)
    torch.cuda.synchronize()
    sc = oracle_scene(tel)
    keep = []
    s = cpu_ref.scene_struct(sc, keep)
    i64, b64, st = cpu_ref.render_band(s, wl.sources, wl.values, wl.source_type,
                                       sc.sensor.n_pixels, 0, 64, band_units(sc.sensor))
    i32 = cpu_ref.render(s, wl.sources, wl.values, wl.source_type, sc.sensor.n_pixels, 0, 32)
    l1, raw, res, nsig = image_vs_fp64(img.cpu().numpy(), i64, b64)
    l1c, rawc, resc, _ = image_vs_fp64(i32, i64, b64)
    print(f"[{name}] vs FP64 run: L1 rel {l1:.2e}; significant px {nsig}: max rel {raw:.2e}, "
          f"band-resolved max rel {res:.2e}; band rays {st['band']} of {st['traced']} "
          f"({st['band'] / max(st['traced'], 1):.2e}); the fp32 CPU port: L1 rel {l1c:.2e}, "
          f"max rel {rawc:.2e}, band-resolved {resc:.2e}", flush=True)
    assert i64.max() > 0
    assert l1 <= 2.0 * l1c + 1e-6
    assert res <= 1e-5
