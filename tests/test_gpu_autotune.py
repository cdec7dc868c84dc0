"""Launch-shape autotuner (GPU): the chosen settings render the same image
as the defaults, and the cache answers the second request."""
import json

import numpy as np
import pytest
import torch

import iactrace_amd as ia
from iactrace_amd import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,ns", [("C2", 200), ("C3", 96)])
def test_autotune_picks_a_valid_launch_and_caches_it(name, ns, tmp_path, monkeypatch):
    monkeypatch.setenv("IACTRACE_AMD_TUNE_CACHE", str(tmp_path / "launch.json"))
    wl = configs.workload(name, n_sources=ns, n_samples=64)
    tel = wl.telescope("cuda", key=2)
    src, vals = torch.as_tensor(wl.sources), torch.as_tensor(wl.values)
    ref = ia.render(tel, src, vals, "point").reshape(-1)
    cfg = ia.autotune(tel, src, vals, "point", reps=1)
    print(f"\n[autotune {name}] {cfg}", flush=True)
    assert tel.launch is cfg
    assert cfg.layout in (1, 2) and cfg.block_threads in (128, 256)
    img = ia.render(tel, src, vals, "point").reshape(-1)
    err = ((img - ref).abs().max() / ref.abs().max()).item()
    assert err < 1e-5, err
    table = json.load(open(tmp_path / "launch.json"))
    assert len(table) == 1 and next(iter(table.values()))["ms"] > 0
    cfg2 = ia.autotune(tel, src, vals, "point", apply=False)
    assert vars(cfg2) == vars(cfg)


def test_autotune_response_mode():
    wl = configs.workload("C4", n_sources=16, n_samples=64)
    tel = wl.telescope("cuda", key=3)
    cfg = ia.autotune(tel, wl.sources, wl.values, "parallel", mode="response", cache=False,
                      apply=False, reps=1)
    rm = ia.render_response_matrix(tel, wl.sources, wl.values, "parallel", launch=cfg)
    rm0 = ia.render_response_matrix(tel, wl.sources, wl.values, "parallel")
    assert cfg.layout != 2
    np.testing.assert_allclose(rm.cpu().numpy(), rm0.cpu().numpy(), rtol=1e-5, atol=1e-7)
