"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py) are reproduced by the oracle from inputs
rebuilt on the CPU -- so a change to the oracle, the samplers or the
configuration generators that moves any result is caught here."""
import numpy as np
import pytest

import oracle.iact_oracle as O
from tests.golden.cases import CASES, build
from tests.helpers import oracle_scene

INPUTS = ("points", "normals", "delta", "weights", "positions", "rotations", "scale")


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_golden(name):
    g = np.load(f"tests/golden/{name}.npz")
    tel, src, vals, stype, mode = build(name)
    sc = oracle_scene(tel)
    for k in INPUTS:
        np.testing.assert_array_equal(getattr(sc, k), g[k], err_msg=k)
    np.testing.assert_array_equal(src, g["sources"])
    x, y, v, pix = O.render(sc, src, vals, stype, "ref32", "debug")
    np.testing.assert_array_equal(pix, g["ray_pix"])
    np.testing.assert_array_equal(v, g["ray_v"])
    np.testing.assert_array_equal(x, g["ray_x"])
    np.testing.assert_allclose(O.render(sc, src, vals, stype, "fp64", mode), g["out_fp64"],
                               rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("name", CASES)
def test_golden_fixtures_are_sane(name):
    g = np.load(f"tests/golden/{name}.npz")
    assert g["band"].sum() < 0.01 * g["band"].size
    assert np.abs(g["out_fp64"]).sum() > 0
    # ref32 and fp64 images agree up to the band rays' worth of flux
    d = np.abs(g["out_ref32"].astype(np.float64) - g["out_fp64"]).sum()
    band_flux = 2 * np.abs(g["ray_v"][g["band"]]).sum() + 1e-5 * np.abs(g["out_fp64"]).sum()
    assert d <= band_flux
