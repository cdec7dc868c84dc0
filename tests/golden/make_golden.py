"""Write the golden fixtures tests/golden/<case>.npz from the oracle.

Run from the repository root:  python tests/golden/make_golden.py
Each file holds the case's inputs (stage-0 samples and poses, sources,
values), the ref32 oracle's per-ray sensor coordinates, values and pixel ids
(facet-major debug order), the edge band of the rays, and the fp64 and ref32
images (or response matrices).
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle.iact_oracle as O  # noqa: E402
from tests.golden.cases import CASES, build  # noqa: E402
from tests.helpers import band_mask, oracle_scene  # noqa: E402


def fixture(name):
    tel, src, vals, stype, mode = build(name)
    sc = oracle_scene(tel)
    x, y, v, pix = O.render(sc, src, vals, stype, "ref32", "debug")
    band, _, _, _ = band_mask(sc, src, vals, stype)
    out = dict(points=sc.points, normals=sc.normals, delta=sc.delta, weights=sc.weights,
               positions=sc.positions, rotations=sc.rotations, scale=sc.scale,
               sources=np.asarray(src, np.float64), values=np.asarray(vals, np.float32),
               ray_x=x, ray_y=y, ray_v=v, ray_pix=pix, band=band,
               out_fp64=O.render(sc, src, vals, stype, "fp64", mode),
               out_ref32=O.render(sc, src, vals, stype, "ref32", mode))
    return out


def main():
    for name in CASES:
        f = fixture(name)
        path = os.path.join(ROOT, "tests", "golden", f"{name}.npz")
        np.savez_compressed(path, **f)
        print(f"{name}: rays {f['ray_v'].size}, band {int(f['band'].sum())}, "
              f"{os.path.getsize(path) / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
