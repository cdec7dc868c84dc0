"""Algorithmic flop counts per ray, as written in the reference expressions.

Counting convention: add, sub, mul, div, sqrt, sin, cos, exp each 1 flop, an
FMA 2; compares, selects, min/max/abs, floor/round and integer work 0.
Quantities that depend only on per-launch, per-facet or per-primitive
constants (rotation matrices, plane n.p, cylinder axis/height, box corners)
are hoisted and not counted per ray.  The stage>=1 Newton iteration is
counted for its analytic-derivative form.

Per-ray costs (render.py:118-157 and callees):
  point-source direction  12 (fp64)     parallel: 0
  primary reflect         12            primary value 3
  cylinder test           86 each       AABB 18 each     sphere 33 each
  shadow multiply          1 per obstruction group
  stage>=1 mirror        453 + 69 K per mirror (Newton 280 + 60 K is fp64)
  sensor plane            31            square bin 11     hex bin 32
  world transform         63 per (facet, sample) per call
"""
from __future__ import annotations

__all__ = ["flops_per_ray", "split_per_ray"]


def split_per_ray(source_type="point", n_cyl=0, n_box=0, n_sph=0, n_mirrors2=0, k_aspheric2=0,
                  sensor="square", n_sources=1):
    """(fp32 flops, fp64 flops) per ray; the world transform (63 per sample)
    is amortised over the sources."""
    fp64 = 12 if source_type == "point" else 0
    fp32 = 12 + 3 + 31 + (11 if sensor == "square" else 32)
    fp32 += 86 * n_cyl + 18 * n_box + 33 * n_sph
    fp32 += sum(1 for n in (n_cyl, n_box, n_sph) if n)
    if n_mirrors2:
        per = 453 + 69 * k_aspheric2
        newton = 280 + 60 * k_aspheric2
        fp64 += n_mirrors2 * newton
        fp32 += n_mirrors2 * (per - newton)
    fp32 += 63.0 / max(n_sources, 1)
    return fp32, fp64


def flops_per_ray(**kw):
    a, b = split_per_ray(**kw)
    return a + b
