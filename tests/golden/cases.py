"""Golden cases: small telescopes built on the CPU from the configuration
generators, with MC samples from the oracle's Philox sampler, so that the
same inputs can be rebuilt anywhere.  make_golden.py writes the oracle's
outputs for them; test_golden.py (CPU) checks the oracle still reproduces
them and test_gpu_golden.py checks the HIP kernels against them.
"""
from __future__ import annotations

import numpy as np
import torch

import iactrace_amd as ia
import oracle.iact_oracle as O
from iactrace_amd import configs
from iactrace_amd.mirrors import AsphericDiskMirrorGroup

CASES = ("dc_square_point", "parabolic_hex_obstructions", "cassegrain_response",
         "dc_splat_square_parallel", "parabolic_splat_hex")


def _sampled(mirrors, m, seed):
    """Group the mirrors (one group expected per stage) and fill stage-0
    samples with the oracle's sampler."""
    out = []
    for gi, g in enumerate(ia.group_mirrors(mirrors)):
        if g.optical_stage == 0:
            surf = O.Surface(g.curvature, g.conic, g.aspheric.double().numpy())
            offs = g.offsets.numpy()
            if isinstance(g, AsphericDiskMirrorGroup):
                p, n, d, w = O.sample_group(seed, 0, surf, offs, m, radii=g.radii.numpy())
            else:
                p, n, d, w = O.sample_group(seed, 0, surf, offs, m, vertices=g.vertices.numpy())
            g = g.replace(points=torch.as_tensor(p), normals=torch.as_tensor(n),
                          perturbation_delta=torch.as_tensor(d),
                          weights=torch.as_tensor(w)[..., None],
                          perturbation_scale=torch.full((len(g),), 2e-3))
        out.append(g)
    return out


def build(name):
    """(telescope on the CPU, sources, values, source type, mode)."""
    p = 15.0 * np.tan(np.deg2rad(0.16))
    if name in ("dc_square_point", "dc_splat_square_parallel"):
        dish = configs.davies_cotton_dish()
        mirrors = [dish[i] for i in (0, 3, 17, 60, 131, 250, 333, 381)]
        groups = _sampled(mirrors, 48, 11)
        if name == "dc_square_point":
            sensor = ia.SquareSensor([0, 0, 15.0], [0, 0, 0], 32, 30,
                                     (-16 * p, 16 * p, -15 * p, 15 * p))
            src, vals = configs.point_source_field(5, 1.0, seed=21)
            return ia.Telescope(groups, [], [sensor]), src, vals, "point", "image"
        sensor = ia.DifferentiableSquareSensor([0, 0, 15.0], [0, 0, 0], 32, 30,
                                               (-16 * p, 16 * p, -15 * p, 15 * p), 0.6, 2)
        src = configs.parallel_grid(3, 2, 0.4)
        vals = np.array([1.0, 2.0, 0.5, 1.5, 0.8, 1.2], np.float32)
        return ia.Telescope(groups, [], [sensor]), src, vals, "parallel", "image"
    if name in ("parabolic_hex_obstructions", "parabolic_splat_hex"):
        dish = configs.parabolic_dish()
        xy = np.array([m.position[:2].numpy() for m in dish])
        # facets near the four support cylinders and near the axis
        pick = sorted(set(np.argsort(np.hypot(xy[:, 0] - 6, xy[:, 1] - 6))[:5].tolist()
                          + np.argsort(np.hypot(xy[:, 0] + 9, xy[:, 1] - 9))[:3].tolist()
                          + [0, 1]))
        groups = _sampled([dish[i] for i in pick], 32, 12)
        obst = ia.group_obstructions(configs._c3_obstructions())
        cam = configs.hex_camera(2048)
        if name == "parabolic_hex_obstructions":
            sensor = ia.HexagonalSensor([0, 0, 36.066], [0, 0, 0], cam)
        else:
            sensor = ia.DifferentiableHexagonalSensor([0, 0, 36.066], [0, 0, 0], cam, 0.6, 1)
        src, vals = configs.point_source_field(4, 1.5, seed=22)
        return ia.Telescope(groups, obst, [sensor]), src, vals, "point", "image"
    if name == "cassegrain_response":
        mirrors, sensor = configs.cassegrain()
        groups = _sampled(mirrors, 256, 13)
        src = configs.parallel_grid(3, 2, 0.05)
        vals = np.ones(len(src), np.float32)
        return ia.Telescope(groups, [], [sensor]), src, vals, "parallel", "response"
    raise KeyError(name)
