"""Functional telescope operations (iactrace/telescope/operations.py).

Each returns a new Telescope; groups are replaced, never mutated.  Units
follow the reference: rotations in Euler degrees, roughness and
misalignment in arcseconds, positions in metres.  Gaussian draws use
torch's generator seeded from the integer key (the reference's jax.random
streams are not reproduced).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .integrators import as_key, split_key
from .telescope import Telescope

__all__ = ["resample_mirrors", "set_mirror_positions", "set_mirror_rotations",
           "scale_mirror_weights", "apply_roughness", "apply_roughness_to_group",
           "apply_misalignment_to_group", "apply_displacement_to_group", "get_mirrors_by_stage",
           "get_mirror_count", "add_sensor", "replace_sensor", "remove_sensor",
           "set_sensor_position", "set_sensor_rotation", "focus", "get_sensor_count",
           "add_obstruction", "remove_obstruction", "clear_obstructions",
           "get_obstruction_count", "clone", "get_info"]


def _with_groups(tel, groups):
    t = Telescope(groups, tel.obstruction_groups, tel.sensors, tel.name)
    t.launch = tel.launch
    return t


def _replace_group(tel, idx, group):
    groups = list(tel.mirror_groups)
    groups[idx] = group
    return _with_groups(tel, groups)


def _normal(key, n, device):
    g = torch.Generator().manual_seed(as_key(key) & ((1 << 63) - 1))
    return torch.randn(n, generator=g, dtype=torch.float64).to(device)


def resample_mirrors(telescope, integrator, key):
    """Resample every group with one key split (operations.py:25-40)."""
    keys = split_key(key, len(telescope.mirror_groups))
    return _with_groups(telescope, [integrator.sample_group(g, k)
                                    for g, k in zip(telescope.mirror_groups, keys)])


def set_mirror_positions(telescope, group_idx, positions):
    """operations.py:43-62."""
    g = telescope.mirror_groups[group_idx]
    return _replace_group(telescope, group_idx, g.replace(
        positions=torch.as_tensor(positions, dtype=torch.float32, device=g.positions.device)
        if not isinstance(positions, torch.Tensor) else positions))


def set_mirror_rotations(telescope, group_idx, rotations):
    """operations.py:65-84 (Euler degrees)."""
    g = telescope.mirror_groups[group_idx]
    return _replace_group(telescope, group_idx, g.replace(
        rotations=torch.as_tensor(rotations, dtype=torch.float32, device=g.rotations.device)
        if not isinstance(rotations, torch.Tensor) else rotations))


def scale_mirror_weights(telescope, group_idx, scale_factors):
    """Multiply the sample weights of each facet (operations.py:87-115).

    Weights divide the ray value, so a factor s scales a facet's light by 1/s.
    """
    g = telescope.mirror_groups[group_idx]
    s = torch.as_tensor(scale_factors, dtype=torch.float32, device=g.weights.device)
    if s.ndim == 0:
        s = s.expand(len(g))
    return _replace_group(telescope, group_idx, g.replace(weights=g.weights * s[:, None, None]))


def apply_roughness(telescope, roughness):
    """Roughness (arcsec) for every group (operations.py:118-135)."""
    sig = roughness * math.pi / (180.0 * 3600.0)
    groups = [g.replace(perturbation_scale=torch.full((len(g),), sig, dtype=torch.float32,
                                                      device=g.positions.device))
              for g in telescope.mirror_groups]
    return _with_groups(telescope, groups)


def apply_roughness_to_group(telescope, group_idx, roughness):
    """operations.py:138-158."""
    g = telescope.mirror_groups[group_idx]
    sig = roughness * math.pi / (180.0 * 3600.0)
    return _replace_group(telescope, group_idx, g.replace(
        perturbation_scale=torch.full((len(g),), sig, dtype=torch.float32,
                                      device=g.positions.device)))


def apply_misalignment_to_group(telescope, group_idx, sigma_h, sigma_v, key):
    """Gaussian tip/tilt errors in arcsec added to rotations[:, 0:2]
    (operations.py:161-198)."""
    g = telescope.mirror_groups[group_idx]
    n = len(g)
    k1, k2 = split_key(key, 2)
    dh = _normal(k1, n, g.rotations.device) * (sigma_h / 3600.0)
    dv = _normal(k2, n, g.rotations.device) * (sigma_v / 3600.0)
    rot = g.rotations.detach().clone()
    rot[:, 0] += dh.to(rot.dtype)
    rot[:, 1] += dv.to(rot.dtype)
    return _replace_group(telescope, group_idx, g.replace(rotations=rot))


def apply_displacement_to_group(telescope, group_idx, sigma_z, key):
    """Gaussian z displacement of facet positions (operations.py:201-229)."""
    g = telescope.mirror_groups[group_idx]
    dz = _normal(key, len(g), g.positions.device) * sigma_z
    pos = g.positions.detach().clone()
    pos[:, 2] += dz.to(pos.dtype)
    return _replace_group(telescope, group_idx, g.replace(positions=pos))


def get_mirrors_by_stage(telescope, stage):
    return [i for i, g in enumerate(telescope.mirror_groups) if g.optical_stage == stage]


def get_mirror_count(telescope):
    return sum(len(g) for g in telescope.mirror_groups)


def _with_sensors(tel, sensors):
    t = Telescope(tel.mirror_groups, tel.obstruction_groups, sensors, tel.name)
    t.launch = tel.launch
    return t


def add_sensor(telescope, sensor):
    return _with_sensors(telescope, list(telescope.sensors) + [sensor])


def replace_sensor(telescope, sensor, idx=0):
    if not (0 <= idx < len(telescope.sensors)):
        raise IndexError(f"Sensor index {idx} out of range for {len(telescope.sensors)} sensors")
    s = list(telescope.sensors)
    s[idx] = sensor
    return _with_sensors(telescope, s)


def remove_sensor(telescope, idx=0):
    if not (0 <= idx < len(telescope.sensors)):
        raise IndexError(f"Sensor index {idx} out of range for {len(telescope.sensors)} sensors")
    s = list(telescope.sensors)
    del s[idx]
    return _with_sensors(telescope, s)


def set_sensor_position(telescope, idx, position):
    return replace_sensor(telescope, telescope.sensors[idx].replace(position=position), idx)


def set_sensor_rotation(telescope, idx, rotation):
    return replace_sensor(telescope, telescope.sensors[idx].replace(rotation=rotation), idx)


def focus(telescope, delta_z, sensor_idx=0):
    """Move a sensor along z by delta_z (operations.py:360-373)."""
    pos = np.asarray(telescope.sensors[sensor_idx].position, np.float32).copy()
    pos[2] += delta_z
    return set_sensor_position(telescope, sensor_idx, pos)


def get_sensor_count(telescope):
    return len(telescope.sensors)


def add_obstruction(telescope, obstruction):
    groups = list(telescope.obstruction_groups or []) + [obstruction]
    t = Telescope(telescope.mirror_groups, groups, telescope.sensors, telescope.name)
    t.launch = telescope.launch
    return t


def remove_obstruction(telescope, group_idx):
    groups = list(telescope.obstruction_groups or [])
    if not groups:
        raise IndexError("No obstruction groups to remove")
    if not (0 <= group_idx < len(groups)):
        raise IndexError(f"Obstruction group index {group_idx} out of range for "
                         f"{len(groups)} groups")
    del groups[group_idx]
    t = Telescope(telescope.mirror_groups, groups, telescope.sensors, telescope.name)
    t.launch = telescope.launch
    return t


def clear_obstructions(telescope):
    t = Telescope(telescope.mirror_groups, [], telescope.sensors, telescope.name)
    t.launch = telescope.launch
    return t


def get_obstruction_count(telescope):
    return sum(len(g) for g in (telescope.obstruction_groups or []))


def clone(telescope):
    groups = [g.replace(**{k: v.clone() for k, v in vars(g).items()
                           if isinstance(v, torch.Tensor)}) for g in telescope.mirror_groups]
    return _with_groups(telescope, groups)


def get_info(telescope):
    return {
        "name": telescope.name,
        "n_mirror_groups": len(telescope.mirror_groups),
        "n_mirrors": get_mirror_count(telescope),
        "mirrors_by_stage": {s: sum(len(g) for g in telescope.mirror_groups
                                    if g.optical_stage == s)
                             for s in sorted({g.optical_stage for g in telescope.mirror_groups})},
        "n_samples": [g.n_samples for g in telescope.mirror_groups],
        "n_obstruction_groups": len(telescope.obstruction_groups or []),
        "n_obstructions": get_obstruction_count(telescope),
        "n_sensors": len(telescope.sensors),
        "sensors": [type(s).__name__ for s in telescope.sensors],
    }
