"""Telescope (iactrace/telescope/telescope.py:22-237).

Holds mirror groups, obstruction groups and sensors.  Calling it renders
sources through the HIP kernels.  Convenience methods delegate to the
functional operations, which return new telescopes.
"""
from __future__ import annotations

from .render import render, render_debug

__all__ = ["Telescope"]


class Telescope:
    """IACT telescope: mirror groups, obstruction groups, sensors, name."""

    def __init__(self, mirror_groups, obstruction_groups=None, sensors=None, name="telescope"):
        self.mirror_groups = list(mirror_groups)
        self.obstruction_groups = obstruction_groups
        self.sensors = list(sensors) if sensors else []
        self.name = name
        self.launch = None

    def __call__(self, sources, values, source_type="point", sensor_idx=0, debug=False):
        """Render sources (telescope.py:56-83): image, or (points, values) if debug."""
        if debug:
            return render_debug(self, sources, values, source_type, sensor_idx)
        return render(self, sources, values, source_type, sensor_idx)

    @classmethod
    def from_yaml(cls, filename, integrator, key=None):
        from .yaml_loader import load_telescope
        return load_telescope(filename, integrator, key)

    def to(self, device):
        """Copy with every mirror group's arrays on ``device``."""
        t = Telescope([g.to(device) for g in self.mirror_groups], self.obstruction_groups,
                      self.sensors, self.name)
        t.launch = self.launch
        return t

    def with_launch(self, launch):
        """Copy that renders with the given LaunchConfig (e.g. from the autotuner)."""
        t = Telescope(self.mirror_groups, self.obstruction_groups, self.sensors, self.name)
        t.launch = launch
        return t

    # -- convenience wrappers (telescope.py:99-237) ---------------------------
    def resample_mirrors(self, integrator, key):
        from .operations import resample_mirrors
        return resample_mirrors(self, integrator, key)

    def set_mirror_positions(self, group_idx, positions):
        from .operations import set_mirror_positions
        return set_mirror_positions(self, group_idx, positions)

    def set_mirror_rotations(self, group_idx, rotations):
        from .operations import set_mirror_rotations
        return set_mirror_rotations(self, group_idx, rotations)

    def scale_mirror_weights(self, group_idx, scale_factors):
        from .operations import scale_mirror_weights
        return scale_mirror_weights(self, group_idx, scale_factors)

    def apply_roughness(self, roughness_arcsec):
        from .operations import apply_roughness
        return apply_roughness(self, roughness_arcsec)

    def apply_roughness_to_group(self, group_idx, roughness):
        from .operations import apply_roughness_to_group
        return apply_roughness_to_group(self, group_idx, roughness)

    def apply_misalignment_to_group(self, group_idx, sigma_h, sigma_v, key):
        from .operations import apply_misalignment_to_group
        return apply_misalignment_to_group(self, group_idx, sigma_h, sigma_v, key)

    def apply_displacement_to_group(self, group_idx, sigma_z, key):
        from .operations import apply_displacement_to_group
        return apply_displacement_to_group(self, group_idx, sigma_z, key)

    def get_mirrors_by_stage(self, stage):
        from .operations import get_mirrors_by_stage
        return get_mirrors_by_stage(self, stage)

    def get_mirror_count(self):
        from .operations import get_mirror_count
        return get_mirror_count(self)

    def add_sensor(self, sensor):
        from .operations import add_sensor
        return add_sensor(self, sensor)

    def replace_sensor(self, sensor, idx=0):
        from .operations import replace_sensor
        return replace_sensor(self, sensor, idx)

    def remove_sensor(self, idx=0):
        from .operations import remove_sensor
        return remove_sensor(self, idx)

    def set_sensor_position(self, idx, position):
        from .operations import set_sensor_position
        return set_sensor_position(self, idx, position)

    def set_sensor_rotation(self, idx, rotation):
        from .operations import set_sensor_rotation
        return set_sensor_rotation(self, idx, rotation)

    def focus(self, delta_z, sensor_idx=0):
        from .operations import focus
        return focus(self, delta_z, sensor_idx)

    def get_sensor_count(self):
        return len(self.sensors)

    def add_obstruction(self, obstruction):
        from .operations import add_obstruction
        return add_obstruction(self, obstruction)

    def remove_obstruction(self, group_idx):
        from .operations import remove_obstruction
        return remove_obstruction(self, group_idx)

    def clear_obstructions(self):
        from .operations import clear_obstructions
        return clear_obstructions(self)

    def get_obstruction_count(self):
        from .operations import get_obstruction_count
        return get_obstruction_count(self)

    def clone(self):
        from .operations import clone
        return clone(self)

    def get_info(self):
        from .operations import get_info
        return get_info(self)
