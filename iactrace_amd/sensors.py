"""Sensors (iactrace/sensors/square.py, hexagonal.py).

The hard-binning sensors and their Gaussian-splat (differentiable) variants.
Set-up work -- the hexagonal lattice detection and axial lookup table -- runs
once on the host in fp32 with the reference's operation order; binning and
the scatter-add run in the HIP kernels (``accumulate`` calls
``iact_sensor_accumulate``).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .geometry import euler_to_matrix

__all__ = ["SquareSensor", "HexagonalSensor", "DifferentiableSquareSensor",
           "DifferentiableHexagonalSensor"]

F32 = np.float32
SQRT3 = 1.7320508075688772
SQRT3_2 = 0.8660254037844386
SQRT3_3 = 0.5773502691896257


class _SensorBase:
    splat = 0

    def _frame(self, s):
        s.position[:] = [float(v) for v in np.asarray(self.position, np.float32).reshape(3)]
        R = euler_to_matrix(np.asarray(self.rotation, np.float32))
        s.rotation[:] = [float(v) for v in R.reshape(-1)]

    def accumulate(self, x, y, values):
        """Scatter-add of ``values`` at sensor-plane points (x, y) on the GPU."""
        x = torch.as_tensor(x, dtype=torch.float32).contiguous()
        dev = x.device
        if dev.type != "cuda":
            raise RuntimeError("Sensor.accumulate runs on the HIP device; move x, y, values there")
        y = torch.as_tensor(y, dtype=torch.float32, device=dev).contiguous()
        v = torch.as_tensor(values, dtype=torch.float32, device=dev).contiguous()
        out = torch.empty(self.n_pixels, dtype=torch.float32, device=dev)
        desc = self.to_abi(dev)
        _lib.check(_lib.lib().iact_sensor_accumulate(
            ctypes.byref(desc), _lib.ptr(x), _lib.ptr(y), _lib.ptr(v), x.numel(), _lib.ptr(out),
            None, _lib.current_stream()), "iact_sensor_accumulate")
        return out.reshape(self.get_accumulator_shape())

    def pixel_ids(self, x, y):
        """Hard pixel id (or -1) of each point, from the same device binning."""
        x = torch.as_tensor(x, dtype=torch.float32).contiguous()
        dev = x.device
        y = torch.as_tensor(y, dtype=torch.float32, device=dev).contiguous()
        v = torch.zeros_like(x)
        out = torch.empty(self.n_pixels, dtype=torch.float32, device=dev)
        pid = torch.empty(x.numel(), dtype=torch.int32, device=dev)
        desc = self.to_abi(dev)
        _lib.check(_lib.lib().iact_sensor_accumulate(
            ctypes.byref(desc), _lib.ptr(x), _lib.ptr(y), _lib.ptr(v), x.numel(), _lib.ptr(out),
            _lib.ptr(pid), _lib.current_stream()), "iact_sensor_accumulate")
        return pid


class SquareSensor(_SensorBase):
    """Square-pixel sensor with hard binning and an edge mask (square.py:28-91)."""

    def __init__(self, position, rotation, width, height, bounds, edge_width=0.0):
        self.position = np.asarray(position, np.float32).reshape(3)
        self.rotation = np.asarray(rotation, np.float32).reshape(3)
        self.width = int(width)
        self.height = int(height)
        self.edge_width = float(edge_width)
        xmin, xmax, ymin, ymax = [float(b) for b in bounds]
        self.bounds = (xmin, xmax, ymin, ymax)
        self.x0 = float(xmin)
        self.y0 = float(ymin)
        self.dx = float((xmax - xmin) / self.width)
        self.dy = float((ymax - ymin) / self.height)

    @property
    def n_pixels(self):
        return self.width * self.height

    def get_accumulator_shape(self):
        return (self.height, self.width)

    def replace(self, **kw):
        args = dict(position=self.position, rotation=self.rotation, width=self.width,
                    height=self.height, bounds=self.bounds, edge_width=self.edge_width)
        args.update(kw)
        return type(self)(**args)

    def to_abi(self, device=None):
        s = _lib.IactSensor()
        s.kind = _lib.SENSOR_SQUARE
        s.splat = self.splat
        s.n_pixels = self.n_pixels
        self._frame(s)
        s.width, s.height = self.width, self.height
        s.x0, s.y0, s.dx, s.dy = self.x0, self.y0, self.dx, self.dy
        s.edge_width = self.edge_width
        s.sigma = getattr(self, "sigma", 0.0)
        s.kernel_size = getattr(self, "kernel_size", 0)
        return s


class DifferentiableSquareSensor(SquareSensor):
    """Gaussian-splat square sensor (square.py:94-173): every ray spreads over a
    (2k+1)^2 neighbourhood with normalised weights exp(-d^2 / 2 sigma^2)."""
    splat = 1

    def __init__(self, position, rotation, width, height, bounds=(-1, 1, -1, 1), sigma=0.1,
                 kernel_size=2):
        super().__init__(position, rotation, width, height, bounds, 0.0)
        self.sigma = float(sigma)
        self.kernel_size = int(kernel_size)

    def replace(self, **kw):
        args = dict(position=self.position, rotation=self.rotation, width=self.width,
                    height=self.height, bounds=self.bounds, sigma=self.sigma,
                    kernel_size=self.kernel_size)
        args.update(kw)
        return type(self)(**args)


def _rotate(x, y, angle):
    c = np.cos(np.float32(angle)).astype(F32)
    s = np.sin(np.float32(angle)).astype(F32)
    return c * x - s * y, s * x + c * y


def detect_hex_grid(centers):
    """(size, rotation, offset) of a hex pixel lattice (hexagonal.py:61-80), fp32."""
    c = np.asarray(centers, F32)
    n = len(c)
    diff = c[:, None, :] - c[None, :, :]
    d2 = diff[..., 0] ** 2 + diff[..., 1] ** 2
    d2[np.arange(n), np.arange(n)] = np.inf
    idx = int(np.argmin(d2))
    min_dist = np.sqrt(d2.reshape(-1)[idx])
    vec = diff[idx // n, idx % n]
    angle = np.mod(np.arctan2(vec[1], vec[0]), F32(np.pi / 3))
    offset = c[int(np.argmin(c[:, 0] ** 2 + c[:, 1] ** 2))]
    return float(min_dist / F32(SQRT3)), float(angle), offset


def build_lookup_table(centers, hex_size, rotation, offset):
    """Axial (q, r) -> pixel index table, -1 = empty (hexagonal.py:83-101)."""
    c = np.asarray(centers, F32)
    x = c[:, 0] - offset[0]
    y = c[:, 1] - offset[1]
    xr, yr = _rotate(x, y, -rotation)
    q = (F32(SQRT3_3) * xr - yr / 3) / F32(hex_size)
    r = (2 * yr / 3) / F32(hex_size)
    qi = np.round(q).astype(np.int32)
    ri = np.round(r).astype(np.int32)
    q_min, q_max, r_min, r_max = int(qi.min()), int(qi.max()), int(ri.min()), int(ri.max())
    table = np.full((q_max - q_min + 1, r_max - r_min + 1), -1, np.int32)
    table[qi - q_min, ri - r_min] = np.arange(len(c), dtype=np.int32)
    return table, q_min, r_min


class HexagonalSensor(_SensorBase):
    """Hexagonal-pixel sensor with hard binning (hexagonal.py:104-194)."""

    def __init__(self, position, rotation, hex_centers, edge_width=0.0):
        self.position = np.asarray(position, np.float32).reshape(3)
        self.rotation = np.asarray(rotation, np.float32).reshape(3)
        self.hex_centers = np.asarray(hex_centers, F32).reshape(-1, 2)
        self.edge_width = float(edge_width)
        size, rot, offset = detect_hex_grid(self.hex_centers)
        self.hex_size = float(size)
        self.hex_inradius = float(size * SQRT3_2)
        self.grid_rotation = float(rot)
        self.grid_offset = (float(offset[0]), float(offset[1]))
        self.lookup_table, self.q_min, self.r_min = build_lookup_table(
            self.hex_centers, self.hex_size, self.grid_rotation, offset)
        self._dev_tables = {}

    @property
    def n_pixels(self):
        return len(self.hex_centers)

    def get_accumulator_shape(self):
        return (self.n_pixels,)

    def replace(self, **kw):
        args = dict(position=self.position, rotation=self.rotation, hex_centers=self.hex_centers,
                    edge_width=self.edge_width)
        args.update(kw)
        return type(self)(**args)

    def device_table(self, device):
        key = str(device)
        if key not in self._dev_tables:
            self._dev_tables[key] = torch.as_tensor(self.lookup_table).to(device).contiguous()
        return self._dev_tables[key]

    def to_abi(self, device=None):
        s = _lib.IactSensor()
        s.kind = _lib.SENSOR_HEX
        s.splat = self.splat
        s.n_pixels = self.n_pixels
        self._frame(s)
        s.hex_size = self.hex_size
        s.hex_inradius = self.hex_inradius
        s.grid_rotation = self.grid_rotation
        s.grid_offset[:] = list(self.grid_offset)
        s.edge_threshold = float(F32(1.0 - self.edge_width / self.hex_inradius))
        s.q_min, s.r_min = self.q_min, self.r_min
        s.table_q, s.table_r = self.lookup_table.shape
        if device is not None and torch.device(device).type == "cuda":
            s.lookup_table = self.device_table(device).data_ptr()
        s.sigma = getattr(self, "sigma", 0.0)
        s.kernel_size = getattr(self, "kernel_size", 0)
        return s


class DifferentiableHexagonalSensor(HexagonalSensor):
    """Gaussian-splat hexagonal sensor over ``kernel_size`` rings
    (hexagonal.py:197-315); sigma is in units of the hex inradius."""
    splat = 1

    def __init__(self, position, rotation, hex_centers, sigma=0.5, kernel_size=1):
        super().__init__(position, rotation, hex_centers, 0.0)
        self.sigma = float(sigma)
        self.kernel_size = int(kernel_size)

    def replace(self, **kw):
        args = dict(position=self.position, rotation=self.rotation, hex_centers=self.hex_centers,
                    sigma=self.sigma, kernel_size=self.kernel_size)
        args.update(kw)
        return type(self)(**args)
