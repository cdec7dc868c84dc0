"""Tensor exchange by DLPack.

The renders take sources, values and batched rotations as any array that
exports DLPack (`__dlpack__` / `__dlpack_device__`: CuPy, JAX, TensorFlow,
torch tensors on ROCm devices ...): the buffer is wrapped without a copy and
handed to the library as a device pointer.  Outputs are torch tensors, which
export DLPack themselves (`torch.utils.dlpack.to_dlpack`, or `__dlpack__` on
the consumer's side).
"""
from __future__ import annotations

import numpy as np
import torch

__all__ = ["as_tensor"]


def as_tensor(x):
    """A torch view of x: torch tensors as they are, objects exporting DLPack
    through torch.from_dlpack (zero-copy, same device), anything else
    (lists, numpy arrays) through torch.as_tensor."""
    if isinstance(x, torch.Tensor):
        return x
    if hasattr(x, "__dlpack__") and not isinstance(x, np.ndarray):
        return torch.from_dlpack(x)
    return torch.as_tensor(x)
