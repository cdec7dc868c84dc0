"""DLPack exchange of render inputs (host side; the GPU variant is in
test_gpu_forward.py)."""
import numpy as np
import torch

from iactrace_amd.dlpack import as_tensor


class Exporter:
    """An array that only exports DLPack (as CuPy or JAX arrays do)."""

    def __init__(self, t):
        self._t = t

    def __dlpack__(self, stream=None):
        return self._t.__dlpack__()

    def __dlpack_device__(self):
        return self._t.__dlpack_device__()


def test_dlpack_inputs_are_wrapped_without_copy():
    t = torch.arange(12, dtype=torch.float64).reshape(4, 3)
    v = as_tensor(Exporter(t))
    assert isinstance(v, torch.Tensor)
    assert v.data_ptr() == t.data_ptr() and v.dtype == t.dtype and torch.equal(v, t)
    assert as_tensor(t) is t
    a = np.linspace(0.0, 1.0, 5, dtype=np.float32)
    assert torch.equal(as_tensor(a), torch.from_numpy(a))
    assert torch.equal(as_tensor([1.0, 2.0]), torch.tensor([1.0, 2.0]))
