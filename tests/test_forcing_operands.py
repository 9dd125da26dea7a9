"""Host logic of the forcing hook (model.py:121-123): the views handed to a
forcing function accept the numpy operands the reference's forcing
experiments use (forced_convection.py:9-23, gyre.py:14-17,
internalwave_maker.py:17-18).  'meta' tensors stand in for device tensors:
like those, they cannot be turned into numpy arrays, so a numpy fallback
would fail here as it fails on the device."""
import warnings
from collections import namedtuple

import numpy as np
import pytest
import torch

from fluids2d_amd.integrators import _host_operands

V = namedtuple("vector", "x y")
S = namedtuple("state", "pv u")


@pytest.mark.parametrize("device", ["cpu", "meta"])
def test_numpy_operands(device):
    s = S(torch.zeros(4, 5, dtype=torch.float32, device=device),
          V(torch.zeros(4, 5, dtype=torch.float64, device=device),
            torch.zeros(4, 5, dtype=torch.float64, device=device)))
    w = _host_operands(s)
    with warnings.catch_warnings():
        warnings.simplefilter("error")  # numpy's __array_wrap__ path warns
        w.pv[:, :] += np.ones((4, 5))
        w.u.y[1:-1, 2:] -= 2.5 * np.tanh(0.3)
        w.u.x[:] += np.arange(5.0) * np.float64(0.5)
        w.u.x[0, :] = np.arange(5.0) * 2
        w.u.y[:, :] += np.ones((4, 5)) * w.u.x     # array on the left
        w.u.y[:, :] -= w.u.x / np.full((4, 5), 2.0)
    if device == "cpu":
        assert torch.equal(s.pv, torch.ones(4, 5))
        ux = np.tile(np.arange(5.0) * 0.5, (4, 1))
        ux[0] = np.arange(5.0) * 2
        uy = np.zeros((4, 5))
        uy[1:-1, 2:] -= 2.5 * np.tanh(0.3)
        uy += ux
        uy -= ux / 2.0
        assert np.array_equal(s.u.x.numpy(), ux)
        assert np.array_equal(s.u.y.numpy(), uy)
    # the views share the state's storage
    assert w.pv.data_ptr() == s.pv.data_ptr()
