"""Failure detection on the device path (no counterpart in the reference,
whose loop has no NaN check and stops only on ctrl-C, model.py:33-45):
f2d_nonfinite flags non-finite values on wet cells, and Model.run checks
the flag at its synchronisation points (history output, host diagnostics,
the end of the run).  A cell poisoned from the host between two steps stops
the run cleanly at the next such point: the loop ends at a step boundary,
the host state is copied back and the diagnostics are finalised."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _model(maxite):
    import fluids2d_amd as f2d
    import fluids2d_amd.param as fp
    fp.QUIET = True
    p = f2d.Param()
    p.model = "eulerpsi"
    p.nx = p.ny = 64
    p.xperiodic = p.yperiodic = False
    p.tend = 1e9
    p.maxite = maxite
    p.nprint = 10 ** 6
    m = f2d.Model(p)
    x, y = m.mesh.xy("c")
    m.state.omega[...] = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / 0.02) * m.mesh.msk \
        * m.mesh.area
    m.integrator.diag(m.state)
    return m


class _Diag:
    """A host diagnostic that poisons one wet cell of the device vorticity
    after step `at` (None: never)."""
    device = False

    def __init__(self, model, at):
        self.model, self.at = model, at
        self.calls, self.finalized = 0, False

    def __call__(self):
        self.calls += 1
        if self.at is not None and self.model.time.ite == self.at:
            self.model.integrator.dstate.omega[20, 20] = float("nan")

    def finalize(self):
        self.finalized = True


def test_nonfinite_flag(dev):
    from fluids2d_amd import _lib
    m = _model(1)
    g = m.mesh.grid_ref()
    a = torch.zeros(m.mesh.shape, dtype=torch.float64, device=dev)
    flag = torch.zeros(1, dtype=torch.int32, device=dev)
    for poison, wet, want in ((None, True, 0), (float("inf"), False, 0),
                              (float("nan"), True, 1), (float("-inf"), True, 1)):
        a.zero_()
        flag.zero_()
        if poison is not None:
            j, i = (20, 20) if wet else (1, 1)  # (1, 1) is a dry halo cell
            assert bool(m.mesh.msk[j, i]) == wet
            a[j, i] = poison
        _lib.call("f2d_nonfinite", _lib.dtype_code(a.dtype), g, _lib.ptr(a),
                  _lib.ptr(m.mesh.d_msk), _lib.ptr(flag), _lib.stream_ptr())
        assert int(flag.item()) == want, (poison, wet)


def test_poisoned_cell_stops_the_run_cleanly(dev):
    m = _model(30)
    d = _Diag(m, at=7)
    m.diags.append(d)
    m.run()
    assert m.stop and m.nonfinite == ["omega"]
    assert m.time.ite == 8  # the next synchronisation point after the poison
    assert d.finalized and d.calls == 7
    assert not np.isfinite(m.state.omega[3:-3, 3:-3]).all()  # the host copy was made


def test_clean_run_reports_nothing(dev):
    m = _model(12)
    d = _Diag(m, at=None)
    m.diags.append(d)
    m.run()
    assert not m.stop and m.nonfinite == [] and m.time.ite == 12
    assert d.finalized and np.isfinite(m.state.omega).all()
