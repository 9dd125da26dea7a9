"""f2d_perpgrad_cfl_noalloc, the allocation-free variant of f2d_perpgrad_ex
with the CFL step, against f2d_perpgrad_ex through the C ABI: velocities,
maxima and dt bit for bit, on two calls in a row, the second on a smaller
grid (the fixed scratch keeps the first call's partials, which the second
must not see), and a grid beyond the scratch's limit refused."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _grid(n, ny, xper, yper):
    import fluids2d_amd._lib as L
    g = L.Grid()
    g.ny, g.nx, g.halowidth = ny, n, 3
    g.xperiodic, g.yperiodic = int(xper), int(yper)
    g.dx, g.dy = 1.0 / n, 1.0 / ny
    g.tile = 0
    return g


def _inputs(n, ny, seed, dtype, dev):
    rng = np.random.default_rng(seed)
    sh = (ny + 6, n + 6)
    psi = torch.from_numpy(rng.standard_normal(sh)).to(dev, dtype)
    mx = torch.from_numpy((rng.random(sh) > 0.1).astype(np.int8)).to(dev)
    my = torch.from_numpy((rng.random(sh) > 0.1).astype(np.int8)).to(dev)
    return psi, mx, my


def _run(name, n, ny, xper, yper, psi, mx, my, dtype):
    import fluids2d_amd._lib as L
    dev = psi.device
    Ux = torch.full_like(psi, 7.0)
    Uy = torch.full_like(psi, -3.0)
    mxa = torch.full((2,), -1.0, dtype=torch.float64, device=dev)
    dt = torch.full((1,), -1.0, dtype=torch.float64, device=dev)
    g = _grid(n, ny, xper, yper)
    L.call(name, L.dtype_code(dtype), ctypes.byref(g), L.ptr(psi), L.ptr(Ux), L.ptr(Uy),
           L.ptr(mx), L.ptr(my), float(n * n), float(ny * ny), 1, L.ptr(mxa), L.ptr(dt),
           0.9, 9e99, L.stream_ptr())
    torch.cuda.synchronize()
    return Ux.cpu(), Uy.cpu(), mxa.cpu(), dt.cpu()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_noalloc_matches_perpgrad_ex(dev, dtype):
    cases = [(1024, 768, True, True, 1), (96, 40, True, False, 2)]  # the second smaller
    for n, ny, xp, yp, seed in cases:
        psi, mx, my = _inputs(n, ny, seed, dtype, dev)
        ref = _run("f2d_perpgrad_ex", n, ny, xp, yp, psi, mx, my, dtype)
        got = _run("f2d_perpgrad_cfl_noalloc", n, ny, xp, yp, psi, mx, my, dtype)
        for a, b, what in zip(got, ref, ("Ux", "Uy", "maxabs", "dt")):
            assert torch.equal(a, b), (n, ny, what)
        assert float(ref[3][0]) > 0


def test_noalloc_refuses_beyond_limit(dev):
    import fluids2d_amd._lib as L
    n = 2900  # (2906)^2 padded points > 8192 * 1024
    g = _grid(n, n, True, True)
    one = torch.zeros(1, dtype=torch.float64, device=dev)
    with pytest.raises(L.F2DError, match="fixed scratch"):
        L.call("f2d_perpgrad_cfl_noalloc", L.dtype_code(torch.float64), ctypes.byref(g),
               L.ptr(one), L.ptr(one), L.ptr(one), L.ptr(one), L.ptr(one), 1.0, 1.0, 1,
               None, L.ptr(one), 0.9, 9e99, L.stream_ptr())


def test_noalloc_needs_a_reduction(dev):
    import fluids2d_amd._lib as L
    g = _grid(16, 16, True, True)
    one = torch.zeros(1, dtype=torch.float64, device=dev)
    with pytest.raises(L.F2DError, match="NULL"):
        L.call("f2d_perpgrad_cfl_noalloc", L.dtype_code(torch.float64), ctypes.byref(g),
               L.ptr(one), L.ptr(one), L.ptr(one), L.ptr(one), L.ptr(one), 1.0, 1.0, 1,
               None, None, 0.9, 9e99, L.stream_ptr())
