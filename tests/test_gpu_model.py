"""Whole-model parity on the GPU: the drop-in Model against the oracle's
FP64 restatement of the reference, on identical seeded inputs.

Tolerances (the FP64 parity bar): relative L2 of the prognostic field
<= 1e-12 after 100 FP64 steps; the reference's own known-answer test
(tests/test_models.py:17-26, advection body rotation: 37 iterations) exact.
"""
import numpy as np
import pytest

import oracle as O
from helpers import P

pytestmark = pytest.mark.gpu

PARITY_RTOL = 1e-12


def _param(**kw):
    import fluids2d_amd as f2d
    import fluids2d_amd.param as fp
    fp.QUIET = True
    p = f2d.Param()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _oracle_param(p):
    keys = ["model", "nx", "ny", "Lx", "Ly", "xperiodic", "yperiodic", "halowidth",
            "noslip", "f0", "beta", "g", "H", "tend", "dt", "maxite", "dtmax",
            "integrator", "cfl", "compflux", "vortexforce", "innerproduct", "maxorder",
            "RAgamma"]
    return P(**{k: getattr(p, k) for k in keys})


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_kat_advection_bodyrotation_37():
    """tests/test_models.py:17-26 with tracer_advection.py:10-48 set-up."""
    import fluids2d_amd as f2d
    p = _param(model="advection", animation=False, tend=10)
    model = f2d.Model(p)
    mesh = model.mesh
    omega = (4 * np.pi / 100) * np.ones(mesh.shape) * mesh.mskv
    f2d.tools.set_uv_from_omega(model, omega, model.state.U)
    x, y = mesh.xy("c")
    q = model.state.q
    q[:, :] = (np.round(x * 8) % 2 + np.round(y * 8) % 2) / 2
    q *= mesh.msk
    model.run()
    assert model.time.ite == 37


def _vortex_ic(mesh, seed, amp_noise=1e-3):
    """Config-1 initial vorticity: Gaussian vortex (amplitude 1, radius 0.1L,
    centred) + seeded uniform noise, times msk*area (vortex.py:23-26)."""
    x, y = mesh.xy("c")
    rng = np.random.default_rng(seed)
    w = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / (2 * 0.1 ** 2))
    w = w + amp_noise * rng.uniform(-1, 1, w.shape)
    return w * mesh.msk * mesh.area


def _run_pair(p, ic, nsteps, msk=None):
    import fluids2d_amd as f2d
    model = f2d.Model(p)
    if msk is not None:
        model.mesh.msk[...] = msk
        model.mesh.finalize()
    w0 = ic(model.mesh)
    getattr(model.state, model.integrator.model.prognostic[0])[...] = w0
    model.integrator.diag(model.state)
    p.maxite = nsteps
    model.run()
    ref = O.Model(_oracle_param(p), msk=model.mesh.msk.copy())
    name = model.integrator.model.prognostic[0]
    getattr(ref.state, name)[...] = w0
    ref.diag(ref.state)
    ref.run(nsteps)
    return model, ref, name


@pytest.mark.parametrize("nx,xper,yper,islands,nsteps", [
    (64, False, False, False, 100), (64, True, True, False, 100),
    (96, True, False, True, 60)])
def test_eulerpsi_fp64_parity_small(nx, xper, yper, islands, nsteps):
    p = _param(model="eulerpsi", nx=nx, ny=nx, xperiodic=xper, yperiodic=yper, cfl=0.8,
               tend=1e9)
    msk = None
    if islands:
        from helpers import random_mask
        msk = random_mask(nx, nx, 1, xper, yper, nisl=4)
    model, ref, name = _run_pair(p, lambda m: _vortex_ic(m, 0), nsteps, msk)
    e = relerr(model.state.omega, ref.state.omega)
    print(f"\neulerpsi {nx}^2 xper={xper} yper={yper} isl={islands}: "
          f"{nsteps} steps rel L2(omega) = {e:.2e}, t={model.time.t:.6f}/{ref.time.t:.6f}")
    assert model.time.ite == ref.time.ite == nsteps
    assert e <= PARITY_RTOL


def test_eulerpsi_fp64_parity_config1():
    """256^2 closed square, Gaussian vortex + 1e-3 noise, CFL 0.8, 100 steps."""
    p = _param(model="eulerpsi", nx=256, ny=256, cfl=0.8, tend=1e9)
    model, ref, name = _run_pair(p, lambda m: _vortex_ic(m, 1234), 100)
    e = relerr(model.state.omega, ref.state.omega)
    ePsi = relerr(model.state.psi, ref.state.psi)
    print(f"\nconfig1 256^2: rel L2 omega {e:.2e} psi {ePsi:.2e} "
          f"t={model.time.t!r} ref t={ref.time.t!r}")
    assert e <= PARITY_RTOL


def test_qg_fp64_parity():
    p = _param(model="qg", nx=64, ny=64, cfl=0.8, tend=1e9, f0=10.0)
    model, ref, name = _run_pair(p, lambda m: _vortex_ic(m, 7), 50)
    e = relerr(model.state.pv, ref.state.pv)
    print(f"\nqg 64^2: 50 steps rel L2(pv) = {e:.2e}")
    assert e <= PARITY_RTOL
