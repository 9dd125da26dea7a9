"""The euler model (momentum-pressure, equations.py:9-24) on the device:
vortex force and kinetic-energy gradient, the pressure projection with the
Neumann cell Poisson solve (direct on all-wet power-of-two domains,
Jacobi-PCG in FP64 otherwise) against the oracle's SuperLU.
The pressure is defined up to a constant (singular Neumann operator), so it
is compared after removing its mean; everything else directly."""
import numpy as np
import pytest

import oracle as O
from helpers import P, random_mask
from test_oracle_operators import euler_dipole

pytestmark = pytest.mark.gpu

KEYS = ["model", "nx", "ny", "Lx", "Ly", "xperiodic", "yperiodic", "halowidth", "noslip",
        "tend", "dt", "maxite", "dtmax", "integrator", "cfl", "compflux", "vortexforce",
        "innerproduct", "maxorder"]


def _device(p_kw, msk=None, precision="fp64"):
    import fluids2d_amd as f2d
    import fluids2d_amd.param as fp
    fp.QUIET = True
    p = f2d.Param()
    for k, v in p_kw.items():
        setattr(p, k, v)
    p.precision = precision
    m = f2d.Model(p)
    if msk is not None:
        m.mesh.msk[...] = msk
        m.mesh.finalize()
        m.integrator.model.mesh = m.mesh
    return m


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_euler_dipole_matches_oracle(dev):
    """The test_euler case (default Param, dipole, tend = 10): same number
    of iterations as the oracle and the same final state."""
    from fluids2d_amd.tools import set_uv_from_omega
    ref, omega = euler_dipole(O, P(model="euler", tend=10))
    ref.run()
    m = _device(dict(model="euler", tend=10))
    m.state.omega[...] = omega
    set_uv_from_omega(m, omega, m.state.u)
    m.integrator.diag(m.state)
    m.run()
    print(f"\neuler dipole: device {m.time.ite} iterations, oracle {ref.time.ite}")
    assert m.time.ite == ref.time.ite
    for c in ("x", "y"):
        e = _rel(getattr(m.state.u, c), getattr(ref.state.u, c))
        print(f"  u.{c} rel L2 {e:.2e}")
        assert e < 1e-9, (c, e)
    assert _rel(m.state.omega, ref.state.omega) < 1e-9
    assert abs(m.time.t - ref.time.t) < 1e-9 * ref.time.t


@pytest.mark.parametrize("nx,ny,nisl", [(48, 40, 3), (64, 32, 0)])
def test_euler_masked_parity(dev, nx, ny, nisl):
    p = dict(model="euler", nx=nx, ny=ny, Lx=nx / ny, Ly=1.0, tend=1e9)
    msk = random_mask(ny, nx, 6, False, False, nisl=nisl) if nisl else None
    m = _device(p, msk)
    ref = O.Model(P(**{k: p.get(k, getattr(P(), k)) for k in KEYS if k in p or hasattr(P(), k)}),
                  msk=None if msk is None else msk.copy())
    rng = np.random.default_rng(2)
    ux = 0.05 * rng.standard_normal(m.mesh.shape) * m.mesh.mskx
    uy = 0.05 * rng.standard_normal(m.mesh.shape) * m.mesh.msky
    for s in (m.state, ref.state):
        s.u.x[...] = ux
        s.u.y[...] = uy
    m.integrator.diag(m.state)
    ref.diag(ref.state)
    m.param.maxite = 12
    m.run()
    ref.run(12)
    assert m.time.ite == ref.time.ite == 12
    wet = m.mesh.msk == 1
    for c in ("x", "y"):
        e = _rel(getattr(m.state.u, c), getattr(ref.state.u, c))
        print(f"\neuler {nx}x{ny} islands={nisl}: u.{c} rel L2 {e:.2e}")
        assert e < 1e-9, (c, e)
    pm = m.state.p[wet] - m.state.p[wet].mean()
    pr = ref.state.p[wet] - ref.state.p[wet].mean()
    assert _rel(pm, pr) < 1e-8
    # the projected velocity is divergence free on the wet cells
    d = np.zeros(m.mesh.shape)
    O.div(ref.mesh, m.state.U, d)
    assert np.abs(d[wet]).max() < 1e-9 * np.abs(m.state.U.x).max()
    st = m.mesh.poisson_centers.read_stats()
    assert st["failed"] == 0 and st["solves"] > 0


def test_euler_fp32_close_to_fp64(dev):
    p = dict(model="euler", nx=48, ny=48, tend=1e9, maxite=10)
    runs = []
    for prec in ("fp64", "fp32"):
        m = _device(p, precision=prec)
        rng = np.random.default_rng(3)
        m.state.u.x[...] = 0.05 * rng.standard_normal(m.mesh.shape) * m.mesh.mskx
        m.state.u.y[...] = 0.05 * rng.standard_normal(m.mesh.shape) * m.mesh.msky
        m.integrator.diag(m.state)
        m.run()
        runs.append(m.state.u.x.copy())
    e = _rel(runs[1], runs[0])
    print(f"\neuler FP32 vs FP64 after 10 steps: {e:.2e}")
    assert e < 1e-5
