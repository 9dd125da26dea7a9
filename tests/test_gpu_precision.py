"""FP32 stencils and solves with FP64 refinement against the FP64 oracle, for
the two workloads that run in that precision, each with its error bar:

* the doubly periodic Euler workload (eulerpsi, WENO5, SSP-RK3, random-phase
  k^-3 exp(-(k/k0)^2) vorticity, k0 = 8, unit RMS) at 128^2 for 1000 steps
  at the fixed step of CFL 0.9 on the initial state: relative L2 of the
  vorticity <= 1e-5 (the bar of the 1024^2 target, there after 1000 steps);
* eulerpsi in a closed smoothed-noise basin (capacitance solve) at 128^2
  for 500 steps: relative L2 of the vorticity <= 1e-5;
* a qgrsw member of the ensemble workload (closed smoothed-noise basin,
  f0 = 10, g = H = 1, Gaussian h bumps balanced by the QG projection,
  geos_adj.py:37-49) at 64^2 and 96x128 for 200 steps: relative L2 of u and
  h <= 1e-5;
* the Euler workload at its full 1024^2 size for 1000 steps against the
  FP64 device path: relative L2 of the vorticity <= 1e-5.

The oracle runs the reference's algorithm in FP64 (SuperLU solves).  The
device runs FP32 arithmetic, stepped two ways: ``Model.run`` (host-checked
solves, each refined in FP64) and ``device_steps`` (graph-replayed steps
whose solves are plain FP32 in the first RK stages and refined once in FP64
at the last, ``Param.poisson_refine = "last"``).
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
from helpers import P

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

KEYS = ["model", "nx", "ny", "Lx", "Ly", "xperiodic", "yperiodic", "halowidth", "noslip",
        "f0", "beta", "g", "H", "tend", "dt", "maxite", "dtmax", "integrator", "cfl",
        "compflux", "vortexforce", "innerproduct", "maxorder", "RAgamma"]


def _param(**kw):
    import fluids2d_amd as f2d
    import fluids2d_amd.param as fp
    fp.QUIET = True
    p = f2d.Param()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _steps(m, nsteps, stepping):
    """nsteps steps by Model.run or by device_steps (then the host state and
    the time pushed forward by the steps' dt, as Model.run leaves them)."""
    if stepping == "run":
        m.run()
        return
    integ = m.integrator
    integ.begin(m.state)
    dts = integ.device_steps(nsteps).cpu().numpy()
    integ.end(m.state)
    for v in dts:
        m.time.dt = float(v)
        m.time.pushforward()


@pytest.mark.parametrize("stepping", ["run", "device_steps"])
def test_euler_periodic_fp32_refined_1000_steps(dev, stepping):
    import fluids2d_amd as f2d
    from bench import initial_omega
    n, nsteps = 128, 1000
    p = _param(model="eulerpsi", nx=n, ny=n, xperiodic=True, yperiodic=True,
               compflux="weno", integrator="rk3", cfl=0.9, tend=1e30, maxite=nsteps)
    ref = O.Model(P(**{k: getattr(p, k) for k in KEYS}))
    w0 = initial_omega(ref.mesh, 2024)
    ref.state.omega[...] = w0
    ref.diag(ref.state)
    ref.set_dt()
    dt = ref.time.dt
    ref.param.dt = dt
    p.dt = dt
    p.precision = "fp32"
    m = f2d.Model(p)
    m.state.omega[...] = w0
    m.integrator.diag(m.state)
    print(f"\n128^2 eulerpsi, dt {dt:.4e}: device FP32 + refinement ({stepping}), "
          "oracle FP64", flush=True)
    _steps(m, nsteps, stepping)
    ref.run(nsteps)
    assert m.time.ite == ref.time.ite == nsteps
    e = _rel(m.state.omega[3:-3, 3:-3], ref.state.omega[3:-3, 3:-3])
    print(f"  vorticity rel L2 after {nsteps} steps: {e:.2e} (bar 1e-5)")
    assert e <= 1e-5


@pytest.mark.parametrize("stepping", ["run", "device_steps"])
def test_euler_basin_fp32_refined(dev, stepping):
    """eulerpsi in a closed smoothed-noise basin (the capacitance solve):
    128^2, Gaussian vortex + noise, 500 steps at the fixed step of CFL 0.8
    on the initial state, vorticity within 1e-5 of the FP64 oracle."""
    import fluids2d_amd as f2d
    from bench import basin_mask
    n, nsteps = 128, 500
    msk = basin_mask(n, 5)
    p = _param(model="eulerpsi", nx=n, ny=n, compflux="weno", integrator="rk3", cfl=0.8,
               tend=1e30, maxite=nsteps)
    ref = O.Model(P(**{k: getattr(p, k) for k in KEYS}), msk=msk)
    x, y = ref.mesh.xy("c")
    rng = np.random.default_rng(5)
    w0 = (np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / (2 * 0.1 ** 2)) +
          1e-3 * rng.uniform(-1, 1, x.shape)) * ref.mesh.msk * ref.mesh.area
    ref.state.omega[...] = w0
    ref.diag(ref.state)
    ref.set_dt()
    dt = ref.time.dt
    ref.param.dt = dt
    p.dt = dt
    p.precision = "fp32"
    m = f2d.Model(p)
    m.mesh.msk[...] = msk
    m.mesh.finalize()
    m.integrator.model.mesh = m.mesh
    assert m.mesh.poisson_vertices.method == "capacitance"
    m.state.omega[...] = w0
    m.integrator.diag(m.state)
    print(f"\n128^2 eulerpsi basin, dt {dt:.4e}: device FP32 + refinement ({stepping}), "
          "oracle FP64", flush=True)
    _steps(m, nsteps, stepping)
    ref.run(nsteps)
    wet = m.mesh.msk == 1
    e = _rel(m.state.omega[wet], ref.state.omega[wet])
    print(f"  vorticity rel L2 after {nsteps} steps: {e:.2e} (bar 1e-5)")
    assert e <= 1e-5


@pytest.mark.parametrize("stepping", ["run", "device_steps"])
@pytest.mark.parametrize("n,ny,seed", [(64, 64, 11), (96, 128, 12)])
def test_qgrsw_basin_fp32_refined(dev, n, ny, seed, stepping):
    import fluids2d_amd as f2d
    from bench import basin_mask
    nsteps = 200
    msk = basin_mask(n, seed, ny=ny)
    p = _param(model="qgrsw", nx=n, ny=ny, f0=10.0, g=1.0, H=1.0, dtmax=0.05,
               tend=1e30, maxite=nsteps, precision="fp32")
    m = f2d.Model(p)
    m.mesh.msk[...] = msk
    m.mesh.finalize()
    m.integrator.model.mesh = m.mesh
    rng = np.random.default_rng(seed + 7)
    x, y = m.mesh.xy("c")
    bumps = np.zeros(m.mesh.shape)
    for _ in range(4):
        x0, y0 = rng.uniform(0.25, 0.75, 2)
        r = rng.uniform(0.03, 0.08)
        bumps += rng.choice([-1, 1]) * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * r * r))
    h0 = (p.H + 0.05 * bumps) * m.mesh.msk * m.mesh.area
    # the balanced initial state, projected in FP64 by the oracle and given
    # to both runs
    ref = O.Model(P(**{k: getattr(p, k) for k in KEYS}), msk=msk)
    rs = ref.state
    rs.h[...] = h0
    with np.errstate(divide="ignore", invalid="ignore"):
        O.qg_projection(ref.mesh, rs.u, rs.h, rs.pv, rs.psi)
    ref.diag(rs)
    s = m.state
    s.u.x[...] = rs.u.x
    s.u.y[...] = rs.u.y
    s.h[...] = rs.h
    m.integrator.diag(s)
    print(f"\nqgrsw {n}x{ny} basin: device FP32 + refinement ({stepping}), oracle FP64",
          flush=True)
    _steps(m, nsteps, stepping)
    with np.errstate(divide="ignore", invalid="ignore"):
        ref.run(nsteps)
    assert m.time.ite == ref.time.ite == nsteps
    wet, wx, wy = m.mesh.msk == 1, m.mesh.mskx == 1, m.mesh.msky == 1
    eh = _rel(s.h[wet], rs.h[wet])
    eu = _rel(np.concatenate([s.u.x[wx], s.u.y[wy]]), np.concatenate([rs.u.x[wx], rs.u.y[wy]]))
    dt_err = abs(m.time.t - ref.time.t) / ref.time.t
    print(f"  after {nsteps} steps: h rel L2 {eh:.2e}, u rel L2 {eu:.2e}, "
          f"time rel {dt_err:.1e} (bar 1e-5)")
    assert eh <= 1e-5 and eu <= 1e-5


def test_euler_periodic_1024_fp32_vs_fp64_1000_steps(dev):
    """The headline workload at its full size (1024^2 doubly periodic
    eulerpsi, k^-3 vorticity, seed 2024): FP32 + refinement against the FP64
    device path (held to the oracle at the sizes the oracle runs) after 1000
    steps at a common fixed dt, relative L2 of the vorticity <= 1e-5, and the
    FP64 bulk diagnostics of the two runs close."""
    import bench
    err, dt, diag = bench.accuracy(1024, 2024, 1000)
    print(f"\n1024^2 FP32 vs FP64 after 1000 steps (dt {dt:.4g}): rel L2 omega {err:.2e} "
          f"(bar 1e-5); diagnostics {diag}")
    assert err <= 1e-5
    assert all(abs(v) <= 1e-6 for v in diag.values())
