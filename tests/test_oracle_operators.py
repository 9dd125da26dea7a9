"""Pins of the oracle's C-grid operators, integrators and time keeping
(operators.py:6-211, integrators.py:82-196, timeline.py:14-35, model.py:71-87)
against discrete identities and closed forms, and the reference's own
known-answer test (tests/test_models.py:17-26) run on the oracle.
"""
import numpy as np
import pytest

import oracle as O
from helpers import P, random_mask


def _vec(shape):
    return O.Vector(np.zeros(shape), np.zeros(shape))


def _psi(mesh, seed):
    """random streamfunction vanishing at the Dirichlet vertices"""
    rng = np.random.default_rng(seed)
    unk = O.unknown_mask(mesh.param, mesh.mskv)
    psi = rng.standard_normal(mesh.shape) * unk
    mesh.fill(psi)
    return psi


@pytest.mark.parametrize("xper,yper,seed", [(True, True, 0), (False, False, 1), (True, False, 2)])
def test_vorticity_of_perpgrad_is_5pt_laplacian(xper, yper, seed):
    p = P(nx=16, ny=12, xperiodic=xper, yperiodic=yper)
    mesh = O.Mesh(p, msk=random_mask(12, 16, seed, xper, yper, nisl=2))
    psi = _psi(mesh, seed)
    u = _vec(mesh.shape)
    O.perpgrad(mesh, psi, u)
    w = np.zeros(mesh.shape)
    O.compute_vorticity(mesh, u, w)
    lap = np.zeros(mesh.shape)
    lap[1:-1, 1:-1] = (psi[2:, 1:-1] + psi[:-2, 1:-1] + psi[1:-1, 2:] + psi[1:-1, :-2]
                       - 4 * psi[1:-1, 1:-1])
    sel = (mesh.slipcoef > 0)
    sel[0, :] = sel[-1, :] = False
    sel[:, 0] = sel[:, -1] = False
    assert np.allclose(w[sel], lap[sel], rtol=0, atol=1e-12)


@pytest.mark.parametrize("xper,seed", [(True, 3), (False, 4)])
def test_divergence_of_contravariant_perpgrad_vanishes(xper, seed):
    """Exact only for dx == dy: the sum is dxdy psi (1/dx^2 - 1/dy^2)."""
    p = P(nx=14, ny=14, xperiodic=xper)
    mesh = O.Mesh(p, msk=random_mask(14, 14, seed, xper, False, nisl=2))
    psi = _psi(mesh, seed)
    U = _vec(mesh.shape)
    O.perpgrad(mesh, psi, U, contravariant=True)
    d = np.zeros(mesh.shape)
    O.div(mesh, U, d)
    scale = np.abs(U.x).max()
    # interior cells (the halo columns are refreshed by the halo fill)
    assert np.abs(d[3:-3, 3:-3]).max() <= 1e-13 * scale
    # and it does not vanish with dx != dy
    p2 = P(nx=14, ny=14, Ly=0.7, xperiodic=xper)
    mesh2 = O.Mesh(p2, msk=mesh.msk)
    O.perpgrad(mesh2, psi, U, contravariant=True)
    O.div(mesh2, U, d)
    assert np.abs(d[3:-3, 3:-3]).max() > 1e-3 * np.abs(U.x).max()


@pytest.mark.parametrize("method", ["weno", "upwind", "centered", "cweno"])
def test_divflux_conserves_and_keeps_constants(method):
    p = P(nx=18, ny=15, Lx=1.2, compflux=method)  # dx == dy
    mesh = O.Mesh(p, msk=random_mask(15, 18, 6, False, False, nisl=3))
    rng = np.random.default_rng(6)
    U = O.Vector(rng.standard_normal(mesh.shape), rng.standard_normal(mesh.shape))
    q = rng.standard_normal(mesh.shape) * mesh.msk
    flx, dq = _vec(mesh.shape), np.zeros(mesh.shape)
    O.divflux(p, mesh, flx, q, U, dq)
    # flux form, zero flux through faces touching dry cells: the total is 0
    assert abs(dq[mesh.msk == 1].sum()) <= 1e-12 * np.abs(dq).sum()
    # a constant is advected unchanged by a divergence-free (dx == dy) flow
    psi = _psi(mesh, 7)
    O.perpgrad(mesh, psi, U, contravariant=True)
    O.divflux(p, mesh, flx, 2.5 * mesh.msk.astype(float), U, dq)
    assert np.abs(dq[3:-3, 3:-3]).max() <= 1e-12 * np.abs(U.x).max()


def test_centers_vertices_constants():
    p = P(nx=10, ny=9)
    mesh = O.Mesh(p, msk=random_mask(9, 10, 8, False, False, nisl=1))
    w = 3.0 * np.ones(mesh.shape)
    v = np.zeros(mesh.shape)
    O.centerstovertices(mesh, w, v)
    assert np.all(v[mesh.mskv == 1] == 3.0) and np.all(v[mesh.mskv == 0] == 0.0)
    h = np.zeros(mesh.shape)
    O.verticestocenters(mesh, v, h)
    touched = np.zeros(mesh.shape, dtype=bool)
    m = mesh.mskv
    touched[:-1, :-1] = (m[:-1, :-1] + m[1:, :-1] + m[:-1, 1:] + m[1:, 1:]) > 0
    sel = touched & (mesh.msk == 1)
    assert np.allclose(h[sel], 3.0, rtol=1e-15)


@pytest.mark.parametrize("name,poly", [
    ("rk3", lambda z: 1 + z + z * z / 2 + z ** 3 / 6),
    ("enrk3", lambda z: 1 + z + z * z / 2 + z ** 3 / 8),
    ("ef", lambda z: 1 + z)])
def test_integrator_amplification(name, poly):
    """Linear test equation s' = lam s: one step multiplies s by the
    stability polynomial b.(I - zA)^-1 1 of the scheme's tableau.  SSP-RK3
    (3rd order) gives 1 + z + z^2/2 + z^3/6.  The incremental enrk3 of
    integrators.py:109-124 has c = (0, 1/3, 5/6), a31 = -5/48, a32 = 15/16,
    b = (1/10, 1/2, 2/5), so its z^3 coefficient is b3 a32 c2 = 1/8."""
    integ, nst = O.INTEGRATORS[name]
    for lam, dt in [(-1.0, 0.3), (2.0, 0.1), (-7.0, 0.05), (0.5, 1.0)]:
        s = (np.array([1.0, -2.0]),)
        scratch = [(np.zeros(2),) for _ in range(nst)]

        def rhs(x, dx):
            dx[0][:] = lam * x[0]

        integ(s, dt, rhs, lambda x: None, scratch)
        g = poly(lam * dt)
        assert np.allclose(s[0], g * np.array([1.0, -2.0]), rtol=1e-14, atol=0)


def _lfra_run(p, gamma, nsteps, y0):
    """The oracle's lfra on the oscillation equation y' = i w y (y as the
    pair re, im); returns the state after each step."""
    s = (np.array(y0, dtype=float),)
    scratch = [(np.zeros(2),) for _ in range(3)]

    def rhs(x, dx):
        dx[0][0] = -x[0][1]
        dx[0][1] = x[0][0]

    out = []
    for n in range(nsteps):
        O.lfra(s, p, rhs, lambda x: None, scratch, first=n == 0, gamma=gamma)
        out.append(s[0][0] + 1j * s[0][1])
    return np.array(out)


def test_lfra_unfiltered_is_leapfrog():
    """gamma = 0: Euler forward first, then y(n+1) = y(n-1) + 2 dt f(y(n))."""
    dt, y0 = 0.3, 1.0 + 0.5j
    ys = _lfra_run(dt, 0.0, 6, [y0.real, y0.imag])
    ref = [y0, y0 + dt * 1j * y0]
    for n in range(1, 6):
        ref.append(ref[n - 1] + 2 * dt * 1j * ref[n])
    assert np.allclose(ys, ref[1:], rtol=1e-14, atol=0)


@pytest.mark.parametrize("p,gamma", [(0.2, 0.1), (0.05, 0.2), (0.5, 0.06)])
def test_lfra_physical_mode_amplification(p, gamma):
    """Leap-frog with the Robert-Asselin filter on y' = i w y (p = w dt):
    the physical mode's amplification factor is
    A = gamma + i p + sqrt((1 - gamma)^2 - p^2) (Durran, Numerical Methods
    for Fluid Dynamics, 2nd ed., sec. 2.3.5); the computational mode,
    gamma + i p - sqrt(...), has |A| < 1 and has died out after a few
    hundred steps, so the step-to-step ratio of the state is A."""
    ys = _lfra_run(p, gamma, 600, [1.0, 0.0])
    A = gamma + 1j * p + np.sqrt((1 - gamma) ** 2 - p * p)
    ratio = ys[-1] / ys[-2]
    assert abs(ratio - A) < 1e-12
    # the filter damps the physical mode (plain leap-frog keeps |A| = 1)
    assert 1 - abs(ratio) > 1e-4


def test_forcing_runs_after_rhs():
    """add_forcing (equations.py:229-238): advection with U = 0 and a
    forcing adding c to the tendency on wet cells grows q as c t there, for
    every integrator; mesh.time is the model's time during the forcing."""
    for integ in ("rk3", "enrk3", "ef", "LFRA"):
        p = P(model="advection", nx=12, ny=8, integrator=integ, dt=0.1, tend=1.0,
              maxite=7)
        m = O.Model(p)
        seen = []

        def forcing(param, mesh, s, ds):
            ds.q[:] += 0.25 * mesh.msk
            seen.append(mesh.time.t)

        m.add_forcing(forcing)
        m.run()
        wet = m.mesh.msk == 1
        assert np.allclose(m.state.q[wet], 0.25 * m.time.t, rtol=1e-13)
        assert seen[0] == 0.0 and max(seen) < m.time.t


def test_time_kahan_summation():
    p = P(dt=0.1, tend=1e9, maxite=10 ** 9)
    t = O.Time(p)
    naive = 0.0
    for _ in range(100000):
        t.pushforward()
        naive += 0.1
    assert abs(t.t - 10000.0) < 1e-9
    assert abs(naive - 10000.0) > 1e-9  # plain summation drifts
    assert t.ite == 100000


def test_kat_advection_bodyrotation_37_on_oracle():
    """tests/test_models.py:17-26 (tracer_advection.py:10-48 set-up): the
    oracle reaches tend = 10 in exactly 37 iterations."""
    p = P(model="advection", tend=10)
    model = O.Model(p)
    mesh = model.mesh
    omega = (4 * np.pi / 100) * np.ones(mesh.shape) * mesh.mskv
    psi = omega * 0
    mesh.poisson_vertices.solve(omega, psi)
    O.perpgrad(mesh, psi, model.state.U)
    x, y = mesh.xy("c")
    q = model.state.q
    q[:, :] = (np.round(x * 8) % 2 + np.round(y * 8) % 2) / 2
    q *= mesh.msk
    model.run()
    assert model.time.ite == 37


def euler_dipole(O, p):
    """experiments/vortex.py:8-21 for the euler model: vertex vorticity of
    the dipole, velocity from set_uv_from_omega (tools.py:6-26), diag."""
    m = O.Model(p)
    mesh = m.mesh
    x, y = mesh.xy("v")

    def g(x0, y0, r):
        return np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * r ** 2))
    x0, y0, r0, d = p.Lx / 2, 0.5, 0.05, 0.05
    omega = (g(x0 + d, y0, r0) - g(x0 - d, y0, r0)) * mesh.mskv * mesh.area
    psi = np.zeros(mesh.shape)
    mesh.poisson_vertices.solve(omega, psi)
    O.perpgrad(mesh, psi, m.state.u)
    m.state.omega[...] = omega
    m.diag(m.state)
    return m, omega


def test_euler_dipole_iterations():
    """The reference's KAT test_euler (tests/test_models.py:8-15: default
    Param, euler, the dipole of experiments/vortex.py:8-42, which ends with a
    diag, then run to tend = 10).  That test asserts 25 iterations, but the
    reference's own code, run exactly as the test runs it, takes 22 (dt from
    0.4456 to 0.4824; tests/golden/reference_vectors.npz, written by
    tests/golden/make_reference_vectors.py, with the dt sequence compared in
    tests/test_reference_vectors.py).  The oracle takes the same 22."""
    import oracle as O
    from helpers import P
    m, _ = euler_dipole(O, P(model="euler", tend=10))
    m.run()
    assert m.time.ite == 22


def test_euler_projection_is_divergence_free():
    """After the pressure projection (operators.py:114-119) the divergence
    of U vanishes on the wet cells to rounding: the Neumann solve is
    consistent (the wall fluxes are masked out)."""
    import oracle as O
    from helpers import P, random_mask
    p = P(model="euler", nx=36, ny=28, Lx=36 / 28)
    m = O.Model(p, msk=random_mask(28, 36, 4, False, False, nisl=2))
    rng = np.random.default_rng(0)
    s = m.state
    s.u.x[...] = rng.standard_normal(m.mesh.shape) * m.mesh.mskx
    s.u.y[...] = rng.standard_normal(m.mesh.shape) * m.mesh.msky
    m.diag(s)
    d = np.zeros(m.mesh.shape)
    O.div(m.mesh, s.U, d)
    assert np.abs(d[m.mesh.msk == 1]).max() < 1e-10 * np.abs(s.U.x).max() * 36
