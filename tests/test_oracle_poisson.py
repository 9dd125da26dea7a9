"""Pins of the oracle's masked Poisson/Helmholtz operator and solve
(elliptic.py:71-203) against closed forms:

* Dirichlet rectangle: sin(k pi i / nx) sin(l pi j / ny) are eigenvectors with
  eigenvalue 2 (dy/dx)(cos(k pi/nx) - 1) + 2 (dx/dy)(cos(l pi/ny) - 1) - maindiag
  (dx != dy, so transposed coefficients fail);
* x-periodic channel: cos(2 pi k i / nx) sin(l pi j / ny), same form;
* Neumann ("c"): constants are mapped to -maindiag times themselves;
* symmetry on random masks;
* the SuperLU solve against a dense solve, and the singular doubly-periodic
  solve against the DFT closed form.
"""
import numpy as np
import pytest

import oracle as O
from helpers import P, random_mask


def _operator(p, m, location="v", maindiag=0.0):
    mesh = O.Mesh(p, msk=m)
    base = mesh.msk if location == "c" else mesh.mskv
    unk = O.unknown_mask(p, base)
    G, N = O.gindex(unk)
    A = O.laplacian(G, N, mesh.dx, mesh.dy, maindiag, location == "c", p.xperiodic,
                    getattr(p, "yperiodic", False), p.halowidth)
    return mesh, unk, G, N, A


@pytest.mark.parametrize("maindiag", [0.0, 0.7])
def test_dirichlet_rectangle_eigenvectors(maindiag):
    nx, ny = 10, 7
    p = P(nx=nx, ny=ny, Lx=1.0, Ly=0.9)
    m = O.default_mask(ny, nx, 3, False, False)
    mesh, unk, G, N, A = _operator(p, m, "v", maindiag)
    assert N == (nx - 1) * (ny - 1)
    jj, ii = np.nonzero(unk)
    iv, jv = ii - 3, jj - 3  # vertex coordinates in cells
    ax, ay = mesh.dy / mesh.dx, mesh.dx / mesh.dy
    for k, l in [(1, 1), (2, 3), (9, 6), (4, 1)]:
        v = np.sin(k * np.pi * iv / nx) * np.sin(l * np.pi * jv / ny)
        lam = (2 * ax * (np.cos(k * np.pi / nx) - 1) + 2 * ay * (np.cos(l * np.pi / ny) - 1)
               - maindiag)
        order = np.argsort(G[unk == 1])
        vv = v[order]
        assert np.allclose(A @ vv, lam * vv, rtol=0, atol=1e-12)


def test_xperiodic_channel_eigenvectors():
    nx, ny = 12, 6
    p = P(nx=nx, ny=ny, Lx=1.3, Ly=1.0, xperiodic=True)
    m = O.default_mask(ny, nx, 3, True, False)
    mesh, unk, G, N, A = _operator(p, m, "v")
    assert N == nx * (ny - 1)
    jj, ii = np.nonzero(unk)
    iv, jv = ii - 3, jj - 3
    ax, ay = mesh.dy / mesh.dx, mesh.dx / mesh.dy
    for k, l in [(0, 1), (1, 2), (5, 5), (6, 3)]:
        v = np.cos(2 * np.pi * k * iv / nx) * np.sin(l * np.pi * jv / ny)
        lam = 2 * ax * (np.cos(2 * np.pi * k / nx) - 1) + 2 * ay * (np.cos(l * np.pi / ny) - 1)
        vv = v[np.argsort(G[unk == 1])]
        assert np.allclose(A @ vv, lam * vv, rtol=0, atol=1e-12)


@pytest.mark.parametrize("seed,xper", [(0, False), (1, True)])
def test_symmetric_and_neumann_constants(seed, xper):
    p = P(nx=14, ny=9, xperiodic=xper)
    m = random_mask(9, 14, seed, xper, False, nisl=2)
    for loc, md in (("v", 0.0), ("v", 0.3), ("c", 0.0), ("c", 0.25)):
        _, _, _, N, A = _operator(p, m, loc, md)
        d = (A - A.T).toarray()
        assert np.abs(d).max() == 0.0
        if loc == "c":
            one = np.ones(N)
            assert np.allclose(A @ one, -md * one, atol=1e-13)


@pytest.mark.parametrize("seed,xper,md", [(3, False, 0.0), (4, True, 0.0), (5, False, 0.4)])
def test_solve_matches_dense(seed, xper, md):
    p = P(nx=12, ny=10, xperiodic=xper)
    m = random_mask(10, 12, seed, xper, False, nisl=2)
    mesh = O.Mesh(p, msk=m)
    ps = O.Poisson2D(p, mesh, "v", maindiag=md)
    rng = np.random.default_rng(seed)
    b = rng.standard_normal(mesh.shape)
    x = np.full(mesh.shape, 7.0)
    ps.solve(b, x)
    A = ps.A.toarray()
    want = np.linalg.solve(A, b[ps.idx])
    assert np.allclose(x[ps.idx], want, rtol=1e-11, atol=1e-12)
    # entries that are not unknowns keep their value unless the halo fill
    # copies an unknown there
    if not xper:
        assert np.all(x[~ps.idx] == 7.0)
    else:
        inner = np.zeros(mesh.shape, dtype=bool)
        inner[:, 3:-3] = True
        assert np.all(x[~ps.idx & inner] == 7.0)
        assert np.array_equal(x[:, :3], x[:, -6:-3])


def test_singular_doubly_periodic_matches_dft():
    nx, ny = 16, 12
    p = P(nx=nx, ny=ny, Lx=1.0, Ly=0.75, xperiodic=True, yperiodic=True)
    mesh = O.Mesh(p)
    ps = O.Poisson2D(p, mesh, "v")
    assert ps.singular
    rng = np.random.default_rng(8)
    b = rng.standard_normal(mesh.shape)
    x = np.zeros(mesh.shape)
    ps.solve(b, x)
    want = O.fft_poisson_periodic(b[3:-3, 3:-3], mesh.dx, mesh.dy)
    assert np.allclose(x[3:-3, 3:-3], want, rtol=0, atol=1e-10 * np.abs(want).max())
    assert abs(x[3:-3, 3:-3].mean()) < 1e-12


def _odd_extension_solve(b, dx, dy, maindiag, xper, yper):
    """The Dirichlet problem of a closed direction as the periodic problem
    on the doubled grid for the odd extension of b (the direct solver's
    method), by numpy's FFT.  b: (ny, nx) vertex values, index 0 (and the
    closed directions' boundary lines) not unknowns."""
    ny, nx = b.shape
    e = b.copy()
    if not xper:
        e[:, 0] = 0
        e = np.concatenate([e, np.zeros((ny, 1)), -e[:, :0:-1]], axis=1)
    if not yper:
        e[0, :] = 0
        e = np.concatenate([e, np.zeros((1, e.shape[1])), -e[:0:-1, :]], axis=0)
    x = O.fft_poisson_periodic(e, dx, dy, maindiag)
    return x[:ny, :nx]


@pytest.mark.parametrize("xper,yper,maindiag", [(False, False, 0.0), (False, False, 0.7),
                                                (True, False, 0.0), (False, True, 0.3)])
def test_closed_directions_by_odd_extension(xper, yper, maindiag):
    """SuperLU on the masked operator of an all-wet closed (or channel)
    domain equals the periodic solve of the odd extension on the doubled
    grid: the reduction the direct GPU solver uses for closed domains."""
    nx, ny = 16, 8
    p = P(nx=nx, ny=ny, Lx=1.0, Ly=0.6, xperiodic=xper, yperiodic=yper)
    mesh = O.Mesh(p)
    ps = O.Poisson2D(p, mesh, "v", maindiag=maindiag)
    unk = O.unknown_mask(p, mesh.mskv)
    assert unk.sum() == (nx - (not xper)) * (ny - (not yper))
    rng = np.random.default_rng(3)
    b = rng.standard_normal(mesh.shape) * unk
    x = np.zeros(mesh.shape)
    ps.solve(b, x)
    want = _odd_extension_solve(b[3:-3, 3:-3], mesh.dx, mesh.dy, maindiag, xper, yper)
    got = x[3:-3, 3:-3]
    if maindiag == 0.0 and xper and yper:
        got = got - got.mean()
    assert np.allclose(got, want, rtol=0, atol=1e-11 * np.abs(want).max())


def _even_extension_solve(b, dx, dy, maindiag, xper, yper):
    """The Neumann cell problem of a closed direction as the periodic
    problem on the doubled grid for the even extension of b about the walls
    (the direct solver's method for cells), by numpy's FFT."""
    e = b.copy()
    if not xper:
        e = np.concatenate([e, e[:, ::-1]], axis=1)
    if not yper:
        e = np.concatenate([e, e[::-1, :]], axis=0)
    x = O.fft_poisson_periodic(e, dx, dy, maindiag)
    return x[:b.shape[0], :b.shape[1]]


@pytest.mark.parametrize("xper,yper,maindiag", [(False, False, 0.0), (False, False, 0.7),
                                                (True, False, 0.0), (False, True, 0.3),
                                                (True, True, 0.0)])
def test_neumann_cells_by_even_extension(xper, yper, maindiag):
    """SuperLU on the Neumann cell operator (elliptic.py:186-190) of an
    all-wet closed (or channel) domain equals the periodic solve of the even
    extension about the walls on the doubled grid: the reduction the direct
    GPU solver uses for the euler model's pressure.  Singular cases (no
    mass term) are compared with zero-mean right-hand side and solutions."""
    nx, ny = 16, 8
    p = P(nx=nx, ny=ny, Lx=1.0, Ly=0.6, xperiodic=xper, yperiodic=yper)
    mesh = O.Mesh(p)
    ps = O.Poisson2D(p, mesh, "c", maindiag=maindiag)
    rng = np.random.default_rng(5)
    b = rng.standard_normal(mesh.shape) * mesh.msk
    if maindiag == 0.0:
        b[3:-3, 3:-3] -= b[3:-3, 3:-3].mean()
    x = np.zeros(mesh.shape)
    ps.solve(b, x)
    want = _even_extension_solve(b[3:-3, 3:-3], mesh.dx, mesh.dy, maindiag, xper, yper)
    got = x[3:-3, 3:-3]
    if maindiag == 0.0:
        got = got - got.mean()
        want = want - want.mean()
    assert np.allclose(got, want, rtol=0, atol=1e-11 * np.abs(want).max())
    # a plausible slip, the odd extension, fails this
    if not (xper and yper):
        odd = _odd_extension_solve(b[3:-3, 3:-3], mesh.dx, mesh.dy, maindiag, xper, yper)
        assert not np.allclose(got, odd - (odd.mean() if maindiag == 0.0 else 0.0),
                               rtol=0, atol=1e-6 * np.abs(want).max())
