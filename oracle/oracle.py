"""CPU oracle for the Fluids2d time-stepping hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker the HIP library is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package
(``fluids2d_amd``) never does.

It restates, in numpy plus the C loops of ``oracle.c``, the reference's
algorithm for the in-scope models (euler, eulerpsi, advection, rsw, qgrsw,
qg):

* masks / stencil orders / slip coefficient      meshes.py:70-186, noslip.py:4-36
* WENO flux loops (C)                            weno.py:346-405
* C-grid operators (numpy, same array algebra)   operators.py:6-211
* masked Poisson/Helmholtz, assembled in C and
  factorised with scipy's SuperLU exactly as
  the reference does                             elliptic.py:71-203
* SSP-RK3 / Euler forward / energy RK3 `addto`   integrators.py:82-196
* leap-frog + Robert-Asselin filter (LFRA)       integrators.py:20-53, 134-137
* forcing hook                                   model.py:121-123, equations.py:229-238
* model closures                                 equations.py:73-157, 190-210
* CFL time step and Kahan time                   model.py:71-87, timeline.py:14-35

Where the reference has no behaviour (y-periodic domains: ``fill_halo_array``
fills x only, ``get_Laplacian_sparse`` fixes ``yperiodic=False``) the oracle
implements the symmetric extension of the x-periodic code: the y halo rows are
filled like the x halo columns, the Laplacian wraps in y as it does in x, and
the singular doubly-periodic operator is solved for the zero-mean part of
the right-hand side with its last unknown removed (the reference's own device
for singular systems, elliptic.py:64-66), the solution being shifted to zero
mean afterwards -- the same solution as the DFT closed form with the k=0 mode
dropped.
"""
import ctypes
import os
from collections import namedtuple

import numpy as np
import scipy.sparse as sparse
import scipy.sparse.linalg as splinalg

_HERE = os.path.dirname(os.path.abspath(__file__))
METHODS = {"weno": 0, "upwind": 1, "centered": 2, "cweno": 3}

_lib = None


def lib():
    """Load oracle/liboracle.so (built by oracle/Makefile)."""
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        d, i8p, dp, i64, i32p = (ctypes.c_double, ctypes.POINTER(ctypes.c_int8),
                                 ctypes.POINTER(ctypes.c_double), ctypes.c_int64,
                                 ctypes.POINTER(ctypes.c_int32))
        L.orc_weno5z.restype = d
        L.orc_weno5z.argtypes = [d] * 5
        L.orc_weno3z.restype = d
        L.orc_weno3z.argtypes = [d] * 3
        L.orc_f1.restype = d
        L.orc_f1.argtypes = [ctypes.c_int] + [d] * 3
        L.orc_f3.restype = d
        L.orc_f3.argtypes = [ctypes.c_int] + [d] * 5
        L.orc_f5.restype = d
        L.orc_f5.argtypes = [ctypes.c_int] + [d] * 7
        L.orc_compflux.argtypes = [dp, dp, dp, i8p, i64, i64, ctypes.c_int, i64, i64]
        L.orc_vortexforce.argtypes = [dp, dp, dp, i8p, i64, i64, i64, ctypes.c_int,
                                      ctypes.c_int, i64, i64]
        L.orc_innerproduct.argtypes = [dp, dp, dp, i8p, i64, i64, ctypes.c_int, i64, i64]
        L.orc_set_order.argtypes = [i8p, i64, i64, ctypes.c_int, i8p]
        L.orc_set_masks.argtypes = [i8p, i64, i64, i8p, i8p, i8p]
        L.orc_gindex.restype = i64
        L.orc_gindex.argtypes = [i8p, i64, i32p]
        L.orc_laplacian_coo.restype = i64
        L.orc_laplacian_coo.argtypes = [i32p, i64, i64, d, d, d, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        i32p, i32p, dp]
        L.orc_set_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def set_threads(n):
    """Threads of the C flat loops (OpenMP; results do not depend on it).
    Returns the count in use."""
    return int(lib().orc_set_threads(int(n)))


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _chk(a, dtype):
    assert a.dtype == dtype and a.flags.c_contiguous, (a.dtype, dtype)
    return a


# --------------------------------------------------------------------------
# pointwise reconstructions and flat loops (weno.py)
# --------------------------------------------------------------------------
def weno5z(qmm, qm, q0, qp, qpp):
    """weno.py:167-197"""
    return lib().orc_weno5z(qmm, qm, q0, qp, qpp)


def weno3z(qm, q0, qp):
    """weno.py:75-100"""
    return lib().orc_weno3z(qm, q0, qp)


def flux5(method, U, *q):
    """flx5 / flxup5 / ce6 / cflx5 selection, weno.py:303-343"""
    return lib().orc_f5(METHODS[method], U, *q)


def flux3(method, U, *q):
    return lib().orc_f3(METHODS[method], U, *q)


def flux1(method, U, *q):
    return lib().orc_f1(METHODS[method], U, *q)


def compflux(flx, U, q, o, s, method):
    """weno.compflux with nthreads=1, weno.py:412-415"""
    n = q.size
    for a in (flx, U, q):
        _chk(a, np.float64)
    _chk(o, np.int8)
    lib().orc_compflux(_p(flx, ctypes.c_double), _p(U, ctypes.c_double),
                       _p(q, ctypes.c_double), _p(o, ctypes.c_int8), n, s,
                       METHODS[method], 0, n)


def vortexforce(du, V, q, o, s, s2, sign, method):
    """weno.vortexforce with nthreads=1, weno.py:423-431"""
    n = du.size
    for a in (du, V, q):
        _chk(a, np.float64)
    _chk(o, np.int8)
    lib().orc_vortexforce(_p(du, ctypes.c_double), _p(V, ctypes.c_double),
                          _p(q, ctypes.c_double), _p(o, ctypes.c_int8), n, s, s2,
                          sign, METHODS[method], 0, n)


def innerproduct(ke, U, q, o, s, method):
    """weno.innerproduct, weno.py:434-436"""
    n = ke.size
    for a in (ke, U, q):
        _chk(a, np.float64)
    _chk(o, np.int8)
    lib().orc_innerproduct(_p(ke, ctypes.c_double), _p(U, ctypes.c_double),
                           _p(q, ctypes.c_double), _p(o, ctypes.c_int8), n, s,
                           METHODS[method], 0, n)


# --------------------------------------------------------------------------
# masks, orders, slip coefficient (meshes.py, noslip.py)
# --------------------------------------------------------------------------
def set_order(m, shift, maxorder):
    """meshes.py:146-186 on the flattened array; returns int8 orders."""
    m8 = np.ascontiguousarray(m, dtype=np.int8)
    o = np.zeros(m8.shape, dtype=np.int8)
    lib().orc_set_order(_p(m8, ctypes.c_int8), m8.size, shift, maxorder,
                        _p(o, ctypes.c_int8))
    return o


def set_masks(msk):
    """meshes.py:77-86"""
    msk = np.ascontiguousarray(msk, dtype=np.int8)
    ny, nx = msk.shape
    mskx = np.zeros_like(msk)
    msky = np.zeros_like(msk)
    mskv = np.zeros_like(msk)
    lib().orc_set_masks(_p(msk, ctypes.c_int8), ny, nx, _p(mskx, ctypes.c_int8),
                        _p(msky, ctypes.c_int8), _p(mskv, ctypes.c_int8))
    return mskx, msky, mskv


def slipcoef(msk, noslip, halowidth):
    """noslip.py:4-36 (returns float64 0/1)."""
    coef = np.zeros(msk.shape)
    coef[:, :] += msk
    coef[:, 1:] += msk[:, :-1]
    coef[1:, :] += msk[:-1, :]
    coef[1:, 1:] += msk[:-1, :-1]

    def freeslip(x):
        return 1.0 * (x == 4)

    def noslip_(x):
        return np.minimum(x, 1)

    if (noslip is None) or (noslip is False):
        return freeslip(coef)
    if noslip is True:
        return noslip_(coef)
    sc = freeslip(coef)
    nh = halowidth
    if "left" in noslip:
        sc[:, nh] = noslip_(coef[:, nh])
    if "right" in noslip:
        sc[:, -nh] = noslip_(coef[:, -nh])
    if "bottom" in noslip:
        sc[nh, :] = noslip_(coef[nh, :])
    if "top" in noslip:
        sc[-nh, :] = noslip_(coef[-nh, :])
    return sc


def default_mask(ny, nx, nh, xperiodic, yperiodic):
    """meshes.py:70-75"""
    msk = np.zeros((ny + 2 * nh, nx + 2 * nh), dtype=np.int8)
    xdx = slice(None) if xperiodic else slice(nh, -nh)
    ydx = slice(None) if yperiodic else slice(nh, -nh)
    msk[ydx, xdx] = 1
    return msk


def fill(param, a):
    """fill_halo_array, meshes.py:135-143, plus the symmetric y extension."""
    n = param.halowidth
    if isinstance(a, tuple):
        for v in a:
            fill(param, v)
        return
    if param.xperiodic:
        a[:, :n] = a[:, -2 * n:-n]
        a[:, -n:] = a[:, n:2 * n]
    if getattr(param, "yperiodic", False):
        a[:n, :] = a[-2 * n:-n, :]
        a[-n:, :] = a[n:2 * n, :]


# --------------------------------------------------------------------------
# Poisson / Helmholtz (elliptic.py)
# --------------------------------------------------------------------------
def unknown_mask(param, mskv):
    """get_msk_from_mesh, elliptic.py:102-111 (+ y extension)."""
    n = param.halowidth
    m = mskv * 1
    if param.xperiodic:
        m[:, :n] = 0
        m[:, -n:] = 0
    if getattr(param, "yperiodic", False):
        m[:n, :] = 0
        m[-n:, :] = 0
    return m.astype(np.int8)


def gindex(m):
    """get_Gindex, elliptic.py:198-203"""
    m = np.ascontiguousarray(m, dtype=np.int8)
    G = np.zeros(m.shape, dtype=np.int32)
    N = lib().orc_gindex(_p(m, ctypes.c_int8), m.size, _p(G, ctypes.c_int32))
    return G, int(N)


def laplacian(G, N, dx, dy, maindiag, neumann, xperiodic, yperiodic, halowidth):
    """get_Laplacian_sparse, elliptic.py:114-195 -> scipy CSC."""
    ny, nx = G.shape
    row = np.zeros(5 * N, dtype=np.int32)
    col = np.zeros(5 * N, dtype=np.int32)
    data = np.zeros(5 * N)
    c = lib().orc_laplacian_coo(_p(G, ctypes.c_int32), ny, nx, dx, dy, maindiag,
                                int(neumann), int(xperiodic), int(yperiodic),
                                halowidth, _p(row, ctypes.c_int32),
                                _p(col, ctypes.c_int32), _p(data, ctypes.c_double))
    A = sparse.coo_matrix((data[:c], (row[:c], col[:c])), shape=(N, N))
    return A.tocsc()


class Poisson2D:
    """elliptic.py:71-87: SuperLU factorisation of the masked 5-point
    operator, solved at the unknowns, x-halo (and y-halo) filled after."""

    def __init__(self, param, mesh, location="v", maindiag=0.0):
        assert location in ("c", "v")
        self.param = param
        base = mesh.msk if location == "c" else mesh.mskv
        self.msk = unknown_mask(param, base)
        self.G, self.N = gindex(self.msk)
        self.A = laplacian(self.G, self.N, mesh.dx, mesh.dy, maindiag,
                           location == "c", param.xperiodic,
                           getattr(param, "yperiodic", False), param.halowidth)
        self.idx = self.G > -1
        # doubly periodic, all wet, no mass term: singular (constant mode)
        self.singular = (location == "v" and maindiag == 0 and param.xperiodic
                         and getattr(param, "yperiodic", False)
                         and self.N == param.nx * param.ny)
        if self.singular:
            self.LU = splinalg.splu(self.A[:-1, :-1].tocsc())
        else:
            self.LU = splinalg.splu(self.A)

    def solve(self, b, x):
        rhs = b[self.idx]
        if self.singular:
            # consistent right-hand side (zero mean), last unknown pinned,
            # zero-mean gauge
            rhs = rhs - rhs.mean()
            sol = np.zeros(self.N)
            sol[:-1] = self.LU.solve(rhs[:-1])
            sol -= sol.mean()
        else:
            sol = self.LU.solve(rhs)
        x[self.idx] = sol
        fill(self.param, x)


def fft_poisson_periodic(b, dx, dy, maindiag=0.0):
    """Closed-form solve of the doubly-periodic 5-point operator on an
    (ny, nx) interior array: the 2-D DFT diagonalises it with eigenvalues
    2(dy/dx)(cos kx - 1) + 2(dx/dy)(cos ky - 1) - maindiag; the k=0 mode of a
    singular operator is set to zero (mean-zero gauge)."""
    ny, nx = b.shape
    kx = 2 * np.pi * np.fft.fftfreq(nx)
    ky = 2 * np.pi * np.fft.fftfreq(ny)
    lam = (2 * (dy / dx) * (np.cos(kx)[None, :] - 1)
           + 2 * (dx / dy) * (np.cos(ky)[:, None] - 1) - maindiag)
    bh = np.fft.fft2(b)
    if maindiag == 0:
        lam[0, 0] = 1.0
        bh[0, 0] = 0.0
    return np.real(np.fft.ifft2(bh / lam))


# --------------------------------------------------------------------------
# mesh
# --------------------------------------------------------------------------
class Mesh:
    """meshes.py:18-52 restated (host, FP64)."""

    def __init__(self, param, msk=None):
        self.param = param
        nh = param.halowidth
        self.nx, self.ny = param.nx, param.ny
        self.shape = (param.ny + 2 * nh, param.nx + 2 * nh)
        self.dx = param.Lx / param.nx
        self.dy = param.Ly / param.ny
        self.area = self.dx * self.dy
        self.xshift = 1
        self.yshift = self.shape[-1]
        if msk is None:
            msk = default_mask(param.ny, param.nx, nh, param.xperiodic,
                               getattr(param, "yperiodic", False))
        self.msk = np.ascontiguousarray(msk, dtype=np.int8)
        self.finalize()

    def finalize(self):
        p = self.param
        self.mskx, self.msky, self.mskv = set_masks(self.msk)
        self.slipcoef = slipcoef(self.msk, p.noslip, p.halowidth)
        S = namedtuple("stencil", ("x", "y"))
        mo = p.maxorder
        self.oc = S(set_order(self.msk, self.xshift, mo),
                    set_order(self.msk, self.yshift, mo))
        mv = (self.slipcoef > 0).astype(np.int8)
        ovx = set_order(mv, -self.xshift, mo) * self.msky
        ovy = set_order(mv, -self.yshift, mo) * self.mskx
        self.ov = S(ovx.astype(np.int8), ovy.astype(np.int8))
        self.ok = S(set_order(self.mskx, -self.xshift, mo),
                    set_order(self.msky, -self.yshift, mo))
        # the solvers (meshes.py:41-47) are factorised on first use: the
        # reference builds them here, but a run that never solves (rsw) need
        # not wait for SuperLU on a large grid; results are the same
        for name in self._SOLVERS:
            self.__dict__.pop(name, None)
        if p.model in ("qg", "qgrsw", "rsw"):
            self.hb = 0
            self.qgcoef = p.f0 / +p.H

    _SOLVERS = ("poisson_vertices", "poisson_centers", "qg_helmholtz")

    def __getattr__(self, name):
        if name not in Mesh._SOLVERS or "param" not in self.__dict__:
            raise AttributeError(name)
        p = self.param
        if name == "poisson_vertices":
            v = Poisson2D(p, self, "v")
        elif name == "poisson_centers":
            # meshes.py:45-47: the Neumann cell operator of the euler projection
            v = Poisson2D(p, self, "c") if p.model == "euler" else None
        elif p.model in ("qg", "qgrsw", "rsw"):
            maindiag = self.area * p.f0 ** 2 / (p.g * p.H)
            v = Poisson2D(p, self, "v", maindiag=maindiag)
        else:
            raise AttributeError(name)
        self.__dict__[name] = v
        return v

    def xy(self, which="c"):
        """meshes.py:57-68"""
        nh = self.param.halowidth
        ix = np.arange(self.nx + 2 * nh) - nh
        iy = np.arange(self.ny + 2 * nh) - nh
        sx = 0.5 if which in ("c", "y") else 0.0
        sy = 0.5 if which in ("c", "x") else 0.0
        return np.meshgrid((ix + sx) * self.dx, (iy + sy) * self.dy)

    def fill(self, a):
        fill(self.param, a)


# --------------------------------------------------------------------------
# operators (operators.py)
# --------------------------------------------------------------------------
Vector = namedtuple("vector", ("x", "y"))


def div(mesh, F, delta):
    """operators.py:104-107"""
    delta[:, :-1] = -np.diff(F.x, axis=1)
    delta[:-1, :] -= np.diff(F.y, axis=0)
    delta *= mesh.msk


def divflux(param, mesh, flx, q, U, dq):
    """operators.py:21-29"""
    compflux(flx.x, U.x, q, mesh.oc.x, mesh.xshift, param.compflux)
    compflux(flx.y, U.y, q, mesh.oc.y, mesh.yshift, param.compflux)
    div(mesh, flx, dq)


def addvortexforce(param, mesh, U, omega, du):
    """operators.py:6-13 (overwrites du)"""
    m = param.vortexforce
    vortexforce(du.x, U.y, omega, mesh.ov.y, mesh.yshift, mesh.xshift, +1, m)
    vortexforce(du.y, U.x, omega, mesh.ov.x, mesh.xshift, mesh.yshift, -1, m)


def addcoriolis(param, mesh, U, du):
    """operators.py:32-39"""
    f = param.f0 * mesh.area * 0.25
    du.x[:-1, 1:-1] += f * (U.y[:-1, :-2] + U.y[1:, :-2] + U.y[:-1, 1:-1] + U.y[1:, 1:-1])
    du.y[1:-1, :-1] -= f * (U.x[:-2, :-1] + U.x[:-2, 1:] + U.x[1:-1, :-1] + U.x[1:-1, 1:])


def addgrad(mesh, phi, du):
    """operators.py:49-56"""
    du.x[:, 1:] -= np.diff(phi, axis=1) * mesh.mskx[:, 1:]
    du.y[1:, :] -= np.diff(phi, axis=0) * mesh.msky[1:, :]


def sharp(mesh, u, U):
    """operators.py:59-64"""
    U.x[:] = u.x * (1 / mesh.dx ** 2)
    U.y[:] = u.y * (1 / mesh.dy ** 2)


def compute_vorticity(mesh, u, omega):
    """operators.py:67-77"""
    omega[1:, :] = -np.diff(u.x, axis=0)
    omega[:, 1:] += np.diff(u.y, axis=1)
    omega *= mesh.slipcoef


def compute_kinetic_energy(param, mesh, u, U, ke):
    """operators.py:80-101"""
    ke[:] = 0.
    if param.innerproduct == "classic":
        ke[:, :-1] = +u.x[:, 1:] * U.x[:, 1:] + u.x[:, :-1] * U.x[:, :-1]
        ke[:-1, :] += u.y[1:, :] * U.y[1:, :] + u.y[:-1, :] * U.y[:-1, :]
        ke *= mesh.msk * 0.25
    else:
        innerproduct(ke, U.x, u.x, mesh.ok.x, mesh.xshift, param.innerproduct)
        innerproduct(ke, U.y, u.y, mesh.ok.y, mesh.yshift, param.innerproduct)
        ke *= mesh.msk * 0.5


def compute_pressure(param, mesh, h, p):
    """operators.py:110-111"""
    p[:] = (param.g / mesh.area) * (h + mesh.hb)


def centerstovertices(mesh, omega, vomega, addto=False):
    """operators.py:126-133"""
    if addto:
        vomega[1:, 1:] += 0.25 * (omega[:-1, :-1] + omega[1:, :-1] + omega[:-1, 1:] + omega[1:, 1:])
    else:
        vomega[1:, 1:] = 0.25 * (omega[:-1, :-1] + omega[1:, :-1] + omega[:-1, 1:] + omega[1:, 1:])
    vomega *= mesh.mskv


def verticestocenters(mesh, vh, h):
    """operators.py:136-141"""
    m = mesh.mskv
    coef = m[:-1, :-1] + m[1:, :-1] + m[:-1, 1:] + m[1:, 1:]
    with np.errstate(divide="ignore", invalid="ignore"):
        h[:-1, :-1] = (1 / coef) * (vh[:-1, :-1] + vh[1:, :-1] + vh[:-1, 1:] + vh[1:, 1:])
    h *= mesh.msk


def perpgrad(mesh, psi, u, contravariant=False):
    """operators.py:144-149"""
    u.x[:-1, :] = -np.diff(psi, axis=0) * mesh.mskx[:-1, :]
    u.y[:, :-1] = np.diff(psi, axis=1) * mesh.msky[:, :-1]
    if contravariant:
        u.x[:] *= (1 / mesh.dy ** 2)
        u.y[:] *= (1 / mesh.dx ** 2)


def add_stretching(mesh, pv, h, anomaly):
    """operators.py:194-201"""
    f0, H = mesh.param.f0, mesh.param.H
    h0 = H * mesh.area - mesh.hb
    if anomaly:
        centerstovertices(mesh, h * (-f0 / H), pv, addto=True)
    else:
        centerstovertices(mesh, (h - h0) * (-f0 / H), pv, addto=True)


def thickness_from_psi(mesh, psi, h, anomaly):
    """operators.py:204-211"""
    f0, H, g = mesh.param.f0, mesh.param.H, mesh.param.g
    verticestocenters(mesh, psi * (f0 * mesh.area / g), h)
    if not anomaly:
        h += (H * mesh.area) - mesh.hb
    h *= mesh.msk


def qg_projection(mesh, u, h, pv, psi, anomaly=False):
    """operators.py:176-183"""
    compute_vorticity(mesh, u, pv)
    add_stretching(mesh, pv, h, anomaly)
    mesh.qg_helmholtz.solve(pv, psi)
    thickness_from_psi(mesh, psi, h, anomaly)
    perpgrad(mesh, psi, u)


def qg_inversion(mesh, pv, work, psi):
    """operators.py:186-191.  With beta != 0 the planetary vorticity
    ``mesh.f`` (a vertex array the caller sets, as
    experiments/inpreparation/gyre.py:24-29 does) is subtracted from the
    whole work array before the solve."""
    pvback = mesh.hb * mesh.qgcoef
    centerstovertices(mesh, pv - pvback, work)
    if mesh.param.beta != 0:
        work -= mesh.f
    mesh.qg_helmholtz.solve(work, psi)


# --------------------------------------------------------------------------
# states, integrators, models
# --------------------------------------------------------------------------
SPECS = {
    "euler": (("u", "U", "omega", "ke", "p", "div", "flx"), ("u",)),
    "eulerpsi": (("omega", "U", "psi", "vomega", "flx"), ("omega",)),
    "advection": (("q", "U", "flx"), ("q",)),
    "rsw": (("u", "h", "U", "omega", "ke", "p", "flx", "pv"), ("u", "h")),
    "qgrsw": (("u", "h", "U", "omega", "ke", "p", "flx", "pv", "psi"), ("u", "h")),
    "qg": (("pv", "U", "h", "flx", "work", "psi"), ("pv",)),
}
VECTORS = ("u", "U", "flx", "v")


def allocate(names, shape):
    """states.py:37-79"""
    T = namedtuple("state", names)
    return T(**{n: (Vector(np.zeros(shape), np.zeros(shape)) if n in VECTORS
                    else np.zeros(shape)) for n in names})


def _leaves(x):
    if isinstance(x, tuple):
        out = []
        for v in x:
            out += _leaves(v)
        return out
    return [x]


def addto(y, *args):
    """integrators.py:154-196: y += sum(c*x) with Python's left-to-right sum."""
    coefs, xs = args[::2], args[1::2]
    ly = _leaves(y)
    lx = [_leaves(x) for x in xs]
    for k, yy in enumerate(ly):
        yy[:] += sum((c * xx[k] for c, xx in zip(coefs, lx)))


def rk3(s, dt, rhs, diag, scratch):
    """integrators.py:92-107"""
    ds1, ds2, ds3 = scratch
    rhs(s, ds1)
    addto(s, dt, ds1)
    diag(s)
    rhs(s, ds2)
    addto(s, -3 * dt / 4, ds1, dt / 4, ds2)
    diag(s)
    rhs(s, ds3)
    addto(s, -dt / 12, ds1, -dt / 12, ds2, 2 * dt / 3, ds3)
    diag(s)


def ef(s, dt, rhs, diag, scratch):
    """integrators.py:82-89"""
    rhs(s, scratch[0])
    addto(s, dt, scratch[0])
    diag(s)


def enrk3(s, dt, rhs, diag, scratch):
    """integrators.py:109-124"""
    ds1, ds2, ds3 = scratch
    rhs(s, ds1)
    addto(s, dt / 3, ds1)
    diag(s)
    rhs(s, ds2)
    addto(s, -dt / 3 - 5 * dt / 48, ds1, 15 * dt / 16, ds2)
    diag(s)
    rhs(s, ds3)
    addto(s, 5 * dt / 48 + dt / 10, ds1, -7 * dt / 16, ds2, 2 * dt / 5, ds3)
    diag(s)


INTEGRATORS = {"rk3": (rk3, 3), "ef": (ef, 1), "enrk3": (enrk3, 3)}


def copyto(x, y):
    """integrators.py:140-151: y[:] = x[:] over the leaves."""
    for xx, yy in zip(_leaves(x), _leaves(y)):
        yy[:] = xx[:]


def lfra(s, dt, rhs, diag, scratch, first=False, gamma=0.1):
    """integrators.py:40-53 with rightpermute (integrators.py:134-137):
    Euler forward on the first iteration, then leap-frog and the
    Robert-Asselin filter of the middle level."""
    sb, sa, ds = scratch
    rhs(s, ds)
    if first:
        copyto(s, sb)
        copyto(s, sa)
        addto(s, dt, ds)
    else:
        addto(sa, 2 * dt, ds)
        addto(s, gamma, sa, gamma, sb, -2 * gamma, s)
        copyto(s, sb)
        copyto(sa, s)
        copyto(sb, sa)
    diag(s)


def get_rhs_and_diag(param, mesh):
    """equations.py:73-157, 190-210 for the in-scope models."""
    m = param.model
    F = mesh.fill

    if m == "euler":
        # equations.py:9-24 with pressure_projection, operators.py:114-119
        def rhs(s, ds):
            addvortexforce(param, mesh, s.U, s.omega, ds.u)
            addgrad(mesh, s.ke, ds.u)
            F(ds.u)

        def diag(s):
            sharp(mesh, s.u, s.U)
            div(mesh, s.U, s.div)
            mesh.poisson_centers.solve(-s.div * mesh.area, s.p)
            addgrad(mesh, s.p, s.u)
            F(s.u)
            sharp(mesh, s.u, s.U)
            compute_vorticity(mesh, s.u, s.omega)
            compute_kinetic_energy(param, mesh, s.u, s.U, s.ke)
            F(s.omega)
            F(s.ke)
    elif m == "eulerpsi":
        def rhs(s, ds):
            divflux(param, mesh, s.flx, s.omega, s.U, ds.omega)
            F(ds.omega)

        def diag(s):
            centerstovertices(mesh, s.omega, s.vomega)
            mesh.poisson_vertices.solve(s.vomega, s.psi)
            perpgrad(mesh, s.psi, s.U, contravariant=True)
            F(s.U)
    elif m == "advection":
        def rhs(s, ds):
            divflux(param, mesh, s.flx, s.q, s.U, ds.q)
            F(ds.q)

        def diag(s):
            pass
    elif m == "rsw":
        def rhs(s, ds):
            addvortexforce(param, mesh, s.U, s.omega, ds.u)
            addcoriolis(param, mesh, s.U, ds.u)
            addgrad(mesh, s.ke, ds.u)
            addgrad(mesh, s.p, ds.u)
            divflux(param, mesh, s.flx, s.h, s.U, ds.h)
            F(ds.u)
            F(ds.h)

        def diag(s):
            sharp(mesh, s.u, s.U)
            compute_vorticity(mesh, s.u, s.omega)
            compute_kinetic_energy(param, mesh, s.u, s.U, s.ke)
            compute_pressure(param, mesh, s.h, s.p)
            F(s.omega)
            F(s.ke)
    elif m == "qgrsw":
        def rhs(s, ds):
            addvortexforce(param, mesh, s.U, s.omega, ds.u)
            addcoriolis(param, mesh, s.U, ds.u)
            divflux(param, mesh, s.flx, s.h, s.U, ds.h)
            qg_projection(mesh, ds.u, ds.h, s.pv, s.psi, anomaly=True)
            F(ds.u)
            F(ds.h)

        def diag(s):
            sharp(mesh, s.u, s.U)
            compute_vorticity(mesh, s.u, s.omega)
            F(s.omega)
    elif m == "qg":
        def rhs(s, ds):
            divflux(param, mesh, s.flx, s.pv, s.U, ds.pv)
            F(ds.pv)

        def diag(s):
            qg_inversion(mesh, s.pv, s.work, s.psi)
            perpgrad(mesh, s.psi, s.U, contravariant=True)
    else:
        raise NotImplementedError(m)
    return rhs, diag


class Time:
    """timeline.py:4-35"""

    def __init__(self, param):
        self.t = 0.0
        self.ite = 0
        self.t0 = 0
        self.ite0 = 0
        self.param = param
        self.dt = param.dt if param.dt > 0 else 0.01
        self._c = 0.0

    @property
    def finished(self):
        return ((self.t >= self.t0 + self.param.tend)
                or (self.ite >= self.ite0 + self.param.maxite))

    def pushforward(self):
        y = self.dt - self._c
        t = self.t + y
        self._c = (t - self.t) - y
        self.t = t
        self.ite += 1


class Model:
    """Host FP64 model: model.py:13-93 minus IO/plots."""

    def __init__(self, param, msk=None):
        self.param = param
        self.mesh = Mesh(param, msk)
        names, prog = SPECS[param.model]
        self.state = allocate(names, self.mesh.shape)
        if param.integrator == "LFRA":
            caller, nst = lfra, 3
        else:
            caller, nst = INTEGRATORS[param.integrator]
        self._integ = caller
        self.scratch = [allocate(prog, self.mesh.shape) for _ in range(nst)]
        self.rhs, self.diag = get_rhs_and_diag(param, self.mesh)
        self.time = Time(param)
        self.prog = prog

    def set_dt(self):
        """model.py:71-87"""
        p = self.param
        if p.dt > 0:
            self.time.dt = p.dt
            return
        if p.model == "rsw":
            c = (p.g * p.H) ** 0.5
            maxU = c / self.mesh.dx + c / self.mesh.dy
        else:
            U = self.state.U
            maxU = np.max(np.abs(U.x)) + np.max(np.abs(U.y)) + 1e-99
        self.time.dt = min(p.cfl / maxU, p.dtmax)

    def add_forcing(self, forcing):
        """model.py:121-123 / equations.py:229-238: forcing(param, mesh, s,
        ds) runs after every rhs.  mesh.time is the model's Time, as the
        reference's forcing experiments set it (forced_convection.py:6)."""
        rhs = self.rhs
        self.mesh.time = self.time

        def newrhs(s, ds):
            rhs(s, ds)
            forcing(self.param, self.mesh, s, ds)

        self.rhs = newrhs

    def step(self):
        """RKIntegrator.step, integrators.py:76-79; LFRAintegrator.step,
        integrators.py:32-36 (first = ite == 0)"""
        args = (_ProgView(self.state, self.prog), self.time.dt, self.rhs,
                self.diag, [_ProgView(x, self.prog) for x in self.scratch])
        if self._integ is lfra:
            self._integ(*args, first=self.time.ite == 0, gamma=self.param.RAgamma)
        else:
            self._integ(*args)
        self.time.pushforward()

    def run(self, nsteps=None):
        """Model.run loop (model.py:45-51) without IO/animation."""
        n = 0
        while not self.time.finished and (nsteps is None or n < nsteps):
            self.set_dt()
            self.step()
            n += 1
        return n


class _ProgView(tuple):
    """A state namedtuple seen as the tuple of its prognostic fields, so that
    `addto` walks only the prognostic variables (the reference passes a
    `State` whose leading fields are the prognostic ones, states.py:50)."""

    def __new__(cls, full, prog):
        obj = super().__new__(cls, tuple(getattr(full, n) for n in prog))
        obj._full = full
        return obj

    def __getattr__(self, name):
        return getattr(object.__getattribute__(self, "_full"), name)
