"""Kernel-level parity of the HIP library against the oracle (GPU).

Integer work (masks, orders, index numbering) must be bit-exact; the FP64
stencils are evaluated in the reference's operation order and must be
bit-exact too; the Poisson solve is iterative and is compared with SuperLU
(the reference's direct solve) and with closed forms within a tolerance
stated per test.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from helpers import P, random_mask, assert_same

pytestmark = pytest.mark.gpu


def T(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def N(t):
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def L():
    from fluids2d_amd import _lib
    return _lib


def sync():
    import torch
    torch.cuda.synchronize()


# ---------------------------------------------------------------------------
# masks and orders
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed,xper,yper", [(0, False, False), (1, True, False),
                                            (2, True, True), (3, False, True)])
def test_masks_orders_bitexact(dev, L, seed, xper, yper):
    ny, nx = 23, 31
    m = random_mask(ny, nx, seed, xper, yper, nisl=4)
    NYP, NXP = m.shape
    dm = T(m, dev)
    import torch
    mx, my, mv = (torch.zeros_like(dm) for _ in range(3))
    L.call("f2d_set_masks", L.ptr(dm), NYP, NXP, L.ptr(mx), L.ptr(my), L.ptr(mv),
           L.stream_ptr())
    ex, ey, ev = O.set_masks(m)
    assert np.array_equal(N(mx), ex) and np.array_equal(N(my), ey)
    assert np.array_equal(N(mv), ev)
    for src in (m, ex, ey, ev):
        for shift in (1, -1, NXP, -NXP):
            for maxorder in (6, 5, 4, 2):
                o = torch.zeros_like(dm)
                dsrc = T(src, dev)
                L.call("f2d_set_order", L.ptr(dsrc), src.size, shift, maxorder,
                       L.ptr(o), L.stream_ptr())
                assert np.array_equal(N(o), O.set_order(src, shift, maxorder)), (shift, maxorder)
    for mode, noslip in ((0, None), (1, True), (2, ["left", "top"]), (2, ["bottom", "right"])):
        sides = 0
        if isinstance(noslip, list):
            sides = sum(b for s, b in (("left", 1), ("right", 2), ("bottom", 4), ("top", 8))
                        if s in noslip)
        c = torch.zeros(m.shape, dtype=torch.float64, device=dev)
        L.call("f2d_slipcoef", L.ptr(dm), NYP, NXP, mode, sides, 3, L.ptr(c),
               L.stream_ptr())
        assert np.array_equal(N(c), O.slipcoef(m, noslip, 3)), noslip
    G = torch.zeros(m.shape, dtype=torch.int32, device=dev)
    cnt = ctypes.c_int64()
    dev_ev = T(ev, dev)
    L.call("f2d_gindex", L.ptr(dev_ev), ev.size, L.ptr(G), ctypes.byref(cnt),
           L.stream_ptr())
    eG, eN = O.gindex(ev)
    assert cnt.value == eN and np.array_equal(N(G), eG)


# ---------------------------------------------------------------------------
# flat WENO loops
# ---------------------------------------------------------------------------
def _fields(shape, seed):
    rng = np.random.default_rng(seed)
    q = rng.standard_normal(shape)
    U = rng.standard_normal(shape)
    U[rng.random(shape) < 0.1] = 0.0  # U == 0 takes the mirrored branch
    V = rng.standard_normal(shape)
    return q, U, V


@pytest.mark.parametrize("method", ["weno", "upwind", "centered", "cweno"])
def test_flat_loops_bitexact(dev, L, method):
    ny, nx = 20, 27
    m = random_mask(ny, nx, 5, True, False)
    mesh = O.Mesh(P(model="eulerpsi", nx=nx, ny=ny, xperiodic=True), msk=m)
    q, U, V = _fields(m.shape, 7)
    n = q.size
    import torch
    dq, dU, dV = T(q, dev), T(U, dev), T(V, dev)
    for o, s in ((mesh.oc.x, 1), (mesh.oc.y, nx + 6)):
        flx = np.zeros_like(q)
        O.compflux(flx, U, q, o, s, method)
        d = torch.zeros(m.shape, dtype=torch.float64, device=dev)
        do = T(o, dev)
        L.call("f2d_compflux", L.F64, L.ptr(d), L.ptr(dU), L.ptr(dq), L.ptr(do), n, s,
               L.METHODS[method], L.stream_ptr())
        assert_same(N(d), flx, f'compflux s={s}')
    for o, s, s2, sg in ((mesh.ov.y, nx + 6, 1, 1), (mesh.ov.x, 1, nx + 6, -1)):
        du = np.zeros_like(q)
        O.vortexforce(du, V, q, o, s, s2, sg, method)
        d = torch.zeros(m.shape, dtype=torch.float64, device=dev)
        do = T(o, dev)
        L.call("f2d_vortexforce", L.F64, L.ptr(d), L.ptr(dV), L.ptr(dq), L.ptr(do), n, s,
               s2, sg, L.METHODS[method], L.stream_ptr())
        assert_same(N(d), du, f'vortexforce s={s}')
    for o, s in ((mesh.ok.x, 1), (mesh.ok.y, nx + 6)):
        ke0 = np.random.default_rng(3).standard_normal(q.shape)
        ke = ke0.copy()
        O.innerproduct(ke, U, q, o, s, method)
        d = T(ke0, dev)
        do = T(o, dev)
        L.call("f2d_innerproduct", L.F64, L.ptr(d), L.ptr(dU), L.ptr(dq), L.ptr(do), n, s,
               L.METHODS[method], L.stream_ptr())
        assert_same(N(d), ke, f'innerproduct s={s}')


# ---------------------------------------------------------------------------
# fused transport stage == divflux + fill + addto
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("method", ["weno", "upwind", "centered", "cweno"])
def test_flat_loops_fp32(dev, L, method):
    """FP32 loops (FMA and hardware-reciprocal forms of the reconstructions)
    against the oracle in FP64 on the same FP32-rounded inputs: within
    3e-6 of the largest value (a few FP32 ulps)."""
    import torch
    ny, nx = 20, 27
    m = random_mask(ny, nx, 5, True, False)
    mesh = O.Mesh(P(model="eulerpsi", nx=nx, ny=ny, xperiodic=True), msk=m)
    q, U, V = (a.astype(np.float32).astype(np.float64) for a in _fields(m.shape, 8))
    n = q.size
    dq, dU, dV = (T(a, dev).float() for a in (q, U, V))
    for o, s in ((mesh.oc.x, 1), (mesh.oc.y, nx + 6)):
        flx = np.zeros_like(q)
        O.compflux(flx, U, q, o, s, method)
        d = torch.zeros(m.shape, dtype=torch.float32, device=dev)
        L.call("f2d_compflux", L.F32, L.ptr(d), L.ptr(dU), L.ptr(dq), L.ptr(T(o, dev)), n, s,
               L.METHODS[method], L.stream_ptr())
        assert_same(N(d), flx, f"compflux fp32 s={s}", rtol=3e-6)
    for o, s, s2, sg in ((mesh.ov.y, nx + 6, 1, 1), (mesh.ov.x, 1, nx + 6, -1)):
        du = np.zeros_like(q)
        O.vortexforce(du, V, q, o, s, s2, sg, method)
        d = torch.zeros(m.shape, dtype=torch.float32, device=dev)
        L.call("f2d_vortexforce", L.F32, L.ptr(d), L.ptr(dV), L.ptr(dq), L.ptr(T(o, dev)), n,
               s, s2, sg, L.METHODS[method], L.stream_ptr())
        assert_same(N(d), du, f"vortexforce fp32 s={s}", rtol=3e-6)


@pytest.mark.parametrize("xper,yper,method,integ,ny,nx,tile", [
    (False, False, "weno", "rk3", 37, 70, 0), (True, False, "weno", "rk3", 37, 70, 0),
    (True, True, "weno", "enrk3", 37, 70, 0), (False, True, "upwind", "rk3", 37, 70, 0),
    (True, False, "centered", "ef", 37, 70, 0), (False, False, "cweno", "rk3", 37, 70, 0),
    (True, True, "weno", "rk3", 64, 256, 4), (False, False, "weno", "rk3", 64, 256, 0),
    (True, False, "weno", "rk3", 128, 128, 3)])
def test_transport_stage_bitexact(dev, L, xper, yper, method, integ, ny, nx, tile):
    """37x70: partial tiles in both directions (k_transport).  The other
    shapes are whole tiles (k_transport_full) whose count is a multiple of 8,
    so the tiles are handed out in XCD order: 16 tiles of 8x128, 16x64 or
    32x32."""
    import torch
    m = random_mask(ny, nx, 11, xper, yper, nisl=3)
    p = P(model="eulerpsi", nx=nx, ny=ny, xperiodic=xper, yperiodic=yper,
          compflux=method, tile=tile)
    mesh = O.Mesh(p, msk=m)
    rng = np.random.default_rng(4)
    q = rng.standard_normal(m.shape)
    Ux, Uy = rng.standard_normal(m.shape), rng.standard_normal(m.shape)
    F = O.Vector(Ux, Uy)
    dt = 0.0123456789
    g = L.make_grid(p)
    dq, dUx, dUy = T(q, dev), T(Ux, dev), T(Uy, dev)
    docx, docy, dm = T(mesh.oc.x, dev), T(mesh.oc.y, dev), T(m, dev)
    ddt = torch.tensor([dt], dtype=torch.float64, device=dev)
    nst = L.NSTAGES[integ]
    # oracle: rhs; fill; addto with the reference coefficients
    dss, qref = [], q.copy()
    dprev = []
    qcur = dq
    for k in range(nst):
        ds = np.zeros_like(q)
        O.divflux(p, mesh, O.Vector(np.zeros_like(q), np.zeros_like(q)), qref, F, ds)
        mesh.fill(ds)
        dss.append(ds)
        coefs = {"rk3": [(dt,), (-3 * dt / 4, dt / 4), (-dt / 12, -dt / 12, 2 * dt / 3)],
                 "ef": [(dt,)],
                 "enrk3": [(dt / 3,), (-dt / 3 - 5 * dt / 48, 15 * dt / 16),
                           (5 * dt / 48 + dt / 10, -7 * dt / 16, 2 * dt / 5)]}[integ][k]
        args = []
        for c, x in zip(coefs, dss):
            args += [c, x]
        O.addto((qref,), *[(a if i % 2 == 0 else (a,)) for i, a in enumerate(args)])
        # device
        qo = torch.empty_like(dq)
        dso = torch.zeros_like(dq)
        L.call("f2d_transport_stage", L.F64, ctypes.byref(g), L.METHODS[method],
               L.ptr(qcur), L.ptr(qo), L.ptr(dUx), L.ptr(dUy), L.ptr(docx), L.ptr(docy),
               L.ptr(dm), L.ptr(dprev[0] if len(dprev) > 0 else None),
               L.ptr(dprev[1] if len(dprev) > 1 else None), L.ptr(dso),
               L.INTEGRATORS[integ], k, L.ptr(ddt), L.stream_ptr())
        dprev.append(dso)
        sync()
        assert_same(N(dso), ds, f"tendency stage {k}")
        assert_same(N(qo), qref, f"update stage {k}")
        qcur = qo


# ---------------------------------------------------------------------------
# c2v, perpgrad + fill, CFL
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("xper,yper", [(False, False), (True, False), (True, True)])
def test_point_operators_bitexact(dev, L, xper, yper):
    import torch
    ny, nx = 30, 44
    p = P(model="eulerpsi", nx=nx, ny=ny, xperiodic=xper, yperiodic=yper, Lx=1.3)
    m = random_mask(ny, nx, 2, xper, yper)
    mesh = O.Mesh(p, msk=m)
    rng = np.random.default_rng(9)
    w = rng.standard_normal(m.shape)
    g = L.make_grid(p)
    for addto in (0, 1):
        v0 = rng.standard_normal(m.shape)
        v = v0.copy()
        O.centerstovertices(mesh, w, v, addto=bool(addto))
        dv, dw, dmv = T(v0, dev), T(w, dev), T(mesh.mskv, dev)
        L.call("f2d_c2v", L.F64, ctypes.byref(g), L.ptr(dw), L.ptr(dv),
               L.ptr(dmv), addto, L.stream_ptr())
        assert_same(N(dv), v, 'c2v')
    psi = rng.standard_normal(m.shape)
    for contra in (False, True):
        U = O.Vector(np.zeros(m.shape), np.zeros(m.shape))
        O.perpgrad(mesh, psi, U, contravariant=contra)
        mesh.fill(U)
        dUx = torch.zeros(m.shape, dtype=torch.float64, device=dev)
        dUy = torch.zeros_like(dUx)
        sx, sy = (1 / mesh.dy ** 2, 1 / mesh.dx ** 2) if contra else (1.0, 1.0)
        ddt = torch.zeros(1, dtype=torch.float64, device=dev)
        dpsi, dmx, dmy = T(psi, dev), T(mesh.mskx, dev), T(mesh.msky, dev)
        L.call("f2d_perpgrad_ex", L.F64, ctypes.byref(g), L.ptr(dpsi), L.ptr(dUx),
               L.ptr(dUy), L.ptr(dmx), L.ptr(dmy), sx, sy, 1,
               None, L.ptr(ddt), 0.8, 9e99, L.stream_ptr())
        assert_same(N(dUx), U.x, 'Ux')
        assert_same(N(dUy), U.y, 'Uy')
        mref = np.max(np.abs(U.x)) + np.max(np.abs(U.y)) + 1e-99
        assert N(ddt)[0] == min(0.8 / mref, 9e99)


# ---------------------------------------------------------------------------
# Poisson / Helmholtz
# ---------------------------------------------------------------------------
def _poisson_case(dev, L, p, m, maindiag, mode, nu=2, seed=0, solver="mg", c2v=False):
    import torch
    mesh = O.Mesh(p, msk=m)
    unk = O.unknown_mask(p, mesh.mskv)
    rng = np.random.default_rng(seed)
    b = rng.standard_normal(m.shape) * mesh.mskv
    if c2v:  # the fused form takes a cell field; the oracle its vertex average
        w = rng.standard_normal(m.shape) * mesh.msk
        O.fill(p, w)
        b = np.zeros(m.shape)
        O.centerstovertices(mesh, w, b)
    x_ref = np.zeros(m.shape)
    ref = O.Poisson2D(p, mesh, "v", maindiag=maindiag)
    ref.solve(b, x_ref)
    g = L.make_grid(p)
    opts = L.PoissonOpts()
    opts.mode, opts.nu = mode, nu
    opts.solver = L.SOLVERS[solver]
    h = ctypes.c_void_p()
    L.call("f2d_poisson_create", unk.ctypes.data_as(ctypes.c_void_p), ctypes.byref(g),
           maindiag, ctypes.byref(opts), ctypes.byref(h))
    dt = torch.float64 if mode == L.POISSON_F64 else torch.float32
    db = T(w if c2v else b, dev).to(dt)
    dx = torch.zeros(m.shape, dtype=dt, device=dev)
    it, rr = ctypes.c_int(), ctypes.c_double()
    rtol = 1e-13 if mode == L.POISSON_F64 else (1e-10 if mode == L.POISSON_F32_REFINE else 1e-5)
    maxit = 200 if mode != L.POISSON_F32_REFINE else 6
    L.call("f2d_poisson_solve_c2v" if c2v else "f2d_poisson_solve", h, L.ptr(db), L.ptr(dx),
           rtol, maxit, ctypes.byref(it), ctypes.byref(rr), L.stream_ptr())
    x = N(dx).astype(np.float64)
    L.call("f2d_poisson_destroy", h)
    err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    return err, it.value, rr.value, x, x_ref, unk


@pytest.mark.parametrize("ny,nx,xper,yper,islands,maindiag", [
    (64, 64, False, False, False, 0.0), (40, 40, False, False, False, 0.0),
    (64, 128, True, False, True, 0.0), (128, 128, True, True, False, 0.0),
    (96, 64, False, False, True, 0.3), (256, 256, False, False, False, 0.0),
    (256, 256, True, True, False, 0.0), (50, 36, True, True, True, 0.0)])
def test_poisson_f64_vs_superlu(dev, L, ny, nx, xper, yper, islands, maindiag):
    p = P(model="eulerpsi", nx=nx, ny=ny, xperiodic=xper, yperiodic=yper,
          Lx=nx / 64.0, Ly=ny / 64.0)
    m = random_mask(ny, nx, 3, xper, yper, nisl=5) if islands else \
        O.default_mask(ny, nx, 3, xper, yper)
    err, it, rr, x, xr, unk = _poisson_case(dev, L, p, m, maindiag, L.POISSON_F64)
    print(f"\n{ny}x{nx} xper={xper} yper={yper} isl={islands} md={maindiag}: "
          f"iters={it} relres={rr:.2e} err={err:.2e}")
    assert rr <= 1e-13
    assert err < 1e-10
    # entries outside the unknowns: untouched except for the halo fill
    inner = np.zeros_like(unk, dtype=bool)
    inner[3:-3, 3:-3] = True
    assert np.all(x[inner & (unk == 0)] == 0)


def test_poisson_periodic_vs_fft(dev, L):
    ny = nx = 128
    p = P(model="eulerpsi", nx=nx, ny=ny, xperiodic=True, yperiodic=True)
    m = O.default_mask(ny, nx, 3, True, True)
    err, it, rr, x, xr, unk = _poisson_case(dev, L, p, m, 0.0, L.POISSON_F64, seed=5)
    rng = np.random.default_rng(5)
    b = rng.standard_normal(m.shape) * O.Mesh(p, msk=m).mskv
    xf = O.fft_poisson_periodic(b[3:-3, 3:-3], 1 / nx, 1 / ny)
    e2 = np.linalg.norm(x[3:-3, 3:-3] - xf) / np.linalg.norm(xf)
    assert e2 < 1e-10, e2


@pytest.mark.parametrize("ny,nx,Ly,maindiag,c2v", [
    (128, 128, 1.3, 0.0, False), (64, 256, 0.7, 0.0, False), (256, 64, 3.0, 0.37, False),
    (8, 16, 0.2, 0.0, False), (512, 512, 8.0, 0.0, True), (32, 1024, 0.9, 1.5, True)])
def test_poisson_fft_f64_vs_superlu(dev, L, ny, nx, Ly, maindiag, c2v):
    """Direct spectral solve of the fully wet doubly periodic operator (and
    its Helmholtz form, dx != dy) against SuperLU on the same matrix."""
    p = P(model="eulerpsi", nx=nx, ny=ny, xperiodic=True, yperiodic=True,
          Lx=nx / 64.0, Ly=Ly)
    m = O.default_mask(ny, nx, 3, True, True)
    err, it, rr, x, xr, unk = _poisson_case(dev, L, p, m, maindiag, L.POISSON_F64,
                                            solver="fft", c2v=c2v)
    print(f"\nfft {ny}x{nx} md={maindiag} c2v={c2v}: relres={rr:.2e} err={err:.2e}")
    assert rr <= 1e-12 and err <= 1e-11
    # halo images are the periodic copies (the reference fills them after solving)
    xx = x.copy()
    O.fill(p, xx)
    assert np.array_equal(xx, x)


@pytest.mark.parametrize("ny,nx,xper,yper,maindiag,mode,c2v", [
    (64, 64, False, False, 0.0, "f64", False), (128, 64, False, False, 0.7, "f64", True),
    (64, 256, True, False, 0.0, "f64", False), (256, 32, False, True, 0.3, "f64", False),
    (512, 512, False, False, 0.0, "f64", True), (256, 256, False, False, 0.0, "refine", False),
    (128, 512, True, False, 0.0, "refine", True)])
def test_poisson_fft_closed_vs_superlu(dev, L, ny, nx, xper, yper, maindiag, mode, c2v):
    """Closed directions in the direct solve (odd extension on the doubled
    periodic grid) against SuperLU on the same masked operator: closed
    rectangles, channels, Helmholtz, the c2v right-hand side, FP64 and FP32
    with refinement.  Entries other than the unknowns are left alone, and a
    periodic direction gets its halo images."""
    p = P(model="eulerpsi", nx=nx, ny=ny, xperiodic=xper, yperiodic=yper,
          Lx=nx / 64.0, Ly=ny / 80.0)
    m = O.default_mask(ny, nx, 3, xper, yper)
    md = L.POISSON_F64 if mode == "f64" else L.POISSON_F32_REFINE
    err, it, rr, x, xr, unk = _poisson_case(dev, L, p, m, maindiag, md, solver="fft",
                                            c2v=c2v)
    print(f"\nfft closed {ny}x{nx} xper={xper} yper={yper} md={maindiag} {mode} c2v={c2v}: "
          f"relres={rr:.2e} err={err:.2e}")
    if mode == "f64":
        assert rr <= 1e-12 and err <= 1e-11
    else:
        assert rr <= 1e-9 and err < 1e-6
    inner = np.zeros(unk.shape, dtype=bool)
    inner[3:-3, 3:-3] = True
    assert np.all(x[inner & (unk == 0)] == 0)
    xx = x.copy()
    O.fill(p, xx)
    assert np.array_equal(xx, x)


@pytest.mark.parametrize("ny,nx,xper,yper,maindiag,mode,solver", [
    (64, 64, False, False, 0.0, "f64", "fft"), (128, 64, False, False, 0.7, "f64", "auto"),
    (64, 256, True, False, 0.0, "f64", "fft"), (256, 32, False, True, 0.3, "f64", "fft"),
    (128, 128, True, True, 0.0, "f64", "fft"), (512, 512, False, False, 0.0, "f64", "auto"),
    (256, 256, False, False, 0.0, "refine", "fft"), (128, 512, True, False, 0.0, "refine", "fft"),
    (64, 64, False, False, 0.0, "f64", "mg")])
def test_poisson_cells_fft_vs_superlu(dev, L, ny, nx, xper, yper, maindiag, mode, solver):
    """The Neumann cell operator (elliptic.py:186-190, the euler model's
    pressure) of an all-wet domain by the direct solve, closed directions as
    even extensions about the walls, against SuperLU; "mg" keeps the
    Jacobi-PCG solver.  Without a mass term the operator is singular: the
    right-hand side is given a nonzero mean (projected out by both) and the
    solutions are compared after removing their means."""
    import torch
    p = P(model="euler", nx=nx, ny=ny, xperiodic=xper, yperiodic=yper,
          Lx=nx / 64.0, Ly=ny / 80.0)
    m = O.default_mask(ny, nx, 3, xper, yper)
    mesh = O.Mesh(p, msk=m)
    unk = O.unknown_mask(p, mesh.msk)  # the cells, without periodic halo images
    rng = np.random.default_rng(11)
    b = (rng.standard_normal(m.shape) + 0.5) * unk
    inner = unk == 1
    bref = b.copy()
    if maindiag == 0.0:
        bref[inner] -= bref[inner].mean()
    x_ref = np.zeros(m.shape)
    O.Poisson2D(p, mesh, "c", maindiag=maindiag).solve(bref, x_ref)
    g = L.make_grid(p)
    opts = L.PoissonOpts()
    opts.mode = L.POISSON_F64 if mode == "f64" else L.POISSON_F32_REFINE
    opts.nu, opts.neumann = 2, 1
    opts.solver = L.SOLVERS[solver]
    h = ctypes.c_void_p()
    L.call("f2d_poisson_create", unk.ctypes.data_as(ctypes.c_void_p), ctypes.byref(g),
           maindiag, ctypes.byref(opts), ctypes.byref(h))
    try:
        method = L.lib().f2d_poisson_method(h)
        assert method == (3 if solver == "mg" else 2), method
        dt = torch.float64 if mode == "f64" else torch.float32
        db = T(b, dev).to(dt)
        dx = torch.zeros(m.shape, dtype=dt, device=dev)
        it, rr = ctypes.c_int(), ctypes.c_double()
        L.call("f2d_poisson_solve", h, L.ptr(db), L.ptr(dx), 1e-13, 2000, ctypes.byref(it),
               ctypes.byref(rr), L.stream_ptr())
        x = N(dx).astype(np.float64)
    finally:
        L.call("f2d_poisson_destroy", h)
    got, want = x[inner], x_ref[inner]
    if maindiag == 0.0:
        got, want = got - got.mean(), want - want.mean()
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"\ncells {solver} {ny}x{nx} xper={xper} yper={yper} md={maindiag} {mode}: "
          f"relres={rr.value:.2e} err={err:.2e}")
    if mode == "f64":
        # the singular Neumann operator's condition number grows as n^2:
        # 512^2 closed reaches 1.5e-11 against SuperLU at relres 6e-14
        assert rr.value <= 1e-12 and err <= 1e-10
    else:
        assert rr.value <= 1e-9 and err < 1e-6
    # periodic directions get their halo images, closed ones are left alone
    xx = x.copy()
    O.fill(p, xx)
    if xper and yper:
        assert np.array_equal(xx, x)
    elif xper:
        assert np.array_equal(xx[3:-3], x[3:-3])
    elif yper:
        assert np.array_equal(xx[:, 3:-3], x[:, 3:-3])


def test_poisson_fft_lds_limit(dev, L):
    """A closed 64 x 2048 basin transforms rows of 4096: FP32 fits the
    workgroup's LDS, FP64 does not, so "auto" takes multigrid in FP64 and an
    explicit "fft" request is refused."""
    p = P(model="eulerpsi", nx=2048, ny=64, Lx=32.0, Ly=1.0)
    m = O.default_mask(64, 2048, 3, False, False)
    mesh = O.Mesh(p, msk=m)
    unk = O.unknown_mask(p, mesh.mskv)
    g = L.make_grid(p)

    def create(mode, solver):
        opts = L.PoissonOpts()
        opts.mode, opts.nu, opts.solver = mode, 2, L.SOLVERS[solver]
        h = ctypes.c_void_p()
        L.call("f2d_poisson_create", unk.ctypes.data_as(ctypes.c_void_p), ctypes.byref(g),
               0.0, ctypes.byref(opts), ctypes.byref(h))
        method = L.lib().f2d_poisson_method(h)
        L.call("f2d_poisson_destroy", h)
        return method

    assert create(L.POISSON_F32_REFINE, "auto") == 2
    assert create(L.POISSON_F64, "auto") == 1
    with pytest.raises(L.F2DError, match="FP64"):
        create(L.POISSON_F64, "fft")


def test_poisson_cells_fft_refuses_dry_cells(dev, L):
    p = P(model="euler", nx=64, ny=64)
    m = random_mask(64, 64, 3, False, False, nisl=3)
    mesh = O.Mesh(p, msk=m)
    unk = O.unknown_mask(p, mesh.msk)
    g = L.make_grid(p)
    opts = L.PoissonOpts()
    opts.mode, opts.nu, opts.neumann, opts.solver = L.POISSON_F64, 2, 1, L.SOLVERS["fft"]
    h = ctypes.c_void_p()
    with pytest.raises(L.F2DError, match="dry cells"):
        L.call("f2d_poisson_create", unk.ctypes.data_as(ctypes.c_void_p), ctypes.byref(g),
               0.0, ctypes.byref(opts), ctypes.byref(h))


@pytest.mark.parametrize("mode,ny,nx", [("refine", 512, 512), ("refine", 64, 128),
                                        ("f32", 256, 256)])
def test_poisson_fft_f32_modes(dev, L, mode, ny, nx):
    p = P(model="eulerpsi", nx=nx, ny=ny, xperiodic=True, yperiodic=True,
          Lx=nx / 256.0, Ly=ny / 256.0)
    m = O.default_mask(ny, nx, 3, True, True)
    md = L.POISSON_F32_REFINE if mode == "refine" else L.POISSON_F32
    err, it, rr, x, xr, unk = _poisson_case(dev, L, p, m, 0.0, md, solver="fft")
    print(f"\nfft {mode} {ny}x{nx}: relres={rr:.2e} err={err:.2e}")
    if mode == "refine":
        assert rr <= 1e-10 and err < 1e-6
    else:
        assert err < 1e-5


@pytest.mark.parametrize("ny,nx,maindiag,nref,bmean,per", [
    (256, 256, 0.0, 1, 0.0, True), (128, 512, 0.0, 1, 3.0, True), (256, 128, 0.4, 1, 0.0, True),
    (512, 512, 0.0, 2, 0.5, True), (256, 256, 0.0, 1, 0.0, False), (128, 256, 0.2, 2, 1.0, False)])
def test_poisson_fft_f32_warm(dev, L, ny, nx, maindiag, nref, bmean, per):
    """Warm-started FP32 refinement (opts.fft_warm): from the handle's
    second solve on, the first refinement starts from the x passed in.
    Solve b1 (cold), then b2 = b1 + 1% noise from the first solution (halo
    of x spoiled: only its interior may be read), then b3 unrelated to b2,
    each against SuperLU on the same b; b may carry a mean, which the
    singular operator drops."""
    import torch
    p = P(model="eulerpsi", nx=nx, ny=ny, xperiodic=per, yperiodic=per,
          Lx=nx / 256.0, Ly=ny / 256.0)
    m = O.default_mask(ny, nx, 3, per, per)
    mesh = O.Mesh(p, msk=m)
    unk = O.unknown_mask(p, mesh.mskv)
    rng = np.random.default_rng(9)
    g = L.make_grid(p)
    opts = L.PoissonOpts()
    opts.mode, opts.solver, opts.fft_refine, opts.fft_warm = \
        L.POISSON_F32_REFINE, L.SOLVERS["fft"], nref, 1
    h = ctypes.c_void_p()
    L.call("f2d_poisson_create", unk.ctypes.data_as(ctypes.c_void_p), ctypes.byref(g),
           maindiag, ctypes.byref(opts), ctypes.byref(h))
    ref = O.Poisson2D(p, mesh, "v", maindiag=maindiag)
    dx = torch.zeros(m.shape, dtype=torch.float32, device=dev)
    b1 = (rng.standard_normal(m.shape) + bmean) * mesh.mskv
    b2 = (b1 + 0.01 * rng.standard_normal(m.shape)) * mesh.mskv
    b3 = (rng.standard_normal(m.shape) + bmean) * mesh.mskv
    errs = []
    for k, b in enumerate((b1, b2, b3)):
        if k == 1 and per:
            dx[:3, :] = 1e6
            dx[:, -3:] = -1e6
        x_ref = np.zeros(m.shape)
        ref.solve(b, x_ref)
        it, rr = ctypes.c_int(), ctypes.c_double()
        L.call("f2d_poisson_solve", h, L.ptr(T(b, dev).float()), L.ptr(dx), 1e-10, 6,
               ctypes.byref(it), ctypes.byref(rr), L.stream_ptr())
        x = N(dx).astype(np.float64)
        xi, xr = x[3:-3, 3:-3], x_ref[3:-3, 3:-3]
        if maindiag == 0.0 and per:  # the singular operator's solution is up to a constant
            xi, xr = xi - xi.mean(), xr - xr.mean()
        err = np.linalg.norm(xi - xr) / np.linalg.norm(xr)
        errs.append((err, rr.value))
        print(f"\nwarm {ny}x{nx} md={maindiag} nref={nref} solve {k}: relres={rr.value:.2e} "
              f"err={err:.2e}")
        if per:  # halo images of the periodic solution
            assert np.array_equal(x[:3, 3:-3], x[-6:-3, 3:-3])
            assert np.array_equal(x[3:-3, :3], x[3:-3, -6:-3])
    L.call("f2d_poisson_destroy", h)
    # k=0: cold FP32 solve + nref FP64-residual refinements; k=1: from the
    # nearby solution the refinement's FP32 error is 1% of a cold solve's;
    # k=2: from an unrelated solution one refinement is an FP32 solve (its
    # FP64 residual is that of FP32 rounding of the high wavenumbers)
    for err, rr in errs:
        assert err < 2e-6, errs
    assert errs[0][1] < 1e-9, errs
    assert errs[1][1] < 3e-7, errs
    assert errs[2][1] < (1e-9 if nref > 1 else 1e-4), errs


@pytest.mark.parametrize("ny,nx,mode,maindiag", [
    (1024, 1024, "f64", 0.0), (1024, 1024, "refine", 0.0), (2048, 256, "f64", 0.5),
    (256, 2048, "refine", 0.0), (1024, 2048, "f64", 0.0), (2048, 1024, "refine", 0.3),
    (256, 4096, "refine", 0.0), (4096, 512, "refine", 0.2)])
def test_poisson_fft_large_vs_dft(dev, L, ny, nx, mode, maindiag):
    """Large direct solves (256..4096 points a transform; 512 threads a
    workgroup from 2048 points, 1024 from 4096) against numpy's 2-D DFT
    closed form of the same operator (oracle.fft_poisson_periodic)."""
    import torch
    p = P(model="eulerpsi", nx=nx, ny=ny, xperiodic=True, yperiodic=True,
          Lx=nx / 512.0, Ly=ny / 700.0)
    m = O.default_mask(ny, nx, 3, True, True)
    unk = np.zeros(m.shape, dtype=np.int8)
    unk[3:-3, 3:-3] = 1
    rng = np.random.default_rng(17)
    b = np.zeros(m.shape)
    b[3:-3, 3:-3] = rng.standard_normal((ny, nx)) + 0.25
    g = L.make_grid(p)
    opts = L.PoissonOpts()
    opts.mode = L.POISSON_F64 if mode == "f64" else L.POISSON_F32_REFINE
    opts.solver = L.SOLVERS["fft"]
    h = ctypes.c_void_p()
    L.call("f2d_poisson_create", unk.ctypes.data_as(ctypes.c_void_p), ctypes.byref(g),
           maindiag, ctypes.byref(opts), ctypes.byref(h))
    dt = torch.float64 if mode == "f64" else torch.float32
    dx = torch.zeros(m.shape, dtype=dt, device=dev)
    it, rr = ctypes.c_int(), ctypes.c_double()
    L.call("f2d_poisson_solve", h, L.ptr(T(b, dev).to(dt)), L.ptr(dx), 1e-10, 6,
           ctypes.byref(it), ctypes.byref(rr), L.stream_ptr())
    x = N(dx).astype(np.float64)
    L.call("f2d_poisson_destroy", h)
    bi = b[3:-3, 3:-3]
    if maindiag == 0.0:
        bi = bi - bi.mean()
    xr = O.fft_poisson_periodic(bi, p.Lx / nx, p.Ly / ny, maindiag)
    xi = x[3:-3, 3:-3]
    if maindiag == 0.0:
        xi = xi - xi.mean()
    err = np.linalg.norm(xi - xr) / np.linalg.norm(xr)
    print(f"\nfft {mode} {ny}x{nx} md={maindiag}: relres={rr.value:.2e} err={err:.2e}")
    assert err <= (1e-11 if mode == "f64" else 1e-6)
    assert rr.value <= (1e-11 if mode == "f64" else 1e-9)
    assert np.array_equal(x[:3, 3:-3], x[-6:-3, 3:-3])
    assert np.array_equal(x[3:-3, :3], x[3:-3, -6:-3])
    assert np.array_equal(x[:3, :3], x[-6:-3, -6:-3])


def test_poisson_fft_matches_mg(dev, L):
    p = P(model="eulerpsi", nx=256, ny=256, xperiodic=True, yperiodic=True)
    m = O.default_mask(256, 256, 3, True, True)
    e1, _, _, x1, _, _ = _poisson_case(dev, L, p, m, 0.0, L.POISSON_F64, solver="fft", seed=4)
    e2, _, _, x2, _, _ = _poisson_case(dev, L, p, m, 0.0, L.POISSON_F64, solver="mg", seed=4)
    assert np.linalg.norm(x1 - x2) / np.linalg.norm(x2) < 1e-11


def test_poisson_fft_refuses_masked(dev, L):
    p = P(model="eulerpsi", nx=64, ny=64, xperiodic=True, yperiodic=True)
    m = random_mask(64, 64, 3, True, True, nisl=2)
    mesh = O.Mesh(p, msk=m)
    unk = O.unknown_mask(p, mesh.mskv)
    g = L.make_grid(p)
    opts = L.PoissonOpts()
    opts.solver = L.SOLVERS["fft"]
    h = ctypes.c_void_p()
    with pytest.raises(L.F2DError, match="FFT solver requested"):
        L.call("f2d_poisson_create", unk.ctypes.data_as(ctypes.c_void_p), ctypes.byref(g),
               0.0, ctypes.byref(opts), ctypes.byref(h))


@pytest.mark.parametrize("mode,ny,nx,xper,yper,islands", [
    ("refine", 256, 256, True, True, False), ("f32", 256, 256, True, True, False),
    ("refine", 256, 256, False, False, False), ("refine", 128, 256, True, False, True),
    ("f32", 128, 256, True, False, True)])
def test_poisson_f32_modes(dev, L, mode, ny, nx, xper, yper, islands):
    p = P(model="eulerpsi", nx=nx, ny=ny, xperiodic=xper, yperiodic=yper,
          Lx=nx / 256.0, Ly=ny / 256.0)
    m = random_mask(ny, nx, 3, xper, yper, nisl=5) if islands else \
        O.default_mask(ny, nx, 3, xper, yper)
    md = L.POISSON_F32_REFINE if mode == "refine" else L.POISSON_F32
    err, it, rr, x, xr, unk = _poisson_case(dev, L, p, m, 0.0, md)
    print(f"\nmode={mode} {ny}x{nx} inner iters={it} relres={rr:.2e} err={err:.2e}")
    if mode == "refine":
        # FP64 residual, FP32 inner solves: error limited by the FP32 output
        assert rr <= 1e-10
        assert err < 1e-6
    else:
        assert err < 1e-3
