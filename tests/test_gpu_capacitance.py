"""The capacitance solve of masked vertex operators (f2d_capacitance.hip)
against SuperLU on the reference's matrix (elliptic.py:114-195, "v"):
closed basins from a smoothed-noise level set, channels and y-periodic
strips with islands, Helmholtz terms, dx != dy, the fused c2v right-hand
side, FP64 and FP32 with FP64 refinement (cold and warm)."""
import ctypes
import os
import sys

import numpy as np
import pytest

import oracle as O
from helpers import P, random_mask

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


@pytest.fixture(scope="module")
def L():
    from fluids2d_amd import _lib
    return _lib


def _mask(kind, ny, nx, xper, yper, seed):
    if kind == "basin":
        from bench import basin_mask
        return basin_mask(nx, seed, ny=ny)
    return random_mask(ny, nx, seed, xper, yper, nisl=6)


def _handle(L, unk, p, maindiag, mode, solver="auto"):
    g = L.make_grid(p)
    opts = L.PoissonOpts()
    opts.mode, opts.nu = mode, 2
    opts.solver = L.SOLVERS[solver]
    opts.fft_warm = 1
    h = ctypes.c_void_p()
    L.call("f2d_poisson_create", unk.ctypes.data_as(ctypes.c_void_p), ctypes.byref(g),
           maindiag, ctypes.byref(opts), ctypes.byref(h))
    return h


def _solve(L, h, db, dx, c2v, rtol=1e-12, maxit=10):
    it, rr = ctypes.c_int(), ctypes.c_double()
    L.call("f2d_poisson_solve_c2v" if c2v else "f2d_poisson_solve", h, L.ptr(db), L.ptr(dx),
           rtol, maxit, ctypes.byref(it), ctypes.byref(rr), L.stream_ptr())
    return it.value, rr.value


@pytest.mark.parametrize("ny,nx,xper,yper,kind,maindiag,mode,c2v", [
    (64, 64, False, False, "basin", 0.0, "f64", False),
    (128, 128, False, False, "basin", 0.0, "f64", True),
    (64, 128, True, False, "islands", 0.0, "f64", False),
    (128, 64, False, True, "islands", 0.3, "f64", False),
    (128, 128, True, True, "islands", 0.5, "f64", True),
    (256, 256, False, False, "basin", 2.0, "f64", False),
    (256, 256, False, False, "basin", 0.0, "refine", False),
    (128, 256, True, False, "islands", 0.0, "refine", True)])
def test_capacitance_vs_superlu(dev, L, ny, nx, xper, yper, kind, maindiag, mode, c2v):
    import torch
    p = P(model="eulerpsi", nx=nx, ny=ny, xperiodic=xper, yperiodic=yper,
          Lx=nx / 64.0, Ly=ny / 80.0)
    m = _mask(kind, ny, nx, xper, yper, 7)
    mesh = O.Mesh(p, msk=m)
    unk = O.unknown_mask(p, mesh.mskv)
    rng = np.random.default_rng(3)
    if c2v:
        w = rng.standard_normal(m.shape) * mesh.msk
        O.fill(p, w)
        b = np.zeros(m.shape)
        O.centerstovertices(mesh, w, b)
    else:
        b = rng.standard_normal(m.shape) * mesh.mskv
    x_ref = np.zeros(m.shape)
    O.Poisson2D(p, mesh, "v", maindiag=maindiag).solve(b, x_ref)
    md = L.POISSON_F64 if mode == "f64" else L.POISSON_F32_REFINE
    h = _handle(L, unk, p, maindiag, md)
    assert L.lib().f2d_poisson_method(h) == L.SOLVERS["cap"]
    dt = torch.float64 if mode == "f64" else torch.float32
    db = torch.from_numpy(w if c2v else b).to(dev, dt)
    dx = torch.zeros(m.shape, dtype=dt, device=dev)
    it, rr = _solve(L, h, db, dx, c2v)
    L.call("f2d_poisson_destroy", h)
    x = dx.double().cpu().numpy()
    err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    print(f"\ncapacitance {ny}x{nx} xper={xper} yper={yper} {kind} md={maindiag} {mode} "
          f"c2v={c2v}: passes={it} relres={rr:.2e} err={err:.2e}")
    if mode == "f64":
        assert rr <= 1e-12 and err <= 1e-11
    else:
        assert rr <= 1e-9 and err <= 1e-7
    # entries other than the unknowns are left alone; periodic halo images
    inner = np.zeros(unk.shape, dtype=bool)
    inner[3:-3, 3:-3] = True
    assert np.all(x[inner & (unk == 0)] == 0)
    xx = x.copy()
    O.fill(p, xx)
    assert np.array_equal(xx, x)


def test_capacitance_warm_fp32(dev, L):
    """FP32 + refinement from the caller's x (the previous solution): one
    refinement pass per solve, to FP64 accuracy of the refined result."""
    import torch
    ny = nx = 128
    p = P(model="rsw", nx=nx, ny=ny, f0=10.0)
    m = _mask("basin", ny, nx, False, False, 11)
    mesh = O.Mesh(p, msk=m)
    unk = O.unknown_mask(p, mesh.mskv)
    maindiag = mesh.area * 10.0 ** 2
    ref = O.Poisson2D(p, mesh, "v", maindiag=maindiag)
    h = _handle(L, unk, p, maindiag, L.POISSON_F32_REFINE)
    dx = torch.zeros(m.shape, dtype=torch.float32, device=dev)
    rng = np.random.default_rng(5)
    b0 = rng.standard_normal(m.shape) * mesh.mskv
    errs = []
    for k in range(4):
        b = b0 + 0.02 * k * rng.standard_normal(m.shape) * mesh.mskv
        x_ref = np.zeros(m.shape)
        ref.solve(b, x_ref)
        it, rr = _solve(L, h, torch.from_numpy(b).to(dev, torch.float32), dx, False)
        x = dx.double().cpu().numpy()
        errs.append(np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref))
        print(f"\nsolve {k}: passes={it} relres={rr:.2e} err={errs[-1]:.2e}")
        assert it == (2 if k == 0 else 1)
    assert max(errs) <= 1e-7
    # the same warm refinement without statistics, fused into the spectral
    # passes (residual in the forward row pass, accumulation in the inverse)
    b = b0 + 0.05 * rng.standard_normal(m.shape) * mesh.mskv
    x_ref = np.zeros(m.shape)
    ref.solve(b, x_ref)
    L.call("f2d_poisson_solve", h, L.ptr(torch.from_numpy(b).to(dev, torch.float32)), L.ptr(dx),
           1e-12, 10, None, None, L.stream_ptr())
    err = np.linalg.norm(dx.double().cpu().numpy() - x_ref) / np.linalg.norm(x_ref)
    print(f"\nfused warm solve: err={err:.2e}")
    assert err <= 1e-7
    # a plain FP32 solve (no refinement): FP32 accuracy of the solution
    L.call("f2d_poisson_set_refine", h, 0)
    b = b0 + 0.1 * rng.standard_normal(m.shape) * mesh.mskv
    x_ref = np.zeros(m.shape)
    ref.solve(b, x_ref)
    it, rr = _solve(L, h, torch.from_numpy(b).to(dev, torch.float32), dx, False)
    err = np.linalg.norm(dx.double().cpu().numpy() - x_ref) / np.linalg.norm(x_ref)
    print(f"\nplain FP32 solve: passes={it} relres={rr:.2e} err={err:.2e}")
    assert it == 1 and err <= 1e-5
    L.call("f2d_poisson_destroy", h)


def test_capacitance_choice(dev, L):
    """auto takes the capacitance solve for a masked operator inside an
    FFT-able rectangle and the FFT solve for the full rectangle; an explicit
    "cap" is refused where it cannot apply (doubly periodic without a mass
    term: singular rectangle); "mg" keeps multigrid."""
    from fluids2d_amd._lib import F2DError
    ny = nx = 64
    pc = P(model="eulerpsi", nx=nx, ny=ny)
    mb = _mask("basin", ny, nx, False, False, 3)
    unk = O.unknown_mask(pc, O.Mesh(pc, msk=mb).mskv)
    for solver, want in (("auto", "cap"), ("mg", "mg"), ("cap", "cap")):
        h = _handle(L, unk, pc, 0.0, L.POISSON_F64, solver)
        assert L.lib().f2d_poisson_method(h) == L.SOLVERS[want], solver
        L.call("f2d_poisson_destroy", h)
    full = O.unknown_mask(pc, O.Mesh(pc).mskv)
    h = _handle(L, full, pc, 0.0, L.POISSON_F64, "auto")
    assert L.lib().f2d_poisson_method(h) == L.SOLVERS["fft"]
    L.call("f2d_poisson_destroy", h)
    pp = P(model="eulerpsi", nx=nx, ny=ny, xperiodic=True, yperiodic=True)
    mi = _mask("islands", ny, nx, True, True, 4)
    unk = O.unknown_mask(pp, O.Mesh(pp, msk=mi).mskv)
    with pytest.raises(F2DError, match="capacitance"):
        _handle(L, unk, pp, 0.0, L.POISSON_F64, "cap")
    h = _handle(L, unk, pp, 0.0, L.POISSON_F64, "auto")
    assert L.lib().f2d_poisson_method(h) == L.SOLVERS["mg"]
    L.call("f2d_poisson_destroy", h)


def test_capacitance_batch(dev, L):
    """f2d_poisson_batch_solve against each member's own f2d_poisson_solve,
    bit for bit (solution and halo images): two groups -- closed 64^2
    basins and x-periodic 64x128 channels with islands -- of different
    masks, cold first solve, then warm solves of changed right-hand sides
    written into the bound arrays; a batch with some members warm and some
    cold; refusals of members that cannot share a batch."""
    import torch
    from fluids2d_amd._lib import F2DError
    mode = L.POISSON_F32_REFINE
    rng = np.random.default_rng(21)

    def group(ny, nx, xper, kind, seeds, maindiag):
        p = P(model="rsw", nx=nx, ny=ny, xperiodic=xper, Lx=nx / 64.0, Ly=ny / 64.0)
        out = []
        for sd in seeds:
            mesh = O.Mesh(p, msk=_mask(kind, ny, nx, xper, False, sd))
            unk = O.unknown_mask(p, mesh.mskv)
            out.append(dict(p=p, mskv=mesh.mskv, hb=_handle(L, unk, p, maindiag, mode),
                            hs=_handle(L, unk, p, maindiag, mode),
                            b=torch.zeros(unk.shape, dtype=torch.float32, device=dev),
                            xb=torch.zeros(unk.shape, dtype=torch.float32, device=dev),
                            xs=torch.zeros(unk.shape, dtype=torch.float32, device=dev)))
        return out

    groups = [group(64, 64, False, "basin", (1, 2, 3), 0.0),
              group(64, 128, True, "islands", (4, 5), 0.4)]

    def batch(ms):
        arr = ctypes.c_void_p * len(ms)
        h = ctypes.c_void_p()
        L.call("f2d_poisson_batch_create", arr(*(m["hb"] for m in ms)), len(ms),
               arr(*(L.ptr(m["b"]) for m in ms)), arr(*(L.ptr(m["xb"]) for m in ms)),
               ctypes.byref(h))
        return h

    batches = [batch(ms) for ms in groups]
    # solves 0-2 with the handles' refinement (cold, then warm), 3-4 plain FP32
    # (nref 0: the batch argument, the solo handles' f2d_poisson_set_refine),
    # 5 warm again from a plain solution
    for rnd in range(6):
        nref = 0 if rnd in (3, 4) else -1
        for ms, bt in zip(groups, batches):
            for m in ms:
                m["b"].copy_(torch.from_numpy(rng.standard_normal(m["mskv"].shape) *
                                              m["mskv"]).to(dev, torch.float32))
                L.call("f2d_poisson_set_refine", m["hs"], nref)
                L.call("f2d_poisson_solve", m["hs"], L.ptr(m["b"]), L.ptr(m["xs"]), 1e-9, 10,
                       None, None, L.stream_ptr())
            L.call("f2d_poisson_batch_solve", bt, nref, L.stream_ptr())
        torch.cuda.synchronize()
        for g, ms in enumerate(groups):
            for i, m in enumerate(ms):
                assert torch.equal(m["xb"], m["xs"]), (rnd, g, i)
                assert float(m["xb"].abs().max()) > 0
    print(f"\nbatched capacitance: {[len(ms) for ms in groups]} members, 6 solves bit-equal")

    # members warm and cold together: each member's own solve
    ms = groups[0]
    fresh = group(64, 64, False, "basin", (1, 2, 3), 0.0)
    for m, f in zip(ms, fresh):
        f["b"].copy_(m["b"])
    L.call("f2d_poisson_solve", fresh[1]["hb"], L.ptr(fresh[1]["b"]), L.ptr(fresh[1]["xb"]),
           1e-9, 10, None, None, L.stream_ptr())
    L.call("f2d_poisson_solve", fresh[1]["hs"], L.ptr(fresh[1]["b"]), L.ptr(fresh[1]["xs"]),
           1e-9, 10, None, None, L.stream_ptr())
    bt = batch(fresh)
    L.call("f2d_poisson_batch_solve", bt, -1, L.stream_ptr())
    for f in fresh:
        L.call("f2d_poisson_solve", f["hs"], L.ptr(f["b"]), L.ptr(f["xs"]), 1e-9, 10,
               None, None, L.stream_ptr())
    torch.cuda.synchronize()
    for i, f in enumerate(fresh):
        assert torch.equal(f["xb"], f["xs"]), i
    L.call("f2d_poisson_batch_destroy", bt)

    # members that cannot share a batch
    with pytest.raises(F2DError, match="share the rectangle"):
        batch([groups[0][0], groups[1][0]])
    p = groups[0][0]["p"]
    unk = O.unknown_mask(p, groups[0][0]["mskv"])
    h64 = _handle(L, unk, p, 0.0, L.POISSON_F64)
    hmg = _handle(L, unk, p, 0.0, mode, "mg")
    for h in (h64, hmg):
        with pytest.raises(F2DError, match="FP32 capacitance"):
            batch([dict(groups[0][0], hb=h)])
        L.call("f2d_poisson_destroy", h)
    for bt in batches:
        L.call("f2d_poisson_batch_destroy", bt)
    for ms in groups + [fresh]:
        for m in ms:
            L.call("f2d_poisson_destroy", m["hb"])
            L.call("f2d_poisson_destroy", m["hs"])
