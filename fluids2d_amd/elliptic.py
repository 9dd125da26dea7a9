"""Poisson2D: masked Poisson/Helmholtz solve on the vertex grid
(elliptic.py:71-203), by the library's multigrid-preconditioned CG, or by its
direct FFT solve when the operator is fully wet and doubly periodic.

The reference factorises the masked 5-point operator with SuperLU once per
``mesh.finalize()`` and solves with it at every stage.  Here the operator is
matrix-free; ``create`` builds the multigrid hierarchy from the mask (and the
spectral tables when they apply) and every iterative solve is warm-started
from the previous solution held in ``x``.
"""
import ctypes

import numpy as np
import torch

from . import _lib

DEFAULT_RTOL = {"fp64": 1e-13, "fp32": 1e-9}
DEFAULT_MAXIT = {"fp64": 100, "fp32": 8}


def _destroy(h, name="f2d_poisson_destroy"):
    try:
        getattr(_lib.lib(), name)(h)
    except Exception:
        pass


# (handle, destructor) pairs released during a graph capture, destroyed at
# the next safe point in release order
_PENDING = []


def _flush_pending():
    """Destroy handles whose owners went away during a graph capture."""
    while _PENDING and not _lib.capturing():
        _destroy(*_PENDING.pop(0))


def _release(h, name):
    """Destroy a library handle now, or after the capture in progress
    (freeing device memory would invalidate it)."""
    if h is None or not h.value:
        return
    try:
        if _lib.capturing():
            _PENDING.append((h, name))
            return
        _destroy(h, name)
    except TypeError:
        # at interpreter shutdown the module's globals may already be gone;
        # the process releases the device memory itself
        pass


class PoissonNotConverged(RuntimeError):
    pass


TUNABLE = ("nu", "nu_coarse", "nu_tail", "tail_points", "tail_threads", "refine_inner",
           "coarse_sweeps")


class Poisson2D:
    def __init__(self, mesh, location, maindiag=0, tuning=None):
        """tuning: solver schedule fields of f2d_poisson_opts (TUNABLE), as
        picked by fluids2d_amd.autotune; default ``param.poisson_opts``."""
        assert location in ["c", "v"]
        self.mesh = mesh
        self.location = location
        self.maindiag = float(maindiag)
        p = mesh.param
        self.precision = getattr(p, "precision", "fp64")
        self.msk = get_msk_from_mesh(mesh, location)
        self.G, self.N = self._gindex()
        opts = _lib.PoissonOpts()
        opts.mode = _lib.POISSON_F64 if self.precision == "fp64" else _lib.POISSON_F32_REFINE
        opts.nu = int(getattr(p, "poisson_nu", 2) or 2)
        opts.solver = _lib.SOLVERS[getattr(p, "poisson_solver", "auto") or "auto"]
        # direct FFT solve in FP32: refine from the previous solution held in
        # x (the state's streamfunction) instead of a cold FP32 solve
        opts.fft_warm = int(bool(getattr(p, "poisson_warm", True)))
        # "c": the Neumann cell operator (elliptic.py:186-190): the direct
        # solve when every cell is wet, else Jacobi-PCG in FP64.  Its
        # right-hand side (the divergence a stage's update left) is not
        # close to the previous stage's: cold FP32 solves, refined once
        opts.neumann = int(location == "c")
        if location == "c":
            opts.fft_warm = 0
        tuning = getattr(p, "poisson_opts", None) if tuning is None else tuning
        self.tuning = dict(tuning or {})
        for k, v in self.tuning.items():
            if k not in TUNABLE:
                raise ValueError(f"poisson_opts: {k} is not a tunable solver option")
            setattr(opts, k, int(v))
        rt = getattr(p, "poisson_rtol", None)
        mi = getattr(p, "poisson_maxit", None)
        self.rtol = DEFAULT_RTOL[self.precision] if rt is None else float(rt)
        self.maxit = DEFAULT_MAXIT[self.precision] if mi is None else int(mi)
        if location == "c":
            # CG iterations (FP64 inside): grows with the grid
            self.rtol = 1e-12 if rt is None else float(rt)
            self.maxit = 8 * max(mesh.nx, mesh.ny) + 100 if mi is None else int(mi)
        self.dtype = torch.float64 if self.precision == "fp64" else torch.float32
        _flush_pending()
        unk = np.ascontiguousarray(self.msk, dtype=np.int8)
        h = ctypes.c_void_p()
        _lib.call("f2d_poisson_create", unk.ctypes.data_as(ctypes.c_void_p), mesh.grid_ref(),
                  self.maindiag, ctypes.byref(opts), ctypes.byref(h))
        self._h = h
        # "fft": direct spectral solve (exact to rounding, no rtol test);
        # "mg": multigrid-preconditioned CG to rtol
        self.method = {2: "fft", 3: "neumann", 4: "capacitance"}.get(
            _lib.lib().f2d_poisson_method(h), "mg")
        self.last_iters = 0
        self.last_relres = 0.0
        # iterations (refinements in FP32) the no-sync solves enqueue:
        # learned from the solves checked on the host, plus a margin
        self.est = None
        self.margin = 2
        # solve calls made from the host (``calls``) and solves replayed by
        # captured HIP graphs that the library's own counter does not see
        # (direct solves enqueued without statistics count on the host)
        self.calls = 0
        self.replayed = 0

    def _gindex(self):
        """get_Gindex (elliptic.py:198-203), computed on the device."""
        d = torch.from_numpy(np.ascontiguousarray(self.msk, dtype=np.int8)).to(self.mesh.device)
        G = torch.zeros(d.shape, dtype=torch.int32, device=d.device)
        n = ctypes.c_int64()
        _lib.call("f2d_gindex", _lib.ptr(d), d.numel(), _lib.ptr(G), ctypes.byref(n),
                  _lib.stream_ptr())
        return G.cpu().numpy(), int(n.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        self._h = None
        _release(h, "f2d_poisson_destroy")


    def handle(self):
        return self._h

    def solve(self, b, x, check=True, stats=True):
        """Poisson2D.solve(b, x) (elliptic.py:80-87): x is written at the
        unknowns, then its halo is filled.  numpy arrays are solved through a
        device copy; device tensors in place.  Raises PoissonNotConverged if
        the relative residual exceeds the tolerance after maxit iterations.

        stats=False (device tensors, once a host-checked solve has set the
        learned count): no host round trip, as in ``solve_c2v``."""
        host = not isinstance(x, torch.Tensor)
        if not host and not stats and self.est is not None:
            self.calls += 1
            n = min(self.est + self.margin, self.maxit)
            _lib.call("f2d_poisson_solve", self._h, _lib.ptr(b), _lib.ptr(x), self.rtol, n,
                      None, None, _lib.stream_ptr())
            return None
        if host:
            db = torch.from_numpy(np.ascontiguousarray(b, dtype=np.float64)).to(
                self.mesh.device).to(self.dtype)
            dx = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(
                self.mesh.device).to(self.dtype)
        else:
            db, dx = b, x
            assert db.dtype == self.dtype and dx.dtype == self.dtype
        it, rr = ctypes.c_int(), ctypes.c_double()
        self.calls += 1
        _lib.call("f2d_poisson_solve", self._h, _lib.ptr(db), _lib.ptr(dx), self.rtol,
                  self.maxit, ctypes.byref(it), ctypes.byref(rr), _lib.stream_ptr())
        self.last_iters, self.last_relres = it.value, rr.value
        if check and not self._accepted(rr.value):
            raise PoissonNotConverged(
                f"Poisson solve reached relres {rr.value:.3e} > {self.rtol:.1e} "
                f"after {it.value} iterations")
        if not host:
            ref = int(self._stats_host()[3]) if self.precision != "fp64" else None
            self._learn(it.value, ref)
        if host:
            x[...] = dx.double().cpu().numpy()
        return None

    def _accepted(self, relres):
        if self.method in ("fft", "capacitance"):
            # direct: exact up to rounding.  FP32 transforms refined once from
            # the previous solution: the FP64 residual of FP32 rounding at
            # the high wavenumbers (the operator amplifies them by k^2)
            return relres <= (1e-6 if self.precision == "fp64" else 1e-4)
        return relres <= self.rtol

    def _learn(self, it, refinements=None):
        n = refinements if self.precision != "fp64" else it
        if n is not None:
            self.est = n if self.est is None else max(self.est, n)

    def solve_c2v(self, omega, x, stats=True):
        """Fused eulerpsi/qg diagnostic: b = centerstovertices(omega) formed on
        the fly, then solve (equations.py:80-82, operators.py:186-191).

        stats=True: the solve checks convergence on the host and raises if it
        failed.  stats=False: no host round trip; the solve enqueues the
        learned iteration count plus a margin (iterations past convergence
        return at once) and leaves its outcome in the device counters that
        ``read_stats()`` returns."""
        self.calls += 1
        if stats or self.est is None:
            it, rr = ctypes.c_int(), ctypes.c_double()
            _lib.call("f2d_poisson_solve_c2v", self._h, _lib.ptr(omega), _lib.ptr(x),
                      self.rtol, self.maxit, ctypes.byref(it), ctypes.byref(rr),
                      _lib.stream_ptr())
            self.last_iters, self.last_relres = it.value, rr.value
            if not self._accepted(rr.value):
                raise PoissonNotConverged(
                    f"Poisson solve reached relres {rr.value:.3e} > {self.rtol:.1e} "
                    f"after {it.value} iterations")
            ref = None
            if self.precision != "fp64":
                ref = int(self._stats_host()[3])
            self._learn(it.value, ref)
        else:
            n = min(self.est + self.margin, self.maxit)
            _lib.call("f2d_poisson_solve_c2v", self._h, _lib.ptr(omega), _lib.ptr(x),
                      self.rtol, n, None, None, _lib.stream_ptr())

    def set_refine(self, nref):
        """FP64 refinements of the following direct (FFT or capacitance)
        solves in FP32 (0: plain FP32 solves, None: the handle's option);
        host state only, so a captured graph keeps the choice each of its
        solves was made with."""
        v = -1 if nref is None else int(nref)
        if getattr(self, "_refine", -1) != v:
            _lib.call("f2d_poisson_set_refine", self._h, v)
            self._refine = v

    def _stats_host(self):
        buf = (ctypes.c_double * 8)()
        _lib.call("f2d_poisson_read_stats", self._h, buf, _lib.stream_ptr())
        return list(buf)

    def reset_stats(self):
        _lib.call("f2d_poisson_reset_stats", self._h, _lib.stream_ptr())
        self.replayed = 0

    def graph_replayed(self, nsolves, captured):
        """Account for the direct solves of a graph replay (nsolves) and of
        its capture (captured, enqueued but never run)."""
        if self.method in ("fft", "capacitance"):
            self.replayed += nsolves - captured

    def read_stats(self):
        """Counters since the last reset: solves, non-converged solves, max
        iterations (refinements in FP32), and the last solve's figures."""
        st = self._stats_host()
        return {"solves": int(st[6]) + self.replayed, "failed": int(st[4]), "max_iters": int(st[5]),
                "last_iters": int(st[0]), "last_relres": st[1],
                "last_refinements": int(st[3])}

    def bench_kernel(self, kind, reps=20):
        """Mean launch time (us) and algorithmic bytes of one fine-level
        kernel (0: update/pre-smooth/restrict, 1: prolongate/post-smooth,
        2: search direction/operator) on the current vectors."""
        us, by = ctypes.c_double(), ctypes.c_double()
        _lib.call("f2d_poisson_bench_kernel", self._h, kind, reps, ctypes.byref(us),
                  ctypes.byref(by), _lib.stream_ptr())
        return us.value, by.value


class PoissonBatch:
    """The solves of many same-size members with one launch per kernel
    (f2d_poisson_batch_*): solvers[i].solve(bs[i], xs[i], stats=False) for
    every i, bit for bit.  Each solver must use the capacitance method in an
    FP32 precision and the solvers must share the rectangle operator (grid,
    spacings, maindiag, periodicity); their masks may differ.  bs and xs
    are fixed device arrays: the batch is bound to them.  The ensembles of
    fluids2d_amd.ensemble use it (the reference solves one model at a time,
    elliptic.py:80-87)."""

    def __init__(self, solvers, bs, xs):
        self.solvers = list(solvers)
        self.arrays = (list(bs), list(xs))  # keeps the bound arrays alive
        n = len(self.solvers)
        arr = ctypes.c_void_p * n
        h = ctypes.c_void_p()
        _flush_pending()
        _lib.call("f2d_poisson_batch_create", arr(*(s.handle() for s in self.solvers)), n,
                  arr(*(_lib.ptr(b) for b in bs)), arr(*(_lib.ptr(x) for x in xs)),
                  ctypes.byref(h))
        self._h = h

    @staticmethod
    def eligible(solver):
        return solver.method == "capacitance" and solver.precision != "fp64"

    @staticmethod
    def key(solver):
        """Solvers with equal keys may share a batch."""
        m = solver.mesh
        p = m.param
        return (m.shape, m.dx, m.dy, solver.maindiag, bool(p.xperiodic),
                bool(getattr(p, "yperiodic", False)), solver.precision,
                tuple(sorted(solver.tuning.items())), bool(getattr(p, "poisson_warm", True)))

    def bound_to(self, bs, xs):
        return all(a.data_ptr() == b.data_ptr() for a, b in zip(self.arrays[0], bs)) and \
            all(a.data_ptr() == b.data_ptr() for a, b in zip(self.arrays[1], xs))

    def solve(self, nref=None):
        """nref: FP64 refinements (0: plain FP32 solves, None: the solvers'
        option), as Poisson2D.set_refine."""
        for s in self.solvers:
            s.calls += 1
        _lib.call("f2d_poisson_batch_solve", self._h, -1 if nref is None else int(nref),
                  _lib.stream_ptr())

    def __del__(self):
        h = getattr(self, "_h", None)
        self._h = None
        _release(h, "f2d_poisson_batch_destroy")


def get_msk_from_mesh(mesh, location):
    """elliptic.py:102-111, plus the y-periodic extension."""
    msk = mesh.msk if location == "c" else mesh.mskv
    n = mesh.param.halowidth
    m = msk * 1
    if mesh.param.xperiodic:
        m[:, :n] = 0
        m[:, -n:] = 0
    if getattr(mesh.param, "yperiodic", False):
        m[:n, :] = 0
        m[-n:, :] = 0
    return m
