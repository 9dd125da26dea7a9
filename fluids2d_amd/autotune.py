"""Autotuner: picks, by timing on the device, the schedule knobs that do not
change what is computed:

* the multigrid smoother schedule of a masked Poisson/Helmholtz operator
  (``Poisson2D`` tuning: RB-GS sweeps per level, size of the
  single-workgroup tail, FP32 inner iterations per FP64 refinement), scored
  by the time of converged solves of a white-noise and a smooth right-hand
  side, with room for the no-sync path's margin under the iteration cap;
* the LDS tile of the fused transport stage (``Param.tile``; results are
  bit-identical for every tile);
* the number of concurrent streams of an ensemble (``Ensemble.tune_streams``).

A candidate that does not converge to the solver's tolerance is rejected.
Choices are cached per device and problem class (grid size, periodicity, wet
fraction, precision, mass term) in a JSON file: ``$F2D_TUNE_CACHE`` or
``~/.cache/fluids2d_amd/tune.json`` (kept in memory when not writable).
"""
import json
import os

import numpy as np
import torch

from . import _lib
from . import operators as op

# nu (= nu_coarse = nu_tail), tail_points, refine_inner (FP32 refinement)
MG_NU = (1, 2)
MG_TAIL = (1024, 4096)
MG_INNER = (3, 5)
TILES = tuple(_lib.TILES)

_memory = {}


def _cache_file():
    return os.environ.get("F2D_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "fluids2d_amd", "tune.json")


def _load():
    if _memory:
        return _memory
    try:
        with open(_cache_file()) as f:
            _memory.update(json.load(f))
    except (OSError, ValueError):
        pass
    return _memory


def _store(key, value):
    c = _load()
    c[key] = value
    try:
        path = _cache_file()
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + f".{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(c, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def _device():
    return torch.cuda.get_device_name(torch.cuda.current_device())


def mg_candidates(precision):
    out = []
    for nu in MG_NU:
        for tp in MG_TAIL:
            inner = MG_INNER if precision != "fp64" else (None,)
            for ri in inner:
                d = dict(nu=nu, nu_coarse=nu, nu_tail=nu, tail_points=tp)
                if ri is not None:
                    d["refine_inner"] = ri
                out.append(d)
    return out


def poisson_key(mesh, msk, maindiag):
    p = mesh.param
    wet = float(np.asarray(msk)[3:-3, 3:-3].mean())
    return (f"mg|{getattr(p, 'poisson_solver', 'auto') or 'auto'}|{_device()}|"
            f"{getattr(p, 'precision', 'fp64')}|{mesh.ny}x{mesh.nx}|"
            f"x{int(bool(p.xperiodic))}y{int(bool(getattr(p, 'yperiodic', False)))}|"
            f"wet{wet:.1f}|md{int(maindiag > 0)}")


def _time_solves(ps, b, x, reps):
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        x.zero_()
        ps.solve(b, x, stats=False)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _test_rhs(shape, msk):
    """Right-hand sides the schedules are scored on: white noise (easy for
    the smoothers) and a smooth field (left to the coarse levels and CG),
    both on the unknowns."""
    from scipy import ndimage
    rng = np.random.default_rng(12345)
    white = rng.standard_normal(shape)
    smooth = ndimage.gaussian_filter(rng.standard_normal(shape), sigma=max(shape) / 16.0)
    smooth /= np.abs(smooth).max()
    return [white * msk, smooth * msk]


def tune_poisson(mesh, maindiag=0.0, reps=4, candidates=None, cache=True, verbose=False):
    """Best multigrid schedule (a dict of Poisson2D tuning options) for the
    vertex operator of `mesh` with mass term `maindiag`; {} when the
    operator is solved directly (FFT or capacitance solve: nothing to
    tune)."""
    from .elliptic import Poisson2D, get_msk_from_mesh
    msk = get_msk_from_mesh(mesh, "v")
    key = poisson_key(mesh, msk, maindiag)
    if cache and key in _load():
        return dict(_load()[key]["opts"])
    precision = getattr(mesh.param, "precision", "fp64")
    cands = candidates if candidates is not None else mg_candidates(precision)
    dtype = torch.float64 if precision == "fp64" else torch.float32
    x = torch.zeros(mesh.shape, dtype=dtype, device=mesh.device)
    best, best_ms, table = None, None, []
    rhs = [torch.from_numpy(r).to(mesh.device, dtype) for r in _test_rhs(mesh.shape, msk)]
    for opts in cands:
        ps = Poisson2D(mesh, "v", maindiag=maindiag, tuning=opts)
        if ps.method in ("fft", "capacitance"):
            return {}  # a direct solve: no multigrid schedule to pick
        ok = True
        for b in rhs:
            x.zero_()
            ps.solve(b, x, check=False)
            ok = ok and ps._accepted(ps.last_relres)
        # the no-sync solves enqueue the learned count plus a margin: it
        # must fit under the cap
        ok = ok and ps.est is not None and ps.est + ps.margin <= ps.maxit
        if not ok:
            table.append((opts, None))
            continue
        ms = 0.0
        for b in rhs:
            ps.solve(b, x, stats=False)  # warm the no-sync path
            ms += _time_solves(ps, b, x, reps)
        table.append((opts, ms))
        if best_ms is None or ms < best_ms:
            best, best_ms = opts, ms
        del ps
    if best is None:
        raise RuntimeError(f"autotune: no multigrid schedule converged for {key}")
    if verbose:
        for opts, ms in table:
            print(f"  {opts}: " + ("not converged" if ms is None else f"{ms:.3f} ms"))
    if cache:
        _store(key, {"opts": best, "ms": best_ms})
    return dict(best)


def tune_tile(model, reps=20, candidates=TILES, cache=True):
    """Fastest LDS tile of the fused transport stage for `model`
    (eulerpsi, advection, qg); sets it on the model's mesh and returns it."""
    m, p = model.mesh, model.param
    key = (f"tile|{_device()}|{p.precision}|{p.model}|{p.compflux}|{m.ny}x{m.nx}|"
           f"x{int(bool(p.xperiodic))}y{int(bool(getattr(p, 'yperiodic', False)))}")
    if cache and key in _load():
        best = int(_load()[key]["tile"])
    else:
        integ = model.integrator
        s = integ.dstate
        q = getattr(s, integ.specs.prognostic[0])
        U = s.U
        q_out = torch.empty_like(q)
        ds = torch.empty_like(q)
        times = {}
        for t in candidates:
            m.grid.tile = int(t)
            op.transport_stage(m, p.compflux, q, q_out, U, [], ds, "rk3", 0, integ.d_dt)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                op.transport_stage(m, p.compflux, q, q_out, U, [], ds, "rk3", 0, integ.d_dt)
            e1.record()
            torch.cuda.synchronize()
            times[int(t)] = e0.elapsed_time(e1) / reps
        best = min(times, key=times.get)
        if cache:
            _store(key, {"tile": best, "ms": times})
    m.grid.tile = best
    p.tile = best
    return best
