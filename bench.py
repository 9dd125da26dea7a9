"""Benchmark: cell-steps/s of the Euler (vorticity-streamfunction) RK3 step on
a 1024^2 doubly periodic grid, FP32 stencils + multigrid with FP64 iterative
refinement vs plain FP64, on one MI355X.

A step = one SSP-RK3 step (3 x [fused WENO-5 transport + RK update,
multigrid-preconditioned CG Poisson solve, perp-gradient + CFL]).
cell-steps/s = nx * ny * steps / (step-loop time), the reciprocal of the
reference's `perf` print (model.py:89-93).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--nx 1024]
                       [--no-cpu-baseline] [--accuracy-steps N]
Prints one JSON line.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def random_phase_vorticity(n, seed, k0=8.0):
    """Seeded random-phase field with E(k) ~ k^-3 exp(-(k/k0)^2), hence
    |w_hat(k)| ~ sqrt(k E(k)); k=0 removed; unit RMS."""
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(n) * n
    ky = np.fft.rfftfreq(n) * n
    k = np.sqrt(kx[:, None] ** 2 + ky[None, :] ** 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        E = np.where(k > 0, k ** -3.0 * np.exp(-(k / k0) ** 2), 0.0)
    amp = np.sqrt(k * E)
    phase = np.exp(2j * np.pi * rng.random(amp.shape))
    w = np.fft.irfft2(amp * phase, s=(n, n))
    w -= w.mean()
    return w / np.sqrt(np.mean(w ** 2))


def fill_periodic(a, h=3):
    a[:, :h] = a[:, -2 * h:-h]
    a[:, -h:] = a[:, h:2 * h]
    a[:h, :] = a[-2 * h:-h, :]
    a[-h:, :] = a[h:2 * h, :]


def initial_omega(mesh, seed):
    n = mesh.nx
    w = np.zeros(mesh.shape)
    w[3:-3, 3:-3] = random_phase_vorticity(n, seed)
    fill_periodic(w)
    return w * mesh.msk * mesh.area


HIPGRAPH = True  # --no-graph: eager launches (for per-dispatch counter passes)
WINDOWS = 2  # timed windows per precision, interleaved


def make_param(nx, precision, dt=0.0):
    import fluids2d_amd.param as fp
    fp.QUIET = True
    from fluids2d_amd import Param
    p = Param()
    p.model = "eulerpsi"
    p.nx = p.ny = nx
    p.Lx = p.Ly = 1.0
    p.xperiodic = p.yperiodic = True
    p.compflux = "weno"
    p.integrator = "rk3"
    p.tend = 1e30
    p.maxite = 10 ** 9
    p.precision = precision
    p.dt = dt
    p.hipgraph = HIPGRAPH
    return p


def build(nx, precision, seed, dt=0.0):
    from fluids2d_amd import Model
    m = Model(make_param(nx, precision, dt))
    m.state.omega[...] = initial_omega(m.mesh, seed)
    m.integrator.diag(m.state)
    m.integrator.begin(m.state)
    return m


def calibrate(model, nsteps=3):
    """A few steps with a host check after every Poisson solve: learn the
    iteration counts the no-sync loop enqueues."""
    integ = model.integrator
    for _ in range(nsteps):
        integ.set_dt(model.time)
        integ.advance(stats=True)
        model.time.pushforward()


def warm(model, warmup, seconds=0.2):
    """`warmup` untimed steps, then the step graph of the buffer assignment
    a timed window starts from is captured (if it is not yet) and replayed,
    untimed, for about `seconds` of GPU work: a timed window replays graphs
    and captures none, and does not start on a GPU that has just been
    idle."""
    import torch
    integ = model.integrator
    if warmup:
        integ.device_steps(warmup)
    integ.prepare_graph()
    t0 = time.perf_counter()
    while True:
        integ.device_steps(8 * integ.graph_steps)
        torch.cuda.synchronize()
        if time.perf_counter() - t0 >= seconds:
            break


def timed_window(model, steps, solver="poisson_vertices"):
    """Exactly `steps` steps between two HIP events, bracketed by device
    synchronisations.  Raises if a step graph was captured inside.  `solver`
    names the mesh's elliptic solver whose statistics the window returns
    (None: a model without one)."""
    import torch
    integ = model.integrator
    pois = getattr(model.mesh, solver) if solver else None
    integ.prepare_graph()  # the buffer assignment this window starts from
    if pois is not None:
        pois.reset_stats()
    ngraphs = integ.graph_count()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    dts = integ.device_steps(steps)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if integ.graph_count() != ngraphs:
        raise RuntimeError("a step graph was captured inside the timed window")
    return (e0.elapsed_time(e1) / steps, wall, pois.read_stats() if pois is not None else None,
            dts)


def timed(model, steps, warmup, solver="poisson_vertices"):
    """warm(), then one timed window."""
    warm(model, warmup)
    return timed_window(model, steps, solver)


def accuracy(nx, seed, nsteps):
    """FP32+refinement vs the FP64 path after nsteps at a common fixed dt."""
    m64 = build(nx, "fp64", seed)
    m64.integrator.set_dt(m64.time)
    dt = m64.time.dt
    m64 = build(nx, "fp64", seed, dt=dt)
    m32 = build(nx, "fp32", seed, dt=dt)
    for m in (m64, m32):
        calibrate(m, 2)
        m.integrator.device_steps(nsteps - 2)
        m.integrator.end(m.state)
    a, b = m32.state.omega[3:-3, 3:-3], m64.state.omega[3:-3, 3:-3]
    return float(np.linalg.norm(a - b) / np.linalg.norm(b)), dt, diag_errors(m32, m64)


def diag_errors(m32, m64):
    """FP64 bulk diagnostics (diagnostics.py:45-66) of the FP32 run against
    the FP64 run's: relative error of ke, enstrophy and angular momentum;
    the domain-mean vorticity (zero up to rounding on this periodic field)
    by its difference scaled by max|omega|/area."""
    import torch
    from fluids2d_amd.diagnostics import Bulk
    out = []
    for m in (m32, m64):
        bk = Bulk(m, ncfile=None, ndiags=1)
        r = torch.zeros(4, dtype=torch.float64, device=m.mesh.device)
        bk.measure(r)
        out.append(r.cpu().numpy())
    a, b = out
    w = float(np.abs(m64.state.omega).max()) / m64.mesh.area
    return {"ke": float(abs(a[0] - b[0]) / abs(b[0])), "ens": float(abs(a[1] - b[1]) / abs(b[1])),
            "angular": float(abs(a[3] - b[3]) / max(abs(b[3]), 1e-300)),
            "vort_over_max": float(abs(a[2] - b[2]) / w)}


def adjoint_cost(n=512, steps=200, checkpoint=20, seed=2024, reps=2):
    """Forward + adjoint time over forward time for the gradient of a
    final-time misfit with respect to the initial vorticity: eulerpsi n^2
    doubly periodic FP32, `steps` RK3 steps at a fixed CFL-0.5 step,
    checkpoints every `checkpoint` steps (fluids2d_amd.adjoint)."""
    import torch
    from fluids2d_amd.adjoint import EulerAdjoint, EulerGradient
    p = make_param(n, "fp32")
    w0 = np.zeros((n + 6, n + 6))
    w0[3:-3, 3:-3] = random_phase_vorticity(n, seed) / (n * n)
    fill_periodic(w0)
    wt = np.zeros_like(w0)
    wt[3:-3, 3:-3] = (random_phase_vorticity(n, seed) +
                      0.3 * random_phase_vorticity(n, seed + 1)) / (n * n)
    fill_periodic(wt)
    probe = EulerAdjoint(p, 1.0, 1)
    dev = probe.mesh.device
    w0 = torch.from_numpy(w0).to(dev)
    wt = torch.from_numpy(wt).to(dev).float()
    runner = EulerAdjoint(p, probe.cfl_dt(w0, 0.5), steps, checkpoint)
    runner.forward(w0)  # warm-up (Poisson set-up, allocator)
    torch.cuda.synchronize()
    tf = []
    tb = []
    for _ in range(reps):
        t0 = time.perf_counter()
        runner.forward(w0)
        torch.cuda.synchronize()
        tf.append(time.perf_counter() - t0)
        x = w0.clone().float().requires_grad_(True)
        t0 = time.perf_counter()
        wT = EulerGradient.apply(x, runner)
        L = 0.5 * ((wT - wt)[3:-3, 3:-3] ** 2).sum()
        L.backward()
        torch.cuda.synchronize()
        tb.append(time.perf_counter() - t0)
    return min(tb) / min(tf), min(tf), min(tb)


ENSEMBLE_SIZES = (128, 256, 512, 1024)


def basin_mask(n, seed, ny=None):
    """Closed basin: Gaussian-filtered white noise above a level, walls on
    the border ring, largest wet component kept; then wet cells that touch
    no wet vertex are dried (repeatedly), since the reference's
    verticestocenters divides by zero there (operators.py:136-141)."""
    from scipy import ndimage
    ny = n if ny is None else ny
    rng = np.random.default_rng(seed)
    sm = ndimage.gaussian_filter(rng.standard_normal((ny, n)), sigma=max(n, ny) / 16.0)
    wet = sm > -0.3 * sm.std()
    wet[0, :] = wet[-1, :] = False
    wet[:, 0] = wet[:, -1] = False
    while True:
        lab, nl = ndimage.label(wet)
        if nl > 1:
            sizes = ndimage.sum(wet, lab, range(1, nl + 1))
            wet = lab == (1 + int(np.argmax(sizes)))
        # vertex (j, i) is wet iff the 4 cells around it are
        v = np.zeros((ny + 1, n + 1), dtype=bool)
        v[1:-1, 1:-1] = wet[1:, 1:] & wet[:-1, 1:] & wet[1:, :-1] & wet[:-1, :-1]
        touch = v[:-1, :-1] | v[1:, :-1] | v[:-1, 1:] | v[1:, 1:]
        bad = wet & ~touch
        if not bad.any():
            break
        wet &= ~bad
    m = np.zeros((ny + 6, n + 6), dtype=np.int8)
    m[3:-3, 3:-3] = wet
    return m


def ensemble_member(n, seed, precision="fp32", model="qgrsw", dtmax=0.05, solver=None):
    """One synthetic member: `model` on an n x n closed basin (Lx = Ly = 1,
    f0 = 10, g = H = 1), h = H + seeded Gaussian bumps, balanced by the QG
    projection (geos_adj.py:37-49).  dtmax keeps f0 dt = 0.5: the CFL step of
    model.py:84-87 sees only the slow advective velocity.  solver: the
    Param.poisson_solver to ask for (None: the default, "auto")."""
    import fluids2d_amd as f2d
    import fluids2d_amd.param as fp
    from fluids2d_amd.rsw import qg_projection
    fp.QUIET = True
    p = f2d.Param()
    p.model = model
    p.nx = p.ny = n
    p.precision = precision
    p.dtmax = dtmax
    p.tend = 1e9
    p.maxite = 10 ** 9
    if solver is not None:
        p.poisson_solver = solver
    m = f2d.Model(p)
    m.mesh.msk[...] = basin_mask(n, seed)
    m.mesh.finalize()
    m.integrator.model.mesh = m.mesh
    rng = np.random.default_rng(seed + 7)
    x, y = m.mesh.xy("c")
    bumps = np.zeros(m.mesh.shape)
    for _ in range(4):
        x0, y0 = rng.uniform(0.25, 0.75, 2)
        r = rng.uniform(0.03, 0.08)
        bumps += rng.choice([-1, 1]) * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * r * r))
    if model in ("qgrsw", "rsw"):
        m.state.h[...] = (p.H + 0.05 * bumps) * m.mesh.msk * m.mesh.area
        qg_projection(m)
    else:  # eulerpsi / qg: vorticity or PV bumps (vortex.py:19-26 convention)
        field = "pv" if model == "qg" else "omega"
        getattr(m.state, field)[...] = bumps * m.mesh.msk * m.mesh.area
        m.integrator.diag(m.state)
    return m


def ensemble_sizes(count, seed):
    rng = np.random.default_rng(seed)
    return [int(v) for v in rng.choice(ENSEMBLE_SIZES, size=count)]


def ensemble_bench(count=64, steps=200, seed=99, precision="fp32", tune=True, log=None,
                   batch=True):
    """The batched-ensemble workload: `count` qgrsw members with sizes drawn
    by seed from ENSEMBLE_SIZES (squared), closed smoothed-noise basins,
    FP32 + refinement, `steps` steps replayed from one captured graph.
    Multigrid schedules per size class and the stream count are picked by
    the autotuner first."""
    import torch
    from fluids2d_amd import autotune
    from fluids2d_amd.ensemble import Ensemble
    sizes = ensemble_sizes(count, seed)
    t0 = time.perf_counter()
    tuned = {}
    members = []
    for k, n in enumerate(sizes):
        m = ensemble_member(n, seed * 1000 + k, precision)
        if tune:
            if n not in tuned:
                tuned[n] = autotune.tune_poisson(m.mesh, maindiag=m.mesh.qg_helmholtz.maindiag)
                if log:
                    log(f"ensemble: tuned {n}^2 multigrid schedule {tuned[n]}")
            m.param.poisson_opts = tuned[n]
            m.mesh.finalize()
            m.integrator.model.mesh = m.mesh
        m.integrator.begin(m.state)
        members.append(m)
        if log and (k + 1) % 8 == 0:
            log(f"ensemble: {k + 1}/{count} members set up "
                f"({time.perf_counter() - t0:.1f} s)")
    setup = time.perf_counter() - t0
    ens = Ensemble(members, nstreams=8, batch=batch)
    ens.steps(2)
    if log:
        log("ensemble: 2 eager steps done")
    streams = ens.tune_streams() if tune else {}
    if log:
        log(f"ensemble: stream counts timed {streams}")
    ens.steps(ens.G)  # capture at the chosen stream count
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    dts = ens.steps(steps)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    cells = sum(n * n for n in sizes)
    st = [m.mesh.qg_helmholtz.read_stats() for m in members]
    # wet cells (dry ones may hold the (1/0)*0 of verticestocenters)
    finite = all(bool(torch.isfinite(m.integrator.dstate.h[m.mesh.d_msk == 1]).all())
                 for m in members)
    return {"workload": f"{count} qgrsw members, sizes from {ENSEMBLE_SIZES} squared "
                        f"(seed {seed}), closed smoothed-noise basins, {precision} with "
                        f"FP64 refinement, {steps} steps from one captured graph",
            "cells": cells, "members": count,
            "cell_steps_per_s": cells * steps / (ms * 1e-3), "ms_per_step": ms / steps,
            "setup_s": setup, "streams": ens.nstreams, "stream_ms_per_step": streams,
            "tuned": {str(k): v for k, v in tuned.items()},
            "failed_solves": int(sum(x["failed"] for x in st)), "finite": finite,
            "dt_range": [float(dts.min()), float(dts.max())]}


def _lib_tiles():
    from fluids2d_amd import _lib
    return _lib.TILES


def _traffic_key(kernel, tile):
    """The transport's traffic depends on its LDS tile (halo re-reads): its
    summary row is keyed 'k_transport <rows>x<columns>'."""
    if kernel.startswith("k_transport") and tile is not None:
        return "k_transport " + "x".join(map(str, _lib_tiles()[tile]))
    return kernel


def traffic_of(kernel, workload, prec):
    """HBM bytes per launch of `kernel` on a workload (an eulerpsi grid size,
    or a name such as "rsw2048x1024"), from the committed rocprofv3 counter
    summary (profiles/traffic_<workload>_<precision>.txt, JSON: FETCH_SIZE +
    WRITE_SIZE per dispatch, scaled by the same counters on a copy kernel of
    known size), or None when no summary covers it."""
    name = f"eulerpsi{workload}" if isinstance(workload, int) else workload
    path = os.path.join(ROOT, "profiles", f"traffic_{name}_{prec}.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    for k, v in t.get("kernels", {}).items():
        if kernel.startswith(k):
            return v
    return None


def host_cores():
    """The host threads this process may use (its CPU affinity), the machine's
    cpu_count and the CPU model name (lscpu's "Model name")."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    model = None
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    return usable, os.cpu_count(), model


# at most this many host threads for the baseline: the CPU share one GPU's
# process is given on the measurement machines
MAX_CPU_THREADS = 16


def cpu_baseline(nx, seed, nsteps, threads=1):
    """The oracle (the reference's algorithm: C restatement of the numba loops,
    -O3, OpenMP over the flat index as the analogue of prange, numpy
    operators, scipy SuperLU) on the host, a few steps of the same workload
    after factorisation, with `threads` threads in the loops (SuperLU and
    numpy run on one)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    p = make_param(nx, "fp64")
    keys = ["model", "nx", "ny", "Lx", "Ly", "xperiodic", "yperiodic", "halowidth",
            "noslip", "f0", "beta", "g", "H", "tend", "dt", "maxite", "dtmax",
            "integrator", "cfl", "compflux", "vortexforce", "innerproduct", "maxorder"]

    class B:
        pass
    b = B()
    for k in keys:
        setattr(b, k, getattr(p, k))
    used = O.set_threads(threads)
    t0 = time.perf_counter()
    m = O.Model(b)
    m.mesh.poisson_vertices  # the SuperLU factorisation (made on first use)
    tfact = time.perf_counter() - t0
    m.state.omega[...] = initial_omega(m.mesh, seed)
    m.diag(m.state)
    t0 = time.perf_counter()
    m.run(nsteps)
    el = time.perf_counter() - t0
    return nx * nx * nsteps / el, el, tfact, used


CLOSED_NX = 256


def closed_param(precision):
    """eulerpsi 256^2 closed square, WENO5, SSP-RK3, CFL 0.8 (the workload the
    reference runs on a CPU)."""
    p = make_param(CLOSED_NX, precision)
    p.xperiodic = p.yperiodic = False
    p.cfl = 0.8
    return p


def vortex_omega(mesh, seed):
    """Gaussian vortex (width 0.1, centred) plus 1e-3 uniform noise at the
    cell centres, times msk * area (vortex.py:23-26's convention)."""
    x, y = mesh.xy("c")
    rng = np.random.default_rng(seed)
    w = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / (2 * 0.1 ** 2))
    w = w + 1e-3 * rng.uniform(-1, 1, w.shape)
    return w * mesh.msk * mesh.area


def closed_bench(seed, steps):
    """GPU step times of the closed 256^2 workload, FP64 and FP32 + refinement
    (direct solve through the odd extension)."""
    import torch
    from fluids2d_amd import Model
    out = {"workload": f"eulerpsi {CLOSED_NX}x{CLOSED_NX} closed square, WENO5, SSP-RK3, "
                       f"CFL 0.8, Gaussian vortex + 1e-3 noise (seed {seed})",
           "steps": steps}
    for prec in ("fp64", "fp32"):
        m = Model(closed_param(prec))
        m.state.omega[...] = vortex_omega(m.mesh, seed)
        m.integrator.diag(m.state)
        m.integrator.begin(m.state)
        calibrate(m, 3)
        ms, wall, st, dts = timed(m, steps, 20)
        out[f"{prec}_ms_per_step"] = ms
        out[f"{prec}_cell_steps_per_s"] = CLOSED_NX ** 2 / (ms * 1e-3)
        del m
        torch.cuda.empty_cache()
    return out


def cpu_baseline_closed(seed, steps, threads=(1,)):
    """The oracle (C loops + SuperLU; the loops on each thread count of
    `threads`) on the closed 256^2 workload for `steps` steps, and the device
    FP64 run of the same steps from the same state checked against it
    (relative L2 of the vorticity)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from fluids2d_amd import Model
    p = closed_param("fp64")
    p.maxite = steps
    m = Model(p)
    w0 = vortex_omega(m.mesh, seed)
    m.state.omega[...] = w0
    m.integrator.diag(m.state)
    m.run()

    class B:
        pass
    b = B()
    for k in ["model", "nx", "ny", "Lx", "Ly", "xperiodic", "yperiodic", "halowidth",
              "noslip", "f0", "beta", "g", "H", "tend", "dt", "maxite", "dtmax",
              "integrator", "cfl", "compflux", "vortexforce", "innerproduct", "maxorder"]:
        setattr(b, k, getattr(p, k))
    runs = {}
    for thr in threads:
        used = O.set_threads(thr)
        t0 = time.perf_counter()
        ref = O.Model(b)
        ref.mesh.poisson_vertices  # the SuperLU factorisation (made on first use)
        tfact = time.perf_counter() - t0
        ref.state.omega[...] = w0
        ref.diag(ref.state)
        t0 = time.perf_counter()
        ref.run(steps)
        el = time.perf_counter() - t0
        err = float(np.linalg.norm(m.state.omega - ref.state.omega) /
                    np.linalg.norm(ref.state.omega))
        runs[used] = dict(value=CLOSED_NX ** 2 * steps / el, seconds=el, fact=tfact, err=err,
                          ite=int(ref.time.ite))
    best = max(runs, key=lambda t: runs[t]["value"])
    r = runs[best]
    return {"value": r["value"], "unit": "cell-steps/s", "cores": best,
            "kind": "port",
            "sample": f"{steps} RK3 steps of the closed {CLOSED_NX}^2 workload through the "
                      f"oracle on {sorted(runs)} threads (value: the faster, "
                      f"{r['seconds']:.1f} s); SuperLU factorisation excluded",
            "by_threads": {str(t): {"value": v["value"], "seconds": round(v["seconds"], 2)}
                           for t, v in runs.items()},
            "gpu_fp64_vs_oracle_rel_l2_omega": max(v["err"] for v in runs.values()),
            "iterations": [int(m.time.ite)] + [v["ite"] for v in runs.values()]}


RSW_NX, RSW_NY = 2048, 1024


def channel_islands_mask(nx, ny, seed, nisl=12):
    """x-periodic channel with walls in y and `nisl` seeded circular
    islands, radius U{20..60} cells, centres at least r + 3 cells from the
    walls; the periodic halo columns equal their interior sources."""
    m = np.ones((ny + 6, nx + 6), dtype=np.int8)
    m[:3, :] = 0
    m[-3:, :] = 0
    rng = np.random.default_rng(seed)
    jj, ii = np.mgrid[0:ny + 6, 0:nx + 6]
    for _ in range(nisl):
        r = int(rng.integers(20, 61))
        cj = 3 + rng.uniform(r + 3, ny - r - 3)
        ci = 3 + rng.uniform(0, nx)
        d2 = (jj - cj) ** 2 + np.minimum(np.abs(ii - ci), nx - np.abs(ii - ci)) ** 2
        m[d2 < r * r] = 0
    m[:, :3] = m[:, nx:nx + 3]
    m[:, nx + 3:] = m[:, 3:6]
    return m


def rsw_model(precision, seed, nx=RSW_NX, ny=RSW_NY):
    """rsw on the nx x ny island channel (Lx = 2, Ly = 1, f0 = g = H = 1,
    WENO5 everywhere, SSP-RK3, CFL 0.9), at rest, h = H (1 + 6 seeded
    Gaussian bumps of amplitude 0.05) on the wet cells, times area."""
    import fluids2d_amd as f2d
    import fluids2d_amd.param as fp
    fp.QUIET = True
    p = f2d.Param()
    p.model = "rsw"
    p.nx, p.ny, p.Lx, p.Ly = nx, ny, 2.0, 1.0
    p.xperiodic, p.yperiodic = True, False
    p.f0 = p.g = p.H = 1.0
    p.cfl = 0.9
    p.compflux = p.vortexforce = p.innerproduct = "weno"
    p.tend, p.maxite = 1e30, 10 ** 9
    p.precision = precision
    p.hipgraph = HIPGRAPH
    m = f2d.Model(p)
    m.mesh.msk[...] = channel_islands_mask(nx, ny, seed)
    m.mesh.finalize()
    m.integrator.model.mesh = m.mesh
    x, y = m.mesh.xy("c")
    rng = np.random.default_rng(seed + 1)
    h = np.ones(m.mesh.shape)
    for _ in range(6):
        x0, y0 = rng.uniform(0, p.Lx), rng.uniform(0, p.Ly)
        r = rng.uniform(0.03, 0.1) * p.Ly
        h += 0.05 * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * r * r))
    h[:, :3] = h[:, nx:nx + 3]
    h[:, nx + 3:] = h[:, 3:6]
    m.state.h[...] = h * p.H * m.mesh.area * m.mesh.msk
    m.integrator.diag(m.state)
    m.integrator.begin(m.state)
    return m


def rsw_bench(seed, steps, warmup=20):
    """Step times of the rsw island channel in FP64 and FP32 (fused stage:
    3 kernels per RK stage; the precisions' timed windows interleaved, as
    for the headline workload), the live roofline of its kernels, and the
    wet-cell mass drift of h over the timed steps (flux form, closed walls:
    conserved to rounding)."""
    import torch
    from fluids2d_amd import roofline as rl
    out = {"workload": f"rsw {RSW_NX}x{RSW_NY} x-periodic channel, walls in y, 12 seeded "
                       f"islands (seed {seed}), f0 = g = H = 1, WENO5, SSP-RK3, CFL 0.9, "
                       f"at rest with Gaussian bumps in h",
           "steps": steps, "windows": WINDOWS}
    precs = ("fp64", "fp32")
    models, wet, m0 = {}, {}, {}
    for prec in precs:
        m = rsw_model(prec, seed)
        w = (m.mesh.d_msk == 1)
        w[:, :3] = False
        w[:, -3:] = False
        wet[prec] = w
        m0[prec] = float(m.integrator.dstate.h[w].double().sum())
        warm(m, warmup)
        models[prec] = m
    windows = {prec: [] for prec in precs}
    for _ in range(WINDOWS):
        for prec in precs:
            windows[prec].append(timed_window(models[prec], steps, solver=None)[0])
    for prec in precs:
        m = models[prec]
        ms = sum(windows[prec]) / len(windows[prec])
        m1 = float(m.integrator.dstate.h[wet[prec]].double().sum())
        prof = rl.sw_timing(m)
        out[f"{prec}_ms_per_step"] = ms
        out[f"{prec}_windows_ms_per_step"] = windows[prec]
        out[f"{prec}_cell_steps_per_s"] = RSW_NX * RSW_NY / (ms * 1e-3)
        out[f"{prec}_mass_drift"] = abs(m1 - m0[prec]) / m0[prec]
        out[f"{prec}_roofline"] = {
            "bound": "hbm", "achieved": prof["achieved_gbs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": prof["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": traffic_of(prof["kernel"].split()[0], f"rsw{RSW_NX}x{RSW_NY}", prec),
            "kernel": prof["kernel"], "bytes_per_launch": prof["bytes"],
            "avg_launch_us": prof["us"], "launches_timed": prof["all"][prof["kernel"]]["launches"],
            "kernels": {k: {"us": round(v["us"], 2), "per_step": v["per_step"],
                            "gbs": round(v["gbs"], 1)} for k, v in prof["all"].items()}}
    del models, m
    torch.cuda.empty_cache()
    return out


def masked_bench(n, seed, solver, steps, warmup=6):
    """eulerpsi in an n x n closed smoothed-noise basin (the ensemble's
    members' masks and vorticity bumps, dt from the CFL condition alone):
    step times in FP64 and FP32 (Param.poisson_refine default) with the
    Poisson solver `solver` ("cap": the capacitance solve; "mg": multigrid
    CG), the precisions' timed windows interleaved, and the solver's
    iteration counts."""
    import torch
    out = {"workload": f"eulerpsi {n}x{n} closed smoothed-noise basin (seed {seed}), "
                       f"Poisson solver {solver!r}", "steps": steps, "windows": WINDOWS}
    models = {}
    for prec in ("fp64", "fp32"):
        m = ensemble_member(n, seed, prec, model="eulerpsi", dtmax=9e99, solver=solver)
        out[f"{prec}_method"] = m.mesh.poisson_vertices.method
        m.integrator.begin(m.state)
        warm(m, warmup)
        models[prec] = m
    windows = {prec: [] for prec in models}
    for _ in range(WINDOWS):
        for prec, m in models.items():
            windows[prec].append(timed_window(m, steps))
    for prec, ws in windows.items():
        ms = sum(w[0] for w in ws) / len(ws)
        st = ws[-1][2]
        out[f"{prec}_ms_per_step"] = ms
        out[f"{prec}_windows_ms_per_step"] = [w[0] for w in ws]
        out[f"{prec}_cell_steps_per_s"] = n * n / (ms * 1e-3)
        out[f"{prec}_solver"] = {k: st[k] for k in ("max_iters", "last_iters", "last_relres",
                                                    "last_refinements", "failed")}
    del models
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--accuracy-steps", type=int, default=1000)
    ap.add_argument("--cpu-steps", type=int, default=16,
                    help="RK3 steps of the CPU baseline sample (about 10 s at 1024^2)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--closed-steps", type=int, default=200,
                    help="steps of the closed 256^2 workload (GPU timing, and the oracle "
                         "with the CPU baseline); 0: skip")
    ap.add_argument("--only", choices=["fp32", "fp64", "both"], default="both")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the forward+adjoint cost measurement")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every kernel eagerly (no captured step graphs)")
    ap.add_argument("--ensemble", type=int, default=64, metavar="MEMBERS",
                    help="also time the batched qgrsw ensemble with this many members "
                         "(0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=200)
    ap.add_argument("--rsw-steps", type=int, default=200,
                    help="steps of the rsw 2048x1024 island channel per precision (0: skip)")
    ap.add_argument("--only-rsw", action="store_true",
                    help="time only the rsw island channel (for profiling its kernels)")
    ap.add_argument("--only-ensemble", action="store_true",
                    help="time only the qgrsw ensemble (for profiling its kernels)")
    ap.add_argument("--only-masked", choices=["cap", "mg"], default=None,
                    help="time only eulerpsi in a closed smoothed-noise basin with this "
                         "Poisson solver (--nx sets the size, --steps the window)")
    ap.add_argument("--no-tune", action="store_true",
                    help="ensemble: default stream count and solver schedules (no tuner "
                         "runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain elementwise copy of MIB MiB three times "
                         "(a kernel of known traffic, to scale rocprofv3 counters)")
    a = ap.parse_args()
    global HIPGRAPH
    HIPGRAPH = not a.no_graph
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the multi-GPU (domain-decomposed) "
              "path is not part of this build", file=sys.stderr)
        sys.exit(2)
    import torch
    from fluids2d_amd import roofline as rl

    if a.copy_calibration:
        n = a.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    if a.only_rsw:
        print(json.dumps({"rsw": rsw_bench(a.seed, a.rsw_steps or a.steps, a.warmup)}))
        return
    if a.only_masked:
        print(json.dumps({"masked": masked_bench(a.nx, 7, a.only_masked, a.steps)}))
        return
    if a.only_ensemble:
        print(json.dumps({"ensemble": ensemble_bench(
            a.ensemble or 64, a.ensemble_steps, tune=not a.no_tune,
            log=lambda m: print(m, file=sys.stderr, flush=True))}))
        return

    from fluids2d_amd import autotune
    nx = a.nx
    res = {}
    tiles = {}
    precs = ["fp64", "fp32"] if a.only == "both" else [a.only]
    models = {}
    for prec in precs:
        m = build(nx, prec, a.seed)
        calibrate(m, 3)
        # the transport tile by measurement (results do not depend on it)
        tiles[prec] = autotune.tune_tile(m, cache=False)
        warm(m, a.warmup)
        models[prec] = m
    # the precisions' windows interleaved, each exactly `steps` steps: a
    # slower first measurement in the process cannot favour one precision
    windows = {prec: [] for prec in precs}
    for _ in range(WINDOWS):
        for prec in precs:
            windows[prec].append(timed_window(models[prec], a.steps))
    for prec in precs:
        m = models[prec]
        ws = windows[prec]
        for w in ws:
            if w[2]["failed"] > 0:
                raise SystemExit(f"{prec}: {w[2]['failed']} Poisson solves did not converge "
                                 f"within the enqueued iterations")
        msps = sum(w[0] for w in ws) / len(ws)
        prof = rl.profile_dominant(m, reps=20)
        res[prec] = dict(ms=msps, windows_ms=[w[0] for w in ws], wall=sum(w[1] for w in ws),
                         stats=ws[-1][2], prof=prof, csps=nx * nx / (msps * 1e-3),
                         t_end=float(ws[-1][3].sum().item()))
    del models, m
    torch.cuda.empty_cache()

    acc_err = acc_diag = None
    if a.only == "both" and a.accuracy_steps > 0:
        acc_err, acc_dt, acc_diag = accuracy(nx, a.seed, a.accuracy_steps)

    head = res.get("fp32") or res.get("fp64")
    head_prec = "fp32" if "fp32" in res else "fp64"
    cpu = None
    if not a.no_cpu_baseline:
        usable, ncpu, cpu_model = host_cores()
        nthr = max(1, min(usable, MAX_CPU_THREADS))
        runs = {}
        for thr in sorted({1, nthr}):
            cps, el, tf, used = cpu_baseline(nx, a.seed, a.cpu_steps, thr)
            runs[thr] = dict(value=cps, seconds=el, threads=used, factorisation_s=tf)
        best = max(runs, key=lambda t: runs[t]["value"])
        cpu = {"value": runs[best]["value"], "unit": "cell-steps/s",
               "cores": runs[best]["threads"], "kind": "port",
               "sample": f"{a.cpu_steps} RK3 steps of the same {nx}^2 doubly periodic "
                         f"Euler workload through the oracle (the reference's algorithm: C "
                         f"restatement of the numba loops at -O3 with OpenMP over the flat "
                         f"index, numpy operators, scipy SuperLU), timed on 1 thread and on "
                         f"{nthr}; value is the faster; the one-time SuperLU "
                         f"factorisation excluded",
               "by_threads": {str(t): {"value": r["value"], "seconds": round(r["seconds"], 2),
                                       "factorisation_s": round(r["factorisation_s"], 2)}
                              for t, r in runs.items()},
               "host": {"usable_cpus": usable, "os_cpu_count": ncpu, "model_name": cpu_model,
                        "thread_cap": MAX_CPU_THREADS}}
    prof = head["prof"]
    out = {
        "metric": "cell-steps/s at 1024² Euler, FP32+refine vs FP64; fwd+adjoint/fwd cost ratio",
        "value": head["csps"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if "fp32" in res else "f64",
        "data": "synthetic",
        "config": {"workload": f"eulerpsi {nx}x{nx} doubly periodic, WENO5 flux, SSP-RK3, "
                               f"CFL 0.9, random-phase k^-3 initial vorticity (k0=8, unit RMS, "
                               f"seed {a.seed}); FP32 stencils/multigrid with FP64 "
                               f"iterative refinement of the Poisson residual",
                   "nx": nx, "ny": nx},
        "roofline": {"bound": "hbm", "achieved": prof["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": prof["achieved_gbs"] / HBM_PEAK_GBS,
                     "traffic": traffic_of(_traffic_key(prof["kernel"], tiles.get(head_prec)),
                                           nx, head_prec),
                     "kernel": prof["kernel"],
                     "bytes_per_launch": prof["bytes"], "avg_launch_us": prof["us"],
                     "kernels": {k: {"us": round(v["us"], 2), "per_step": v["per_step"],
                                     "gbs": round(v["gbs"], 1)}
                                 for k, v in prof["all"].items()}},
        "cpu_baseline": cpu,
    }
    out["timing"] = {"windows": WINDOWS, "steps_per_window": a.steps,
                     "note": "value and ms_per_step are the mean over the precision's "
                             "windows; FP64 and FP32 windows alternate; every window "
                             "replays step graphs captured before it",
                     "windows_ms_per_step": {k: v["windows_ms"] for k, v in res.items()}}
    if "fp64" in res:
        out["fp64_cell_steps_per_s"] = res["fp64"]["csps"]
        out["fp64_ms_per_step"] = res["fp64"]["ms"]
    if "fp64" in res and "fp32" in res:
        out["fp32_refine_over_fp64"] = res["fp32"]["csps"] / res["fp64"]["csps"]
    if acc_err is not None:
        out["fp32_vs_fp64_rel_l2_omega"] = acc_err
        out["fp32_vs_fp64_diagnostics"] = acc_diag
        out["accuracy_steps"] = a.accuracy_steps
    out["poisson"] = {k: v["stats"] for k, v in res.items()}
    out["transport_tile"] = {k: "x".join(map(str, _lib_tiles()[v])) for k, v in tiles.items()}
    out["adjoint_cost_ratio"] = None
    if not a.no_adjoint:
        ratio, tfw, tfb = adjoint_cost()
        out["adjoint_cost_ratio"] = ratio
        out["adjoint"] = {"workload": "eulerpsi 512x512 doubly periodic FP32, 200 RK3 steps, "
                                      "checkpoints every 20, gradient of a final misfit "
                                      "w.r.t. the initial vorticity",
                          "forward_s": tfw, "forward_plus_adjoint_s": tfb}
    if a.closed_steps > 0:
        out["closed256"] = closed_bench(a.seed, a.closed_steps)
        if cpu is not None:
            cpu["closed256"] = cpu_baseline_closed(a.seed, a.closed_steps,
                                                   sorted({1, cpu["cores"]} | {int(t) for t in cpu["by_threads"]}))
    if a.rsw_steps > 0:
        out["rsw"] = rsw_bench(a.seed, a.rsw_steps)
    if a.ensemble > 0:
        out["ensemble"] = ensemble_bench(a.ensemble, a.ensemble_steps,
                                         log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
