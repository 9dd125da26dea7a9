"""Live roofline figures of the hot loop's kernels (HIP events on the stream
the kernels run on).

Algorithmic bytes per launch count every array a kernel reads once and every
array it writes once (int8 masks at 1 B), over the grid points one launch
covers; achieved bandwidth = those bytes / mean launch time.
"""
import torch

from . import _lib
from . import operators as op

KINDS = {0: "k_down<CG> (CG update + pre-smooth + residual + restrict, fine level)",
         1: "k_up<FINE> (prolongate + post-smooth + (r,z), fine level)",
         2: "k_pupdate (p = z + beta p, q = K p, (p,q))"}


def transport_timing(model, reps=20):
    """Mean launch time and algorithmic bytes of the fused transport stage
    (stage 2 of RK3: reads q, Ux, Uy, ds1, two orders + mask; writes q, ds2)."""
    integ = model.integrator
    s, alt = integ.dstate, integ.dalt
    m = model.mesh
    name = integ.model.prognostic[0]
    q, qo = getattr(s, name), getattr(alt, name)
    ds0, ds1 = integ.dscratch[0][0], integ.dscratch[1][0]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    args = (m, model.param.compflux, q, qo, s.U, [ds0], ds1, model.param.integrator, 1,
            integ.d_dt)
    op.transport_stage(*args)
    torch.cuda.synchronize()
    # back-to-back launches replayed from a graph, as the step runs them:
    # no host launch gap between dispatches
    g = torch.cuda.CUDAGraph()
    with _lib.capture(g):
        for _ in range(reps):
            op.transport_stage(*args)
    g.replay()
    torch.cuda.synchronize()
    e0.record()
    g.replay()
    e1.record()
    torch.cuda.synchronize()
    us = 1e3 * e0.elapsed_time(e1) / reps
    t = q.element_size()
    npts = q.numel()
    byts = npts * (6 * t + 3)
    return us, byts


FFT_KINDS = {3: "k_fft_rows_fwd (row FFTs of the right-hand side)",
             4: "k_fft_rows_fwd (row FFTs of the FP64 refinement residual)",
             5: "k_fft_cols (column FFT, eigenvalue division, inverse FFT)",
             6: "k_fft_rows_inv (inverse row FFTs, halo images)",
             7: "k_fft_rows_inv (inverse row FFTs accumulating the refinement)",
             8: "k_fft_rows_fwd (row FFTs of the FP64 residual of the previous solution)",
             9: "k_fft_rows_inv (inverse row FFTs accumulating onto the previous solution)"}


def profile_dominant(model, reps=20):
    """Time the kernels of one RK3 step live (HIP events on the stream they
    run on) and pick the dominant one: the largest mean launch time x
    launches per step.  Launch counts come from the solver's schedule (FFT:
    one solve per stage; in FP32 one FP64-residual refinement, from the
    previous solution or after a cold FP32 solve; MG: the iterations of the
    last solves)."""
    pois = model.mesh.poisson_vertices
    prec = model.param.precision
    rows = {}
    if pois.method == "fft":
        if prec == "fp64":
            per = {3: 3, 5: 3, 6: 3}
        elif getattr(model.param, "poisson_warm", True):
            # one FP64-residual refinement from the previous solution
            per = {8: 3, 5: 3, 9: 3}
        else:
            per = {3: 3, 4: 3, 5: 6, 6: 3, 7: 3}
        for kind, n in per.items():
            us, by = pois.bench_kernel(kind, reps)
            rows[FFT_KINDS[kind]] = dict(us=us, bytes=by, per_step=n)
    else:
        st = pois.read_stats()
        if prec == "fp64":
            iters = max(st["max_iters"], 1)
            per_solve = {0: iters + 1, 1: iters, 2: iters}
        else:
            nref = max(st["max_iters"], 1)
            k = 4
            per_solve = {0: nref * (k + 1), 1: nref * k, 2: nref * k}
        for kind in (0, 1, 2):
            us, by = pois.bench_kernel(kind, reps)
            rows[KINDS[kind]] = dict(us=us, bytes=by, per_step=3 * per_solve[kind])
    us, by = transport_timing(model, reps)
    rows["k_transport (fused WENO5 flux + div + RK update)"] = dict(us=us, bytes=by,
                                                                     per_step=3)
    name = max(rows, key=lambda k: rows[k]["us"] * rows[k]["per_step"])
    r = rows[name]
    return {"kernel": name, "us": r["us"], "bytes": r["bytes"],
            "achieved_gbs": r["bytes"] / (r["us"] * 1e-6) / 1e9,
            "all": {k: dict(v, gbs=v["bytes"] / (v["us"] * 1e-6) / 1e9)
                    for k, v in rows.items()}}


def sw_timing(model, reps=20):
    """Mean launch time and algorithmic bytes of the fused rsw stage kernels
    at stage 1 of RK3, each replayed back to back from a graph: the momentum
    kernel (reads U.x, U.y, omega, ke, p, the stage-0 tendency and u, two
    orders and two masks; writes u and the stage-1 tendency), the transport
    stage of h (reads h, U.x, U.y, the stage-0 tendency, two orders and msk;
    writes h and the tendency) and the diag kernel (reads u.x, u.y, h, the
    float64 slip coefficient, two orders and msk; writes U.x, U.y, omega, ke,
    p).  The updates go to scratch copies: the model's state is not
    advanced (the diag rewrites the diagnostics it already holds)."""
    integ = model.integrator
    sw = integ.model
    s = integ.dstate
    m, p = model.mesh, model.param
    d0, d1 = integ.dscratch[0], integ.dscratch[1]
    ux, uy, h = s.u.x.clone(), s.u.y.clone(), torch.empty_like(s.h)
    st = sw._stage_terms([d0], d1, 1, integ.d_dt, (ux, uy, None))
    t = s.h.element_size()
    npts = s.h.numel()
    launches = {
        "k_sw_momentum (vortex force + Coriolis + gradients + fill + RK update)":
            (lambda: sw._momentum(s, st, gradients=True, fill=True), npts * (13 * t + 4)),
        "k_transport (fused WENO5 flux + div + RK update of h)":
            (lambda: op.transport_stage(m, p.compflux, s.h, h, s.U, [d0.h], d1.h,
                                        p.integrator, 1, integ.d_dt), npts * (6 * t + 3)),
        "k_sw_diag (sharp + vorticity + kinetic energy + pressure + fills)":
            (lambda: sw._diag_fused(s, energy=True), npts * (8 * t + 11)),
    }
    rows = {}
    for name, (fn, byts) in launches.items():
        fn()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with _lib.capture(g):
            for _ in range(reps):
                fn()
        g.replay()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        g.replay()
        e1.record()
        torch.cuda.synchronize()
        us = 1e3 * e0.elapsed_time(e1) / reps
        rows[name] = dict(us=us, bytes=byts, per_step=3, gbs=byts / (us * 1e-6) / 1e9)
    name = max(rows, key=lambda k: rows[k]["us"] * rows[k]["per_step"])
    r = rows[name]
    return {"kernel": name, "us": r["us"], "bytes": r["bytes"], "achieved_gbs": r["gbs"],
            "all": rows}
