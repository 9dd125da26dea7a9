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
    previous solution or after a cold FP32 solve, at every stage or only at
    the last with plain FP32 solves before it; MG: the iterations of the
    last solves)."""
    pois = model.mesh.poisson_vertices
    prec = model.param.precision
    rows = {}
    if pois.method == "fft":
        if prec == "fp64":
            per = {3: 3, 5: 3, 6: 3}
        elif getattr(model.param, "poisson_warm", True) and \
                getattr(model.param, "poisson_refine", "every") == "last":
            # plain FP32 solves in stages 0 and 1; the last stage's refines
            # once in FP64 from the previous stage's solution
            per = {3: 2, 8: 1, 5: 3, 6: 2, 9: 1}
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


def sw_timing(model, reps=4):
    """Mean launch time and algorithmic bytes of the fused rsw stage kernels
    within steps: `reps` times, the step graph is replayed three times (12
    steps) and two eager steps are enqueued behind it with HIP events around
    each of their kernels, so that the device never waits for the host and
    an interval is the kernel's duration in the step (the caches hold what
    the step's earlier kernels left).  Bytes per launch are averaged over
    the three RK3 stages: the momentum kernel reads u.x, u.y, omega, ke, h
    and the earlier tendencies (0, 1, 2 pairs), two orders and two masks,
    and writes the new u and the stage's tendency (stages 0 and 1); the
    transport of h reads h, u.x, u.y and the earlier tendencies, two orders
    and msk, and writes h and the tendency (stages 0, 1); the diag reads
    u.x, u.y, the float64 slip coefficient, two orders and msk, and writes
    omega and ke.  The model advances by the steps taken."""
    integ = model.integrator
    sw = integ.model
    s = integ.dstate
    t = s.h.element_size()
    npts = s.h.numel()
    per_stage = {  # arrays of the element type read + written, int8 / float64 bytes
        "k_sw_momentum": ([5 + 4, 5 + 2 + 4, 5 + 4 + 2], 4),
        "k_transport": ([3 + 2, 4 + 2, 5 + 1], 3),
        "k_sw_diag": ([4, 4, 4], 11),
    }
    names = {"k_sw_momentum": "k_sw_momentum (vortex force + Coriolis + gradients + fill + "
                              "RK update)",
             "k_transport": "k_transport (fused WENO5 flux + div + RK update of h)",
             "k_sw_diag": "k_sw_diag (vorticity + kinetic energy + fills)"}
    integ.prepare_graph()
    sw.events = []
    try:
        for _ in range(reps):
            g = integ._graph()
            for _ in range(3):
                g.replay()
            for _ in range(2):
                integ._select_dt()
                integ.advance(stats=False)
        torch.cuda.synchronize()
        ev = sw.events
    finally:
        sw.events = None
    rows = {}
    for key, (arrays, extra) in per_stage.items():
        durs = [e0.elapsed_time(e1) * 1e3 for (n, k, e0, e1) in ev if n == key]
        byts = npts * (t * sum(arrays) / len(arrays) + extra)
        us = sum(durs) / len(durs)
        rows[names[key]] = dict(us=us, bytes=byts, per_step=3, gbs=byts / (us * 1e-6) / 1e9,
                                launches=len(durs))
    name = max(rows, key=lambda k: rows[k]["us"] * rows[k]["per_step"])
    r = rows[name]
    return {"kernel": name, "us": r["us"], "bytes": r["bytes"], "achieved_gbs": r["gbs"],
            "all": rows}
