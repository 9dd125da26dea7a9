"""The CFL step of model.py:84-87 split in two (f2d_perpgrad_part /
f2d_cfl_finish / f2d_transport_stage_cfl): the step's last perpgrad leaves
its block maxima and the next step's first transport forms the dt from
them.  The dt is bit for bit the one f2d_perpgrad_ex writes, the stage is
bit for bit f2d_transport_stage_log given that dt, and eulerpsi runs that
defer the dt (eager and graph-replayed, solo and in an ensemble) match runs
that do not, bit for bit, including the host path Model.set_dt in between."""
import ctypes

import numpy as np
import pytest
import torch

from test_gpu_graph import _model

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_cfl_parts_match_fused_cfl(dev, dtype):
    from fluids2d_amd import _lib
    from fluids2d_amd import operators as op
    m = _model("eulerpsi", 96, 64, True, False, 2, "fp64" if dtype == torch.float64 else "fp32",
               5, False)
    mesh, p = m.mesh, m.param
    s = m.integrator.dstate
    npart = op.cfl_partials(mesh)
    part = torch.full((npart,), float("nan"), dtype=torch.float64, device=dev)
    Ua = (torch.zeros_like(s.U.x), torch.zeros_like(s.U.y))
    Ub = (torch.zeros_like(s.U.x), torch.zeros_like(s.U.y))
    dt_a = torch.zeros(1, dtype=torch.float64, device=dev)
    dt_b = torch.zeros(1, dtype=torch.float64, device=dev)
    sx, sy = 1 / mesh.dy ** 2, 1 / mesh.dx ** 2
    g = mesh.grid_ref()
    _lib.call("f2d_perpgrad_ex", _lib.dtype_code(dtype), g, _lib.ptr(s.psi), _lib.ptr(Ua[0]),
              _lib.ptr(Ua[1]), _lib.ptr(mesh.d_mskx), _lib.ptr(mesh.d_msky), sx, sy, 1, None,
              _lib.ptr(dt_a), p.cfl, p.dtmax, _lib.stream_ptr())
    _lib.call("f2d_perpgrad_part", _lib.dtype_code(dtype), g, _lib.ptr(s.psi), _lib.ptr(Ub[0]),
              _lib.ptr(Ub[1]), _lib.ptr(mesh.d_mskx), _lib.ptr(mesh.d_msky), sx, sy, 1,
              _lib.ptr(part), npart, _lib.stream_ptr())
    op.cfl_finish(part, dt_b, p.cfl, p.dtmax)
    torch.cuda.synchronize()
    assert torch.isfinite(part).all()  # every block wrote its maxima
    assert torch.equal(Ua[0], Ub[0]) and torch.equal(Ua[1], Ub[1])
    assert dt_a.item() == dt_b.item() and 0 < dt_a.item() < float("inf")
    # the stage forming dt from the maxima against the stage reading it
    q = s.omega
    outs = []
    for mode in ("log", "cfl"):
        qo = torch.zeros_like(q)
        ds = torch.zeros_like(q)
        log = torch.zeros(1, dtype=torch.float64, device=dev)
        dtb = dt_a.clone() if mode == "log" else torch.zeros_like(dt_a)
        op.transport_stage(mesh, p.compflux, q, qo, m.integrator.dstate.U, [], ds, "rk3", 0,
                           dtb, log, cfl_part=part if mode == "cfl" else None, cfl=p.cfl,
                           dtmax=p.dtmax)
        torch.cuda.synchronize()
        outs.append((qo, ds, log.item(), dtb.item()))
    (qa, da, la, ta), (qb, db, lb, tb) = outs
    assert torch.equal(qa, qb) and torch.equal(da, db)
    assert la == lb == ta == tb == dt_a.item()
    # a partial array of another grid's length is refused
    with pytest.raises(_lib.F2DError, match="npart"):
        op.transport_stage(mesh, p.compflux, q, torch.zeros_like(q), m.integrator.dstate.U, [],
                           torch.zeros_like(q), "rk3", 0, dt_b, None, cfl_part=part[:-2],
                           cfl=p.cfl, dtmax=p.dtmax)


def _state(m):
    return m.integrator.dstate.omega.clone()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("hipgraph", [False, True])
def test_deferred_cfl_matches_fused(dev, precision, hipgraph):
    runs = []
    for defer in (True, False):
        m = _model("eulerpsi", 64, 64, True, True, 0, precision, 11, hipgraph)
        integ = m.integrator
        if not defer:
            integ._defers = lambda: False
        logs = [integ.device_steps(n).cpu().numpy() for n in (3, 8, 5)]
        # the host path in between: Model.set_dt finishes a pending step
        integ.set_dt(m.time)
        h_dt = m.time.dt
        integ.advance(stats=False)
        logs.append(integ.device_steps(8).cpu().numpy())
        torch.cuda.synchronize()
        assert (integ._cfl is not None) == defer
        runs.append((np.concatenate(logs), h_dt, _state(m)))
    (la, ha, sa), (lb, hb, sb) = runs
    assert np.array_equal(la, lb) and np.all(la > 0)
    assert ha == hb
    assert torch.equal(sa, sb)


def test_deferred_cfl_in_ensemble(dev):
    from fluids2d_amd.ensemble import Ensemble
    runs = []
    for defer in (True, False):
        ms = [_model("eulerpsi", 64, 64, True, True, 0, "fp32", s, True) for s in (3, 4, 5)]
        if not defer:
            for m in ms:
                m.integrator._defers = lambda: False
        e = Ensemble(ms, nstreams=2)
        d = np.concatenate([e.steps(3).cpu().numpy(), e.steps(8).cpu().numpy()], axis=1)
        runs.append((d, [_state(m) for m in ms], e.graph is not None))
    (da, sa, ga), (db, sb, gb) = runs
    assert ga and gb
    assert np.array_equal(da, db) and np.all(da > 0)
    for a, b in zip(sa, sb):
        assert torch.equal(a, b)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_cfl_dt_part_recorded_batch(dev, dtype):
    """f2d_cfl_dt_part gives f2d_cfl_dt's dt, and a recorded batch of three
    members' calls (one launch per kernel) gives each member's own."""
    from fluids2d_amd import _lib
    from fluids2d_amd import operators as op
    from fluids2d_amd.ensemble import LaunchBatch
    from fluids2d_amd.states import Vector
    rng = np.random.default_rng(8)
    n = 262 * 262
    Us, parts, dts, refs = [], [], [], []
    for i in range(3):
        ux = torch.from_numpy(rng.standard_normal(n) * (i + 1)).to(dev, dtype)
        uy = torch.from_numpy(rng.standard_normal(n) * 0.5).to(dev, dtype)
        Us.append(Vector(ux, uy))
        parts.append(torch.zeros(op.CFL_PARTS, dtype=torch.float64, device=dev))
        dts.append(torch.zeros(1, dtype=torch.float64, device=dev))
        refs.append(torch.zeros(1, dtype=torch.float64, device=dev))
    for U, r in zip(Us, refs):
        op.cfl_dt(U, 0.9, 1e9, r)
    for U, pt, d in zip(Us, parts, dts):
        op.cfl_dt(U, 0.9, 1e9, d, part=pt)
    torch.cuda.synchronize()
    assert [d.item() for d in dts] == [r.item() for r in refs]
    assert len({r.item() for r in refs}) == 3
    for d in dts:
        d.zero_()
    lb = LaunchBatch([lambda U=U, pt=pt, d=d: op.cfl_dt(U, 0.9, 1e9, d, part=pt)
                      for U, pt, d in zip(Us, parts, dts)])
    torch.cuda.synchronize()
    assert all(d.item() == 0.0 for d in dts)  # recording launched nothing
    lb.run()
    torch.cuda.synchronize()
    assert [d.item() for d in dts] == [r.item() for r in refs]
    with pytest.raises(_lib.F2DError, match="npart"):
        op.cfl_dt(Us[0], 0.9, 1e9, dts[0], part=parts[0][:100])
