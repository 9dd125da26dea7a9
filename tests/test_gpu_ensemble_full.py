"""The qgrsw ensemble workload at its members' full sizes (closed
smoothed-noise basins of 1024^2 and 512^2, f0 = 10, g = H = 1, Gaussian h
bumps balanced by the QG projection, geos_adj.py:37-49; members seeded as
bench.py's ensemble seeds them), through properties that do not depend on
the size:

* member batching is exact: the batched ensemble (one launch per kernel and
  one batched capacitance solve per same-size group, its 2048-point odd
  extensions included) gives each member's state bit for bit what the same
  members stepped one by one give;
* FP32 with FP64 refinement stays within its stated bar of the FP64 path
  (which is held to the oracle at the sizes the oracle runs,
  test_gpu_sw.py) on a 1024^2 member, at a common fixed dt."""
import numpy as np
import pytest
import torch

import bench

pytestmark = pytest.mark.gpu

SEED = 99


def _members(sizes, precision="fp32", dt=None):
    out = []
    for k, n in sizes:
        m = bench.ensemble_member(n, SEED * 1000 + k, precision)
        if dt is not None:
            m.param.dt = dt
        m.integrator.begin(m.state)
        out.append(m)
    return out


def test_full_size_batched_matches_member_by_member(dev):
    from fluids2d_amd.ensemble import Ensemble
    sizes = [(k, n) for k, n in enumerate(bench.ensemble_sizes(64, SEED))
             if n >= 512][:5]
    assert sum(n == 1024 for _, n in sizes) >= 2  # a batched 1024^2 group
    states = []
    for batch in (True, False):
        ms = _members(sizes)
        e = Ensemble(ms, nstreams=2, batch=batch)
        if batch:
            assert e.groups() and max(len(g) for g, _ in e.groups()) >= 2
        d = np.concatenate([e.steps(2).cpu().numpy(), e.steps(8).cpu().numpy()], axis=1)
        assert np.isfinite(d).all() and (d > 0).all()
        assert e.graph is not None
        states.append((d, [(m.integrator.dstate.h.clone(), m.integrator.dstate.u.x.clone(),
                            m.integrator.dstate.u.y.clone()) for m in ms]))
        for m in ms:
            st = m.mesh.qg_helmholtz.read_stats()
            assert st["failed"] == 0
    (da, sa), (db, sb) = states
    assert np.array_equal(da, db)
    for a, b in zip(sa, sb):
        for x, y in zip(a, b):
            # (wet cells without a wet vertex hold NaN, as in the reference:
            # test_gpu_sw.py::test_qgrsw_nan_cells_like_reference)
            assert torch.equal(torch.isnan(x), torch.isnan(y))
            assert torch.equal(torch.nan_to_num(x), torch.nan_to_num(y))


def test_full_size_fp32_vs_fp64(dev):
    k, n = next((k, n) for k, n in enumerate(bench.ensemble_sizes(64, SEED)) if n == 1024)
    nsteps = 200
    runs = {}
    for prec in ("fp64", "fp32"):
        m = _members([(k, n)], prec, dt=0.02)[0]
        m.integrator.device_steps(nsteps)
        torch.cuda.synchronize()
        s = m.integrator.dstate
        runs[prec] = [a.double().cpu().numpy() for a in (s.h, s.u.x, s.u.y)]
    for a, b in zip(runs["fp32"], runs["fp64"]):
        assert np.array_equal(np.isnan(a), np.isnan(b))
    errs = [np.linalg.norm(np.nan_to_num(a - b)) / np.linalg.norm(np.nan_to_num(b))
            for a, b in zip(runs["fp32"], runs["fp64"])]
    print(f"\n1024^2 qgrsw member, {nsteps} steps at dt 0.02: FP32 vs FP64 rel L2 "
          f"h {errs[0]:.2e}, u.x {errs[1]:.2e}, u.y {errs[2]:.2e} (bar 1e-5)")
    assert max(errs) <= 1e-5
