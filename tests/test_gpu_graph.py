"""Captured HIP graphs of device steps replay exactly the eager launches:
the state after a run with graphs equals, bit for bit, the state after the
same run without them (eulerpsi with the direct FFT solve and with the
masked multigrid CG, qg, qgrsw, euler with the direct Neumann solve; FP64
and FP32+refinement).  The graph is
replayed several times with eager steps and host reads between replays,
and every replay's dt log is checked against the eager run's."""
import numpy as np
import pytest

from helpers import random_mask

pytestmark = pytest.mark.gpu


def _model(model, nx, ny, xper, yper, nisl, precision, seed, hipgraph):
    import fluids2d_amd as f2d
    import fluids2d_amd.param as fp
    fp.QUIET = True
    p = f2d.Param()
    p.model, p.nx, p.ny = model, nx, ny
    p.Lx, p.Ly = nx / ny, 1.0
    p.xperiodic, p.yperiodic = xper, yper
    p.precision = precision
    p.hipgraph = hipgraph
    p.tend = 1e9
    if model == "qgrsw":
        p.dt = 0.005
    m = f2d.Model(p)
    if nisl:
        m.mesh.msk[...] = random_mask(ny, nx, seed, xper, yper, nisl=nisl)
        m.mesh.finalize()
    rng = np.random.default_rng(seed)
    if model == "qgrsw":
        s = m.state
        s.h[...] = (1 + 0.05 * rng.standard_normal(m.mesh.shape)) * m.mesh.area * m.mesh.msk
        s.u.x[...] = 1e-3 * rng.standard_normal(m.mesh.shape) * m.mesh.mskx
        s.u.y[...] = 1e-3 * rng.standard_normal(m.mesh.shape) * m.mesh.msky
    elif model == "euler":
        s = m.state
        s.u.x[...] = 0.05 * rng.standard_normal(m.mesh.shape) * m.mesh.mskx
        s.u.y[...] = 0.05 * rng.standard_normal(m.mesh.shape) * m.mesh.msky
    else:
        x, y = m.mesh.xy("c")
        w = np.exp(-((x - 0.4 * p.Lx) ** 2 + (y - 0.5) ** 2) / 0.02) - \
            np.exp(-((x - 0.6 * p.Lx) ** 2 + (y - 0.5) ** 2) / 0.02)
        w += 1e-3 * rng.uniform(-1, 1, w.shape)
        field = "pv" if model == "qg" else "omega"
        getattr(m.state, field)[...] = w * m.mesh.msk * m.mesh.area
    m.integrator.diag(m.state)
    m.integrator.begin(m.state)
    return m


def _run(m, chunks):
    import torch
    integ = m.integrator
    logs = []
    for n in chunks:
        logs.append(integ.device_steps(n).cpu().numpy())
        torch.cuda.synchronize()
        # eager host work between replays
        _ = float(integ.d_dt.cpu()[0])
    integ.end(m.state)
    return np.concatenate(logs)


CASES = [
    ("eulerpsi", 64, 64, True, True, 0, "fp64"),
    ("eulerpsi", 64, 64, True, True, 0, "fp32"),
    ("eulerpsi", 48, 40, False, False, 3, "fp64"),
    ("eulerpsi", 48, 40, True, False, 2, "fp32"),
    ("qg", 48, 48, False, False, 2, "fp64"),
    ("qgrsw", 48, 48, False, False, 0, "fp32"),
    ("euler", 64, 32, False, False, 0, "fp64"),
    ("euler", 32, 32, True, False, 0, "fp32"),
]


@pytest.mark.parametrize("model,nx,ny,xper,yper,nisl,precision", CASES)
def test_graph_replay_matches_eager(dev, model, nx, ny, xper, yper, nisl, precision):
    chunks = [11, 3, 9]  # 2 eager + 2 graphs + 1 eager; eager; 2 graphs + 1 eager
    print(f"\ngraph case {model} {nx}x{ny} xper={xper} yper={yper} islands={nisl} "
          f"{precision}", flush=True)
    a = _model(model, nx, ny, xper, yper, nisl, precision, 5, True)
    b = _model(model, nx, ny, xper, yper, nisl, precision, 5, False)
    la = _run(a, chunks)
    lb = _run(b, chunks)
    assert len(a.integrator._graphs) >= 1
    assert np.array_equal(la, lb), (la, lb)
    names = {"qgrsw": ["h", "u"], "qg": ["pv"], "euler": ["u", "p"]}.get(model, ["omega"])
    for name in names:
        x, y = getattr(a.state, name), getattr(b.state, name)
        for u, v in (zip(x, y) if name == "u" else [(x, y)]):
            assert np.array_equal(np.nan_to_num(u), np.nan_to_num(v)), name
    solver = "poisson_centers" if model == "euler" else "poisson_vertices"
    pa = getattr(a.mesh, solver).read_stats()
    pb = getattr(b.mesh, solver).read_stats()
    assert pa["failed"] == 0 and pb["failed"] == 0
