"""Bulk diagnostics on the device (fluids2d_amd.diagnostics.Bulk) against
the reference's formulas (diagnostics.py:45-66) evaluated with numpy on the
host copy of the same states: domain averages over the padded arrays,
every third iteration, FP64 sums (also for FP32 runs)."""
import numpy as np
import pytest

from helpers import random_mask

pytestmark = pytest.mark.gpu


def _model(model, nx, ny, xper, yper, nisl, precision, seed=3):
    import fluids2d_amd as f2d
    import fluids2d_amd.param as fp
    fp.QUIET = True
    p = f2d.Param()
    p.model, p.nx, p.ny = model, nx, ny
    p.Lx, p.Ly = nx / ny, 1.0
    p.xperiodic, p.yperiodic = xper, yper
    p.precision = precision
    p.tend, p.maxite = 1e9, 9
    if model == "rsw":
        p.f0, p.g, p.H = 1.0, 1.0, 1.0
    m = f2d.Model(p)
    if nisl:
        m.mesh.msk[...] = random_mask(ny, nx, seed, xper, yper, nisl=nisl)
        m.mesh.finalize()
        m.integrator.model.mesh = m.mesh
    rng = np.random.default_rng(seed)
    x, y = m.mesh.xy("c")
    bump = np.exp(-((x - 0.45 * p.Lx) ** 2 + (y - 0.5) ** 2) / 0.02)
    if model == "rsw":
        m.state.h[...] = (1 + 0.05 * bump) * m.mesh.msk * m.mesh.area
        m.state.u.x[...] = 1e-3 * rng.standard_normal(m.mesh.shape) * m.mesh.mskx
    else:
        w = bump - np.exp(-((x - 0.6 * p.Lx) ** 2 + (y - 0.5) ** 2) / 0.02)
        m.state.omega[...] = (w + 1e-3 * rng.uniform(-1, 1, w.shape)) * m.mesh.msk * m.mesh.area
    m.integrator.diag(m.state)
    return m


def _numpy_bulk(m):
    """diagnostics.py:45-66 on the host state (eulerpsi: the cell-centred
    1/2 u.U that stands in for the missing s.ke)."""
    s, mesh = m.state, m.mesh
    msk = mesh.msk.astype(float)

    def domavg(a):
        return np.sum(a) / np.sum(msk)
    if hasattr(s, "ke"):
        ke = s.ke
    else:
        uxe = np.zeros_like(s.U.x)
        uxe[:, :-1] = s.U.x[:, 1:]
        uyn = np.zeros_like(s.U.y)
        uyn[:-1, :] = s.U.y[1:, :]
        ke = 0.25 * (mesh.dx ** 2 * (s.U.x ** 2 + uxe ** 2) +
                     mesh.dy ** 2 * (s.U.y ** 2 + uyn ** 2)) * msk
    xv, yv = mesh.xy("y")
    xu, yu = mesh.xy("x")
    return np.array([domavg(ke), 0.5 * domavg(s.omega ** 2) / mesh.area ** 2,
                     domavg(s.omega) / mesh.area,
                     domavg(s.U.y * xv) * mesh.dx - domavg(s.U.x * yu) * mesh.dy])


@pytest.mark.parametrize("model,nx,ny,xper,yper,nisl,precision", [
    ("eulerpsi", 64, 64, True, True, 0, "fp64"), ("eulerpsi", 48, 40, False, False, 2, "fp64"),
    ("eulerpsi", 64, 64, True, True, 0, "fp32"), ("rsw", 48, 32, True, False, 2, "fp64")])
def test_bulk_matches_reference_formulas(dev, model, nx, ny, xper, yper, nisl, precision):
    from fluids2d_amd.diagnostics import Bulk
    print(f"\nbulk {model} {nx}x{ny} {precision}", flush=True)
    a = _model(model, nx, ny, xper, yper, nisl, precision)
    bulk = Bulk(a, ncfile=None)
    a.diags.append(bulk)
    a.run()
    rec = bulk.read()
    assert len(rec.time) == 3  # iterations 3, 6, 9
    b = _model(model, nx, ny, xper, yper, nisl, precision)
    k = 0
    for ite in range(1, 10):
        b.set_dt()  # Model.run's CFL step (Model.step keeps time.dt)
        b.step()
        if ite % 3:
            continue
        ref = _numpy_bulk(b)
        got = np.array([rec.ke[k], rec.ens[k], rec.vort[k], rec.angular[k]])
        scale = np.maximum(np.abs(ref), 1e-300)
        err = np.abs(got - ref) / scale
        print(f"  ite {ite}: {got} rel err {err}")
        # vort (a zero-mean quantity in periodic runs) by absolute error
        assert err[0] < 1e-12 and err[1] < 1e-12, (ite, got, ref)
        assert abs(got[2] - ref[2]) <= 1e-12 * max(abs(ref[2]), np.abs(b.state.omega).max()
                                                 / b.mesh.area)
        assert abs(got[3] - ref[3]) <= 1e-12 * max(abs(ref[3]), 1e-300) + 1e-15
        assert rec.time[k] == b.time.t
        k += 1
