"""Checkpoint and resume (fluids2d_amd.restart, restart.py:1-58): a model
restored from a checkpoint continues exactly as the uninterrupted run
(FP64: every state array bit for bit, the time and iteration count)."""
import numpy as np
import pytest

from helpers import random_mask

pytestmark = pytest.mark.gpu


def _model(model, nx, ny, xper, yper, nisl, seed=4):
    import fluids2d_amd as f2d
    import fluids2d_amd.param as fp
    fp.QUIET = True
    p = f2d.Param()
    p.model, p.nx, p.ny, p.Lx, p.Ly = model, nx, ny, nx / ny, 1.0
    p.xperiodic, p.yperiodic = xper, yper
    p.tend, p.maxite = 1e9, 8
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
        m.state.omega[...] = (bump + 1e-3 * rng.uniform(-1, 1, bump.shape)) * m.mesh.msk * m.mesh.area
    m.integrator.diag(m.state)
    return m


def _flat(state):
    out = {}
    for n in state._fields:
        v = getattr(state, n)
        if hasattr(v, "_fields"):
            for c in v._fields:
                out[n + "." + c] = np.array(getattr(v, c))
        else:
            out[n] = np.array(v)
    return out


@pytest.mark.parametrize("model,nx,ny,xper,yper,nisl", [
    ("eulerpsi", 64, 64, True, True, 0), ("eulerpsi", 48, 40, False, False, 2),
    ("rsw", 48, 32, True, False, 2)])
def test_checkpoint_resume_bitexact(dev, tmp_path, model, nx, ny, xper, yper, nisl):
    from fluids2d_amd.restart import load_checkpoint, save_checkpoint
    print(f"\nrestart {model} {nx}x{ny} islands={nisl}", flush=True)
    a = _model(model, nx, ny, xper, yper, nisl)
    a.run()
    path = tmp_path / "ckpt.npz"
    save_checkpoint(a, path)
    a.param.maxite = 16
    a.run()
    b = load_checkpoint(path)
    assert b.time.ite == 8 and b.param.model == model
    b.param.maxite = 16
    b.run()
    assert a.time.ite == b.time.ite == 16
    assert a.time.t == b.time.t
    fa, fb = _flat(a.state), _flat(b.state)
    for k in fa:
        assert np.array_equal(np.nan_to_num(fa[k]), np.nan_to_num(fb[k])), k


def test_history_file_and_netcdf_restart(dev, tmp_path):
    """History output (io.py:1-92: a record every nhis iterations and at the
    end, float32 variables) and the restart from it
    (restart.py:53-57): the restarted model holds the last record and its
    time, and steps on.  NetCDF-3 through scipy when netCDF4 is absent."""
    import fluids2d_amd as f2d
    import fluids2d_amd.param as fp
    from fluids2d_amd.ncfile import Dataset
    from fluids2d_amd.restart import get_model_from_restart
    fp.QUIET = True
    path = str(tmp_path / "history.nc")
    p = f2d.Param()
    p.model, p.nx, p.ny, p.Lx = "eulerpsi", 48, 40, 1.2
    p.tend, p.maxite, p.nhis = 1e9, 10, 4
    p.var_to_store, p.outputfile = ["omega", "psi", "U"], path
    m = f2d.Model(p)
    x, y = m.mesh.xy("c")
    m.state.omega[...] = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / 0.02) * m.mesh.msk * m.mesh.area
    m.integrator.diag(m.state)
    m.run()
    with Dataset(path) as nc:
        ites = [int(v) for v in nc.variables["ite"][:]]
        assert ites == [0, 4, 8, 10]
        last = np.array(nc.variables["omega"][-1])
        assert np.array_equal(last, m.state.omega.astype(np.float32))
        assert np.array_equal(np.array(nc.variables["Ux"][-1]), m.state.U.x.astype(np.float32))
        assert float(nc.variables["t"][-1]) == np.float32(m.time.t)
    # as in the reference, the restarted Model creates its own history file
    # (the stored outputfile), so the restart reads a copy
    import shutil
    rpath = str(tmp_path / "restart.nc")
    shutil.copy(path, rpath)
    r = get_model_from_restart(rpath)
    assert (r.param.model, r.param.nx, r.param.ny, r.param.nhis) == ("eulerpsi", 48, 40, 4)
    assert r.param.var_to_store == ["omega", "psi", "U"]
    assert r.time.ite == 10 and r.time.ite0 == 10
    assert np.array_equal(r.state.omega, last.astype(np.float64))
    # diag ran on the restored vorticity: psi solves it (float32 storage aside)
    e = np.linalg.norm(r.state.psi - m.state.psi) / np.linalg.norm(m.state.psi)
    print(f"\nrestart from history: psi rel L2 vs run {e:.2e}")
    assert e < 1e-6
    r.param.maxite = 3
    r.run()
    assert r.time.ite == 13 and np.isfinite(r.state.omega).all()


def test_bulk_diagnostics_file(dev, tmp_path):
    """Bulk diagnostics written to their netCDF file (diagnostics.py:79-101)
    read back as recorded (float32 variables)."""
    from fluids2d_amd.diagnostics import Bulk
    path = str(tmp_path / "bulk.nc")
    m = _model("eulerpsi", 64, 64, True, True, 0)
    m.param.maxite = 12
    b = Bulk(m, ncfile=path)
    m.diags.append(b)
    m.run()
    disk = b.read()  # from the file
    b.ncfile = None
    mem = b.read()  # the same records, kept in memory
    assert len(disk.time) == len(mem.time) == 4
    for got, want in zip(disk, mem):
        assert np.array_equal(np.asarray(got), np.asarray(want).astype(np.float32))
