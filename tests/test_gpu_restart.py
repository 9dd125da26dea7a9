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
