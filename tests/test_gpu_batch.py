"""Member batching of kernel launches (f2d_record_begin / f2d_launch_batch_*,
f2d_batch.hip): the recorded launches of several members, replayed as one
launch per kernel, give each member's plain launches bit for bit; entry
points that cannot be recorded refuse to run while a recording is open;
recordings that differ (kernels, grids) are refused; an ensemble whose
group cannot be recorded (x-periodic qgrsw: the unfilled divergence goes
through f2d_compflux) steps its members one by one, with the same
results."""
import ctypes

import numpy as np
import pytest
import torch

import bench

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from fluids2d_amd import _lib
    return _lib


def _grid(L, n):
    from helpers import P
    return L.make_grid(P(model="qgrsw", nx=n, ny=n))


def _pv_call(L, g, dux, duy, dh, slip, mv, pv):
    L.call("f2d_qg_pv", L.dtype_code(pv.dtype), ctypes.byref(g), L.ptr(dux), L.ptr(duy),
           L.ptr(dh), L.ptr(slip), L.ptr(mv), -10.0, L.ptr(pv), L.stream_ptr())


def _record(L, fn):
    L.call("f2d_record_begin")
    h = ctypes.c_void_p()
    try:
        fn()
    finally:
        L.call("f2d_record_end", ctypes.byref(h))
    return h


def _batch(L, recs):
    arr = ctypes.c_void_p * len(recs)
    h = ctypes.c_void_p()
    L.call("f2d_launch_batch_create", arr(*recs), len(recs), ctypes.byref(h))
    return h


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_recorded_batch_matches_plain_launches(dev, L, dtype):
    n, nm = 64, 3
    g = _grid(L, n)
    shape = (n + 6, n + 6)
    rng = np.random.default_rng(4)
    members = []
    for i in range(nm):
        arrs = [torch.from_numpy(rng.standard_normal(shape)).to(dev, dtype) for _ in range(3)]
        slip = torch.from_numpy(rng.uniform(0, 1, shape)).to(dev)
        mv = torch.from_numpy((rng.uniform(0, 1, shape) > 0.2).astype(np.int8)).to(dev)
        pv0 = torch.from_numpy(rng.standard_normal(shape)).to(dev, dtype)
        members.append((arrs, slip, mv, pv0.clone(), pv0.clone()))
    recs = []
    for arrs, slip, mv, pvb, _ in members:
        recs.append(_record(L, lambda: _pv_call(L, g, *arrs, slip, mv, pvb)))
    assert [L.lib().f2d_recording_size(r) for r in recs] == [1] * nm
    torch.cuda.synchronize()
    for arrs, slip, mv, pvb, pvp in members:
        assert torch.equal(pvb, pvp)  # recording launched nothing
    bt = _batch(L, recs)
    for r in recs:
        L.lib().f2d_recording_destroy(r)
    for rep in range(2):
        L.call("f2d_launch_batch_run", bt, L.stream_ptr())
        for arrs, slip, mv, _, pvp in members:
            _pv_call(L, g, *arrs, slip, mv, pvp)
        torch.cuda.synchronize()
        for i, (_, _, _, pvb, pvp) in enumerate(members):
            assert torch.equal(pvb, pvp), (rep, i)
    L.call("f2d_launch_batch_destroy", bt)


def test_recording_refusals(dev, L):
    from fluids2d_amd._lib import F2DError
    g = _grid(L, 32)
    a = torch.zeros((38, 38), dtype=torch.float64, device=dev)
    # an entry point that cannot be recorded fails and launches nothing
    L.call("f2d_record_begin")
    with pytest.raises(F2DError, match="cannot be recorded"):
        L.call("f2d_fill_halo", L.dtype_code(a.dtype), ctypes.byref(g), L.ptr(a),
               L.stream_ptr())
    with pytest.raises(F2DError, match="already open"):
        L.call("f2d_record_begin")
    h = ctypes.c_void_p()
    L.call("f2d_record_end", ctypes.byref(h))
    assert L.lib().f2d_recording_size(h) == 0
    L.lib().f2d_recording_destroy(h)
    with pytest.raises(F2DError, match="no recording"):
        L.call("f2d_record_end", ctypes.byref(h))
    # recordings of different grids do not merge
    sl = torch.ones((38, 38), dtype=torch.float64, device=dev)
    mv = torch.ones((38, 38), dtype=torch.int8, device=dev)
    r1 = _record(L, lambda: _pv_call(L, g, a, a, a, sl, mv, a))
    g2 = _grid(L, 128)
    b = torch.zeros((134, 134), dtype=torch.float64, device=dev)
    s2 = torch.ones((134, 134), dtype=torch.float64, device=dev)
    m2 = torch.ones((134, 134), dtype=torch.int8, device=dev)
    r2 = _record(L, lambda: _pv_call(L, g2, b, b, b, s2, m2, b))
    with pytest.raises(F2DError, match="different launches"):
        _batch(L, [r1, r2])
    r3 = _record(L, lambda: None)
    with pytest.raises(F2DError, match="different numbers"):
        _batch(L, [r1, r3])
    for r in (r1, r2, r3):
        L.lib().f2d_recording_destroy(r)


def _channel_member(n, seed):
    import fluids2d_amd as f2d
    import fluids2d_amd.param as fp
    from fluids2d_amd.rsw import qg_projection
    fp.QUIET = True
    p = f2d.Param()
    p.model = "qgrsw"
    p.nx = p.ny = n
    p.xperiodic = True
    p.precision = "fp32"
    p.dtmax = 0.05
    p.tend, p.maxite = 1e9, 10 ** 9
    m = f2d.Model(p)
    # one (convex) island: every wet cell touches a wet vertex, so the
    # projection's verticestocenters never divides by zero on a wet cell
    m.mesh.msk[...] = bench.channel_islands_mask(n, n, seed, nisl=1)
    m.mesh.finalize()
    m.integrator.model.mesh = m.mesh
    x, y = m.mesh.xy("c")
    rng = np.random.default_rng(seed + 7)
    bumps = np.zeros(m.mesh.shape)
    for _ in range(3):
        x0, y0 = rng.uniform(0.2, 0.8, 2)
        bumps += np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * 0.05 ** 2))
    bumps[:, :3] = bumps[:, n:n + 3]
    bumps[:, n + 3:] = bumps[:, 3:6]
    m.state.h[...] = (p.H + 0.05 * bumps) * m.mesh.msk * m.mesh.area
    qg_projection(m)
    return m


def test_unrecordable_group_steps_member_by_member(dev):
    from fluids2d_amd.ensemble import Ensemble
    runs = [[_channel_member(128, s) for s in (3, 4)] for _ in range(2)]
    for ms in runs:
        for m in ms:
            assert m.mesh.qg_helmholtz.method == "capacitance"
            m.integrator.begin(m.state)
    eb = Ensemble(runs[0], nstreams=2, batch=True)
    eu = Ensemble(runs[1], nstreams=2, batch=False)
    assert [len(mm) for mm, _ in eb.groups()] == [2]
    d = [np.concatenate([e.steps(3).cpu().numpy(), e.steps(4).cpu().numpy()], axis=1)
         for e in (eb, eu)]
    # the parts with the pre-solve kernels (k < 3) fell back to member-by-member
    # launches; the last stage's post-solve part (k = 3) and the CFL steps
    # were recorded
    assert eb.graph is not None and eb._lb
    for key, v in eb._lb.items():
        if key[1] == "cfl":
            assert v is not None
        else:
            assert (v is None) == (key[1] < 3), key
    assert np.isfinite(d[0]).all()
    assert np.array_equal(d[0], d[1])
    for mb, mu in zip(*runs):
        for name in ("h", "psi"):
            a = getattr(mb.integrator.dstate, name)
            b = getattr(mu.integrator.dstate, name)
            assert torch.equal(torch.nan_to_num(a), torch.nan_to_num(b)), name
