"""Batched ensembles (fluids2d_amd.ensemble) and the autotuner
(fluids2d_amd.autotune).  An ensemble of members with uneven sizes and
masks, bin-packed onto concurrent streams and replayed from one captured
graph, gives each member's solo run bit for bit; tuned schedules change
speed, not results (transport tiles) or only within the solver tolerance
(multigrid schedules)."""
import numpy as np
import pytest
import torch

import bench

pytestmark = pytest.mark.gpu


def _members(specs, precision):
    out = []
    for n, model, seed in specs:
        out.append(bench.ensemble_member(n, seed, precision, model=model))
    return out


def _state(m, names):
    m.integrator.end(m.state)
    out = {}
    for name in names:
        v = getattr(m.state, name)
        if hasattr(v, "_fields"):
            for c in v._fields:
                out[name + "." + c] = np.array(getattr(v, c))
        else:
            out[name] = np.array(v)
    return out


SPECS = [(32, "qgrsw", 1), (64, "qgrsw", 2), (48, "qgrsw", 3), (40, "qg", 4),
         (64, "eulerpsi", 5), (32, "qgrsw", 6)]
NAMES = {"qgrsw": ["h", "u"], "qg": ["pv"], "eulerpsi": ["omega"]}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_ensemble_matches_solo_runs(dev, precision):
    from fluids2d_amd.ensemble import Ensemble
    print(f"\nensemble {precision}: {SPECS}", flush=True)
    ens_members = _members(SPECS, precision)
    solo = _members(SPECS, precision)
    for m in ens_members + solo:
        m.integrator.begin(m.state)
    ens = Ensemble(ens_members, nstreams=3, graph_steps=2)
    d1 = ens.steps(7).cpu().numpy()    # 2 eager, 2 replays, 1 eager
    d2 = ens.steps(4).cpu().numpy()    # 2 replays
    assert ens.graph is not None
    for k, (m, ms) in enumerate(zip(solo, ens_members)):
        dts = np.concatenate([m.integrator.device_steps(7).cpu().numpy(),
                              m.integrator.device_steps(4).cpu().numpy()])
        assert np.array_equal(dts, np.concatenate([d1[k], d2[k]])), k
        a = _state(ms, NAMES[m.param.model])
        b = _state(m, NAMES[m.param.model])
        wet = ms.mesh.msk == 1
        for name in a:
            assert np.array_equal(np.nan_to_num(a[name]), np.nan_to_num(b[name])), (k, name)
            assert np.array_equal(np.isnan(a[name]), np.isnan(b[name])), (k, name)
            # dry cells may hold the reference's (1/0)*0 of verticestocenters
            assert np.isfinite(a[name][wet]).all(), (k, name)


BATCH_SPECS = [(64, "qgrsw", 11), (128, "qgrsw", 12), (64, "qgrsw", 13), (32, "qgrsw", 14),
               (64, "qgrsw", 15), (40, "qg", 16), (128, "qgrsw", 17), (48, "qgrsw", 18)]


def test_ensemble_batched_solves(dev):
    """Same-size qgrsw members stepped stage by stage as groups -- one
    launch per kernel before and after the solves, the Helmholtz solves
    batched (FP32 capacitance solves: groups 64^2 x3, 128^2 x2, 32^2 x1;
    the 48^2 qgrsw member (multigrid) and the qg member take whole steps)
    -- give the unbatched ensemble's and each member's solo run, bit for
    bit, eager and replayed."""
    from fluids2d_amd.ensemble import Ensemble
    print(f"\nbatched ensemble: {BATCH_SPECS}", flush=True)
    runs = [_members(BATCH_SPECS, "fp32") for _ in range(3)]
    for ms in runs:
        for m in ms:
            m.integrator.begin(m.state)
    eb = Ensemble(runs[0], nstreams=4, graph_steps=2, batch=True)
    eu = Ensemble(runs[1], nstreams=4, graph_steps=2, batch=False)
    sizes = sorted(len(members) for members, _ in eb.groups())
    assert sizes == [1, 2, 3], sizes
    d = []
    for ens in (eb, eu):
        d.append(np.concatenate([ens.steps(5).cpu().numpy(), ens.steps(4).cpu().numpy()],
                                axis=1))
    assert eb.graph is not None
    # every stage part of every group ran as recorded launch batches
    assert eb._lb and all(v is not None and v.launches > 0 for v in eb._lb.values())
    assert np.array_equal(d[0], d[1])
    for k, (mb, mu, ms) in enumerate(zip(*runs)):
        dts = np.concatenate([ms.integrator.device_steps(5).cpu().numpy(),
                              ms.integrator.device_steps(4).cpu().numpy()])
        assert np.array_equal(dts, d[0][k]), k
        names = NAMES[ms.param.model]
        a, b, c = _state(mb, names), _state(mu, names), _state(ms, names)
        wet = ms.mesh.msk == 1
        for name in a:
            for other in (b, c):
                assert np.array_equal(np.nan_to_num(a[name]), np.nan_to_num(other[name])), \
                    (k, name)
            assert np.isfinite(a[name][wet]).all(), (k, name)


def test_ensemble_stream_tuning(dev):
    from fluids2d_amd.ensemble import Ensemble
    ms = _members(SPECS[:4], "fp32")
    for m in ms:
        m.integrator.begin(m.state)
    ens = Ensemble(ms, nstreams=2, graph_steps=2)
    times = ens.tune_streams(candidates=(1, 2, 4), reps=1)
    assert set(times) == {1, 2, 4} and ens.nstreams in times
    ens.steps(4)
    for m in ms:
        s = m.integrator.dstate
        v = s.h if m.param.model == "qgrsw" else s.pv
        assert bool(torch.isfinite(v[m.mesh.d_msk == 1]).all())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_transport_tiles_bitexact(dev, precision):
    """Every LDS tile of the fused transport stage gives the same state;
    the tuner picks one of them and leaves it on the mesh."""
    from fluids2d_amd import autotune
    import fluids2d_amd._lib as L
    ref = None
    for tile in sorted(L.TILES):
        m = bench.ensemble_member(96, 11, precision, model="eulerpsi")
        m.mesh.grid.tile = tile
        m.integrator.begin(m.state)
        m.integrator.device_steps(6)
        w = _state(m, ["omega"])["omega"]
        if ref is None:
            ref = w
        assert np.array_equal(w, ref), tile
    m = bench.ensemble_member(96, 11, precision, model="eulerpsi")
    m.integrator.begin(m.state)
    t = autotune.tune_tile(m, reps=3, cache=False)
    assert t in L.TILES and m.mesh.grid.tile == t


def test_autotune_poisson_converges_and_caches(dev, tmp_path, monkeypatch):
    """Multigrid schedules (solver "mg": the basin's operator is otherwise
    solved directly, which tune_poisson reports as {})."""
    from fluids2d_amd import autotune
    from fluids2d_amd.elliptic import Poisson2D
    monkeypatch.setenv("F2D_TUNE_CACHE", str(tmp_path / "tune.json"))
    autotune._memory.clear()
    m = bench.ensemble_member(128, 21, "fp32")
    md = m.mesh.qg_helmholtz.maindiag
    assert m.mesh.qg_helmholtz.method == "capacitance"
    assert autotune.tune_poisson(m.mesh, maindiag=md, cache=False) == {}
    m.param.poisson_solver = "mg"
    opts = autotune.tune_poisson(m.mesh, maindiag=md, reps=2, verbose=True)
    assert opts and set(opts) <= set(autotune.mg_candidates("fp32")[0]) | {"refine_inner"}
    assert (tmp_path / "tune.json").exists()
    autotune._memory.clear()
    assert autotune.tune_poisson(m.mesh, maindiag=md) == opts  # from the file
    ps = Poisson2D(m.mesh, "v", maindiag=md, tuning=opts)
    rng = np.random.default_rng(0)
    b = torch.from_numpy(rng.standard_normal(m.mesh.shape) * ps.msk).to(dev, torch.float32)
    x = torch.zeros_like(b)
    ps.solve(b, x)
    assert ps.last_relres <= ps.rtol
    autotune._memory.clear()
