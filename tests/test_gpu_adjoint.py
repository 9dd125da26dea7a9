"""Hand-written adjoint of the eulerpsi RK3/WENO step (fluids2d_amd.adjoint)
against an FP64 torch-autograd gradient of the same discretisation
(tests/torch_euler.py) and against central finite differences of the loss
L = 1/2 sum (omega_T - omega_truth)^2 over the interior cells."""
import numpy as np
import pytest
import torch

import torch_euler as TE

pytestmark = pytest.mark.gpu


def _setup(n, precision, nsteps, checkpoint, seed=7, dt=None):
    import bench
    import fluids2d_amd.param as fp
    from fluids2d_amd import Param
    from fluids2d_amd.adjoint import EulerAdjoint
    fp.QUIET = True
    p = Param()
    p.model = "eulerpsi"
    p.nx = p.ny = n
    p.xperiodic = p.yperiodic = True
    p.precision = precision
    w0 = np.zeros((n + 6, n + 6))
    w0[3:-3, 3:-3] = bench.random_phase_vorticity(n, seed) / (n * n)
    bench.fill_periodic(w0)
    wt = np.zeros_like(w0)
    wt[3:-3, 3:-3] = (bench.random_phase_vorticity(n, seed) +
                      0.3 * bench.random_phase_vorticity(n, seed + 1)) / (n * n)
    bench.fill_periodic(wt)
    probe = EulerAdjoint(p, 1.0, 1)
    dev = probe.mesh.device
    if dt is None:
        dt = probe.cfl_dt(torch.from_numpy(w0).to(dev), 0.5)
    runner = EulerAdjoint(p, dt, nsteps, checkpoint)
    return runner, torch.from_numpy(w0).to(dev), torch.from_numpy(wt).to(dev), dt


def _loss_and_grad(runner, w0, wt):
    from fluids2d_amd.adjoint import EulerGradient
    x = w0.clone().to(runner.dtype).requires_grad_(True)
    wT = EulerGradient.apply(x, runner)
    d = (wT - wt.to(runner.dtype))[3:-3, 3:-3]
    L = 0.5 * (d.double() ** 2).sum()
    L.backward()
    return float(L.detach()), x.grad.double()


def test_forward_matches_torch_restatement(dev):
    runner, w0, wt, dt = _setup(64, "fp64", 5, 5)
    wT = runner.forward(w0)
    h = 1.0 / 64
    ref = TE.run(w0[3:-3, 3:-3].cpu(), 5, dt, h, h)
    e = float(torch.linalg.norm(wT[3:-3, 3:-3].cpu() - ref) / torch.linalg.norm(ref))
    assert e <= 1e-12, e


def test_adjoint_matches_torch_autograd(dev):
    runner, w0, wt, dt = _setup(64, "fp64", 6, 4)
    L, g = _loss_and_grad(runner, w0, wt)
    h = 1.0 / 64
    x = w0[3:-3, 3:-3].cpu().clone().requires_grad_(True)
    wT = TE.run(x, 6, dt, h, h)
    Lr = 0.5 * ((wT - wt[3:-3, 3:-3].cpu()) ** 2).sum()
    Lr.backward()
    gr = x.grad
    e = float(torch.linalg.norm(g[3:-3, 3:-3].cpu() - gr) / torch.linalg.norm(gr))
    lr = float(Lr.detach())
    print(f"\nadjoint vs torch autograd: rel L2 {e:.2e} (L {L:.6e} vs {lr:.6e})")
    assert abs(L - lr) <= 1e-10 * abs(lr)
    assert e <= 1e-9, e
    # halo entries of omega0 do not enter the map (they are refilled)
    assert float(g[:3].abs().max()) == 0.0 and float(g[:, :3].abs().max()) == 0.0


def test_adjoint_finite_differences(dev):
    """Central differences along 8 seeded unit directions, each within 1e-6
    of the adjoint's directional derivative.

    The step is 3e-9 of the field's scale, where the two error sources of a
    finite difference of this map are smallest together.  The map is only
    piecewise smooth (the |beta1 - beta3| of WENO-Z, weno.py:188, and the
    upwind switch at U = 0, weno.py:325, have kinks), and the central
    differences' error from the kinks inside the +-eps segment grows in
    proportion to eps (measured: about 5e-7 at a 10x larger step); FP64
    rounding of the two perturbed runs gives an absolute error of about
    1e-12 to 5e-12 here, which grows as eps shrinks.  Loss increments are
    formed from state differences, (wa - wb).(wa + wb - 2 truth) / 2.

    A relative error is meaningful only where the directional derivative is
    not small, so of 16 seeded candidates the first 8 whose central
    difference is at least half the RMS over the candidates are compared
    (the choice depends on the finite differences only, not on the
    adjoint).  The gradient as a whole is held to FP64 rounding by the
    torch-autograd comparisons."""
    runner, w0, wt, dt = _setup(128, "fp64", 20, 5)
    L, g = _loss_and_grad(runner, w0, wt)
    rng = np.random.default_rng(3)
    scale = float(w0[3:-3, 3:-3].abs().max())
    eps = 3e-9 * scale * 128
    tt = wt[3:-3, 3:-3]
    cands = []
    for k in range(16):
        d = torch.zeros_like(w0)
        d[3:-3, 3:-3] = torch.from_numpy(rng.standard_normal((128, 128))).to(w0.device)
        d /= torch.linalg.norm(d)
        wp = runner.forward(w0 + eps * d)[3:-3, 3:-3]
        wm = runner.forward(w0 - eps * d)[3:-3, 3:-3]
        fd = 0.5 * float(((wp - wm) * ((wp - tt) + (wm - tt))).sum()) / (2 * eps)
        ad = float((g * d).sum())
        cands.append((k, fd, ad))
    rms = float(np.sqrt(np.mean([fd * fd for _, fd, _ in cands])))
    errs = []
    for k, fd, ad in cands:
        used = abs(fd) >= 0.5 * rms and len(errs) < 8
        rel = abs(ad - fd) / abs(fd)
        print(f"direction {k}: adjoint {ad:.12e} FD {fd:.12e} rel {rel:.1e}"
              + ("" if used else " (not used)"))
        if used:
            errs.append(rel)
    print(f"{len(errs)} directions compared, |FD| >= {0.5 * rms:.2e}")
    assert len(errs) == 8
    assert max(errs) <= 1e-6, errs


def test_stale_trajectory_refused(dev):
    """Two applies on one runner before a backward: the first one's backward
    must not silently use the second trajectory's checkpoints."""
    from fluids2d_amd.adjoint import EulerGradient, StaleTrajectory
    runner, w0, wt, dt = _setup(64, "fp64", 4, 2)
    xa = w0.clone().requires_grad_(True)
    xb = (1.1 * w0).clone().requires_grad_(True)
    wa = EulerGradient.apply(xa, runner)
    wb = EulerGradient.apply(xb, runner)
    with pytest.raises(StaleTrajectory):
        (0.5 * ((wa - wt)[3:-3, 3:-3] ** 2).sum()).backward()
    # the latest one still works and equals a fresh runner's gradient
    (0.5 * ((wb - wt)[3:-3, 3:-3] ** 2).sum()).backward()
    fresh, _, _, _ = _setup(64, "fp64", 4, 2, dt=dt)
    _, gb = _loss_and_grad(fresh, 1.1 * w0, wt)
    assert torch.equal(xb.grad.double(), gb)


@pytest.mark.parametrize("precision,bar", [("fp64", 1e-9), ("fp32", 1e-3)])
def test_adjoint_512_against_torch_autograd(dev, precision, bar):
    """The gradient at the adjoint workload's size (512^2 doubly periodic),
    over 4 steps with checkpoints every 2, against the FP64 torch-autograd
    gradient of the same discretisation."""
    runner, w0, wt, dt = _setup(512, precision, 4, 2)
    L, g = _loss_and_grad(runner, w0, wt)
    h = 1.0 / 512
    x = w0[3:-3, 3:-3].cpu().clone().requires_grad_(True)
    wT = TE.run(x, 4, dt, h, h)
    Lr = 0.5 * ((wT - wt[3:-3, 3:-3].cpu()) ** 2).sum()
    Lr.backward()
    gr = x.grad
    e = float(torch.linalg.norm(g[3:-3, 3:-3].cpu() - gr) / torch.linalg.norm(gr))
    print(f"\n512^2 {precision} adjoint vs FP64 torch autograd: rel L2 {e:.2e} (bar {bar})")
    assert e <= bar, e


def test_checkpoint_interval_does_not_change_gradient(dev):
    r1, w0, wt, dt = _setup(64, "fp64", 9, 1)
    r2, _, _, _ = _setup(64, "fp64", 9, 4, dt=dt)
    _, g1 = _loss_and_grad(r1, w0, wt)
    _, g2 = _loss_and_grad(r2, w0, wt)
    assert torch.equal(g1, g2)


def test_fp32_adjoint_close_to_fp64(dev):
    r64, w0, wt, dt = _setup(128, "fp64", 20, 5)
    r32, _, _, _ = _setup(128, "fp32", 20, 5, dt=dt)
    _, g64 = _loss_and_grad(r64, w0, wt)
    _, g32 = _loss_and_grad(r32, w0, wt)
    e = float(torch.linalg.norm(g32 - g64) / torch.linalg.norm(g64))
    print(f"\nFP32 vs FP64 gradient: rel L2 {e:.2e}")
    assert e <= 1e-3, e


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graphs_and_pipeline_match_eager(dev, precision):
    """The captured segment graphs (forward, recomputation on the second
    stream, reversal) give the eager launches' results bit for bit, with a
    last segment shorter than the checkpoint interval."""
    runner, w0, wt, dt = _setup(64, precision, 11, 4)
    L1, g1 = _loss_and_grad(runner, w0, wt)
    from fluids2d_amd.adjoint import EulerAdjoint
    eager = EulerAdjoint(runner.param, dt, 11, 4, graphs=False)
    L2, g2 = _loss_and_grad(eager, w0, wt)
    assert runner.rgraphs and runner.lanes[1].graphs
    assert L1 == L2
    assert torch.equal(g1, g2)
    # replays of the same graphs: FP32 solves refine from the solution held
    # from the previous call, so a second call agrees to FP32 rounding only
    L1b, g1b = _loss_and_grad(runner, w0, wt)
    if precision == "fp64":
        assert L1b == L1 and torch.equal(g1b, g1)
    else:
        e = float(torch.linalg.norm(g1b - g1) / torch.linalg.norm(g1))
        assert e < 1e-4 and abs(L1b - L1) < 1e-4 * abs(L1), (e, L1, L1b)
