"""Gradient of a final-time vorticity misfit with respect to the initial
vorticity, for the eulerpsi model on a fully wet, doubly periodic grid with
SSP-RK3, WENO fluxes and a fixed time step.

The forward map is the library's own time step: per RK stage the fused
transport + update (`f2d_transport_stage`), then U = perpgrad(A^-1 c2v(s))
(Poisson solve + perp-gradient with halo fill).  The reverse sweep is
hand-written:

* per RK3 step (integrators.py:92-107, incremental form
      s1 = s0 + dt ds1,  s2 = s1 - 3dt/4 ds1 + dt/4 ds2,
      s3 = s2 - dt/12 ds1 - dt/12 ds2 + 2dt/3 ds3,  ds_k = R(s_{k-1}))
  the adjoint walks the stages backwards.  Each stage is R'(s)^T applied to
  the RK combination of the stage outputs' adjoints: `f2d_adj_stage`
  (WENO-Z5 derivative, face gather, perpgrad^T), the Poisson solve (A^-1 is
  self-adjoint) and `f2d_adj_combine` (c2v^T, plus the adjoint carried over);
* checkpoints: the state at the start of every `checkpoint`-step segment.
  The forward run writes each segment's stage states and velocities into
  segment storage and keeps the last segment there; the reverse sweep
  recomputes every other segment from its checkpoint, once, before
  reversing it.

Layout for the device: a segment is one (K, 9, ny+6, nx+6) array (per step
s0, s1, s2 and the velocities U0, U1, U2).  Two segment stores alternate:
while the reverse sweep reads one on the main stream, the recomputation of
the segment before it writes the other on a second stream, with its own
Poisson solver and scratch (the two streams share no buffer).  Each
segment's forward, recomputation and reversal is captured once as a HIP
graph (per store and length) and replayed.

`EulerGradient` wraps it as a torch.autograd.Function.
"""
import torch

from . import _lib
from . import operators as op
from .elliptic import Poisson2D
from .meshes import Mesh
from .states import Vector


class _Lane:
    """A stream with its own Poisson solver and stage scratch: forward
    segments (main stream) or recomputed ones (second stream)."""

    def __init__(self, mesh, dtype, stream):
        self.stream = stream
        self.pois = Poisson2D(mesh, "v")
        z = lambda: torch.zeros(mesh.shape, dtype=dtype, device=mesh.device)  # noqa: E731
        self.psi = z()
        self.ds = [z(), z()]
        self.s_end = z()
        self.U_end = Vector(x=z(), y=z())
        self.graphs = {}


class _Segment:
    """Stage states and velocities of up to K steps."""

    def __init__(self, K, mesh, dtype):
        self.a = torch.zeros((K, 9) + mesh.shape, dtype=dtype, device=mesh.device)

    def s(self, n, k):
        return self.a[n, k]

    def U(self, n, k):
        return Vector(x=self.a[n, 3 + 2 * k], y=self.a[n, 4 + 2 * k])


class EulerAdjoint:
    """Forward run and reverse sweep of `nsteps` RK3 steps of size `dt`."""

    def __init__(self, param, dt, nsteps, checkpoint=20, graphs=None):
        p = param
        if p.model != "eulerpsi" or p.integrator != "rk3" or p.compflux != "weno":
            raise NotImplementedError("adjoint: eulerpsi with rk3 and weno fluxes only")
        if not (p.xperiodic and getattr(p, "yperiodic", False)):
            raise NotImplementedError("adjoint: doubly periodic grids only")
        if p.maxorder < 5:
            raise NotImplementedError("adjoint: 5-point WENO faces only (maxorder >= 5)")
        self.param = p
        self.mesh = Mesh(p)
        m = self.mesh
        if not bool((m.d_msk == 1).all()):
            raise NotImplementedError("adjoint: fully wet grids only")
        self.dt = float(dt)
        self.nsteps = int(nsteps)
        self.K = max(1, min(int(checkpoint), self.nsteps))
        self.checkpoint = self.K
        self.nseg = (self.nsteps + self.K - 1) // self.K
        self.graphs = getattr(p, "hipgraph", True) if graphs is None else graphs
        self.dtype = torch.float64 if p.precision == "fp64" else torch.float32
        self.dcode = _lib.dtype_code(self.dtype)
        self.shape = m.shape
        dev = m.device
        z = lambda: torch.zeros(self.shape, dtype=self.dtype, device=dev)  # noqa: E731
        self.d_dt = torch.full((1,), self.dt, dtype=torch.float64, device=dev)
        main = torch.cuda.current_stream()
        self.lanes = [_Lane(m, self.dtype, main)]
        if self.nseg > 1:
            self.lanes.append(_Lane(m, self.dtype, torch.cuda.Stream()))
        self.segs = [_Segment(self.K, m, self.dtype)]
        if self.nseg > 1:
            self.segs.append(_Segment(self.K, m, self.dtype))
        self.ckpt = torch.zeros((self.nseg,) + self.shape, dtype=self.dtype, device=dev)
        # reverse sweep: the adjoint of the step's input (in) and of the
        # stage outputs s2 and s1 (t1, t2); adjoint scratch
        self.lam, self.t1, self.t2 = z(), z(), z()
        self.lUx, self.lUy, self.lq, self.lpsi, self.lb = z(), z(), z(), z(), z()
        self.work = torch.zeros((12,) + self.shape, dtype=self.dtype, device=dev)
        self.rgraphs = {}
        self._warm = False
        # bumped by every forward: the reverse sweep reads the checkpoints
        # and the kept segment of the latest forward only
        self.generation = 0
        dt_ = self.dt
        # integrators.py:98-106, in Python's evaluation order
        self.c1 = (dt_,)
        self.c2 = ((-3 * dt_) / 4, dt_ / 4)
        self.c3 = ((-dt_) / 12, (-dt_) / 12, (2 * dt_) / 3)

    # -- forward ------------------------------------------------------------------
    def _U(self, lane, s, U):
        """U = perpgrad(A^-1 c2v(s)), contravariant, halos filled (the
        eulerpsi diag, equations.py:80-84)."""
        lane.pois.solve_c2v(s, lane.psi, stats=False)
        op.perpgrad(self.mesh, lane.psi, U, contravariant=True, fill=True)

    def _segment_ops(self, lane, seg, L):
        """U0 of the segment's first state, then L RK3 steps; the stage
        states and velocities go to `seg`, the final state to lane.s_end."""
        m, p = self.mesh, self.param
        self._U(lane, seg.s(0, 0), seg.U(0, 0))
        for n in range(L):
            for k in range(3):
                last = k == 2
                q_in, U = seg.s(n, k), seg.U(n, k)
                if not last:
                    q_out, U_out = seg.s(n, k + 1), seg.U(n, k + 1)
                elif n + 1 < L:
                    q_out, U_out = seg.s(n + 1, 0), seg.U(n + 1, 0)
                else:
                    q_out, U_out = lane.s_end, lane.U_end
                op.transport_stage(m, p.compflux, q_in, q_out, U, lane.ds[:k],
                                   lane.ds[k] if k < 2 else None, "rk3", k, self.d_dt)
                self._U(lane, q_out, U_out)

    def _warmup(self):
        """One eager solve per lane: the solvers' first (host-checked) call
        and lazy set-up must not happen under graph capture."""
        if self._warm:
            return
        gen = torch.Generator(device=self.mesh.device).manual_seed(0)
        for lane in self.lanes:
            w = torch.rand(self.shape, generator=gen, device=self.mesh.device,
                           dtype=self.dtype)
            op.fill(self.mesh, w)
            lane.stream.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(lane.stream):
                lane.pois.solve_c2v(w, lane.psi, stats=True)
            lane.stream.synchronize()
        self._warm = True

    def _run_segment(self, lane, seg, L):
        if not self.graphs:
            self._segment_ops(lane, seg, L)
            return
        key = (id(seg), L)
        g = lane.graphs.get(key)
        if g is None:
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with _lib.capture(g):
                self._segment_ops(lane, seg, L)
            lane.graphs[key] = g
        g.replay()

    def _length(self, c):
        return min(self.K, self.nsteps - c * self.K)

    def forward(self, omega0):
        """omega0: padded device tensor (halo ignored); returns omega_T."""
        self._warmup()
        self.generation += 1
        lane, seg = self.lanes[0], self.segs[0]
        s0 = seg.s(0, 0)
        s0.copy_(omega0.detach().to(self.dtype))
        op.fill(self.mesh, s0)
        for c in range(self.nseg):
            self.ckpt[c].copy_(s0)
            self._run_segment(lane, seg, self._length(c))
            if c + 1 < self.nseg:
                s0.copy_(lane.s_end)
        self._last = seg
        return lane.s_end.clone()

    def cfl_dt(self, omega0, cfl):
        """The CFL step of model.py:84-87 for the initial state (a fixed dt
        for the whole run keeps the map differentiable)."""
        self._warmup()
        lane = self.lanes[0]
        s = omega0.detach().to(self.dtype).clone()
        op.fill(self.mesh, s)
        U = lane.U_end
        self._U(lane, s, U)
        mx = float(U.x.abs().max()) + float(U.y.abs().max()) + 1e-99
        return cfl / mx

    # -- reverse ------------------------------------------------------------------
    def _stage_back(self, s, U, lams, coefs, base, out):
        """out = base + R'(s)^T (sum coefs * lams)."""
        m = self.mesh
        g = m.grid_ref()
        st = _lib.stream_ptr()
        l = list(lams) + [None] * (3 - len(lams))
        c = list(coefs) + [0.0] * (3 - len(coefs))
        _lib.call("f2d_adj_stage", self.dcode, g, _lib.ptr(s), _lib.ptr(U.x), _lib.ptr(U.y),
                  _lib.ptr(l[0]), _lib.ptr(l[1]), _lib.ptr(l[2]), c[0], c[1], c[2],
                  _lib.ptr(self.lUx), _lib.ptr(self.lUy), _lib.ptr(self.work),
                  _lib.ptr(self.lq), _lib.ptr(self.lpsi), st)
        # A^-1 is self-adjoint (the zero-mean pseudo-inverse when singular)
        pois = self.lanes[0].pois
        _lib.call("f2d_poisson_solve", pois.handle(), _lib.ptr(self.lpsi),
                  _lib.ptr(self.lb), pois.rtol, pois.maxit, None, None, st)
        _lib.call("f2d_adj_combine", self.dcode, g, _lib.ptr(base), _lib.ptr(self.lq),
                  _lib.ptr(self.lb), _lib.ptr(out), st)

    def _reverse_ops(self, seg, L):
        """The adjoint in self.lam is carried back over the L steps of seg."""
        c1, c2, c3 = self.c1, self.c2, self.c3
        lam, t1, t2 = self.lam, self.t1, self.t2
        for n in range(L - 1, -1, -1):
            # stage 3: ds3 = R(s2), s3 = s2 + ... + c3[2] ds3
            self._stage_back(seg.s(n, 2), seg.U(n, 2), [lam], [c3[2]], lam, t1)
            # stage 2: ds2 = R(s1) enters s2 (c2[1]) and s3 (c3[1])
            self._stage_back(seg.s(n, 1), seg.U(n, 1), [lam, t1], [c3[1], c2[1]], t1, t2)
            # stage 1: ds1 = R(s0) enters s1 (c1[0]), s2 (c2[0]), s3 (c3[0])
            self._stage_back(seg.s(n, 0), seg.U(n, 0), [lam, t1, t2],
                             [c3[0], c2[0], c1[0]], t2, lam)

    def _reverse_segment(self, seg, L):
        if not self.graphs:
            self._reverse_ops(seg, L)
            return
        key = (id(seg), L)
        g = self.rgraphs.get(key)
        if g is None:
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with _lib.capture(g):
                self._reverse_ops(seg, L)
            self.rgraphs[key] = g
        g.replay()

    def backward(self, lamT):
        """Adjoint of the forward map applied to lamT (padded; its interior
        is the derivative of the loss with respect to omega_T)."""
        main = torch.cuda.current_stream()
        self.lam.zero_()
        self.lam[3:-3, 3:-3] = lamT.detach()[3:-3, 3:-3].to(self.dtype)
        cur = self._last
        other = self.segs[1] if len(self.segs) > 1 else None
        if other is not None and other is cur:
            other = self.segs[0]
        side = self.lanes[1] if len(self.lanes) > 1 else None
        if side is not None:
            side.stream.wait_stream(main)
        for c in range(self.nseg - 1, -1, -1):
            done = None
            if c > 0:
                # recompute segment c-1 on the second stream meanwhile
                with torch.cuda.stream(side.stream):
                    other.s(0, 0).copy_(self.ckpt[c - 1])
                    self._run_segment(side, other, self._length(c - 1))
                    done = torch.cuda.Event()
                    done.record(side.stream)
            self._reverse_segment(cur, self._length(c))
            if c > 0:
                main.wait_event(done)
                read = torch.cuda.Event()
                read.record(main)
                side.stream.wait_event(read)  # cur is overwritten next
                cur, other = other, cur
        out = torch.zeros_like(self.lam)
        out[3:-3, 3:-3] = self.lam[3:-3, 3:-3]
        return out


class StaleTrajectory(RuntimeError):
    """The runner ran another forward after the one a backward belongs to."""


class EulerGradient(torch.autograd.Function):
    """omega_T = EulerGradient.apply(omega0, runner): the forward run of an
    EulerAdjoint; its backward is the hand-written reverse sweep.

    A runner holds the checkpoints of one trajectory, its latest forward.
    The backward of an earlier apply on the same runner (another apply, or
    a forward, came in between) raises StaleTrajectory instead of returning
    the latest trajectory's gradient; use one runner per live graph."""

    @staticmethod
    def forward(ctx, omega0, runner):
        out = runner.forward(omega0)
        ctx.runner = runner
        ctx.generation = runner.generation
        return out

    @staticmethod
    def backward(ctx, grad_out):
        if ctx.runner.generation != ctx.generation:
            raise StaleTrajectory(
                f"the adjoint runner's checkpoints are of forward {ctx.runner.generation}; "
                f"this backward belongs to forward {ctx.generation}")
        return ctx.runner.backward(grad_out), None
