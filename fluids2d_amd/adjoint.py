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
  the adjoint walks the stages backwards, with R'(s)^T from
  `f2d_adj_transport` (WENO-Z5 derivative, face gather), `f2d_adj_perpgrad`,
  the Poisson solve again (A^-1 is self-adjoint) and `f2d_adj_c2v`;
* checkpoints: the forward run keeps the state every `checkpoint` steps; the
  reverse sweep recomputes each segment once, keeping its stage states and
  velocities, and reverses it.

`EulerGradient` wraps it as a torch.autograd.Function.
"""
import torch

from . import _lib
from . import operators as op
from .meshes import Mesh


class EulerAdjoint:
    """Forward run and reverse sweep of `nsteps` RK3 steps of size `dt`."""

    def __init__(self, param, dt, nsteps, checkpoint=20):
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
        self.checkpoint = max(1, int(checkpoint))
        self.dtype = torch.float64 if p.precision == "fp64" else torch.float32
        self.dcode = _lib.dtype_code(self.dtype)
        self.shape = m.shape
        dev = m.device
        z = lambda: torch.zeros(self.shape, dtype=self.dtype, device=dev)  # noqa: E731
        self._z = z
        self.d_dt = torch.full((1,), self.dt, dtype=torch.float64, device=dev)
        self.psi = z()
        self.vtmp = z()
        self.lUx, self.lUy, self.lpsi, self.lb = z(), z(), z(), z()
        self.work = torch.zeros((12,) + self.shape, dtype=self.dtype, device=dev)
        self.ds = [z(), z(), z()]
        self.pois = m.poisson_vertices
        self.checkpoints = []

    # -- forward ------------------------------------------------------------------
    def _U(self, s, U):
        """U = perpgrad(A^-1 c2v(s)), contravariant, halos filled (the
        eulerpsi diag, equations.py:80-84)."""
        self.pois.solve_c2v(s, self.psi, stats=False)
        op.perpgrad(self.mesh, self.psi, U, contravariant=True, fill=True)

    def _vec(self):
        return op_vector(self._z(), self._z())

    def _step(self, s, U, keep=None):
        """One RK3 step from (s, U = U(s)); returns (s3, U3).  `keep` (a list)
        receives (s0, s1, s2, U0, U1, U2)."""
        m, p = self.mesh, self.param
        states, vels = [s], [U]
        for k in range(3):
            nxt = self._z()
            prev = self.ds[:k]
            op.transport_stage(m, p.compflux, states[k], nxt, vels[k], prev,
                               self.ds[k] if k < 2 else None, "rk3", k, self.d_dt)
            Un = self._vec()
            self._U(nxt, Un)
            states.append(nxt)
            vels.append(Un)
        if keep is not None:
            keep.append((states[0], states[1], states[2], vels[0], vels[1], vels[2]))
        return states[3], vels[3]

    def forward(self, omega0):
        """omega0: padded device tensor (halo ignored); returns omega_T."""
        s = omega0.detach().to(self.dtype).clone()
        op.fill(self.mesh, s)
        U = self._vec()
        self._U(s, U)
        self.checkpoints = []
        for n in range(self.nsteps):
            if n % self.checkpoint == 0:
                self.checkpoints.append(s.clone())
            s, U = self._step(s, U)
        return s

    def cfl_dt(self, omega0, cfl):
        """The CFL step of model.py:84-87 for the initial state (a fixed dt
        for the whole run keeps the map differentiable)."""
        s = omega0.detach().to(self.dtype).clone()
        op.fill(self.mesh, s)
        U = self._vec()
        self._U(s, U)
        mx = float(U.x.abs().max()) + float(U.y.abs().max()) + 1e-99
        return cfl / mx

    # -- reverse ------------------------------------------------------------------
    def _Rt(self, s, U, lam):
        """R'(s)^T lam (interior of a padded array)."""
        m = self.mesh
        g = m.grid_ref()
        st = _lib.stream_ptr()
        op.fill(m, lam)
        lq = self._z()
        _lib.call("f2d_adj_transport", self.dcode, g, _lib.ptr(s), _lib.ptr(U.x),
                  _lib.ptr(U.y), _lib.ptr(lam), _lib.ptr(self.lUx), _lib.ptr(self.lUy),
                  _lib.ptr(self.work), _lib.ptr(lq), st)
        op.fill(m, self.lUx)
        op.fill(m, self.lUy)
        _lib.call("f2d_adj_perpgrad", self.dcode, g, _lib.ptr(self.lUx), _lib.ptr(self.lUy),
                  _lib.ptr(self.lpsi), st)
        # A^-1 is self-adjoint (the zero-mean pseudo-inverse when singular)
        _lib.call("f2d_poisson_solve", self.pois.handle(), _lib.ptr(self.lpsi),
                  _lib.ptr(self.lb), self.pois.rtol, self.pois.maxit, None, None, st)
        _lib.call("f2d_adj_c2v", self.dcode, g, _lib.ptr(self.lb), _lib.ptr(lq), st)
        return lq

    def _step_back(self, lam3, saved):
        s0, s1, s2, U0, U1, U2 = saved
        dt = self.dt
        c1 = (dt,)
        c2 = ((-3 * dt) / 4, dt / 4)
        c3 = ((-dt) / 12, (-dt) / 12, (2 * dt) / 3)
        l_s2 = lam3.clone()
        l_ds3 = c3[2] * lam3
        l_ds2 = c3[1] * lam3
        l_ds1 = c3[0] * lam3
        l_s2 += self._Rt(s2, U2, l_ds3)
        l_s1 = l_s2.clone()
        l_ds1 += c2[0] * l_s2
        l_ds2 += c2[1] * l_s2
        l_s1 += self._Rt(s1, U1, l_ds2)
        l_s0 = l_s1.clone()
        l_ds1 += c1[0] * l_s1
        l_s0 += self._Rt(s0, U0, l_ds1)
        return l_s0

    def backward(self, lamT):
        """Adjoint of the forward map applied to lamT (padded; its interior
        is the derivative of the loss with respect to omega_T)."""
        lam = lamT.detach().to(self.dtype).clone()
        interior = torch.zeros_like(lam)
        interior[3:-3, 3:-3] = 1
        lam *= interior
        K = self.checkpoint
        for c in range(len(self.checkpoints) - 1, -1, -1):
            n0 = c * K
            n1 = min(n0 + K, self.nsteps)
            s = self.checkpoints[c].clone()
            U = self._vec()
            self._U(s, U)
            saved = []
            for _ in range(n0, n1):
                s, U = self._step(s, U, keep=saved)
            for rec in reversed(saved):
                lam = self._step_back(lam, rec)
                lam *= interior
        return lam


def op_vector(x, y):
    from .states import Vector
    return Vector(x=x, y=y)


class EulerGradient(torch.autograd.Function):
    """omega_T = EulerGradient.apply(omega0, runner): the forward run of an
    EulerAdjoint; its backward is the hand-written reverse sweep."""

    @staticmethod
    def forward(ctx, omega0, runner):
        ctx.runner = runner
        return runner.forward(omega0)

    @staticmethod
    def backward(ctx, grad_out):
        return ctx.runner.backward(grad_out), None
