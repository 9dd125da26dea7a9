"""Device models: rhs/diag of each model on the GPU (equations.py:73-238).

Each model class offers
  * ``rhs(s, ds)`` / ``diag(s)`` -- the reference's closures, operator by
    operator, on device states (used when user code calls them);
  * ``stage(...)`` -- one fused RK stage used by the integrator's hot loop:
    the tendency, the halo fill and the `addto` update in one kernel, then the
    diagnostic fields (Poisson solve with the right-hand side formed inside
    the solver, perp-gradient + fill + next CFL step in one kernel).
"""
from . import _lib
from . import operators as op
from .elliptic import PoissonNotConverged  # noqa: F401  (re-export)


class _Base:
    prognostic = ()
    vector_prognostic = ()

    def __init__(self, param, mesh):
        self.param = param
        self.mesh = mesh

    def materialize(self, s, names=None):
        """Fill the diagnostic fields the fused path does not store, before a
        device -> host copy or a device read (names: only these are needed;
        None: all)."""

    def cfl_velocity(self, s):
        return s.U


class EulerPsi(_Base):
    """Euler in vorticity-streamfunction form (equations.py:269-282)."""
    logs_dt = True  # stage(..., dt_log=) records the step's dt
    # stage(..., cfl_out=) leaves the next step's CFL step as block maxima,
    # stage(..., cfl_in=) finishes it inside the first transport
    defers_cfl = True
    prognostic = ("omega",)

    def rhs(self, s, ds):
        m, p = self.mesh, self.param
        op.compflux(s.flx.x, s.U.x, s.omega, m.d_oc.x, m.xshift, p.compflux)
        op.compflux(s.flx.y, s.U.y, s.omega, m.d_oc.y, m.yshift, p.compflux)
        op.transport_rhs(m, p.compflux, s.omega, s.U, ds.omega)

    def diag(self, s):
        m = self.mesh
        op.centerstovertices(m, s.omega, s.vomega)
        m.poisson_vertices.solve(s.vomega, s.psi)
        op.perpgrad(m, s.psi, s.U, contravariant=True, fill=True)

    def diag_fused(self, s, dt_next=None, stats=True):
        m, p = self.mesh, self.param
        m.poisson_vertices.solve_c2v(s.omega, s.psi, stats=stats)
        op.perpgrad(m, s.psi, s.U, contravariant=True, fill=True, dt_dev=dt_next,
                    cfl=p.cfl, dtmax=p.dtmax)

    def stage(self, s, alt, ds_prev, ds_out, k, dt, dt_next, stats=True, dt_log=None,
              cfl_in=None, cfl_out=None):
        """One RK stage: transport + update, then the diag.  U is stored
        only by the step's last stage (it feeds the CFL step and the state):
        a later stage forms U = perpgrad(psi) inside its transport kernel,
        bit for bit the same velocity, and the earlier stages skip the
        perpgrad pass.  dt_log (stage 0): the transport kernel also copies
        the step's dt there.  cfl_out (last stage, instead of dt_next): the
        CFL step's block maxima are left in that array; cfl_in (stage 0):
        the transport forms this step's dt from such maxima and stores it in
        dt (the same value, without a launch of its own)."""
        m, p = self.mesh, self.param
        if k > 0:
            op.transport_stage_psi(m, p.compflux, s.omega, alt.omega, s.psi, True, True,
                                   ds_prev, ds_out, p.integrator, k, dt)
        else:
            op.transport_stage(m, p.compflux, s.omega, alt.omega, s.U, ds_prev, ds_out,
                               p.integrator, k, dt, dt_log, cfl_part=cfl_in, cfl=p.cfl,
                               dtmax=p.dtmax)
        s, alt = s._replace(omega=alt.omega), alt._replace(omega=s.omega)
        last = k == _lib.NSTAGES[p.integrator] - 1
        if p.precision != "fp64" and getattr(p, "poisson_refine", "every") == "last" and \
                m.poisson_vertices.method == "fft":
            # plain FP32 solves inside the step; the last one refines in FP64
            # from the previous stage's solution.  (A host-checked solve
            # refines too: the FP64 residual of a plain FP32 solution is
            # that of x rounded to FP32, about 1e-3 at 1024^2.)  The masked
            # (capacitance) solve refines at every stage: its plain FP32
            # solutions cost the vorticity 5e-6 over 500 steps at 128^2,
            # against 6e-7.
            m.poisson_vertices.set_refine(None if (last or stats) else 0)
        m.poisson_vertices.solve_c2v(s.omega, s.psi, stats=stats)
        if last:
            if cfl_out is not None:
                op.perpgrad(m, s.psi, s.U, contravariant=True, fill=True, cfl_part=cfl_out)
            else:
                op.perpgrad(m, s.psi, s.U, contravariant=True, fill=True, dt_dev=dt_next,
                            cfl=p.cfl, dtmax=p.dtmax)
        return s, alt

    def materialize(self, s, names=None):
        if names is None or "vomega" in names:
            op.centerstovertices(self.mesh, s.omega, s.vomega)


class Advection(_Base):
    """Tracer advection by a fixed velocity (equations.py:356-366)."""
    logs_dt = True  # stage(..., dt_log=) records the step's dt
    prognostic = ("q",)

    def rhs(self, s, ds):
        m, p = self.mesh, self.param
        op.compflux(s.flx.x, s.U.x, s.q, m.d_oc.x, m.xshift, p.compflux)
        op.compflux(s.flx.y, s.U.y, s.q, m.d_oc.y, m.yshift, p.compflux)
        op.transport_rhs(m, p.compflux, s.q, s.U, ds.q)

    def diag(self, s):
        pass

    def stage(self, s, alt, ds_prev, ds_out, k, dt, dt_next, stats=True, dt_log=None):
        p = self.param
        op.transport_stage(self.mesh, p.compflux, s.q, alt.q, s.U, ds_prev, ds_out,
                           p.integrator, k, dt, dt_log if k == 0 else None)
        s, alt = s._replace(q=alt.q), alt._replace(q=s.q)
        if dt_next is not None:
            op.cfl_dt(s.U, p.cfl, p.dtmax, dt_next)
        return s, alt


class QG(_Base):
    """Quasi-geostrophic PV transport (equations.py:285-299).

    With beta != 0 the inversion subtracts the planetary vorticity
    ``mesh.f``, a vertex array the caller sets on the mesh, as
    experiments/inpreparation/gyre.py:24-29 does (operators.py:189-190).
    It is copied to the device the first time a solve needs it, and again
    whenever ``mesh.f`` is rebound to another array."""
    logs_dt = True  # stage(..., dt_log=) records the step's dt
    prognostic = ("pv",)

    def __init__(self, param, mesh):
        super().__init__(param, mesh)
        self._f_src = None
        self._f_dev = None

    def _planetary(self, like):
        f = getattr(self.mesh, "f", None)
        if f is None:
            raise AttributeError("qg with beta != 0 needs mesh.f (the planetary "
                                 "vorticity at the vertices, times the cell area)")
        if f is not self._f_src or self._f_dev is None or self._f_dev.dtype != like.dtype:
            import numpy as np
            import torch
            if isinstance(f, torch.Tensor):
                d = f.to(device=like.device, dtype=like.dtype)
            else:
                a = np.broadcast_to(np.asarray(f, dtype=np.float64), tuple(like.shape))
                d = torch.from_numpy(np.array(a)).to(like.device).to(like.dtype)  # a writable copy
            self._f_src, self._f_dev = f, d.contiguous()
        return self._f_dev

    def _inversion_rhs(self, s):
        """work = centerstovertices(pv - hb*qgcoef) - f (operators.py:186-190;
        hb = 0 on the device, so pv - 0 is pv bit for bit)."""
        m = self.mesh
        op.centerstovertices(m, s.pv, s.work)
        if self.param.beta != 0:
            # work + (-1 * f) is work - f exactly
            op.addto_terms(s.work, [self._planetary(s.work)], [-1.0])

    def rhs(self, s, ds):
        m, p = self.mesh, self.param
        op.compflux(s.flx.x, s.U.x, s.pv, m.d_oc.x, m.xshift, p.compflux)
        op.compflux(s.flx.y, s.U.y, s.pv, m.d_oc.y, m.yshift, p.compflux)
        op.transport_rhs(m, p.compflux, s.pv, s.U, ds.pv)

    def diag(self, s):
        # qg_inversion (operators.py:186-191) with hb = 0, then perpgrad
        m = self.mesh
        self._inversion_rhs(s)
        m.qg_helmholtz.solve(s.work, s.psi)
        op.perpgrad(m, s.psi, s.U, contravariant=True)

    def stage(self, s, alt, ds_prev, ds_out, k, dt, dt_next, stats=True, dt_log=None):
        """As EulerPsi.stage: U stored by the last stage only, formed from psi
        inside the later stages' transport.  With beta = 0 the solve forms
        its right-hand side itself (centerstovertices on the fly)."""
        m, p = self.mesh, self.param
        if k > 0:
            op.transport_stage_psi(m, p.compflux, s.pv, alt.pv, s.psi, True, False,
                                   ds_prev, ds_out, p.integrator, k, dt)
        else:
            op.transport_stage(m, p.compflux, s.pv, alt.pv, s.U, ds_prev, ds_out,
                               p.integrator, k, dt, dt_log)
        s, alt = s._replace(pv=alt.pv), alt._replace(pv=s.pv)
        if p.beta != 0:
            self._inversion_rhs(s)
            m.qg_helmholtz.solve(s.work, s.psi, stats=stats)
        else:
            m.qg_helmholtz.solve_c2v(s.pv, s.psi, stats=stats)
        if k == _lib.NSTAGES[p.integrator] - 1:
            op.perpgrad(m, s.psi, s.U, contravariant=True, fill=False, dt_dev=dt_next,
                        cfl=p.cfl, dtmax=p.dtmax)
        return s, alt

    def materialize(self, s, names=None):
        if names is None or "work" in names:
            self._inversion_rhs(s)


MODELS = {"eulerpsi": EulerPsi, "advection": Advection, "qg": QG}


def get_device_model(param, mesh):
    if param.model in MODELS:
        return MODELS[param.model](param, mesh)
    try:
        from . import rsw
        return rsw.get_device_model(param, mesh)
    except (ImportError, KeyError):
        raise NotImplementedError(f"model {param.model!r} has no GPU path") from None
