"""Shallow-water device models: rsw (equations.py:139-157) and qgrsw, the
RSW tendency projected onto the QG manifold (equations.py:106-136).

The RK stage is the reference's: rhs into the stage's tendency, `addto` on
every prognostic array (u.x, u.y, h), then diag.  ``rhs`` and ``diag`` run
operator by operator in the reference's order (the host API and forced
stages use them); ``stage`` runs the same sequence fused (f2d_sw.hip and
the transport stage of f2d_stencil.hip), bit for bit the same results:

* rsw: the momentum kernel (vortex force, Coriolis, both gradients, fill,
  RK update of u in place), the transport stage for h (divflux, fill, RK
  update into the other h buffer), the diag kernel (sharp, vorticity,
  kinetic energy, pressure, fills) -- 3 launches per stage;
* qgrsw: the unprojected, unfilled tendencies (momentum kernel, then
  compflux and div as the reference leaves them), the pv kernel, the
  Helmholtz solve, the projection kernel (v2c, perpgrad, fill, RK update
  of u and h) and the diag kernel.
"""
import ctypes
from collections import namedtuple

import torch

from . import _lib
from . import operators as op
from .equations import _Base
from .states import allocate_device, leaves


def _ops(name, *args):
    _lib.call(name, *args, _lib.stream_ptr())


class _SW(_Base):
    prognostic = ("u", "h")
    fused = True  # stage() through the fused kernels (False: operator by operator)

    @property
    def logs_dt(self):
        """stage(..., dt_log=) records the step's dt (the fused momentum
        kernel copies it)."""
        return self.fused

    def __init__(self, param, mesh):
        super().__init__(param, mesh)
        self._ds_last = None
        self._pre = None
        self._cflpart = None  # block maxima of stage_split's CFL step
        self._stats = True  # host-checked solves unless a stage says otherwise

    def _grid(self):
        return self.mesh.grid_ref()

    def _dt(self, t):
        return _lib.dtype_code(t.dtype)

    def _vortexforce_coriolis(self, s, ds):
        m, p = self.mesh, self.param
        op.vortexforce(ds.u.x, s.U.y, s.omega, m.d_ov.y, m.yshift, m.xshift, +1,
                       p.vortexforce)
        op.vortexforce(ds.u.y, s.U.x, s.omega, m.d_ov.x, m.xshift, m.yshift, -1,
                       p.vortexforce)
        f = p.f0 * m.area * 0.25
        _ops("f2d_coriolis", self._dt(s.h), self._grid(), _lib.ptr(s.U.x), _lib.ptr(s.U.y),
             _lib.ptr(ds.u.x), _lib.ptr(ds.u.y), float(f))

    def _divflux(self, s, dh):
        """divflux (operators.py:21-29): face fluxes into s.flx, then div."""
        m, p = self.mesh, self.param
        op.compflux(s.flx.x, s.U.x, s.h, m.d_oc.x, m.xshift, p.compflux)
        op.compflux(s.flx.y, s.U.y, s.h, m.d_oc.y, m.yshift, p.compflux)
        _ops("f2d_div", self._dt(s.h), self._grid(), _lib.ptr(s.flx.x), _lib.ptr(s.flx.y),
             _lib.ptr(m.d_msk), _lib.ptr(dh))

    def _sharp_vorticity(self, s):
        m = self.mesh
        g = self._grid()
        dt = self._dt(s.u.x)
        _ops("f2d_sharp", dt, g, _lib.ptr(s.u.x), _lib.ptr(s.u.y), _lib.ptr(s.U.x),
             _lib.ptr(s.U.y))
        _ops("f2d_vorticity", dt, g, _lib.ptr(s.u.x), _lib.ptr(s.u.y),
             _lib.ptr(m.d_slipcoef), _lib.ptr(s.omega))

    def _stage_terms(self, ds_prev, ds_out, k, dt, y, dt_log=None, y_in=(None,) * 3):
        """The f2d_sw_stage of RK stage k: earlier tendencies, this stage's
        tendency (None on the last stage: not kept), fields updated (from
        y_in where given)."""
        prev = [(x.u.x, x.u.y, x.h) for x in ds_prev]
        out = (ds_out.u.x, ds_out.u.y, ds_out.h) if ds_out is not None else (None,) * 3
        return _lib.sw_stage(self.param.integrator, k, dt, prev, out, y, dt_log, y_in)

    def _momentum(self, s, st, gradients, fill):
        m, p = self.mesh, self.param
        ptr = _lib.ptr
        _ops("f2d_sw_momentum", self._dt(s.h), self._grid(), _lib.METHODS[p.vortexforce],
             ptr(s.U.x), ptr(s.U.y), ptr(s.omega), ptr(m.d_ov.x), ptr(m.d_ov.y),
             float(p.f0 * m.area * 0.25), ptr(s.ke) if gradients else None,
             ptr(s.p) if gradients else None, ptr(m.d_mskx), ptr(m.d_msky), int(fill),
             ctypes.byref(st))

    def _diag_fused(self, s, energy, store_U=True, pressure=None):
        """sharp (unless store_U is False), vorticity, kinetic energy (with
        energy) and pressure (default: with energy), and the fills."""
        m, p = self.mesh, self.param
        ptr = _lib.ptr
        pressure = energy if pressure is None else pressure
        method = _lib.METHODS.get(p.innerproduct, 0)
        _ops("f2d_sw_diag", self._dt(s.h), self._grid(), method, ptr(s.u.x), ptr(s.u.y),
             ptr(s.h), ptr(m.d_slipcoef), ptr(m.d_ok.x), ptr(m.d_ok.y), ptr(m.d_msk),
             float(p.g / m.area), 0.0, ptr(s.U.x) if store_U else None,
             ptr(s.U.y) if store_U else None, ptr(s.omega),
             ptr(s.ke) if energy else None, ptr(s.p) if pressure else None)

    def cfl_velocity(self, s):
        return s.U

    def stage(self, s, alt, ds_prev, ds_out, k, dt, dt_next, stats=True, dt_log=None):
        """One stage of the reference's RK (integrators.py:92-124): rhs,
        addto on u.x, u.y and h, diag.  dt_log: the momentum kernel copies
        the step's dt there."""
        if self.fused:
            s, alt = self._stage_fused(s, alt, ds_prev, ds_out, k, dt, stats, dt_log)
            if dt_next is not None:
                op.cfl_dt(self.cfl_velocity(s), self.param.cfl, self.param.dtmax, dt_next)
            return s, alt
        p = self.param
        if ds_out is None:
            if self._ds_last is None:
                self._ds_last = allocate_device(self.prognostic, self.mesh.shape,
                                                s.h.dtype, s.h.device)
            ds = self._ds_last
        else:
            ds = ds_out
        self._stats = stats
        if hasattr(self, "refine_for"):  # the fused stage's solve schedule
            self.mesh.qg_helmholtz.set_refine(self.refine_for(k, stats))
        try:
            self.rhs(s, ds)
        finally:
            self._stats = True
        ys = leaves(s.u) + [s.h]
        dsl = leaves(ds.u) + [ds.h]
        prev = [leaves(x.u) + [x.h] for x in ds_prev]
        for n, y in enumerate(ys):
            op.rk_addto(y, [pv[n] for pv in prev] + [dsl[n]], p.integrator, k, dt)
        self.diag(s)
        if dt_next is not None:
            op.cfl_dt(s.U, p.cfl, p.dtmax, dt_next)
        return s, alt


class RSW(_SW):
    """Rotating shallow water (equations.py:139-157).

    The fused stage stores neither U nor p: the momentum kernel and the
    transport of h form U = sharp(u) and p = (g/area) h from u and h (the
    same arithmetic, so the same bits), u and h go to the other buffer of a
    ping-pong pair, and the diag keeps omega and ke.  ``materialize`` (before
    any read of the device state) and ``cfl_velocity`` store U and p."""

    def _momentum_u(self, s, st):
        """The momentum kernel from u and h (U and p formed on the fly)."""
        m, p = self.mesh, self.param
        ptr = _lib.ptr
        _ops("f2d_sw_momentum_u", self._dt(s.h), self._grid(), _lib.METHODS[p.vortexforce],
             ptr(s.u.x), ptr(s.u.y), ptr(s.omega), ptr(m.d_ov.x), ptr(m.d_ov.y),
             float(p.f0 * m.area * 0.25), ptr(s.ke), ptr(s.h), float(p.g / m.area), 0.0,
             ptr(m.d_mskx), ptr(m.d_msky), 1, ctypes.byref(st))

    def _transport_u(self, s, h_out, dh_prev, dh_out, k, dt):
        """The transport stage of h with U = sharp(u) formed on the fly."""
        m, p = self.mesh, self.param
        ptr = _lib.ptr
        d0 = dh_prev[0] if len(dh_prev) > 0 else None
        d1 = dh_prev[1] if len(dh_prev) > 1 else None
        _ops("f2d_transport_stage_u", self._dt(s.h), self._grid(), _lib.METHODS[p.compflux],
             ptr(s.h), ptr(h_out), ptr(s.u.x), ptr(s.u.y), ptr(m.d_oc.x), ptr(m.d_oc.y),
             ptr(m.d_msk), ptr(d0), ptr(d1), ptr(dh_out), _lib.INTEGRATORS[p.integrator], k,
             ptr(dt))

    def _stage_fused(self, s, alt, ds_prev, ds_out, k, dt, stats, dt_log=None):
        st = self._stage_terms(ds_prev, ds_out, k, dt, (alt.u.x, alt.u.y, None), dt_log,
                               y_in=(s.u.x, s.u.y, None))
        self._timed("k_sw_momentum", k, lambda: self._momentum_u(s, st))
        self._timed("k_transport", k, lambda: self._transport_u(
            s, alt.h, [x.h for x in ds_prev], ds_out.h if ds_out is not None else None, k, dt))
        s, alt = s._replace(u=alt.u, h=alt.h), alt._replace(u=s.u, h=s.h)
        self._timed("k_sw_diag", k, lambda: self._diag_fused(s, energy=True, store_U=False,
                                                             pressure=False))
        return s, alt

    events = None  # a list: eager stages record (kernel, stage, start, end) HIP events

    def _timed(self, name, k, fn):
        if self.events is None:
            return fn()
        import torch
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        self.events.append((name, k, e0, e1))

    def _store_U_p(self, s):
        m, p = self.mesh, self.param
        dt, g = self._dt(s.h), self._grid()
        _ops("f2d_sharp", dt, g, _lib.ptr(s.u.x), _lib.ptr(s.u.y), _lib.ptr(s.U.x),
             _lib.ptr(s.U.y))
        _ops("f2d_pressure", dt, g, _lib.ptr(s.h), _lib.ptr(s.p), float(p.g / m.area), 0.0)

    def materialize(self, s, names=None):
        if self.fused and (names is None or "U" in names or "p" in names):
            self._store_U_p(s)

    def cfl_velocity(self, s):
        if self.fused:
            self._store_U_p(s)
        return s.U

    def rhs(self, s, ds):
        m = self.mesh
        self._vortexforce_coriolis(s, ds)
        dt, g = self._dt(s.h), self._grid()
        for phi in (s.ke, s.p):  # addgrad(ke), addgrad(p)
            _ops("f2d_addgrad", dt, g, _lib.ptr(phi), _lib.ptr(m.d_mskx),
                 _lib.ptr(m.d_msky), _lib.ptr(ds.u.x), _lib.ptr(ds.u.y))
        self._divflux(s, ds.h)
        for a in (ds.u.x, ds.u.y, ds.h):
            op.fill(m, a)

    def diag(self, s):
        m, p = self.mesh, self.param
        self._sharp_vorticity(s)
        # compute_kinetic_energy (operators.py:80-101)
        s.ke.zero_()
        op.innerproduct(s.ke, s.U.x, s.u.x, m.d_ok.x, m.xshift, p.innerproduct)
        op.innerproduct(s.ke, s.U.y, s.u.y, m.d_ok.y, m.yshift, p.innerproduct)
        dt, g = self._dt(s.h), self._grid()
        _ops("f2d_ke_finish", dt, g, _lib.ptr(s.ke), _lib.ptr(m.d_msk))
        # compute_pressure (operators.py:110-111), hb = 0
        _ops("f2d_pressure", dt, g, _lib.ptr(s.h), _lib.ptr(s.p), float(p.g / m.area),
             0.0)
        op.fill(m, s.omega)
        op.fill(m, s.ke)


class QGRSW(_SW):
    """RSW tendency projected onto the QG manifold (equations.py:106-136,
    Thiry et al. 2024)."""

    def _stage_fused(self, s, alt, ds_prev, ds_out, k, dt, stats, dt_log=None):
        self._stage_pre(s, k, dt, dt_log)
        solver = self.mesh.qg_helmholtz
        solver.set_refine(self.refine_for(k, stats))
        solver.solve(s.pv, s.psi, stats=stats)
        self._stage_post(s, ds_prev, ds_out, k, dt)
        return s, alt

    def refine_for(self, k, stats=False):
        """FP64 refinements of stage k's Helmholtz solve (Poisson2D.set_refine):
        in FP32 with ``poisson_refine`` "last", plain FP32 solves before the
        last stage (a host-checked solve always refines); None: the
        solver's option."""
        p = self.param
        if p.precision == "fp64" or getattr(p, "poisson_refine", "every") != "last":
            return None
        return None if (stats or k == _lib.NSTAGES[p.integrator] - 1) else 0

    def stage_split(self, s, ds_prev, ds_out, k, dt, dt_next, dt_log=None):
        """stage() cut at its Helmholtz solve, for callers that batch the
        solves of many models: (pre, post, cfl) callables (cfl None unless
        dt_next is given); pre, the solve of s.pv into s.psi by
        mesh.qg_helmholtz (with ``refine_for(k)`` refinements), post, then
        cfl run the stage.  pre and post launch only kernels that
        f2d_record_begin can record."""
        cfl = None
        if dt_next is not None:
            if self._cflpart is None:
                if _lib.capturing():
                    raise RuntimeError("stage_split: CFL array not allocated before a capture")
                self._cflpart = torch.zeros(op.CFL_PARTS, dtype=torch.float64,
                                            device=s.h.device)
            part = self._cflpart

            def cfl():
                # its own block-maxima array: recordable into a member batch
                op.cfl_dt(self.cfl_velocity(s), self.param.cfl, self.param.dtmax, dt_next,
                          part=part)
        return ((lambda: self._stage_pre(s, k, dt, dt_log)),
                (lambda: self._stage_post(s, ds_prev, ds_out, k, dt)), cfl)

    def _stage_pre(self, s, k, dt, dt_log):
        """The unprojected tendencies and the pv of the projection."""
        m, p = self.mesh, self.param
        ptr = _lib.ptr
        if self._pre is None:
            # the unprojected tendencies (the projection reads them while it
            # writes the stage's tendency)
            self._pre = allocate_device(self.prognostic, m.shape, s.h.dtype, s.h.device)
        pre = self._pre
        dt_, g = self._dt(s.h), self._grid()
        self._momentum(s, _lib.sw_stage(None, 0, dt, out=(pre.u.x, pre.u.y, None),
                                        dt_log=dt_log), gradients=False, fill=False)
        if p.xperiodic or getattr(p, "yperiodic", False):
            self._divflux(s, pre.h)  # the unfilled divergence the projection reads
        else:
            # closed domain: the fill does nothing and the halo cells are dry
            # (their divergence is 0, as div's product with msk leaves it),
            # so the transport stage's divergence is div's, entry for entry
            op.transport_rhs(m, p.compflux, s.h, s.U, pre.h)
        _ops("f2d_qg_pv", dt_, g, ptr(pre.u.x), ptr(pre.u.y), ptr(pre.h), ptr(m.d_slipcoef),
             ptr(m.d_mskv), float(-p.f0 / p.H), ptr(s.pv))

    def _stage_post(self, s, ds_prev, ds_out, k, dt):
        """The projection (RK update of u and h) and the diag, from s.psi."""
        m, p = self.mesh, self.param
        ptr = _lib.ptr
        pre = self._pre
        dt_, g = self._dt(s.h), self._grid()
        st = self._stage_terms(ds_prev, ds_out, k, dt, (s.u.x, s.u.y, s.h))
        _ops("f2d_qg_project", dt_, g, ptr(s.psi), ptr(m.d_mskx), ptr(m.d_msky),
             ptr(m.d_mskv), ptr(m.d_msk), float(p.f0 * m.area / p.g), ptr(pre.u.x),
             ptr(pre.u.y), ptr(pre.h), ctypes.byref(st))
        self._diag_fused(s, energy=False)

    def rhs(self, s, ds):
        m, p = self.mesh, self.param
        self._vortexforce_coriolis(s, ds)
        self._divflux(s, ds.h)
        dt, g = self._dt(s.h), self._grid()
        # qg_projection(mesh, ds.u, ds.h, s.pv, s.psi, anomaly=True),
        # operators.py:176-183
        _ops("f2d_vorticity", dt, g, _lib.ptr(ds.u.x), _lib.ptr(ds.u.y),
             _lib.ptr(m.d_slipcoef), _lib.ptr(s.pv))
        _ops("f2d_c2v_scaled", dt, g, _lib.ptr(ds.h), float(-p.f0 / p.H), _lib.ptr(s.pv),
             _lib.ptr(m.d_mskv), 1)
        m.qg_helmholtz.solve(s.pv, s.psi, stats=self._stats)
        _ops("f2d_v2c_scaled", dt, g, _lib.ptr(s.psi), float(p.f0 * m.area / p.g),
             _lib.ptr(m.d_mskv), _lib.ptr(m.d_msk), _lib.ptr(ds.h))
        op.perpgrad(m, s.psi, ds.u, contravariant=False)
        for a in (ds.u.x, ds.u.y, ds.h):
            op.fill(m, a)

    def diag(self, s):
        self._sharp_vorticity(s)
        op.fill(self.mesh, s.omega)


_Du = namedtuple("prognostic", ("u",))


class Euler(_SW):
    """Euler in momentum-pressure form (equations.py:9-24): vortex force and
    kinetic-energy gradient, then the pressure projection
    (operators.py:114-119) with the Neumann cell Poisson solve.  With the
    direct solve (all-wet domains, power-of-two sizes) stages may skip the
    host check and steps are graph-captured; the Jacobi-PCG solve always
    checks convergence on the host, so its steps are not."""
    prognostic = ("u",)
    logs_dt = False

    @property
    def capturable(self):
        return self.mesh.poisson_centers.method == "fft"

    def rhs(self, s, ds):
        m, p = self.mesh, self.param
        op.vortexforce(ds.u.x, s.U.y, s.omega, m.d_ov.y, m.yshift, m.xshift, +1,
                       p.vortexforce)
        op.vortexforce(ds.u.y, s.U.x, s.omega, m.d_ov.x, m.xshift, m.yshift, -1,
                       p.vortexforce)
        _ops("f2d_addgrad", self._dt(s.u.x), self._grid(), _lib.ptr(s.ke), _lib.ptr(m.d_mskx),
             _lib.ptr(m.d_msky), _lib.ptr(ds.u.x), _lib.ptr(ds.u.y))
        op.fill(m, ds.u.x)
        op.fill(m, ds.u.y)

    def diag(self, s):
        m, p = self.mesh, self.param
        dt, g = self._dt(s.u.x), self._grid()
        # pressure_projection (operators.py:114-119)
        _ops("f2d_sharp", dt, g, _lib.ptr(s.u.x), _lib.ptr(s.u.y), _lib.ptr(s.U.x),
             _lib.ptr(s.U.y))
        _ops("f2d_div", dt, g, _lib.ptr(s.U.x), _lib.ptr(s.U.y), _lib.ptr(m.d_msk),
             _lib.ptr(s.div))
        rhs = s.div * (-m.area)
        pc = m.poisson_centers
        pc.solve(rhs, s.p, stats=self._stats if pc.method == "fft" else True)
        _ops("f2d_addgrad", dt, g, _lib.ptr(s.p), _lib.ptr(m.d_mskx), _lib.ptr(m.d_msky),
             _lib.ptr(s.u.x), _lib.ptr(s.u.y))
        op.fill(m, s.u.x)
        op.fill(m, s.u.y)
        # sharp, vorticity, kinetic energy, fill (equations.py:17-22)
        self._sharp_vorticity(s)
        s.ke.zero_()
        op.innerproduct(s.ke, s.U.x, s.u.x, m.d_ok.x, m.xshift, p.innerproduct)
        op.innerproduct(s.ke, s.U.y, s.u.y, m.d_ok.y, m.yshift, p.innerproduct)
        _ops("f2d_ke_finish", dt, g, _lib.ptr(s.ke), _lib.ptr(m.d_msk))
        op.fill(m, s.omega)
        op.fill(m, s.ke)

    def stage(self, s, alt, ds_prev, ds_out, k, dt, dt_next, stats=True):
        p = self.param
        if ds_out is None:
            if self._ds_last is None:
                self._ds_last = allocate_device(self.prognostic, self.mesh.shape,
                                                s.u.x.dtype, s.u.x.device)
            ds = self._ds_last
        else:
            # the integrator hands a one-field prognostic as the field itself
            ds = ds_out if hasattr(ds_out, "u") else _Du(u=ds_out)
        self.rhs(s, ds)
        ys = leaves(s.u)
        dsl = leaves(ds.u)
        prev = [leaves(x.u if hasattr(x, "u") else x) for x in ds_prev]
        for n, y in enumerate(ys):
            op.rk_addto(y, [pv[n] for pv in prev] + [dsl[n]], p.integrator, k, dt)
        self._stats = stats
        try:
            self.diag(s)
        finally:
            self._stats = True
        if dt_next is not None:
            op.cfl_dt(s.U, p.cfl, p.dtmax, dt_next)
        return s, alt


def qg_projection(model):
    """Geostrophic balance of a model's initial state on the device, as
    geos_adj.py:43-44 sets it up: qg_projection (operators.py:176-183) with
    anomaly=False -- pv = vorticity(u) + c2v((h - h0) (-f0/H)) with
    h0 = H area - hb (add_stretching, operators.py:194-201), psi = Helmholtz
    solve of pv, h = v2c(psi f0 area/g) + h0, times msk
    (thickness_from_psi, operators.py:204-211), u = perpgrad(psi)
    (covariant) -- then the model's diag (geos_adj.py:55).  Works on the
    host state: it is copied in and back."""
    m, p = model.mesh, model.param
    if p.model not in ("qgrsw", "rsw"):
        raise NotImplementedError("qg_projection: rsw / qgrsw states (u, h)")
    integ = model.integrator
    integ.begin(model.state)
    s = integ.dstate
    dt, g = _lib.dtype_code(s.h.dtype), m.grid_ref()
    _ops("f2d_vorticity", dt, g, _lib.ptr(s.u.x), _lib.ptr(s.u.y), _lib.ptr(m.d_slipcoef),
         _lib.ptr(s.pv))
    h0 = p.H * m.area - m.hb
    dh = s.h - h0
    _ops("f2d_c2v_scaled", dt, g, _lib.ptr(dh), float(-p.f0 / p.H), _lib.ptr(s.pv),
         _lib.ptr(m.d_mskv), 1)
    m.qg_helmholtz.solve(s.pv, s.psi)
    _ops("f2d_v2c_scaled", dt, g, _lib.ptr(s.psi), float(p.f0 * m.area / p.g),
         _lib.ptr(m.d_mskv), _lib.ptr(m.d_msk), _lib.ptr(s.h))
    s.h.add_(h0)
    s.h.mul_(m.d_msk)
    op.perpgrad(m, s.psi, s.u, contravariant=False)
    integ.end(model.state)
    integ.diag(model.state)


MODELS = {"rsw": RSW, "qgrsw": QGRSW, "euler": Euler}


def get_device_model(param, mesh):
    return MODELS[param.model](param, mesh)
