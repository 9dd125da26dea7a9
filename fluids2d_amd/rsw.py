"""Shallow-water device models: rsw (equations.py:139-157) and qgrsw, the
RSW tendency projected onto the QG manifold (equations.py:106-136), operator
by operator on the device in the reference's order.

The RK stage is the reference's: rhs into the stage's tendency, `addto` on
every prognostic array (u.x, u.y, h), then diag.
"""
from collections import namedtuple

from . import _lib
from . import operators as op
from .equations import _Base
from .states import allocate_device, leaves


def _ops(name, *args):
    _lib.call(name, *args, _lib.stream_ptr())


class _SW(_Base):
    prognostic = ("u", "h")

    def __init__(self, param, mesh):
        super().__init__(param, mesh)
        self._ds_last = None
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

    def stage(self, s, alt, ds_prev, ds_out, k, dt, dt_next, stats=True):
        """One stage of the reference's RK (integrators.py:92-124): rhs,
        addto on u.x, u.y and h, diag."""
        p = self.param
        if ds_out is None:
            if self._ds_last is None:
                self._ds_last = allocate_device(self.prognostic, self.mesh.shape,
                                                s.h.dtype, s.h.device)
            ds = self._ds_last
        else:
            ds = ds_out
        self._stats = stats
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
    """Rotating shallow water (equations.py:139-157)."""

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
