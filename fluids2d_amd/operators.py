"""Device operators: thin wrappers over the library's C ABI, named after the
reference's operators (operators.py, weno.py).  Every argument is a device
tensor in the padded layout; every wrapper raises on a library error.
"""
import ctypes

import torch

from . import _lib
from ._lib import call, ptr, stream_ptr


def _dt(t):
    return _lib.dtype_code(t.dtype)


def compflux(flx, U, q, o, s, method):
    """weno.compflux (weno.py:412-420)"""
    call("f2d_compflux", _dt(q), ptr(flx), ptr(U), ptr(q), ptr(o), q.numel(), s,
         _lib.METHODS[method], stream_ptr())


def vortexforce(du, V, q, o, s, s2, sign, method):
    """weno.vortexforce (weno.py:423-431)"""
    call("f2d_vortexforce", _dt(q), ptr(du), ptr(V), ptr(q), ptr(o), q.numel(), s, s2,
         sign, _lib.METHODS[method], stream_ptr())


def innerproduct(ke, U, q, o, s, method):
    """weno.innerproduct (weno.py:434-436)"""
    call("f2d_innerproduct", _dt(q), ptr(ke), ptr(U), ptr(q), ptr(o), q.numel(), s,
         _lib.METHODS[method], stream_ptr())


def transport_rhs(mesh, method, q, U, dq):
    """divflux + fill (operators.py:21-29, equations.py:273-274):
    dq = fill(-div(flux(q, U)) * msk), without storing the face fluxes."""
    call("f2d_transport_stage", _dt(q), mesh.grid_ref(), _lib.METHODS[method], ptr(q),
         None, ptr(U.x), ptr(U.y), ptr(mesh.d_oc.x), ptr(mesh.d_oc.y), ptr(mesh.d_msk),
         None, None, ptr(dq), -1, 0, None, stream_ptr())


def transport_stage(mesh, method, q_in, q_out, U, ds_prev, ds_out, integrator, stage,
                    dt_dev, dt_log=None, cfl_part=None, cfl=0.0, dtmax=0.0):
    """Fused rhs + fill + RK addto of one stage (see f2d_transport_stage);
    with dt_log (a one-element float64 device tensor) the kernel also
    copies dt there (f2d_transport_stage_log).  cfl_part: the block maxima
    perpgrad(..., cfl_part=) left; the kernel forms the CFL step of
    model.py:84-87 from them (with cfl, dtmax), uses it and stores it in
    dt_dev (f2d_transport_stage_cfl)."""
    d0 = ds_prev[0] if len(ds_prev) > 0 else None
    d1 = ds_prev[1] if len(ds_prev) > 1 else None
    args = (_dt(q_in), mesh.grid_ref(), _lib.METHODS[method],
            ptr(q_in), ptr(q_out), ptr(U.x), ptr(U.y), ptr(mesh.d_oc.x), ptr(mesh.d_oc.y),
            ptr(mesh.d_msk), ptr(d0), ptr(d1), ptr(ds_out), _lib.INTEGRATORS[integrator],
            stage, ptr(dt_dev))
    if cfl_part is not None:
        call("f2d_transport_stage_cfl", *args, ptr(cfl_part), cfl_part.numel(), float(cfl),
             float(dtmax), ptr(dt_log), stream_ptr())
    elif dt_log is None:
        call("f2d_transport_stage", *args, stream_ptr())
    else:
        call("f2d_transport_stage_log", *args, ptr(dt_log), stream_ptr())


def transport_stage_psi(mesh, method, q_in, q_out, psi, contravariant, fill, ds_prev, ds_out,
                        integrator, stage, dt_dev):
    """transport_stage with U = perpgrad(psi) formed in the kernel
    (f2d_transport_stage_psi): bit for bit the U perpgrad would store."""
    d0 = ds_prev[0] if len(ds_prev) > 0 else None
    d1 = ds_prev[1] if len(ds_prev) > 1 else None
    sx, sy = (1 / mesh.dy ** 2, 1 / mesh.dx ** 2) if contravariant else (1.0, 1.0)
    call("f2d_transport_stage_psi", _dt(q_in), mesh.grid_ref(), _lib.METHODS[method],
         ptr(q_in), ptr(q_out), ptr(psi), ptr(mesh.d_mskx), ptr(mesh.d_msky), sx, sy,
         int(fill), ptr(mesh.d_oc.x), ptr(mesh.d_oc.y), ptr(mesh.d_msk), ptr(d0), ptr(d1),
         ptr(ds_out), _lib.INTEGRATORS[integrator], stage, ptr(dt_dev), stream_ptr())


def rk_addto(y, xs, integrator, stage, dt_dev):
    """addto (integrators.py:154-196) with the stage coefficients from dt."""
    x0 = xs[0]
    x1 = xs[1] if len(xs) > 1 else None
    x2 = xs[2] if len(xs) > 2 else None
    call("f2d_rk_addto", _dt(y), ptr(y), ptr(x0), ptr(x1), ptr(x2), y.numel(),
         _lib.INTEGRATORS[integrator], stage, ptr(dt_dev), stream_ptr())


def addto_terms(y, xs, coefs):
    """y += ((c0 x0 + c1 x1) + c2 x2) over every entry, for 1 to 3 terms
    (integrators.py:154-196 with constant coefficients)."""
    n = len(xs)
    assert 1 <= n <= 3 and len(coefs) == n
    xs = list(xs) + [None] * (3 - n)
    cs = [float(c) for c in coefs] + [0.0] * (3 - n)
    call("f2d_addto", _dt(y), ptr(y), ptr(xs[0]), ptr(xs[1]), ptr(xs[2]), y.numel(), n,
         cs[0], cs[1], cs[2], None, 0, stream_ptr())


def centerstovertices(mesh, omega, vomega, addto=False):
    """operators.py:126-133"""
    call("f2d_c2v", _dt(omega), mesh.grid_ref(), ptr(omega), ptr(vomega),
         ptr(mesh.d_mskv), int(addto), stream_ptr())


def perpgrad(mesh, psi, U, contravariant=False, fill=False, dt_dev=None, cfl=0.0,
             dtmax=0.0, cfl_part=None):
    """operators.py:144-149 (+ fill, + the CFL step of model.py:84-87 into
    dt_dev; or, with cfl_part, its block maxima left in that device array
    of cfl_partials(mesh) entries, for cfl_finish or the next stage)."""
    if contravariant:
        sx, sy = 1 / mesh.dy ** 2, 1 / mesh.dx ** 2
    else:
        sx, sy = 1.0, 1.0
    if cfl_part is not None:
        assert dt_dev is None
        call("f2d_perpgrad_part", _dt(psi), mesh.grid_ref(), ptr(psi), ptr(U.x), ptr(U.y),
             ptr(mesh.d_mskx), ptr(mesh.d_msky), sx, sy, int(fill), ptr(cfl_part),
             cfl_part.numel(), stream_ptr())
        return
    call("f2d_perpgrad_ex", _dt(psi), mesh.grid_ref(), ptr(psi), ptr(U.x), ptr(U.y),
         ptr(mesh.d_mskx), ptr(mesh.d_msky), sx, sy, int(fill), None, ptr(dt_dev),
         float(cfl), float(dtmax), stream_ptr())


def cfl_partials(mesh):
    """Entries of the CFL block maxima that perpgrad(..., cfl_part=) writes."""
    n = ctypes.c_int64()
    call("f2d_cfl_partials", mesh.grid_ref(), ctypes.byref(n))
    return int(n.value)


def cfl_finish(part, dt_dev, cfl, dtmax):
    """The CFL step of model.py:84-87 into dt_dev from perpgrad's block
    maxima (bit for bit perpgrad(..., dt_dev=))."""
    call("f2d_cfl_finish", ptr(part), part.numel(), None, ptr(dt_dev), float(cfl),
         float(dtmax), stream_ptr())


def fill(mesh, a):
    """meshes.py:106-111 on a device tensor"""
    call("f2d_fill_halo", _dt(a), mesh.grid_ref(), ptr(a), stream_ptr())


def cfl_dt(U, cfl, dtmax, dt_dev, part=None):
    """Model.set_dt (model.py:84-87) from device velocities.  part (a
    float64 device array of CFL_PARTS entries): the block maxima go there,
    and the call may be recorded into a member batch (f2d_cfl_dt_part)."""
    if part is not None:
        call("f2d_cfl_dt_part", _dt(U.x), ptr(U.x), ptr(U.y), U.x.numel(), float(cfl),
             float(dtmax), ptr(part), part.numel(), ptr(dt_dev), stream_ptr())
        return
    call("f2d_cfl_dt", _dt(U.x), ptr(U.x), ptr(U.y), U.x.numel(), float(cfl),
         float(dtmax), ptr(dt_dev), stream_ptr())


CFL_PARTS = 2048  # entries of cfl_dt's part array (any grid)


def set_scalar(dst, value):
    call("f2d_set_scalar", ptr(dst), float(value), stream_ptr())
