"""The reference vectors catch plausible mistakes: each case plants one in
an oracle operator (a swapped shift, a dropped factor, a wrong sign or
offset, the wrong velocity, mask or flag) and checks that the planted version
disagrees with the reference's own output (tests/golden/reference_vectors.npz)
on the same inputs, while the oracle's operator, checked the same way,
agrees with it.  The operators are those of weno.py:367-405 and
operators.py:6-211; the comparison is the one the vector tests use."""
import numpy as np
import pytest

import oracle as O
from test_reference_vectors import OP_CASES, V, _case


def _vectors(pre, key):
    return V[pre + key + "_x"], V[pre + key + "_y"]


def _vf(p, mesh, ins, pre, swap=False, sign=+1):
    m = p.vortexforce
    du = O.Vector(V[pre + f"vf_{m}_pre_x"].copy(), V[pre + f"vf_{m}_pre_y"].copy())
    s, s2 = (mesh.xshift, mesh.yshift) if swap else (mesh.yshift, mesh.xshift)
    O.vortexforce(du.x, ins["Uy"], ins["omega"], mesh.ov.y, s, s2, sign, m)
    return du.x, V[pre + f"vf_{m}_x"]


def _coriolis(p, mesh, ins, pre, shift=0, ysign=-1.0):
    du = O.Vector(V[pre + "cor_pre_x"].copy(), V[pre + "cor_pre_y"].copy())
    U = O.Vector(ins["Ux"], ins["Uy"])
    if shift == 0 and ysign == -1.0:
        O.addcoriolis(p, mesh, U, du)
    else:
        # shift: the four-point average of U.y taken one column to the right
        f = p.f0 * mesh.area * 0.25
        Uy = np.roll(U.y, -shift, axis=1)
        du.x[:-1, 1:-1] += f * (Uy[:-1, :-2] + Uy[1:, :-2] + Uy[:-1, 1:-1] + Uy[1:, 1:-1])
        du.y[1:-1, :-1] += ysign * f * (U.x[:-2, :-1] + U.x[:-2, 1:] + U.x[1:-1, :-1]
                                        + U.x[1:-1, 1:])
    return np.stack([du.x, du.y]), np.stack(_vectors(pre, "cor"))


def _grad(p, mesh, ins, pre, swap_masks=False):
    du = O.Vector(V[pre + "grad_pre_x"].copy(), V[pre + "grad_pre_y"].copy())
    if swap_masks:
        du.x[:, 1:] -= np.diff(ins["phi"], axis=1) * mesh.msky[:, 1:]
        du.y[1:, :] -= np.diff(ins["phi"], axis=0) * mesh.mskx[1:, :]
    else:
        O.addgrad(mesh, ins["phi"], du)
    return np.stack([du.x, du.y]), np.stack(_vectors(pre, "grad"))


def _ke(p, mesh, ins, pre, half=True, swap_uU=False):
    p.innerproduct = "weno"
    ke = np.full(mesh.shape, 123.0)
    u = O.Vector(ins["ux"], ins["uy"])
    U = O.Vector(ins["Ux"], ins["Uy"])
    if half and not swap_uU:
        O.compute_kinetic_energy(p, mesh, u, U, ke)
    else:
        a, b = (u, U) if swap_uU else (U, u)
        ke[:] = 0.0
        O.innerproduct(ke, a.x, b.x, mesh.ok.x, mesh.xshift, "weno")
        O.innerproduct(ke, a.y, b.y, mesh.ok.y, mesh.yshift, "weno")
        ke *= mesh.msk * (0.5 if half else 1.0)
    return ke, V[pre + "ke_weno"]


def _vort(p, mesh, ins, pre, ysign=+1.0):
    w = np.zeros(mesh.shape)
    u = O.Vector(ins["ux"], ins["uy"])
    if ysign > 0:
        O.compute_vorticity(mesh, u, w)
    else:
        w[1:, :] = -np.diff(u.x, axis=0)
        w[:, 1:] -= np.diff(u.y, axis=1)
        w *= mesh.slipcoef
    return w, V[pre + "vort"]


def _pressure(p, mesh, ins, pre, per_area=True):
    pr = np.zeros(mesh.shape)
    if per_area:
        O.compute_pressure(p, mesh, ins["h"], pr)
    else:
        pr[:] = p.g * (ins["h"] + mesh.hb)
    return pr, V[pre + "pressure"]


def _stretch(p, mesh, ins, pre, sign=-1.0):
    pv = ins["pv"].copy()
    if sign < 0:
        O.add_stretching(mesh, pv, ins["h"], True)
    else:
        O.centerstovertices(mesh, ins["h"] * (p.f0 / p.H), pv, addto=True)
    return pv, V[pre + "stretch_1"]


def _thick(p, mesh, ins, pre, area=True, add_h0=False):
    hh = V[pre + "thick_1_pre"].copy()
    with np.errstate(divide="ignore", invalid="ignore"):
        if area and not add_h0:
            O.thickness_from_psi(mesh, ins["psi"], hh, True)
        else:
            scale = p.f0 * (mesh.area if area else 1.0) / p.g
            O.verticestocenters(mesh, ins["psi"] * scale, hh)
            if add_h0:  # the anomaly flag ignored
                hh += p.H * mesh.area - mesh.hb
            hh *= mesh.msk
    return hh, V[pre + "thick_1"]


def _projection(p, mesh, ins, pre, contravariant=False):
    u = O.Vector(ins["ux"].copy(), ins["uy"].copy())
    h = ins["h"].copy()
    pv = V[pre + "proj_1_pre_pv"].copy()
    psi = V[pre + "proj_1_pre_psi"].copy()
    with np.errstate(divide="ignore", invalid="ignore"):
        O.compute_vorticity(mesh, u, pv)
        O.add_stretching(mesh, pv, h, True)
        mesh.qg_helmholtz.solve(pv, psi)
        O.thickness_from_psi(mesh, psi, h, True)
        O.perpgrad(mesh, psi, u, contravariant=contravariant)
    return np.stack([u.x, u.y]), np.stack([V[pre + "proj_1_ux"], V[pre + "proj_1_uy"]])


# (name, function, kwargs of the planted mistake)
PLANTS = [
    ("vortexforce: s and s2 swapped in the V average", _vf, dict(swap=True)),
    ("vortexforce: wrong sign", _vf, dict(sign=-1)),
    ("addcoriolis: x average one column off", _coriolis, dict(shift=1)),
    ("addcoriolis: wrong sign on du.y", _coriolis, dict(ysign=+1.0)),
    ("addgrad: mskx and msky swapped", _grad, dict(swap_masks=True)),
    ("kinetic energy: 0.5 dropped", _ke, dict(half=False)),
    ("kinetic energy: u and U swapped", _ke, dict(swap_uU=True)),
    ("vorticity: wrong sign of the u.y difference", _vort, dict(ysign=-1.0)),
    ("pressure: g instead of g/area", _pressure, dict(per_area=False)),
    ("add_stretching: +f0/H", _stretch, dict(sign=+1.0)),
    ("thickness_from_psi: area dropped", _thick, dict(area=False)),
    ("thickness_from_psi: h0 added to the anomaly", _thick, dict(add_h0=True)),
    ("qg_projection: contravariant perpgrad", _projection, dict(contravariant=True)),
]


def _differs(got, ref):
    got, ref = np.asarray(got, dtype=float), np.asarray(ref, dtype=float)
    if not np.array_equal(np.isnan(got), np.isnan(ref)):
        return True
    d = np.abs(np.where(np.isnan(ref), 0.0, got - ref))
    return d.max() > 1e-9 * max(np.nanmax(np.abs(ref)), 1e-300)


@pytest.mark.parametrize("name", list(OP_CASES))
@pytest.mark.parametrize("label,fn,kw", PLANTS, ids=[p[0] for p in PLANTS])
def test_planted_mistake_is_caught(name, label, fn, kw):
    p, mesh, ins, pre = _case(name)
    got, ref = fn(p, mesh, ins, pre)
    assert not _differs(got, ref), f"{label}: the oracle's own operator must agree"
    p, mesh, ins, pre = _case(name)
    bad, ref = fn(p, mesh, ins, pre, **kw)
    assert _differs(bad, ref), f"{label}: the planted mistake was not caught"
