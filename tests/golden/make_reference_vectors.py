"""Input/output vectors of the reference's own code, for the functions that no
closed form, invariant or other fixture pins.

    python -B tests/golden/make_reference_vectors.py

Writes tests/golden/reference_vectors.npz.  It runs in the build container
only: it imports the reference package from /root/reference/src as pure
Python, with two stand-in modules in sys.modules and nothing written to the
reference tree (no bytecode):

* numba: ``njit``/``jit`` return the function unchanged, ``prange`` is
  ``range``.  The loops of weno.py then run as CPython loops over numpy
  float64 scalars: the same IEEE operations in the same order as numba's
  code (no contraction; ``x**2`` is a correctly rounded square either way;
  negative indices wrap as numba's do).
* netCDF4: a ``Dataset`` that raises; no vector here writes history.

Recorded (each operator's output array is pre-filled with seeded junk, so the
vectors also pin which entries an operator leaves untouched):

* operators on two small grids with dx != dy and non-default f0, g, H (a
  closed basin with an island, free slip; an x-periodic channel with an
  island, no slip): vortexforce for every method (weno.py:367-385,
  operators.py:6-18), kinetic energy for every method (weno.py:388-405,
  operators.py:80-101), addcoriolis, addgrad, sharp, compute_vorticity, div,
  compute_pressure, centerstovertices, verticestocenters, perpgrad,
  add_stretching, thickness_from_psi, qg_projection, qg_inversion (beta = 0
  and beta != 0 with the user-set ``mesh.f`` of
  experiments/inpreparation/gyre.py:24-29), pressure_projection
  (operators.py:21-211);
* model runs: the euler dipole of tests/test_models.py:8-15 run exactly as
  that test runs it (its dt sequence and final velocity), the advection body
  rotation of tests/test_models.py:17-26, and short rsw, qgrsw, qg (beta !=
  0) and masked euler runs (equations.py:9-24, 89-157).
"""
import os
import sys
import types

sys.dont_write_bytecode = True

import numpy as np  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SRC = "/root/reference/src"
OUT = os.path.join(HERE, "reference_vectors.npz")


def _stub_modules():
    nb = types.ModuleType("numba")

    def deco(*a, **k):
        if len(a) == 1 and callable(a[0]) and not k:
            return a[0]
        return lambda f: f
    nb.njit = deco
    nb.jit = deco
    nb.prange = range
    nb.set_num_threads = lambda n: None
    cfg = types.ModuleType("numba.config")
    cfg.NUMBA_DEFAULT_NUM_THREADS = 1
    nb.config = cfg
    sys.modules["numba"] = nb
    sys.modules["numba.config"] = cfg

    nc = types.ModuleType("netCDF4")

    class Dataset:
        def __init__(self, *a, **k):
            raise RuntimeError("no NetCDF output while writing vectors")
    nc.Dataset = Dataset
    sys.modules["netCDF4"] = nc
    import matplotlib
    matplotlib.use("Agg")


def _quiet(fn, *a, **k):
    import contextlib
    import io
    with contextlib.redirect_stdout(io.StringIO()):
        return fn(*a, **k)


def island_mask(ny, nx, xper, centre, radius):
    """Default mask (meshes.py:70-75) with one circular island."""
    nh = 3
    m = np.zeros((ny + 2 * nh, nx + 2 * nh), dtype=np.int8)
    xdx = slice(None) if xper else slice(nh, -nh)
    m[nh:-nh, xdx] = 1
    jj, ii = np.mgrid[0:ny + 2 * nh, 0:nx + 2 * nh]
    cj, ci = centre
    m[(jj - cj) ** 2 + (ii - ci) ** 2 < radius ** 2] = 0
    return m


SEEDS = {"closed": 101, "channel": 202}

# grid cases for the operator vectors: (name, param overrides, mask args)
OP_CASES = [
    ("closed", dict(nx=14, ny=10, Lx=1.0, Ly=1.3, xperiodic=False, noslip=None,
                    f0=2.5, g=1.7, H=0.8), dict(xper=False, centre=(8, 9), radius=2.2)),
    ("channel", dict(nx=16, ny=9, Lx=1.6, Ly=0.81, xperiodic=True, noslip=True,
                     f0=3.0, g=0.6, H=1.4), dict(xper=True, centre=(7, 6), radius=1.9)),
]


def make_param(fd, model, **kw):
    p = _quiet(fd.Param)
    p.model = model
    p.animation = False
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def make_model(fd, p, msk=None):
    m = _quiet(fd.Model, p)
    if msk is not None:
        m.mesh.msk[...] = msk
        m.mesh.finalize()
    return m


def operator_vectors(fd, out):
    from fluids2d import operators as op
    from fluids2d.states import Vector

    for name, kw, mk in OP_CASES:
        p = make_param(fd, "rsw", **kw)
        msk = island_mask(kw["ny"], kw["nx"], **mk)
        model = make_model(fd, p, msk)
        mesh = model.mesh
        sh = mesh.shape
        rng = np.random.default_rng(SEEDS[name])
        pre = f"op_{name}_"
        out[pre + "msk"] = msk

        def rnd():
            return rng.standard_normal(sh)

        def junk():
            return rng.uniform(-9, 9, sh)

        ins = dict(ux=rnd(), uy=rnd(), Ux=rnd(), Uy=rnd(), omega=rnd(),
                   h=rnd(), psi=rnd(), phi=rnd(), pv=rnd())
        for k, v in ins.items():
            out[pre + "in_" + k] = v
        u = Vector(ins["ux"], ins["uy"])
        U = Vector(ins["Ux"], ins["Uy"])

        def vec_junk():
            return Vector(junk(), junk())

        def save(key, arr):
            out[pre + key] = np.array(arr, copy=True)

        for meth in ("weno", "upwind", "centered", "cweno"):
            p.vortexforce = meth
            du = vec_junk()
            save(f"vf_{meth}_pre_x", du.x)
            save(f"vf_{meth}_pre_y", du.y)
            op.addvortexforce(p, mesh, U, ins["omega"], du)
            save(f"vf_{meth}_x", du.x)
            save(f"vf_{meth}_y", du.y)
        p.vortexforce = "weno"

        for meth in ("weno", "upwind", "centered", "cweno", "classic"):
            p.innerproduct = meth
            ke = junk()
            op.compute_kinetic_energy(p, mesh, u, U, ke)
            save(f"ke_{meth}", ke)
        p.innerproduct = "weno"

        du = vec_junk()
        save("cor_pre_x", du.x)
        save("cor_pre_y", du.y)
        op.addcoriolis(p, mesh, U, du)
        save("cor_x", du.x)
        save("cor_y", du.y)

        du = vec_junk()
        save("grad_pre_x", du.x)
        save("grad_pre_y", du.y)
        op.addgrad(mesh, ins["phi"], du)
        save("grad_x", du.x)
        save("grad_y", du.y)

        Uo = vec_junk()
        op.sharp(mesh, u, Uo)
        save("sharp_x", Uo.x)
        save("sharp_y", Uo.y)

        w = junk()
        op.compute_vorticity(mesh, u, w)
        save("vort", w)

        d = junk()
        save("div_pre", d)
        op.div(mesh, U, d)
        save("div", d)

        pr = junk()
        op.compute_pressure(p, mesh, ins["h"], pr)
        save("pressure", pr)

        for addto in (False, True):
            v = junk()
            save(f"c2v_{int(addto)}_pre", v)
            op.centerstovertices(mesh, ins["omega"], v, addto=addto)
            save(f"c2v_{int(addto)}", v)

        hh = junk()
        save("v2c_pre", hh)
        with np.errstate(divide="ignore", invalid="ignore"):
            op.verticestocenters(mesh, ins["psi"], hh)
        save("v2c", hh)

        for contra in (False, True):
            uo = vec_junk()
            save(f"perp_{int(contra)}_pre_x", uo.x)
            save(f"perp_{int(contra)}_pre_y", uo.y)
            op.perpgrad(mesh, ins["psi"], uo, contravariant=contra)
            save(f"perp_{int(contra)}_x", uo.x)
            save(f"perp_{int(contra)}_y", uo.y)

        for an in (False, True):
            pv = ins["pv"].copy()
            op.add_stretching(mesh, pv, ins["h"], an)
            save(f"stretch_{int(an)}", pv)

            hh = junk()
            save(f"thick_{int(an)}_pre", hh)
            with np.errstate(divide="ignore", invalid="ignore"):
                op.thickness_from_psi(mesh, ins["psi"], hh, an)
            save(f"thick_{int(an)}", hh)

            uu = Vector(ins["ux"].copy(), ins["uy"].copy())
            h2 = ins["h"].copy()
            pv = junk()
            ps = junk()
            save(f"proj_{int(an)}_pre_pv", pv)
            save(f"proj_{int(an)}_pre_psi", ps)
            with np.errstate(divide="ignore", invalid="ignore"):
                op.qg_projection(mesh, uu, h2, pv, ps, anomaly=an)
            save(f"proj_{int(an)}_ux", uu.x)
            save(f"proj_{int(an)}_uy", uu.y)
            save(f"proj_{int(an)}_h", h2)
            save(f"proj_{int(an)}_pv", pv)
            save(f"proj_{int(an)}_psi", ps)

        for beta in (0.0, 7.0):
            p.beta = beta
            if beta:
                xv, yv = mesh.xy("v")
                mesh.f = beta * (yv - p.Ly / 2) * mesh.area
                save("inv_f", mesh.f)
            work = junk()
            ps = junk()
            save(f"inv_{int(beta)}_pre_work", work)
            save(f"inv_{int(beta)}_pre_psi", ps)
            op.qg_inversion(mesh, ins["pv"], work, ps)
            save(f"inv_{int(beta)}_work", work)
            save(f"inv_{int(beta)}_psi", ps)
        p.beta = 0.0

        # euler projection: the cell Neumann operator (singular, SuperLU)
        pe = make_param(fd, "euler", **{k: v for k, v in kw.items()
                                        if k not in ("f0", "g", "H")})
        me = make_model(fd, pe, msk)
        uu = Vector(ins["ux"] * me.mesh.mskx, ins["uy"] * me.mesh.msky)
        Uo = vec_junk()
        d = junk()
        save("pproj_div_pre", d)
        pr = np.zeros(sh)
        op.pressure_projection(me.mesh, Uo, d, pr, uu)
        save("pproj_ux", uu.x)
        save("pproj_uy", uu.y)
        save("pproj_div", d)
        save("pproj_p", pr)


def run_recording(model, nmax=None):
    """Model.run's loop (model.py:45-51) with each step's dt recorded."""
    dts = []
    orig = model.set_dt

    def set_dt():
        orig()
        dts.append(model.time.dt)
    model.set_dt = set_dt
    if nmax is not None:
        model.param.maxite = nmax
    _quiet(model.run)
    return np.array(dts)


def model_vectors(fd, out):
    import experiments.vortex as vortex
    import experiments.tracer_advection as tadv
    from fluids2d import operators as op

    # tests/test_models.py:8-15 exactly
    p = make_param(fd, "euler", tend=10)
    m = make_model(fd, p)
    vortex.set_initial_dipole(m)
    out["kat_euler_dts"] = run_recording(m)
    out["kat_euler_ite"] = np.array(m.time.ite)
    out["kat_euler_ux"] = m.state.u.x.copy()
    out["kat_euler_uy"] = m.state.u.y.copy()

    # tests/test_models.py:17-26 exactly
    p = make_param(fd, "advection", tend=10)
    m = make_model(fd, p)
    tadv.set_initial_velocity(m, flow="bodyrotation")
    tadv.set_initial_tracer(m)
    out["kat_advection_dts"] = run_recording(m)
    out["kat_advection_ite"] = np.array(m.time.ite)

    def bumps(x, y, rng, n, r):
        f = np.zeros_like(x)
        for _ in range(n):
            x0, y0 = rng.uniform(0.2, 0.8, 2) * [x.max(), y.max()]
            f += rng.choice([-1, 1]) * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * r ** 2))
        return f

    # rsw: x-periodic channel with an island, WENO everywhere
    kw = dict(nx=24, ny=16, Lx=1.5, Ly=1.0, xperiodic=True, f0=1.0, g=1.0, H=1.0,
              tend=1e9, cfl=0.9)
    msk = island_mask(16, 24, True, (10, 14), 3.1)
    m = make_model(fd, make_param(fd, "rsw", **kw), msk)
    rng = np.random.default_rng(31)
    x, y = m.mesh.xy("c")
    m.state.h[...] = (1 + 0.05 * bumps(x, y, rng, 3, 0.12)) * m.mesh.area * m.mesh.msk
    m.integrator.diag(m.state)
    out["rsw_msk"] = msk
    out["rsw_h0"] = m.state.h.copy()
    out["rsw_dts"] = run_recording(m, 12)
    out["rsw_ux"] = m.state.u.x.copy()
    out["rsw_uy"] = m.state.u.y.copy()
    out["rsw_h"] = m.state.h.copy()

    # qgrsw: closed basin with an island, geos_adj.py:37-49 set-up
    kw = dict(nx=20, ny=20, f0=10.0, tend=1e9, cfl=0.9)
    msk = island_mask(20, 20, False, (12, 9), 3.0)
    m = make_model(fd, make_param(fd, "qgrsw", **kw), msk)
    rng = np.random.default_rng(41)
    x, y = m.mesh.xy("c")
    s = m.state
    s.h[...] = (1 + 0.1 * bumps(x, y, rng, 3, 0.1)) * m.mesh.msk * m.mesh.area
    op.qg_projection(m.mesh, s.u, s.h, s.pv, s.psi)
    m.integrator.diag(s)
    out["qgrsw_msk"] = msk
    out["qgrsw_ux0"] = s.u.x.copy()
    out["qgrsw_uy0"] = s.u.y.copy()
    out["qgrsw_h0"] = s.h.copy()
    out["qgrsw_dts"] = run_recording(m, 10)
    out["qgrsw_ux"] = s.u.x.copy()
    out["qgrsw_uy"] = s.u.y.copy()
    out["qgrsw_h"] = s.h.copy()

    # qg with beta != 0: gyre.py:24-34 set-up (mesh.f on vertices, pv += f)
    kw = dict(nx=24, ny=16, Lx=1.5, Ly=1.0, f0=25.0, beta=10.0, tend=1e9, cfl=0.9)
    msk = island_mask(16, 24, False, (9, 15), 2.5)
    p = make_param(fd, "qg", **kw)
    m = make_model(fd, p, msk)
    xv, yv = m.mesh.xy("v")
    m.mesh.f = p.beta * (yv - p.Ly / 2) * m.mesh.area
    x, y = m.mesh.xy("c")
    rng = np.random.default_rng(51)
    pv0 = (p.beta * (y - p.Ly / 2) * m.mesh.area
           + 0.5 * bumps(x, y, rng, 2, 0.1) * m.mesh.area) * m.mesh.msk
    m.state.pv[...] = pv0
    m.integrator.diag(m.state)
    out["qg_msk"] = msk
    out["qg_pv0"] = pv0
    out["qg_f"] = m.mesh.f.copy()
    out["qg_dts"] = run_recording(m, 10)
    out["qg_pv"] = m.state.pv.copy()
    out["qg_psi"] = m.state.psi.copy()

    # euler, masked basin, dipole of vortex.py:8-21 moved off centre
    kw = dict(nx=24, ny=20, Lx=1.2, Ly=1.0, tend=1e9, cfl=0.9)
    msk = island_mask(20, 24, False, (15, 19), 2.4)
    m = make_model(fd, make_param(fd, "euler", **kw), msk)
    vortex.set_initial_dipole(m)
    out["euler_msk"] = msk
    out["euler_ux0"] = m.state.u.x.copy()
    out["euler_uy0"] = m.state.u.y.copy()
    out["euler_dts"] = run_recording(m, 10)
    out["euler_ux"] = m.state.u.x.copy()
    out["euler_uy"] = m.state.u.y.copy()


def main():
    if not os.path.isdir(REF_SRC):
        raise SystemExit(f"{REF_SRC} is not here: vectors are written in the build container")
    _stub_modules()
    sys.path.insert(0, REF_SRC)
    import fluids2d as fd
    out = {}
    operator_vectors(fd, out)
    model_vectors(fd, out)
    np.savez_compressed(OUT, **out)
    n = sum(v.nbytes for v in out.values())
    print(f"wrote {OUT}: {len(out)} arrays, {n / 1e3:.0f} kB raw; "
          f"euler KAT ite = {int(out['kat_euler_ite'])}, "
          f"advection KAT ite = {int(out['kat_advection_ite'])}")


if __name__ == "__main__":
    main()
