"""Committed fixtures (tests/golden, written by make_fixtures.py): the oracle
must keep reproducing them (CPU), and the HIP path must match them (GPU)
without running the oracle."""
import json
import os

import numpy as np
import pytest

import oracle as O
from helpers import P

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return dict(np.load(os.path.join(G, name), allow_pickle=False))


def test_kat_fixture_is_the_references():
    """tests/test_models.py asserts 37 (advection) and 25 (euler); the
    reference's code, run as its test runs it, reaches the euler tend in 22
    (reference_vectors.npz), the count the oracle and the device follow."""
    k = json.load(open(os.path.join(G, "kat.json")))
    assert k["advection_bodyrotation_tend10_ite"] == 37
    assert k["euler_dipole_tend10_ite"] == 25
    v = np.load(os.path.join(G, "reference_vectors.npz"))
    assert k["euler_dipole_tend10_ite_as_executed"] == int(v["kat_euler_ite"]) == 22
    assert int(v["kat_advection_ite"]) == 37


def test_oracle_reproduces_weno_points():
    f = _load("weno_points.npz")
    w5 = np.array([O.weno5z(*r) for r in f["q"]])
    w3 = np.array([O.weno3z(*r[1:4]) for r in f["q"]])
    assert np.array_equal(w5, f["weno5z"]) and np.array_equal(w3, f["weno3z"])


def test_oracle_reproduces_eulerpsi_32():
    f = _load("eulerpsi_32.npz")
    p = P(model="eulerpsi", nx=32, ny=32, cfl=0.8, tend=1e9, maxite=20)
    m = O.Model(p)
    m.state.omega[...] = f["omega0"]
    m.diag(m.state)
    m.run(20)
    e = np.linalg.norm(m.state.omega - f["omega20"]) / np.linalg.norm(f["omega20"])
    assert e <= 1e-13 and m.time.ite == int(f["ite"])


def test_oracle_reproduces_poisson_v():
    f = _load("poisson_v.npz")
    p = P(nx=24, ny=20, xperiodic=True)
    mesh = O.Mesh(p, msk=f["msk"])
    x = np.zeros(mesh.shape)
    mesh.poisson_vertices.solve(f["b"], x)
    assert np.abs(x - f["x"]).max() <= 1e-12 * np.abs(f["x"]).max()


@pytest.mark.gpu
def test_gpu_eulerpsi_32_matches_fixture(dev):
    import fluids2d_amd as f2d
    import fluids2d_amd.param as fp
    fp.QUIET = True
    f = _load("eulerpsi_32.npz")
    p = f2d.Param()
    p.model, p.nx, p.ny, p.cfl, p.tend, p.maxite = "eulerpsi", 32, 32, 0.8, 1e9, 20
    model = f2d.Model(p)
    model.state.omega[...] = f["omega0"]
    model.integrator.diag(model.state)
    model.run()
    e = np.linalg.norm(model.state.omega - f["omega20"]) / np.linalg.norm(f["omega20"])
    assert model.time.ite == int(f["ite"])
    assert e <= 1e-12, e


@pytest.mark.gpu
def test_gpu_poisson_matches_fixture(dev):
    import fluids2d_amd as f2d
    import fluids2d_amd.param as fp
    from fluids2d_amd.meshes import Mesh
    fp.QUIET = True
    f = _load("poisson_v.npz")
    p = f2d.Param()
    p.model, p.nx, p.ny, p.xperiodic = "eulerpsi", 24, 20, True
    mesh = Mesh(p)
    mesh.msk[...] = f["msk"]
    mesh.finalize()
    ps = mesh.poisson_vertices
    x = np.zeros(f["b"].shape)
    ps.solve(f["b"], x)
    err = np.abs(x - f["x"]).max() / np.abs(f["x"]).max()
    assert err <= 1e-10, err
