"""Writes the committed fixtures of tests/golden from the oracle.

    python tests/golden/make_fixtures.py

* kat.json          the reference's own known answers
                    (tests/test_models.py:15,26: iteration counts 25, 37),
                    and the euler count its code gives when run (22)
* weno_points.npz   seeded stencils and their WENO-Z5/Z3 values
* eulerpsi_32.npz   32x32 closed eulerpsi, Gaussian vortex + seeded noise,
                    omega after 20 FP64 RK3 steps (CFL 0.8)
* poisson_v.npz     masked vertex Poisson: rhs, solution (SuperLU)

They freeze the oracle: tests/test_golden.py checks that the oracle still
reproduces them, and the GPU tests compare the HIP path with them directly.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, ".."))

import oracle as O  # noqa: E402
from helpers import P, random_mask  # noqa: E402


def weno_points():
    rng = np.random.default_rng(11)
    q = rng.standard_normal((64, 5))
    q[:8] = np.cumsum(rng.uniform(-1, 1, (8, 5)), axis=1)  # smooth-ish
    q[8:16, 3:] += 10.0  # jumps
    w5 = np.array([O.weno5z(*r) for r in q])
    w3 = np.array([O.weno3z(*r[1:4]) for r in q])
    return dict(q=q, weno5z=w5, weno3z=w3)


def eulerpsi_32():
    p = P(model="eulerpsi", nx=32, ny=32, cfl=0.8, tend=1e9, maxite=20)
    m = O.Model(p)
    x, y = m.mesh.xy("c")
    rng = np.random.default_rng(2024)
    w = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / (2 * 0.1 ** 2))
    w = (w + 1e-3 * rng.uniform(-1, 1, w.shape)) * m.mesh.msk * m.mesh.area
    m.state.omega[...] = w
    m.diag(m.state)
    m.run(20)
    return dict(omega0=w, omega20=m.state.omega.copy(), t=np.array(m.time.t),
                ite=np.array(m.time.ite))


def poisson_v():
    p = P(nx=24, ny=20, xperiodic=True)
    msk = random_mask(20, 24, 5, True, False, nisl=3)
    mesh = O.Mesh(p, msk=msk)
    rng = np.random.default_rng(5)
    b = rng.standard_normal(mesh.shape) * mesh.mskv
    x = np.zeros(mesh.shape)
    mesh.poisson_vertices.solve(b, x)
    return dict(msk=msk, b=b, x=x)


def main():
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump({"source": "tests/test_models.py:15,26",
                   "euler_dipole_tend10_ite": 25,
                   "euler_dipole_tend10_ite_as_executed": 22,
                   "advection_bodyrotation_tend10_ite": 37,
                   "note": "test_euler asserts 25; the reference's code, run as "
                           "that test runs it, takes 22 "
                           "(tests/golden/make_reference_vectors.py)"}, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "weno_points.npz"), **weno_points())
    np.savez_compressed(os.path.join(HERE, "eulerpsi_32.npz"), **eulerpsi_32())
    np.savez_compressed(os.path.join(HERE, "poisson_v.npz"), **poisson_v())


if __name__ == "__main__":
    main()
