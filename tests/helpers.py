"""Shared test helpers: a minimal attribute bag for the oracle and seeded
inputs.  Tests compare the HIP path with the oracle on identical arrays."""
import numpy as np


class P:
    """Parameter bag with the reference's defaults (param.py:13-59)."""

    def __init__(self, **kw):
        d = dict(model="eulerpsi", nx=40, ny=40, Lx=1.0, Ly=1.0, xperiodic=False,
                 yperiodic=False, halowidth=3, noslip=None, f0=10.0, beta=0.0, g=1,
                 H=1, tend=1.0, dt=0.0, maxite=100, dtmax=9e99, integrator="rk3",
                 cfl=0.9, compflux="weno", vortexforce="weno", innerproduct="weno",
                 maxorder=6, nthreads=1, RAgamma=0.1)
        d.update(kw)
        for k, v in d.items():
            setattr(self, k, v)


def random_mask(ny, nx, seed, xper=False, yper=False, nisl=3):
    """Default mask (meshes.py:70-75) with a few circular islands."""
    import oracle as O
    rng = np.random.default_rng(seed)
    m = O.default_mask(ny, nx, 3, xper, yper)
    jj, ii = np.mgrid[0:ny + 6, 0:nx + 6]
    for _ in range(nisl):
        r = rng.uniform(1.5, max(2.0, min(nx, ny) / 6))
        cj, ci = rng.uniform(6, ny), rng.uniform(6, nx)
        m[(jj - cj) ** 2 + (ii - ci) ** 2 < r * r] = 0
    return m


def assert_same(got, ref, name="", rtol=0.0):
    """Bit-exact (rtol == 0) or relative comparison with a compact report."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert got.shape == ref.shape, (name, got.shape, ref.shape)
    if rtol == 0.0:
        bad = ~((got == ref) | (np.isnan(got) & np.isnan(ref)))
    else:
        scale = np.max(np.abs(ref)) if ref.size else 1.0
        bad = np.abs(got - ref) > rtol * max(scale, 1e-300)
    if not bad.any():
        return
    idx = np.argwhere(bad)
    d = np.abs(got - ref)[bad]
    rel = d / np.maximum(np.abs(ref[bad]), 1e-300)
    lines = [f"{name}: {bad.sum()} of {bad.size} differ; max abs {d.max():.3e}, "
             f"max rel {rel.max():.3e}"]
    for k in idx[:6]:
        k = tuple(k)
        lines.append(f"  at {k}: got {got[k]!r} ref {ref[k]!r}")
    raise AssertionError("\n".join(lines))
