"""Pins of the oracle's mask, stencil-order, slip and index logic
(meshes.py:70-186, noslip.py:4-36, elliptic.py:102-111, 198-203) against
definitions written in a different form from the oracle's loops:

* masks: a face or vertex is wet when every cell that touches it is wet;
* orders: the order of a face is 2k for the widest k in (3, 2, 1) whose
  2k-cell window around the face lies inside the flat array and is all wet,
  capped by maxorder (the reference's running sums s2/s4/s6 compute this);
* G: row-major numbering of the unknowns;
* halo fill: periodic copies.
"""
import numpy as np
import pytest

import oracle as O
from helpers import P, random_mask


def _order_by_windows(m, shift, maxorder):
    f = m.reshape(-1)
    n = f.size
    s = shift
    out = np.zeros(n, dtype=np.int8)
    for i in range(n):
        best = 0
        for k in (3, 2, 1):
            # cells i - k s, ..., i + (k - 1) s (the face sits between i - s and i)
            idx = [i + (j - k) * s for j in range(2 * k)]
            if all(0 <= t < n for t in idx) and all(f[t] == 1 for t in idx):
                best = 2 * k
                break
        out[i] = min(best, maxorder)
    return out.reshape(m.shape)


@pytest.mark.parametrize("seed,maxorder,xper,yper", [
    (0, 6, False, False), (1, 4, True, False), (2, 5, True, True), (3, 2, False, True)])
def test_set_order_matches_window_definition(seed, maxorder, xper, yper):
    ny, nx = 11, 13
    m = random_mask(ny, nx, seed, xper, yper, nisl=3)
    # islands touching the halo too
    m[0, 5] = 1
    m[-1, 2] = 0
    for shift in (1, nx + 6, -1, -(nx + 6)):
        got = O.set_order(m, shift, maxorder)
        want = _order_by_windows(m, shift, maxorder)
        assert np.array_equal(got, want), shift


def test_set_masks_definition():
    rng = np.random.default_rng(9)
    m = (rng.uniform(size=(9, 12)) > 0.3).astype(np.int8)
    mx, my, mv = O.set_masks(m)
    for j in range(m.shape[0]):
        for i in range(m.shape[1]):
            west = m[j, i - 1] if i > 0 else 0
            south = m[j - 1, i] if j > 0 else 0
            sw = m[j - 1, i - 1] if (i > 0 and j > 0) else 0
            assert mx[j, i] == m[j, i] * west
            assert my[j, i] == m[j, i] * south
            assert mv[j, i] == m[j, i] * west * south * sw


def test_slipcoef_free_and_noslip():
    """noslip.py:4-36: free slip keeps vertices whose four cells are wet;
    no slip keeps every vertex touching a wet cell."""
    m = O.default_mask(6, 7, 3, False, False)
    free = O.slipcoef(m, None, 3)
    nos = O.slipcoef(m, True, 3)
    _, _, mv = O.set_masks(m)
    assert np.array_equal(free, mv.astype(float))
    for j in range(m.shape[0]):
        for i in range(m.shape[1]):
            touch = sum(m[jj, ii] for jj in (j - 1, j) for ii in (i - 1, i)
                        if 0 <= jj < m.shape[0] and 0 <= ii < m.shape[1])
            assert nos[j, i] == (1.0 if touch > 0 else 0.0)
    part = O.slipcoef(m, ["left"], 3)
    assert np.array_equal(part[:, 3], nos[:, 3])
    assert np.array_equal(part[:, 4:], free[:, 4:])


def test_gindex_row_major():
    rng = np.random.default_rng(2)
    m = (rng.uniform(size=(7, 9)) > 0.4).astype(np.int8)
    G, N = O.gindex(m)
    assert N == int(m.sum())
    wet = np.argwhere(m == 1)
    assert np.array_equal(G[m == 1], np.arange(N))
    assert np.all(G[m == 0] == -1)
    # row-major order: argwhere enumerates rows first
    assert [tuple(G[j, i] for j, i in [w]) for w in wet] == [(k,) for k in range(N)]


def test_fill_halo_periodic():
    p = P(nx=8, ny=5, xperiodic=True, yperiodic=True)
    a = np.arange(11 * 14, dtype=float).reshape(11, 14)
    O.fill(p, a)
    n = 3
    core = a[n:-n, n:-n]
    # every halo entry equals the interior entry one period away
    for j in range(11):
        for i in range(14):
            jj = (j - n) % 5 + n
            ii = (i - n) % 8 + n
            assert a[j, i] == a[jj, ii]
    assert core.shape == (5, 8)
    # x only (the reference has no y fill)
    p2 = P(nx=8, ny=5, xperiodic=True, yperiodic=False)
    b = np.arange(11 * 14, dtype=float).reshape(11, 14)
    b0 = b.copy()
    O.fill(p2, b)
    assert np.array_equal(b[:n, n:-n], b0[:n, n:-n])
    assert np.array_equal(b[:, :n], b[:, 8:8 + n])


def test_unknown_mask_drops_periodic_halo():
    p = P(nx=6, ny=5, xperiodic=True, yperiodic=False)
    m = O.default_mask(5, 6, 3, True, False)
    _, _, mv = O.set_masks(m)
    u = O.unknown_mask(p, mv)
    assert u[:, :3].sum() == 0 and u[:, -3:].sum() == 0
    # walls in y: the first wet vertex row is one row inside the wall
    assert u[3].sum() == 0 and u[4, 3:-3].sum() == 6
