"""Pins of the oracle's WENO reconstructions (weno.py:75-100, 167-197,
303-343) against properties the mathematics fixes, not against a retyped
formula:

* linear data gives equal smoothness indicators, so WENO-Z5 reduces to the
  linear 5th-order upwind value (weno.py:42-43 coefficients 2,-13,47,27,-3
  over 60) and WENO-Z3 to the 3rd-order one (weno.py:37-38);
* mirror symmetry of the upwind selection (U < 0 takes the reversed stencil);
* 5th-order convergence of WENO-Z5 on smooth cell averages;
* a jump inside the stencil gives a non-oscillatory value.
"""
import math

import numpy as np
import pytest

import oracle as O


def up5(qmm, qm, q0, qp, qpp):
    return (2 * qmm - 13 * qm + 47 * q0 + 27 * qp - 3 * qpp) / 60


def up3(qm, q0, qp):
    return (5 * q0 + 2 * qp - qm) / 6


@pytest.mark.parametrize("a,b", [(0.3, 1.7), (-2.0, 0.01), (5.0, -3.0), (1e-3, 1e3)])
def test_weno5z_linear_data_is_up5(a, b):
    q = [a + b * i for i in range(5)]
    v = O.weno5z(*q)
    assert abs(v - up5(*q)) <= 1e-13 * max(1.0, abs(v))
    # and the exact face value of linear data (midpoint between q0 and qp)
    assert abs(v - (a + 2.5 * b)) <= 1e-12 * max(1.0, abs(v))


@pytest.mark.parametrize("a,b", [(0.3, 1.7), (-2.0, 0.01), (5.0, -3.0)])
def test_weno3z_linear_data_is_up3(a, b):
    q = [a + b * i for i in range(3)]
    v = O.weno3z(*q)
    assert abs(v - up3(*q)) <= 1e-13 * max(1.0, abs(v))
    assert abs(v - (a + 1.5 * b)) <= 1e-12 * max(1.0, abs(v))


def test_constant_data_exact():
    for c in (0.0, 1.0, -3.25, 7e5):
        assert O.weno5z(c, c, c, c, c) == c
        assert O.weno3z(c, c, c) == c


@pytest.mark.parametrize("method", ["weno", "upwind", "centered"])
def test_mirror_symmetry(method):
    """Upwinded methods select the mirrored stencil for U < 0, so the value
    is bit-identical; the centred one does not depend on U at all and is
    mirror-symmetric up to the order of its additions."""
    rng = np.random.default_rng(3)
    tol = 1e-14 if method == "centered" else 0.0
    for _ in range(200):
        q = list(rng.standard_normal(6))
        U = abs(rng.standard_normal()) + 0.1
        pairs = [(O.flux5(method, U, *q), O.flux5(method, -U, *q[::-1])),
                 (O.flux3(method, U, *q[1:5]), O.flux3(method, -U, *q[1:5][::-1])),
                 (O.flux1(method, U, *q[2:4]), O.flux1(method, -U, *q[2:4][::-1]))]
        for a, b in pairs:
            assert abs(a - b) <= tol * max(1.0, abs(a)), (method, a, b)


def test_cweno_properties():
    """cweno (weno.py:104-134, 256-300): the 3-point member mirrors like the
    upwinded ones (its smoothness indicators swap with the stencil); the
    5-point member takes the same left-biased indicators for both signs of
    U, so its value does not depend on U; both are exact on linear data."""
    rng = np.random.default_rng(5)
    for _ in range(100):
        q = list(rng.standard_normal(6))
        U = abs(rng.standard_normal()) + 0.1
        assert O.flux3("cweno", U, *q[1:5]) == O.flux3("cweno", -U, *q[1:5][::-1])
        assert O.flux5("cweno", U, *q) == O.flux5("cweno", -U, *q)
    q = [0.5 + 0.25 * i for i in range(6)]
    assert abs(O.flux5("cweno", 1.0, *q) - (0.5 + 0.25 * 2.5)) < 1e-14
    assert abs(O.flux3("cweno", 1.0, *q[1:5]) - (0.5 + 0.25 * 2.5)) < 1e-14


def test_upwind_selection_zero_velocity_takes_mirrored_branch():
    """weno.py:305: U == 0 is not > 0, so the right-biased stencil is used."""
    rng = np.random.default_rng(4)
    q = list(rng.standard_normal(6))
    assert O.flux5("weno", 0.0, *q) == O.weno5z(q[5], q[4], q[3], q[2], q[1])
    assert O.flux5("weno", 1.0, *q) == O.weno5z(q[0], q[1], q[2], q[3], q[4])


def _cell_averages(F, x0, h, n):
    return [(F(x0 + (i + 1) * h) - F(x0 + i * h)) / h for i in range(n)]


def test_weno5z_fifth_order_on_smooth_averages():
    """Face value between cells 2 and 3 of 5 cell averages of sin(2 pi x):
    the error of WENO-Z5 falls like h^5 (Borges et al. 2008)."""
    f = lambda x: math.sin(2 * math.pi * x)
    F = lambda x: -math.cos(2 * math.pi * x) / (2 * math.pi)
    errs = []
    hs = [0.02, 0.01, 0.005]
    for h in hs:
        xf = 0.3  # face position, f' != 0 there
        q = _cell_averages(F, xf - 3 * h, h, 5)
        errs.append(abs(O.weno5z(*q) - f(xf)))
    orders = [math.log(errs[i] / errs[i + 1], 2) for i in range(2)]
    assert min(orders) > 4.5, (errs, orders)


def test_weno3z_third_order_on_smooth_averages():
    f = lambda x: math.sin(2 * math.pi * x)
    F = lambda x: -math.cos(2 * math.pi * x) / (2 * math.pi)
    errs = []
    for h in (0.01, 0.005, 0.0025):
        xf = 0.3
        q = _cell_averages(F, xf - 2 * h, h, 3)
        errs.append(abs(O.weno3z(*q) - f(xf)))
    orders = [math.log(errs[i] / errs[i + 1], 2) for i in range(2)]
    assert min(orders) > 2.5, (errs, orders)


@pytest.mark.parametrize("jump", [1.0, -4.0, 1e-3])
def test_non_oscillatory_jump(jump):
    """A jump between q0 and qp: the upwind smooth stencil dominates, the
    value stays at the upwind level (the linear scheme gives 0.4 jump)."""
    v = O.weno5z(0.0, 0.0, 0.0, jump, jump)
    assert abs(v) <= 1e-6 * abs(jump)
    assert abs(up5(0.0, 0.0, 0.0, jump, jump) - 0.4 * jump) < 1e-15
    # a jump upstream: value is the downstream constant level
    v = O.weno5z(0.0, jump, jump, jump, jump)
    assert abs(v - jump) <= 1e-6 * abs(jump)
    v3 = O.weno3z(0.0, 0.0, jump)
    assert abs(v3) <= 1e-3 * abs(jump)


def test_compflux_orders_select_stencils():
    """compflux_on_interval (weno.py:346-364): order 5/6 -> 5-point, 3/4 ->
    3-point, 1/2 -> 1-point upwind, 0 -> zero flux; flux = value * U."""
    rng = np.random.default_rng(7)
    n, s = 40, 1
    q = rng.standard_normal(n)
    U = rng.standard_normal(n)
    o = rng.choice(np.array([0, 1, 2, 3, 4, 5, 6], dtype=np.int8), n)
    o[:3] = 0
    o[-3:] = 0
    flx = np.zeros(n)
    O.compflux(flx, U, q, o, s, "weno")
    for i in range(3, n - 3):
        u = U[i]
        if o[i] > 4:
            want = O.flux5("weno", u, *q[i - 3:i + 3]) * u
        elif o[i] > 2:
            want = O.flux3("weno", u, *q[i - 2:i + 2]) * u
        elif o[i] > 0:
            want = O.flux1("weno", u, *q[i - 1:i + 1]) * u
        else:
            want = 0.0
        assert flx[i] == want, i
    # 1-point upwind is the donor cell
    i = 10
    o1 = np.zeros(n, dtype=np.int8)
    o1[i] = 1
    f1 = np.zeros(n)
    O.compflux(f1, U, q, o1, s, "weno")
    donor = q[i - 1] if U[i] > 0 else q[i]
    assert f1[i] == donor * U[i]
