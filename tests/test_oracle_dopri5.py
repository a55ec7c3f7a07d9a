"""Pins of the Dormand-Prince restatement (oracle/dopri5_oracle.py).

* the tableau, one step and its error estimate against scipy's RK45, which implements the same
  published Dormand-Prince 5(4) pair (a transposed or dropped coefficient fails it);
* the continuous extension against the order conditions of a fourth-order interpolant, in exact
  rational arithmetic (a wrong digit in any of its constants fails it), plus its end values;
* the step-size adjustments on hand-worked errors;
* integrate_const / integrate_adaptive on an uncoupled system against its closed form
  (omega = omega0 e^{-xi t}, phi = phi0 + omega0 (1 - e^{-xi t}) / xi, x and y by quadrature),
  with the observation times odeint's integrate_const produces.
"""
from fractions import Fraction as F

import numpy as np
import pytest
from scipy import integrate

from oracle import dopri5_oracle as D


def test_tableau_matches_scipy_rk45():
    rk = integrate.RK45
    ours_a = [[D.A21], [D.A31, D.A32], [D.A41, D.A42, D.A43], [D.A51, D.A52, D.A53, D.A54],
              [D.A61, D.A62, D.A63, D.A64, D.A65]]
    for i, row in enumerate(ours_a, start=1):
        np.testing.assert_allclose(rk.A[i, :len(row)], row, rtol=1e-15, atol=0)
    np.testing.assert_allclose(rk.B, [D.B1, 0.0, D.B3, D.B4, D.B5, D.B6], rtol=1e-15, atol=0)
    # scipy's E is (fourth-order minus fifth-order) weights over k1..k7: ours is minus that
    np.testing.assert_allclose(-np.asarray(rk.E), [D.D1, 0.0, D.D3, D.D4, D.D5, D.D6, D.D7], rtol=1e-13, atol=1e-17)


def _fun(t, y):  # a smooth nonlinear test system
    return np.array([np.cos(y[1]) * 0.3, y[2], -0.7 * y[2] + 0.2 * np.sin(y[0] - y[1]), 0.1 * y[0]])


def test_one_step_matches_scipy_rk_step():
    from scipy.integrate._ivp.rk import rk_step
    rk = integrate.RK45
    y = np.array([0.3, -1.2, 0.5, 2.0])
    f = _fun(0.0, y)
    h = 0.137
    K = np.empty((rk.n_stages + 1, y.size))
    y_new, f_new = rk_step(_fun, 0.0, y, f, h, rk.A, rk.B, rk.C, K)
    out, dout, xerr, _ = D.dopri5_step(lambda x: _fun(0.0, x), y, f, h)
    np.testing.assert_allclose(out, y_new, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(dout, f_new, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(xerr, -(K.T @ rk.E) * h, rtol=1e-9, atol=1e-18)


# ---- the continuous extension: exact rational checks of the order conditions
_A = {(2, 1): F(1, 5), (3, 1): F(3, 40), (3, 2): F(9, 40), (4, 1): F(44, 45), (4, 2): F(-56, 15),
      (4, 3): F(32, 9), (5, 1): F(19372, 6561), (5, 2): F(-25360, 2187), (5, 3): F(64448, 6561),
      (5, 4): F(-212, 729), (6, 1): F(9017, 3168), (6, 2): F(-355, 33), (6, 3): F(46732, 5247),
      (6, 4): F(49, 176), (6, 5): F(-5103, 18656)}
_B5 = {1: F(35, 384), 2: F(0), 3: F(500, 1113), 4: F(125, 192), 5: F(-2187, 6784), 6: F(11, 84), 7: F(0)}
for _j in range(1, 7):
    _A[(7, _j)] = _B5[_j]  # first same as last: stage 7 is the new point
_C = {i: sum((_A.get((i, j), F(0)) for j in range(1, i)), F(0)) for i in range(1, 8)}


def _dense_exact(theta):
    """Boost's polynomial form of the weights, in exact arithmetic"""
    t = F(theta)
    X1 = 5 * (F(2558722523) - 31403016 * t) / 11282082432
    X3 = 100 * (F(882725551) - 15701508 * t) / 32700410799
    X4 = 25 * (F(443332067) - 31403016 * t) / 1880347072
    X5 = 32805 * (F(23143187) - 3489224 * t) / 199316789632
    X6 = 55 * (F(29972135) - 7076736 * t) / 822651844
    X7 = 10 * (F(7414447) - 829305 * t) / 29380423
    Aq, Bq, Cq, Dq = t * t * (3 - 2 * t), t * t * (t - 1), t * t * (t - 1) ** 2, t * (t - 1) ** 2
    return {1: Aq * _B5[1] - Cq * X1 + Dq, 2: F(0), 3: Aq * _B5[3] + Cq * X3, 4: Aq * _B5[4] - Cq * X4,
            5: Aq * _B5[5] + Cq * X5, 6: Aq * _B5[6] - Cq * X6, 7: Bq + Cq * X7}


@pytest.mark.parametrize("theta", [F(0), F(1, 7), F(1, 3), F(1, 2), F(5, 6), F(1)])
def test_dense_output_order_conditions(theta):
    b = _dense_exact(theta)
    rng = range(1, 8)
    a = lambda i, j: _A.get((i, j), F(0))  # noqa: E731
    t = theta
    assert sum(b[i] for i in rng) == t
    assert sum(b[i] * _C[i] for i in rng) == t ** 2 / 2
    assert sum(b[i] * _C[i] ** 2 for i in rng) == t ** 3 / 3
    assert sum(b[i] * a(i, j) * _C[j] for i in rng for j in rng) == t ** 3 / 6
    assert sum(b[i] * _C[i] ** 3 for i in rng) == t ** 4 / 4
    assert sum(b[i] * _C[i] * a(i, j) * _C[j] for i in rng for j in rng) == t ** 4 / 8
    assert sum(b[i] * a(i, j) * _C[j] ** 2 for i in rng for j in rng) == t ** 4 / 12
    assert sum(b[i] * a(i, j) * a(j, k) * _C[k] for i in rng for j in rng for k in rng) == t ** 4 / 24
    if theta == 1:  # the end of the step is the fifth-order point
        assert all(b[i] == _B5[i] for i in rng)
    # the float form the oracle (and the device) evaluates
    fl = D.dense_coefficients(float(theta))
    ex = [b[1], b[3], b[4], b[5], b[6], b[7]]
    np.testing.assert_allclose(fl, [float(v) for v in ex], rtol=0, atol=1e-15)


def test_step_size_adjustments():
    assert D.decrease_step(1.0, 8.0) == pytest.approx(0.45, rel=1e-15)   # 0.9 * 8^(-1/3)
    assert D.decrease_step(1.0, 1e9) == 0.2                                 # floor 1/5
    assert D.increase_step(1.0, 0.7) == 1.0                                 # 0.5 <= err <= 1: unchanged
    assert D.increase_step(2.0, 0.25) == pytest.approx(2.0 * 0.9 * 4.0 ** 0.2, rel=1e-15)
    assert D.increase_step(1.0, 1e-12) == pytest.approx(4.5, rel=1e-14)   # err floored at 5^-5


def _uncoupled(xi=0.7, v0=0.3):
    def rhs(state):
        x = state.reshape(-1, 4)
        out = np.empty_like(x)
        out[:, 0] = v0 * np.cos(x[:, 2])
        out[:, 1] = v0 * np.sin(x[:, 2])
        out[:, 2] = x[:, 3]
        out[:, 3] = -xi * x[:, 3]
        return out.reshape(-1)

    def exact(x0, t):
        x0 = x0.reshape(-1, 4)
        res = np.empty_like(x0)
        for i, (px, py, ph, om) in enumerate(x0):
            phi = lambda s: ph + om * (1 - np.exp(-xi * s)) / xi  # noqa: E731
            q = dict(epsabs=1e-13, epsrel=1e-13, limit=200)
            res[i, 0] = px + v0 * integrate.quad(lambda s: np.cos(phi(s)), 0, t, **q)[0]
            res[i, 1] = py + v0 * integrate.quad(lambda s: np.sin(phi(s)), 0, t, **q)[0]
            res[i, 2] = phi(t)
            res[i, 3] = om * np.exp(-xi * t)
        return res.reshape(-1)
    return rhs, exact


def test_integrate_const_against_closed_form():
    rhs, exact = _uncoupled()
    rng = np.random.default_rng(4)
    x0 = np.column_stack([rng.random(5), rng.random(5), rng.uniform(0, 6.28, 5), rng.normal(0, 2.0, 5)]).reshape(-1)
    final, seen, st = D.integrate_const(rhs, x0, 0.0, 3.0, 0.25, 1e-10, 1e-10)
    times = [t for t, _ in seen]
    np.testing.assert_allclose(times, 0.25 * np.arange(13), rtol=0, atol=1e-15)
    for t, x in seen[1:]:
        np.testing.assert_allclose(x, exact(x0, t), rtol=0, atol=5e-9)
    np.testing.assert_array_equal(final, seen[-1][1])
    assert st.steps > 0


def test_integrate_adaptive_against_closed_form():
    rhs, exact = _uncoupled()
    x0 = np.array([0.2, 0.4, 1.0, 1.5, 0.9, 0.1, 4.0, -2.5])
    final, seen, st = D.integrate_adaptive(rhs, x0, 0.0, 2.0, 0.1, 1e-9, 1e-9)
    assert abs(seen[-1][0] - 2.0) < 1e-14
    np.testing.assert_allclose(final, exact(x0, 2.0), rtol=0, atol=5e-8)
    # the loose tolerance takes fewer, longer steps and is less accurate
    final6, _, st6 = D.integrate_adaptive(rhs, x0, 0.0, 2.0, 0.1, 1e-6, 1e-6)
    assert st6.steps < st.steps
    assert np.max(np.abs(final6 - exact(x0, 2.0))) > np.max(np.abs(final - exact(x0, 2.0)))
