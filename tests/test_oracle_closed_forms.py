"""Closed-form and cross-model pins for the oracle functions the reference's golden vectors
leave open (no reference counterpart, or no fixture):

* model 3 (first-order phase model, m = 0): its RHS equals model 4's coupling slot with
  xi = 0 (same neighbour sets and sums) for every method, and the two-particle closed form
  sigma*sin(phi_j - phi_i - alpha) + omega_i;
* natural frequencies in the local branches (natfreq_local): k[3] gains exactly omega_i,
  also for isolated particles, where the reference-compatible mode leaves k[3] unwritten;
* steppers on isolated particles (no neighbours): RK4 amplification of omega
  1 + z + z^2/2 + z^3/6 + z^4/24, z = -xi*dt; Euler and SRK with the zero-filled derivative
  (Q1 of the reference) move omega by the noise term only;
* order parameter: R = 1 for identical phases, R = 0 for a balanced pair, rotation covariance.
Each check is chosen so that a dropped term, a wrong sign or a swapped operand fails it.
"""
import numpy as np
import pytest

import oracle
from oracle import pyoracle

METHODS = [1, 2, 3, 4, 5]  # global field, all pairs, linked list, Verlet, linked list + Verlet


def _state(n, s, seed, omega=True):
    rng = np.random.default_rng(seed)
    st = np.zeros((n, s))
    st[:, 0], st[:, 1] = rng.random(n), rng.random(n)
    st[:, 2] = rng.random(n) * 2 * np.pi
    if s == 4 and omega:
        st[:, 3] = rng.normal(0, 0.4, n)
    return st


@pytest.mark.parametrize("method", METHODS)
def test_model3_is_model4_coupling_with_zero_friction(method):
    n, rho = 400, 0.12 if method != 1 else 0.8
    s4 = _state(n, 4, 1)
    s3 = np.ascontiguousarray(s4[:, :3]).reshape(-1)
    natf = np.random.default_rng(2).normal(0, 0.3, n)
    kw = dict(n=n, v0=0.03, xi=0.0, sigma=1.7, rho=rho, alpha=0.25, natfreq_local=1)
    o4 = oracle.Oracle(model=4, **kw)
    o3 = oracle.Oracle(model=3, **kw)
    o4.set_natural_frequencies(natf)
    o3.set_natural_frequencies(natf)
    k4 = np.zeros(4 * n)
    k3 = np.zeros(3 * n)
    o4.evaluate_rhs(method, s4.reshape(-1), np.zeros(4 * n), k4, 0.0, 0.01)
    o3.evaluate_rhs(method, s3, np.zeros(3 * n), k3, 0.0, 0.01)
    k4 = k4.reshape(n, 4)
    k3 = k3.reshape(n, 3)
    assert np.array_equal(k3[:, :2], k4[:, :2])
    assert np.max(np.abs(k3[:, 2] - k4[:, 3])) <= 1e-14


@pytest.mark.parametrize("model", [3, 4])
def test_two_particle_closed_form(model):
    sigma, alpha, xi, v0 = 1.3, 0.4, 0.7, 0.05
    st = np.zeros((2, model))
    st[0, :3] = (0.50, 0.50, 0.3)
    st[1, :3] = (0.53, 0.49, 1.9)  # distance ~0.0316 < rho
    if model == 4:
        st[:, 3] = (0.2, -0.1)
    natf = np.array([0.05, -0.02])
    o = oracle.Oracle(model=model, n=2, v0=v0, xi=xi, sigma=sigma, rho=0.05, alpha=alpha, natfreq_local=1)
    o.set_natural_frequencies(natf)
    k = np.zeros(2 * model)
    o.evaluate_rhs(2, st.reshape(-1), np.zeros(2 * model), k, 0.0, 0.01)
    k = k.reshape(2, model)
    for i, j in ((0, 1), (1, 0)):
        c = sigma * np.sin(st[j, 2] - st[i, 2] - alpha)
        assert abs(k[i, 0] - v0 * np.cos(st[i, 2])) <= 1e-15
        assert abs(k[i, 1] - v0 * np.sin(st[i, 2])) <= 1e-15
        if model == 3:
            assert abs(k[i, 2] - (natf[i] + c)) <= 1e-15
        else:
            assert k[i, 2] == st[i, 3]
            assert abs(k[i, 3] - (-xi * st[i, 3] + natf[i] + c)) <= 1e-15


def test_natfreq_local_adds_omega_and_fills_isolated():
    n = 300
    st = _state(n, 4, 5)
    st[0, :2] = (0.5, 0.5)
    st[1:, 0] = np.where(np.hypot(st[1:, 0] - 0.5, st[1:, 1] - 0.5) < 0.1, st[1:, 0] + 0.3, st[1:, 0])
    natf = np.random.default_rng(6).normal(0, 0.3, n)
    kw = dict(model=4, n=n, v0=0.02, xi=0.9, sigma=1.1, rho=0.06, alpha=0.1)
    k_on, k_off, k_zero = (np.full(4 * n, 7.0) for _ in range(3))
    on = oracle.Oracle(natfreq_local=1, **kw)
    on.set_natural_frequencies(natf)
    zero = oracle.Oracle(natfreq_local=1, **kw)
    zero.set_natural_frequencies(np.zeros(n))
    off = oracle.Oracle(natfreq_local=0, **kw)
    off.set_natural_frequencies(natf)
    for o, k in ((on, k_on), (zero, k_zero), (off, k_off)):
        o.evaluate_rhs(3, st.reshape(-1), np.zeros(4 * n), k, 0.0, 0.01)
    k_on, k_zero, k_off = (k.reshape(n, 4) for k in (k_on, k_zero, k_off))
    assert np.max(np.abs(k_on[:, 3] - k_zero[:, 3] - natf)) <= 1e-14
    # particle 0 is isolated: compatible mode leaves k[3] as it was (7.0), local mode writes -xi*omega + nat
    assert k_off[0, 3] == 7.0
    assert abs(k_on[0, 3] - (-kw["xi"] * st[0, 3] + natf[0])) <= 1e-15


def _isolated(n, s=4):
    st = np.zeros((n, s))
    g = int(np.ceil(np.sqrt(n)))
    st[:, 0] = (np.arange(n) % g + 0.5) / g
    st[:, 1] = (np.arange(n) // g + 0.5) / g
    st[:, 2] = np.linspace(0, 6, n)
    if s == 4:
        st[:, 3] = np.linspace(-1, 1, n)
    return st


def test_rk4_isolated_amplification():
    n, xi, dt = 16, 1.0, 0.01
    st = _isolated(n)
    o = oracle.Oracle(model=4, n=n, v0=0.0, xi=xi, sigma=1.0, rho=0.01, natfreq_local=1)
    s = st.reshape(-1).copy()
    o.rk4_step(2, s, dt)
    z = -xi * dt
    amp = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    assert abs(amp - 0.99004983375) < 1e-11
    assert np.max(np.abs(s.reshape(n, 4)[:, 3] - amp * st[:, 3])) <= 1e-15
    # phase integrates omega with the same Butcher weights: dt*(1 + z/2 + z^2/6 + z^3/24)*omega
    dphi = dt * (1 + z / 2 + z * z / 6 + z ** 3 / 24) * st[:, 3]
    assert np.max(np.abs(s.reshape(n, 4)[:, 2] - st[:, 2] - dphi)) <= 1e-15


def test_euler_and_srk_isolated_follow_the_noise_only():
    n, xi, dt, D = 16, 0.8, 0.01, 0.05
    st = _isolated(n)
    z = np.random.default_rng(9).normal(0, 1, 2 * n)
    o = oracle.Oracle(model=4, n=n, v0=0.0, xi=xi, sigma=1.0, rho=0.01, D_phi=D)
    s = st.reshape(-1).copy()
    o.euler_step(2, s, dt, z[:n])
    s = s.reshape(n, 4)
    # zero-filled derivative: omega_dot = 0 for n_i = 0 (not -xi*omega)
    assert np.max(np.abs(s[:, 3] - (st[:, 3] + np.sqrt(2 * D) * np.sqrt(dt) * z[:n]))) <= 1e-15
    assert np.max(np.abs(s[:, 2] - (st[:, 2] + st[:, 3] * dt))) <= 1e-15
    o2 = oracle.Oracle(model=4, n=n, v0=0.0, xi=xi, sigma=1.0, rho=0.01, D_phi=D)
    s2 = st.reshape(-1).copy()
    o2.srk_step(2, s2, dt, z)
    s2 = s2.reshape(n, 4)
    u1, u2 = z[0::2], z[1::2]
    dW = u1 * np.sqrt(dt)
    dZ = 0.5 * dt ** 1.5 * (u1 + u2 / np.sqrt(3.0))
    sq = np.sqrt(2 * D)
    assert np.max(np.abs(s2[:, 3] - (st[:, 3] + sq * (dW - xi * dZ)))) <= 1e-15
    assert np.max(np.abs(s2[:, 2] - (st[:, 2] + st[:, 3] * dt + sq * dZ))) <= 1e-15


def test_order_parameter_closed_forms():
    n = 1000
    st = _state(n, 4, 3)
    st[:, 2] = 1.234
    re, im = pyoracle.order_parameter(st.reshape(-1), n, 4, 2)
    assert abs(np.hypot(re, im) - 1.0) <= 1e-15 and abs(np.arctan2(im, re) - 1.234) <= 1e-14
    st[::2, 2] = 0.4
    st[1::2, 2] = 0.4 + np.pi
    re, im = pyoracle.order_parameter(st.reshape(-1), n, 4, 2)
    assert np.hypot(re, im) <= 1e-15
    rng = np.random.default_rng(4)
    st[:, 2] = rng.random(n) * 2 * np.pi
    re, im = pyoracle.order_parameter(st.reshape(-1), n, 4, 2)
    st[:, 2] += 0.7
    re2, im2 = pyoracle.order_parameter(st.reshape(-1), n, 4, 2)
    rot = complex(re, im) * np.exp(0.7j)
    assert abs(re2 - rot.real) <= 1e-14 and abs(im2 - rot.imag) <= 1e-14
