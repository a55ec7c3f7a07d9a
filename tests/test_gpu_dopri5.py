"""The adaptive Dormand-Prince integration with dense output (spss_dopri5_integrate, the
reference's Boost.Odeint integrate_const / integrate_adaptive) against the restatement in
oracle/dopri5_oracle.py driven by the oracle's FP64 RHS, and against the closed form of an
uncoupled system.

Device and oracle run the same algorithm on RHS evaluations that agree to ~1e-13 relative, so
they take the same steps (counts compared exactly) and their states agree to 1e-10 relative.
The systems are dense enough that no particle is without neighbours (the RHS leaves k[3] of such
a particle untouched, as the reference's does, and the oracle's fresh buffers would differ there).
"""
import numpy as np
import pytest

from oracle import dopri5_oracle as D

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def spss():
    import spsslangevinintegration_amd as m
    return m


def _state(model, n, seed):
    rng = np.random.default_rng(seed)
    st = np.zeros((n, model))
    st[:, 0], st[:, 1] = rng.random(n), rng.random(n)
    st[:, 2] = rng.random(n) * 2 * np.pi
    if model == 4:
        st[:, 3] = rng.normal(0, 0.5, n)
    return st.reshape(-1)


def _oracle_rhs(o, method, m):
    zeros = np.zeros(m)

    def rhs(x):
        out = np.zeros(m)
        o.evaluate_rhs(method, np.ascontiguousarray(x, dtype=np.float64), zeros, out, 0.0, 0.0)
        return out
    return rhs


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


# Neighbour sets change as particles cross the interaction radius, and the controller shortens
# the step around each such jump of the RHS: the cases keep the oracle's share to seconds.
CASES = [
    # model, n, rho, method, mode, t1, dt, eps
    (4, 400, 0.2, 2, 0, 1.0, 0.1, 1e-6),
    (4, 600, 0.15, 3, 0, 0.15, 0.05, 1e-7),
    (3, 400, 0.18, 2, 0, 0.6, 0.2, 1e-6),
    (4, 300, 0.2, 3, 1, 0.15, 0.05, 1e-7),
]


@pytest.mark.parametrize("model,n,rho,method,mode,t1,dt,eps", CASES)
def test_dopri5_matches_oracle(spss, model, n, rho, method, mode, t1, dt, eps):
    import oracle
    print(f"[gpu] dopri5 model={model} n={n} rho={rho} method={method} mode={mode}", flush=True)
    kw = dict(model=model, n=n, v0=0.1, xi=0.7, sigma=1.3, rho=rho, alpha=0.3, D_phi=0.0)
    st = _state(model, n, n + model)
    o = oracle.Oracle(**kw)
    g = spss.ParticleSystem(**kw)
    g.set_state(st)
    seen = []
    steps, rejected = g.integrate_dopri5(0.0, t1, dt, eps, eps, method=method, mode=mode,
                                         observer=lambda t, x: seen.append((t, x)))
    integ = D.integrate_const if mode == 0 else D.integrate_adaptive
    want_final, want_seen, ost = integ(_oracle_rhs(o, method, model * n), st, 0.0, t1, dt, eps, eps)
    print(f"  steps {steps}/{ost.steps} rejected {rejected}/{ost.rejected} observations {len(seen)}", flush=True)
    assert (steps, rejected) == (ost.steps, ost.rejected)
    assert len(seen) == len(want_seen)
    for (t, x), (tw, xw) in zip(seen, want_seen):
        # observation times of integrate_const are t0 + j dt exactly; accepted-step times carry
        # the controller's step sizes, whose error norms differ in the last bits
        assert t == (tw if mode == 0 else pytest.approx(tw, rel=1e-12, abs=1e-15))
        assert _rel(x, xw) <= TOL, (t, _rel(x, xw))
    assert _rel(g.get_state(), want_final) <= TOL


def test_dopri5_closed_form_uncoupled(spss):
    """sigma = 0 with the global field: omega = omega0 e^{-xi t}, phi = phi0 + omega0 (1 - e^{-xi t})
    / xi exactly for every particle; x, y by quadrature for a subset"""
    from scipy import integrate
    n, xi, v0 = 3000, 0.7, 0.3
    print(f"[gpu] dopri5 closed form n={n}", flush=True)
    g = spss.ParticleSystem(model=4, n=n, v0=v0, xi=xi, sigma=0.0, rho=0.8)
    st = _state(4, n, 11)
    g.set_state(st)
    seen = []
    g.integrate_dopri5(0.0, 3.0, 0.25, 1e-10, 1e-10, method=spss.METHOD_GLOBAL_FIELD,
                       observer=lambda t, x: seen.append((t, x)))
    assert [t for t, _ in seen] == pytest.approx(list(0.25 * np.arange(13)), abs=1e-15)
    x0 = st.reshape(-1, 4)
    for t, x in seen[1:]:
        x = x.reshape(-1, 4)
        np.testing.assert_allclose(x[:, 3], x0[:, 3] * np.exp(-xi * t), rtol=0, atol=2e-9)
        np.testing.assert_allclose(x[:, 2], x0[:, 2] + x0[:, 3] * (1 - np.exp(-xi * t)) / xi, rtol=0, atol=2e-9)
    t, x = seen[-1]
    x = x.reshape(-1, 4)
    for i in range(0, n, 150):
        px, py, ph, om = x0[i]
        phi = lambda s: ph + om * (1 - np.exp(-xi * s)) / xi  # noqa: E731
        q = dict(epsabs=1e-13, epsrel=1e-13, limit=200)
        assert abs(x[i, 0] - (px + v0 * integrate.quad(lambda s: np.cos(phi(s)), 0, t, **q)[0])) < 2e-9
        assert abs(x[i, 1] - (py + v0 * integrate.quad(lambda s: np.sin(phi(s)), 0, t, **q)[0])) < 2e-9
    np.testing.assert_array_equal(g.get_state(), seen[-1][1])


def test_dopri5_observer_stop_and_errors(spss):
    g = spss.ParticleSystem(model=4, n=300, rho=0.2)
    g.set_state(_state(4, 300, 5))
    calls = []

    def stop_after_three(t, x):
        calls.append(t)
        return len(calls) >= 3

    with pytest.raises(spss.SpssError):
        g.integrate_dopri5(0.0, 1.0, 0.1, observer=stop_after_three)
    assert len(calls) == 3
    with pytest.raises(spss.SpssError):
        g.integrate_dopri5(0.0, 1.0, -0.1)
    with pytest.raises(spss.SpssError):
        g.integrate_dopri5(0.0, 1.0, 0.1, eps_abs=0.0, eps_rel=0.0)
    # an empty interval: no steps, the state unchanged
    before = g.get_state()
    assert g.integrate_dopri5(1.0, 1.0, 0.1) == (0, 0)
    np.testing.assert_array_equal(g.get_state(), before)
