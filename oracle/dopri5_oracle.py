"""Restatement of the reference's adaptive Dormand-Prince 5(4) integration (numpy).

TEST INFRASTRUCTURE ONLY: the checker of spss_dopri5_integrate in tests/ (the product never
imports it).

The reference integrates with Boost.Odeint (ControlEngines/SimulationEngineFirstOrder.cpp:158-165:
integrate_const(make_dense_output(1e-6, 1e-6, runge_kutta_dopri5<std::vector<Real>>()),
particle_system, system_state_, t_0, t_1, dt, boost::ref(binary_observer)); :273-281:
integrate_adaptive(make_dense_output(1e-10, 1e-10, ...), particle_system, system_state_, t_0,
t_1, dt)).  Boost is not part of the reference tree (an external dependency of the build, Boost
1.6x; Odeint's algorithms are unchanged across those releases), so this file restates its
published algorithm:

* runge_kutta_dopri5: the Dormand-Prince (1980) 5(4) pair with first-same-as-last derivative;
  stage and output states as Boost's scale_sum operations, 1.0 * x + (dt a_1) k_1 + ... added
  left to right; error estimate dt * sum_i (b_i - b^_i) k_i over k_1, k_3 .. k_7.
* controlled_runge_kutta with default_error_checker(eps_abs, eps_rel, a_x = 1, a_dxdt = 1):
  err = max_i |xerr_i| / (eps_abs + eps_rel (|x_old,i| + dt |dxdt_old,i|)); default_step_adjuster:
  rejected when err > 1 with dt *= max(0.9 err^(-1/3), 1/5) (error order 4); accepted otherwise,
  and when err < 0.5, dt *= 0.9 max(err, 5^-5)^(-1/5) (stepper order 5).
* dense_output_runge_kutta: trial steps until one is accepted (at most 500); calc_state with the
  continuous extension of Hairer's DOPRI5 dense output in Boost's polynomial form.
* integrate_const / integrate_adaptive for dense-output steppers, with odeint's time comparisons
  (a tolerance of one machine epsilon in the direction of dt).

The system is the reference functor (DynamicalSystems/ParticleSystemFirstOrder.cpp:114-120):
EvaluateRhs(state, zeros, derivative, 0, t); here any callable rhs(x) -> dxdt.

Pinned in tests/test_oracle_dopri5.py: the tableau and one step against scipy's RK45 (the same
published pair), the continuous extension against the order conditions (exact rational
arithmetic), hand-worked step-size adjustments, and integrations of an uncoupled system against
its closed form.  The step-size controller's agreement with Boost itself is parity unpinned
(Boost absent).
"""
import numpy as np

EPS = np.finfo(np.float64).eps

A21 = 1.0 / 5.0
A31, A32 = 3.0 / 40.0, 9.0 / 40.0
A41, A42, A43 = 44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0
A51, A52, A53, A54 = 19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0
A61, A62, A63, A64, A65 = 9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0
B1, B3, B4, B5, B6 = 35.0 / 384.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0
D1 = B1 - 5179.0 / 57600.0
D3 = B3 - 7571.0 / 16695.0
D4 = B4 - 393.0 / 640.0
D5 = B5 - (-92097.0 / 339200.0)
D6 = B6 - 187.0 / 2100.0
D7 = -1.0 / 40.0


def combine(x, terms):
    """1.0 * x + c_1 k_1 + c_2 k_2 + ..., left to right"""
    v = 1.0 * x
    for c, k in terms:
        v = v + c * k
    return v


def dopri5_step(rhs, x, dxdt, dt):
    """one Dormand-Prince step: (out, dxdt_out, xerr, (k3, k4, k5, k6))"""
    k2 = rhs(combine(x, [(dt * A21, dxdt)]))
    k3 = rhs(combine(x, [(dt * A31, dxdt), (dt * A32, k2)]))
    k4 = rhs(combine(x, [(dt * A41, dxdt), (dt * A42, k2), (dt * A43, k3)]))
    k5 = rhs(combine(x, [(dt * A51, dxdt), (dt * A52, k2), (dt * A53, k3), (dt * A54, k4)]))
    k6 = rhs(combine(x, [(dt * A61, dxdt), (dt * A62, k2), (dt * A63, k3), (dt * A64, k4), (dt * A65, k5)]))
    out = combine(x, [(dt * B1, dxdt), (dt * B3, k3), (dt * B4, k4), (dt * B5, k5), (dt * B6, k6)])
    dout = rhs(out)
    xerr = dt * D1 * dxdt
    for c, k in ((D3, k3), (D4, k4), (D5, k5), (D6, k6), (D7, dout)):
        xerr = xerr + (dt * c) * k
    return out, dout, xerr, (k3, k4, k5, k6)


def error_norm(x_old, dxdt_old, xerr, dt, eps_abs, eps_rel):
    return float(np.max(np.abs(xerr) / (eps_abs + eps_rel * (1.0 * np.abs(x_old) + (1.0 * dt) * np.abs(dxdt_old)))))


def decrease_step(dt, err):
    return dt * max(0.9 * err ** (-1.0 / 3.0), 0.2)


def increase_step(dt, err):
    if err < 0.5:
        err = max(5.0 ** -5.0, err)
        dt = dt * (0.9 * err ** (-1.0 / 5.0))
    return dt


def dense_coefficients(theta):
    """the continuous-extension weights (b1, b3, b4, b5, b6, b7)(theta)"""
    X1 = 5.0 * (2558722523.0 - 31403016.0 * theta) / 11282082432.0
    X3 = 100.0 * (882725551.0 - 15701508.0 * theta) / 32700410799.0
    X4 = 25.0 * (443332067.0 - 31403016.0 * theta) / 1880347072.0
    X5 = 32805.0 * (23143187.0 - 3489224.0 * theta) / 199316789632.0
    X6 = 55.0 * (29972135.0 - 7076736.0 * theta) / 822651844.0
    X7 = 10.0 * (7414447.0 - 829305.0 * theta) / 29380423.0
    tm1 = theta - 1.0
    tsq = theta * theta
    Aq = tsq * (3.0 - 2.0 * theta)
    Bq = tsq * tm1
    Cq = tsq * tm1 * tm1
    Dq = theta * tm1 * tm1
    return (Aq * B1 - Cq * X1 + Dq, Aq * B3 + Cq * X3, Aq * B4 - Cq * X4, Aq * B5 + Cq * X5, Aq * B6 - Cq * X6,
            Bq + Cq * X7)


class DenseDopri5:
    """dense_output_runge_kutta<controlled_runge_kutta<runge_kutta_dopri5>>"""

    def __init__(self, rhs, eps_abs, eps_rel):
        self.rhs, self.eps_abs, self.eps_rel = rhs, eps_abs, eps_rel
        self.steps = self.rejected = 0

    def initialize(self, x0, t0, dt0):
        self.x = np.array(x0, dtype=np.float64, copy=True)
        self.t, self.dt = float(t0), float(dt0)
        self.dxdt = None

    def do_step(self):
        if self.dxdt is None:
            self.dxdt = self.rhs(self.x)
        self.t_old = self.t
        for fails in range(501):
            if fails == 500:
                raise RuntimeError("500 failed step size adjustments")
            out, dout, xerr, ks = dopri5_step(self.rhs, self.x, self.dxdt, self.dt)
            err = error_norm(self.x, self.dxdt, xerr, self.dt, self.eps_abs, self.eps_rel)
            if err > 1.0:
                self.dt = decrease_step(self.dt, err)
                self.rejected += 1
                continue
            self.t += self.dt
            self.dt = increase_step(self.dt, err)
            break
        self.x_old, self.dxdt_old, self.ks = self.x, self.dxdt, ks
        self.x, self.dxdt = out, dout
        self.steps += 1

    def calc_state(self, t):
        dt = self.t - self.t_old
        b = dense_coefficients((t - self.t_old) / dt)
        k3, k4, k5, k6 = self.ks
        return combine(self.x_old, [(dt * b[0], self.dxdt_old), (dt * b[1], k3), (dt * b[2], k4), (dt * b[3], k5),
                                    (dt * b[4], k6), (dt * b[5], self.dxdt)])


def less_with_sign(t1, t2, dt):
    return t2 - t1 > EPS if dt > 0 else t1 - t2 > EPS


def less_eq_with_sign(t1, t2, dt):
    return t1 - t2 <= EPS if dt > 0 else t2 - t1 <= EPS


def integrate_const(rhs, x0, t0, t1, dt, eps_abs, eps_rel, obs=None):
    """returns (final state = last observed state, [(t, state)], stepper)"""
    st = DenseDopri5(rhs, eps_abs, eps_rel)
    seen = []

    def observe(x, t):
        seen.append((t, np.array(x)))
        if obs:
            obs(t, x)

    x = np.array(x0, dtype=np.float64, copy=True)
    if not less_with_sign(t0, t1, dt):
        return x, seen, st
    st.initialize(x, t0, dt)
    observe(x, t0)
    obs_step = 1
    time = t0 + obs_step * dt
    while less_eq_with_sign(time + dt, t1, dt):
        while less_eq_with_sign(time, st.t, dt):
            x = st.calc_state(time)
            observe(x, time)
            obs_step += 1
            time = t0 + obs_step * dt
        if less_with_sign(st.t + st.dt, t1, st.dt):
            while less_eq_with_sign(st.t, time, dt):
                st.do_step()
        elif less_with_sign(st.t, t1, st.dt):
            st.initialize(st.x, st.t, t1 - st.t)
            st.do_step()
    if less_eq_with_sign(time, t1, dt):
        x = st.calc_state(time)
        observe(x, time)
    return x, seen, st


def integrate_adaptive(rhs, x0, t0, t1, dt, eps_abs, eps_rel, obs=None):
    """returns (state at t1, [(t, state)], stepper)"""
    st = DenseDopri5(rhs, eps_abs, eps_rel)
    seen = []

    def observe(x, t):
        seen.append((t, np.array(x)))
        if obs:
            obs(t, x)

    st.initialize(x0, t0, dt)
    observe(st.x, t0)
    while less_with_sign(st.t, t1, st.dt):
        while less_eq_with_sign(st.t + st.dt, t1, st.dt):
            st.do_step()
            observe(st.x, st.t)
        st.initialize(st.x, st.t, t1 - st.t)
    return np.array(st.x), seen, st
