"""Adapter giving the HIP library (through its C ABI) the method names of the
oracle, so tests/golden_replay.py can drive either one.  Host state arrays are
uploaded before and downloaded after every step (small test sizes only)."""
import numpy as np

import spsslangevinintegration_amd as spss


class GpuEngine:
    def __init__(self, **kw):
        self.sys = spss.ParticleSystem(**kw)
        self.n = kw["n"]
        self.s = kw["model"]

    def set_natural_frequencies(self, w):
        self.sys.set_natural_frequencies(w)

    def set_higher_order(self, on):
        self.sys.set_higher_order(on)

    def complementary_force(self):
        return self.sys.complementary_force()

    def evaluate_rhs(self, method, state, k_prev, k_next, k_coef, dt):
        self.sys.evaluate_rhs(method, state, k_prev, k_next, k_coef, dt)

    def rk4_step(self, method, state, dt):
        self.sys.set_state(state)
        self.sys.rk4_step(method, dt)
        self.sys.get_state(state)

    def euler_step(self, method, state, dt, normals):
        self.sys.set_state(state)
        self.sys.euler_step(method, dt, spss.NOISE_HOST, normals)
        self.sys.get_state(state)

    def srk_step(self, method, state, dt, normals):
        self.sys.set_state(state)
        self.sys.srk_step(method, dt, spss.NOISE_HOST, normals)
        self.sys.get_state(state)

    def verlet_lists(self):
        return self.sys.verlet_lists()

    def wrap(self, state, n, s):
        self.sys.set_state(state)
        self.sys.apply_periodic_boundary_conditions()
        self.sys.get_state(state)


def factory(**kw):
    return GpuEngine(**kw)


def replay_wrap_for(engine_holder):
    return lambda st, n, s: engine_holder[0].wrap(st, n, s)


def relerr_slots(got, want, s):
    return [float(np.max(np.abs(got[k::s] - want[k::s])) / max(np.max(np.abs(want[k::s])), 1e-300))
            for k in range(s)]
