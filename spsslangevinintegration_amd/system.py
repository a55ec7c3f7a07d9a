"""Python view of one device-resident particle system (the C ABI behind it).

Method names follow the reference classes: EvaluateRhs, the three steppers'
DoStep, SetNaturalFrequencies, CalculateMaxDisplacement, ... in snake case.
"""
import ctypes as C
import os

import numpy as np

from . import _capi as A


class ParticleSystem:
    def __init__(self, model=4, n=0, v0=0.01, xi=1.0, sigma=1.0, rho=0.1, alpha=0.0, D_phi=0.0, xi_r=1.0,
                 D_r=0.0, x_size=1.0, y_size=1.0, verlet_skin=0.05, verlet_mode=A.VERLET_COMPAT,
                 natfreq_local=0, precision=A.PRECISION_FP64, cell_subdivision=1, seed=0x5EED0000):
        self.lib = A.load()
        p = A.SpssParams()
        A.check(self.lib.spss_params_default(C.byref(p)))
        for k, v in dict(model=model, n=n, v0=v0, xi=xi, sigma=sigma, rho=rho, alpha=alpha, D_phi=D_phi,
                         xi_r=xi_r, D_r=D_r, x_size=x_size, y_size=y_size, verlet_skin=verlet_skin,
                         verlet_mode=verlet_mode, natfreq_local=natfreq_local, precision=precision,
                         cell_subdivision=cell_subdivision, seed=seed).items():
            setattr(p, k, v)
        self.params = p
        self.s, self.n = model, n
        h = C.c_void_p()
        A.check(self.lib.spss_create(C.byref(p), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.spss_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- state
    def set_state(self, state):
        st = np.ascontiguousarray(state, dtype=np.float64)
        assert st.size == self.s * self.n
        A.check(self.lib.spss_set_state(self.h, A.ptr(st)))

    def get_state(self, out=None):
        out = np.empty(self.s * self.n) if out is None else out
        A.check(self.lib.spss_get_state(self.h, A.ptr(out)))
        return out

    def state_device_ptr(self):
        return self.lib.spss_state_device(self.h)

    def set_natural_frequencies(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        A.check(self.lib.spss_set_natural_frequencies(self.h, A.ptr(w), int(w.size)))

    def set_higher_order(self, on):
        A.check(self.lib.spss_set_higher_order(self.h, int(bool(on))))

    def complementary_force(self):
        out = np.empty(self.n)
        A.check(self.lib.spss_get_complementary_force(self.h, A.ptr(out)))
        return out

    # ---- RHS and steppers
    def evaluate_rhs(self, method, state, k_prev, k_next, k_coef, dt):
        A.check(self.lib.spss_evaluate_rhs(self.h, method, A.ptr(state), A.ptr(k_prev), A.ptr(k_next),
                                           k_coef, dt))

    def evaluate_rhs_device(self, method, d_state, d_kprev, d_knext, k_coef, dt):
        A.check(self.lib.spss_evaluate_rhs_device(self.h, method, d_state, d_kprev, d_knext, k_coef, dt))

    def rk4_step(self, method, dt):
        A.check(self.lib.spss_rk4_step(self.h, method, dt))

    def euler_step(self, method, dt, noise_mode=A.NOISE_NONE, noise=None):
        nz = None if noise is None else np.ascontiguousarray(noise, dtype=np.float64)
        A.check(self.lib.spss_euler_step(self.h, method, dt, noise_mode, A.ptr(nz)))

    def srk_step(self, method, dt, noise_mode=A.NOISE_NONE, noise=None):
        nz = None if noise is None else np.ascontiguousarray(noise, dtype=np.float64)
        A.check(self.lib.spss_srk_step(self.h, method, dt, noise_mode, A.ptr(nz)))

    def run(self, stepper, method, nsteps, dt, noise_mode=A.NOISE_NONE):
        A.check(self.lib.spss_run(self.h, stepper, method, int(nsteps), dt, noise_mode))

    def prepare_run(self, stepper, method, dt, noise_mode=A.NOISE_NONE):
        """capture the step graphs later runs of this stepper replay (FP32C Euler/SRK, Philox)"""
        A.check(self.lib.spss_prepare_run(self.h, stepper, method, dt, noise_mode))

    def run_stats(self, reset=False):
        """(eager steps, graph-replayed steps, graphs captured) since the last reset"""
        e, r, c = C.c_long(), C.c_long(), C.c_long()
        A.check(self.lib.spss_run_stats(self.h, C.byref(e), C.byref(r), C.byref(c), int(bool(reset))))
        return e.value, r.value, c.value

    @property
    def step(self):
        v = C.c_long()
        A.check(self.lib.spss_get_step(self.h, C.byref(v)))
        return v.value

    @step.setter
    def step(self, v):
        A.check(self.lib.spss_set_step(self.h, int(v)))

    def synchronize(self):
        A.check(self.lib.spss_synchronize(self.h))

    # ---- Verlet bookkeeping and neighbour sets
    def calculate_max_displacement(self, total):
        t = np.ascontiguousarray(total, dtype=np.float64)
        A.check(self.lib.spss_calculate_max_displacement(self.h, A.ptr(t)))

    def verlet_status(self):
        su, uv = C.c_int(), C.c_int()
        acc = C.c_double()
        A.check(self.lib.spss_verlet_status(self.h, C.byref(su), C.byref(acc), C.byref(uv)))
        return bool(su.value), acc.value, bool(uv.value)

    def verlet_lists(self):
        counts = np.zeros(self.n, dtype=np.int32)
        total = C.c_long()
        A.check(self.lib.spss_get_verlet_lists(self.h, A.ptr(counts), None, 0, C.byref(total)))
        flat = np.zeros(max(total.value, 1), dtype=np.int32)
        A.check(self.lib.spss_get_verlet_lists(self.h, A.ptr(counts), A.ptr(flat), total.value, C.byref(total)))
        out, pos = [], 0
        for c in counts:
            out.append(flat[pos:pos + c].copy())
            pos += c
        return out

    def neighbour_pairs(self, method, stage):
        st = np.ascontiguousarray(stage, dtype=np.float64)
        cnt = C.c_long()
        A.check(self.lib.spss_neighbour_pairs(self.h, method, A.ptr(st), None, 0, C.byref(cnt)))
        out = np.zeros(2 * max(cnt.value, 1), dtype=np.int32)
        A.check(self.lib.spss_neighbour_pairs(self.h, method, A.ptr(st), A.ptr(out), cnt.value, C.byref(cnt)))
        return out[: 2 * cnt.value].reshape(-1, 2)

    # ---- output files of the reference's observers, restart
    def frame(self, t):
        """the frame the reference writes: float t, then the s*n state values as floats
        (models 3/4 wrapped into the box and phi into [0, 2 pi))"""
        out = np.empty(1 + self.s * self.n, dtype=np.float32)
        A.check(self.lib.spss_frame(self.h, float(t), A.ptr(out)))
        return out

    def append_frame(self, path, t):
        A.check(self.lib.spss_frame_append(self.h, os.fsencode(path), float(t)))

    def summary(self):
        """(|Z|, arg Z) as floats, Z accumulated in complex<float> in particle order"""
        a, g = C.c_float(), C.c_float()
        A.check(self.lib.spss_summary(self.h, C.byref(a), C.byref(g)))
        return a.value, g.value

    def append_summary(self, path, t):
        A.check(self.lib.spss_summary_append(self.h, os.fsencode(path), float(t)))

    def load_last_frame(self, path):
        """restart from the last frame of a trajectory file; returns its time"""
        t = C.c_double()
        A.check(self.lib.spss_load_last_frame(self.h, os.fsencode(path), C.byref(t)))
        return t.value

    # ---- adaptive Dormand-Prince 5(4) with dense output
    def integrate_dopri5(self, t0, t1, dt, eps_abs=1e-6, eps_rel=1e-6, method=None, mode=None, observer=None):
        """the reference's Boost.Odeint integration (SimulationEngineFirstOrder.cpp:158-165 and
        :273-281): mode DOPRI5_CONST observes t0 + j*dt from the dense output and ends at the last
        observed state, DOPRI5_ADAPTIVE observes every accepted step and ends at t1.  observer(t,
        state) gets a host copy of the s*n state; returning True stops the integration.
        Returns (accepted steps, rejected trials)."""
        method = A.METHOD_AUTO if method is None else method
        mode = A.DOPRI5_CONST if mode is None else mode
        cb = A.OBSERVER_FN(0)
        if observer is not None:
            def _obs(user, t, state, n, s):
                try:
                    return 1 if observer(t, np.ctypeslib.as_array(state, shape=(n * s,)).copy()) else 0
                except Exception:
                    return 1
            cb = A.OBSERVER_FN(_obs)
        steps, rejected = C.c_long(), C.c_long()
        A.check(self.lib.spss_dopri5_integrate(self.h, int(method), int(mode), float(t0), float(t1), float(dt),
                                               float(eps_abs), float(eps_rel), cb, None, 1, C.byref(steps),
                                               C.byref(rejected)))
        return steps.value, rejected.value

    # ---- cell-size autotuner (FP32C)
    def autotune_cells(self, stepper, dt, probe_steps=10):
        """times the FP32C step for every cell layout of CELL_LAYOUTS and keeps the fastest; returns
        (ms per step for each, -1 where the grid is too fine; the chosen subdivision; cell_layout
        gives the chosen (subdivision, x split))"""
        table = np.zeros(len(A.CELL_LAYOUTS))
        k = C.c_int()
        A.check(self.lib.spss_autotune_cells(self.h, int(stepper), float(dt), int(probe_steps), A.ptr(table),
                                             table.size, C.byref(k)))
        return table, k.value

    def set_cell_layout(self, k, xsplit):
        """FP32C cells 1/(k*xsplit*int(1/rho)) wide and 1/(k*int(1/rho)) high (the state is kept)"""
        A.check(self.lib.spss_set_cell_layout(self.h, int(k), int(xsplit)))

    @property
    def cell_layout(self):
        """(subdivision, x split) of the FP32C cell grid"""
        k, x = C.c_int(), C.c_int()
        A.check(self.lib.spss_get_cell_layout(self.h, C.byref(k), C.byref(x)))
        return k.value, x.value

    def autotune_all_pairs(self, probe_evals=3):
        """times the all-pairs kernel's workgroup / j-tile shapes on the current state and keeps the
        fastest; returns (ms per evaluation for each shape, chosen index)"""
        table = np.zeros(6)
        chosen = C.c_int()
        A.check(self.lib.spss_autotune_all_pairs(self.h, int(probe_evals), A.ptr(table), 6, C.byref(chosen)))
        return table, chosen.value

    def set_all_pairs_config(self, index):
        A.check(self.lib.spss_set_all_pairs_config(self.h, int(index)))

    @property
    def all_pairs_config(self):
        """(particles per workgroup, j per LDS tile) of the all-pairs kernel"""
        b, j = C.c_int(), C.c_int()
        A.check(self.lib.spss_get_all_pairs_config(self.h, C.byref(b), C.byref(j)))
        return b.value, j.value

    @property
    def cell_subdivision(self):
        k = C.c_int()
        A.check(self.lib.spss_get_cell_subdivision(self.h, C.byref(k)))
        return k.value

    # ---- instrumentation
    def kernel_timer_start(self):
        A.check(self.lib.spss_kernel_timer_start(self.h))

    def kernel_timer_stop(self):
        ms, cnt = C.c_double(), C.c_long()
        A.check(self.lib.spss_kernel_timer_stop(self.h, C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    # ---- observers
    def order_parameter(self, slot=None):
        slot = (4 if self.s == 6 else 2) if slot is None else slot
        re, im = C.c_double(), C.c_double()
        A.check(self.lib.spss_order_parameter(self.h, slot, C.byref(re), C.byref(im)))
        return re.value, im.value

    def apply_periodic_boundary_conditions(self):
        A.check(self.lib.spss_apply_periodic_boundary_conditions(self.h))

    def averaging_reset(self):
        A.check(self.lib.spss_averaging_reset(self.h))

    def averaging_accumulate(self):
        A.check(self.lib.spss_averaging_accumulate(self.h))

    def averaging_get(self):
        out = np.empty(self.s * self.n)
        cnt = C.c_long()
        A.check(self.lib.spss_averaging_get(self.h, A.ptr(out), C.byref(cnt)))
        return out, cnt.value
