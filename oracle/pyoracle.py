"""ctypes bindings to the C oracle (oracle/spss_oracle.c).

TEST INFRASTRUCTURE ONLY: used as the checker by tests/, smoke() and the CPU
baseline in bench.py.  Arrays are numpy float64 in the reference's AoS layout.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

METHOD_AUTO, METHOD_GLOBAL_FIELD, METHOD_ALL_PAIRS, METHOD_LINKED_LIST, METHOD_VERLET, \
    METHOD_LINKED_LIST_VERLET = range(6)
VERLET_COMPAT, VERLET_EXACT = 0, 1


class OracleParams(C.Structure):
    _fields_ = [
        ("model", C.c_int), ("n", C.c_int),
        ("v0", C.c_double), ("xi", C.c_double), ("sigma", C.c_double), ("rho", C.c_double),
        ("alpha", C.c_double), ("D_phi", C.c_double), ("xi_r", C.c_double), ("D_r", C.c_double),
        ("x_size", C.c_double), ("y_size", C.c_double), ("verlet_skin", C.c_double),
        ("verlet_mode", C.c_int), ("natfreq_local", C.c_int),
    ]


def load_library(build=True):
    """Load oracle/liboracle.so, building it with make if it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path) and build:
        subprocess.run(["make", "-s", "-C", _HERE, "liboracle"], check=True)
    lib = C.CDLL(path)
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    d = C.c_double
    lib.orc_create.restype = C.c_void_p
    lib.orc_create.argtypes = [C.POINTER(OracleParams)]
    lib.orc_destroy.argtypes = [C.c_void_p]
    lib.orc_set_natural_frequencies.argtypes = [C.c_void_p, dp]
    lib.orc_set_higher_order.argtypes = [C.c_void_p, C.c_int]
    lib.orc_set_threads.argtypes = [C.c_void_p, C.c_int]
    lib.orc_uses_verlet.argtypes = [C.c_void_p]
    lib.orc_should_update_lists.argtypes = [C.c_void_p]
    lib.orc_accumulated_displacement.argtypes = [C.c_void_p]
    lib.orc_accumulated_displacement.restype = d
    lib.orc_num_cells_x.argtypes = [C.c_void_p]
    lib.orc_complementary_force.argtypes = [C.c_void_p]
    lib.orc_complementary_force.restype = C.POINTER(C.c_double)
    pd = C.POINTER(C.c_double)
    for nm in ("orc_pbc_class_a", "orc_pbc_class_b_signed", "orc_pbc_class_b_unsigned",
               "orc_pbc_class_c_signed", "orc_pbc_class_c_unsigned"):
        getattr(lib, nm).argtypes = [d, d, d, d, d, d, pd, pd]
    lib.orc_pbc_apply.argtypes = [d, d, d, d, pd, pd]
    lib.orc_pbc_apply_state.argtypes = [d, d, dp, C.c_int, C.c_int]
    lib.orc_evaluate_rhs.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, d, d]
    lib.orc_calculate_max_displacement.argtypes = [C.c_void_p, dp]
    lib.orc_rk4_step.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, dp, dp, dp, d]
    lib.orc_euler_step.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, d, dp]
    lib.orc_srk_step.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, d, dp]
    lib.orc_pairs_bruteforce.restype = C.c_long
    lib.orc_pairs_bruteforce.argtypes = [dp, C.c_int, C.c_int, d, d, d, ip, C.c_long]
    lib.orc_pairs_linked_list.restype = C.c_long
    lib.orc_pairs_linked_list.argtypes = [C.c_void_p, dp, ip, C.c_long]
    lib.orc_verlet_lists.restype = C.c_long
    lib.orc_verlet_lists.argtypes = [C.c_void_p, ip, ip, C.c_long]
    lib.orc_order_parameter.argtypes = [dp, C.c_int, C.c_int, C.c_int, pd, pd]
    lib.orc_order_parameter_ref_float.argtypes = [dp, C.c_int, C.c_int, C.c_int,
                                                  C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.orc_averaging_accumulate.argtypes = [dp, dp, C.c_long]
    lib.orc_polar_fields.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_int, pd, pd, pd, pd]
    _LIB = lib
    return lib


class _Pbc:
    """Scalar PBC routines (PeriodicBoundaryConditions.cpp)."""

    def _call(self, name, xs, ys, xi, yi, xj, yj):
        lib = load_library()
        dx, dy = C.c_double(), C.c_double()
        getattr(lib, name)(xs, ys, xi, yi, xj, yj, C.byref(dx), C.byref(dy))
        return dx.value, dy.value

    def class_a(self, *a):
        return self._call("orc_pbc_class_a", *a)

    def class_b_signed(self, *a):
        return self._call("orc_pbc_class_b_signed", *a)

    def class_b_unsigned(self, *a):
        return self._call("orc_pbc_class_b_unsigned", *a)

    def class_c_signed(self, *a):
        return self._call("orc_pbc_class_c_signed", *a)

    def class_c_unsigned(self, *a):
        return self._call("orc_pbc_class_c_unsigned", *a)

    def apply(self, xs, ys, x, y):
        lib = load_library()
        a, b = C.c_double(), C.c_double()
        lib.orc_pbc_apply(xs, ys, x, y, C.byref(a), C.byref(b))
        return a.value, b.value


pbc = _Pbc()


class Oracle:
    """One particle system of the oracle with its persistent lists and flags."""

    def __init__(self, model=4, n=0, v0=0.01, xi=1.0, sigma=1.0, rho=0.1, alpha=0.0, D_phi=0.0,
                 xi_r=1.0, D_r=0.0, verlet_skin=0.05, verlet_mode=VERLET_COMPAT, natfreq_local=0,
                 x_size=1.0, y_size=1.0):
        self.lib = load_library()
        self.p = OracleParams(model, n, v0, xi, sigma, rho, alpha, D_phi, xi_r, D_r, x_size, y_size,
                              verlet_skin, verlet_mode, natfreq_local)
        self.s = model
        self.n = n
        self.h = self.lib.orc_create(C.byref(self.p))
        m = model * n
        self.k = [np.zeros(m) for _ in range(4)]
        self.total = np.zeros(m)
        self.deriv = np.zeros(m)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def set_natural_frequencies(self, w):
        self.lib.orc_set_natural_frequencies(self.h, np.ascontiguousarray(w, dtype=np.float64))

    def set_higher_order(self, on):
        self.lib.orc_set_higher_order(self.h, int(on))

    def set_threads(self, threads):
        """worker threads of the linked-list particle loop (results equal the serial loop's)"""
        self.lib.orc_set_threads(self.h, int(threads))

    @property
    def should_update_lists(self):
        return bool(self.lib.orc_should_update_lists(self.h))

    @property
    def accumulated_displacement(self):
        return self.lib.orc_accumulated_displacement(self.h)

    @property
    def num_cells_x(self):
        return self.lib.orc_num_cells_x(self.h)

    def complementary_force(self):
        ptr = self.lib.orc_complementary_force(self.h)
        return np.ctypeslib.as_array(ptr, shape=(self.n,)).copy()

    def evaluate_rhs(self, method, state, k_prev, k_next, k_coef, dt):
        self.lib.orc_evaluate_rhs(self.h, method, state, k_prev, k_next, k_coef, dt)

    def rk4_step(self, method, state, dt):
        self.lib.orc_rk4_step(self.h, method, state, self.k[0], self.k[1], self.k[2], self.k[3],
                              self.total, dt)

    def euler_step(self, method, state, dt, normals):
        self.lib.orc_euler_step(self.h, method, state, self.deriv, self.total, dt,
                                np.ascontiguousarray(normals, dtype=np.float64))

    def srk_step(self, method, state, dt, normals):
        self.lib.orc_srk_step(self.h, method, state, self.deriv, self.total, dt,
                              np.ascontiguousarray(normals, dtype=np.float64))

    def calculate_max_displacement(self, total):
        self.lib.orc_calculate_max_displacement(self.h, np.ascontiguousarray(total, dtype=np.float64))

    def pairs_linked_list(self, stage, cap=None):
        cap = cap or self.n * 64
        out = np.zeros(2 * cap, dtype=np.int32)
        cnt = self.lib.orc_pairs_linked_list(self.h, np.ascontiguousarray(stage), out, cap)
        if cnt > cap:
            return self.pairs_linked_list(stage, cap=int(cnt))
        return out[: 2 * cnt].reshape(-1, 2)

    def verlet_lists(self, cap=None):
        cap = cap or self.n * 256
        oi = np.zeros(cap, dtype=np.int32)
        oj = np.zeros(cap, dtype=np.int32)
        cnt = self.lib.orc_verlet_lists(self.h, oi, oj, cap)
        if cnt > cap:
            return self.verlet_lists(cap=int(cnt))
        return oi[:cnt].copy(), oj[:cnt].copy()


def pairs_bruteforce(stage, n, s, rho, xs=1.0, ys=1.0, cap=None):
    lib = load_library()
    cap = cap or n * 64
    out = np.zeros(2 * cap, dtype=np.int32)
    cnt = lib.orc_pairs_bruteforce(np.ascontiguousarray(stage), n, s, rho, xs, ys, out, cap)
    if cnt > cap:
        return pairs_bruteforce(stage, n, s, rho, xs, ys, cap=int(cnt))
    return out[: 2 * cnt].reshape(-1, 2)


def order_parameter(state, n, s, slot):
    lib = load_library()
    re, im = C.c_double(), C.c_double()
    lib.orc_order_parameter(np.ascontiguousarray(state), n, s, slot, C.byref(re), C.byref(im))
    return re.value, im.value


def order_parameter_ref_float(state, n, s, slot):
    lib = load_library()
    re, im = C.c_float(), C.c_float()
    lib.orc_order_parameter_ref_float(np.ascontiguousarray(state), n, s, slot, C.byref(re), C.byref(im))
    return re.value, im.value


def polar_fields(stage, n, s, slot_phi, slot_omega):
    lib = load_library()
    v = [C.c_double() for _ in range(4)]
    lib.orc_polar_fields(np.ascontiguousarray(stage), n, s, slot_phi, slot_omega, *[C.byref(x) for x in v])
    return tuple(x.value for x in v)
