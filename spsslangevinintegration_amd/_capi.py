"""ctypes declarations of the C ABI in include/spss_hip.h (libspss_hip.so).

Loading fails loudly: there is no CPU fallback behind this package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspss_hip.so")

SPSS_OK = 0
METHOD_AUTO, METHOD_GLOBAL_FIELD, METHOD_ALL_PAIRS, METHOD_LINKED_LIST, METHOD_VERLET, \
    METHOD_LINKED_LIST_VERLET = range(6)
PRECISION_FP64, PRECISION_FP32C = 0, 1
VERLET_COMPAT, VERLET_EXACT = 0, 1
STEPPER_RK4, STEPPER_EULER, STEPPER_SRK = 0, 1, 2
NOISE_NONE, NOISE_HOST, NOISE_DEVICE, NOISE_PHILOX = 0, 1, 2, 3
DOPRI5_CONST, DOPRI5_ADAPTIVE = 0, 1
# FP32C cell layouts (subdivision k, x split) in the autotuners' table order (SPSS_CELL_LAYOUTS)
CELL_LAYOUTS = ((1, 1), (2, 1), (3, 1), (1, 2), (1, 4))
# int observer(void *user, double t, const double *state, int n, int s)
OBSERVER_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.POINTER(C.c_double), C.c_int, C.c_int)

# every entry point include/spss_hip.h declares (checked by tests/test_capi_symbols.py)
EXPORTS = [
    "spss_params_default", "spss_create", "spss_destroy", "spss_last_error", "spss_abi_version",
    "spss_set_stream", "spss_synchronize", "spss_set_state", "spss_get_state", "spss_set_state_device",
    "spss_state_device", "spss_set_natural_frequencies", "spss_set_higher_order",
    "spss_get_complementary_force", "spss_evaluate_rhs", "spss_evaluate_rhs_device", "spss_rk4_step",
    "spss_euler_step", "spss_srk_step", "spss_run", "spss_prepare_run", "spss_run_stats", "spss_get_step",
    "spss_set_step",
    "spss_calculate_max_displacement", "spss_verlet_status", "spss_get_verlet_lists",
    "spss_neighbour_pairs", "spss_order_parameter", "spss_apply_periodic_boundary_conditions",
    "spss_averaging_reset", "spss_averaging_accumulate", "spss_averaging_get",
    "spss_kernel_timer_start", "spss_kernel_timer_stop", "spss_device_copy_bandwidth",
    "spss_philox4x32_10", "spss_philox_normals", "spss_autotune_cells", "spss_get_cell_subdivision",
    "spss_set_cell_layout", "spss_get_cell_layout",
    "spss_autotune_all_pairs", "spss_get_all_pairs_config", "spss_set_all_pairs_config",
    "spss_batch_create", "spss_batch_destroy", "spss_batch_set_state", "spss_batch_get_state",
    "spss_batch_set_natural_frequencies", "spss_batch_reset", "spss_batch_run", "spss_batch_order_stats",
    "spss_batch_get_tile_widths", "spss_batch_set_tile_widths", "spss_batch_autotune",
    "spss_batch_autotune_cells", "spss_batch_get_cell_subdivision",
    "spss_batch_set_cell_layout", "spss_batch_get_cell_layout",
    "spss_mlp_create", "spss_mlp_destroy", "spss_mlp_param_count", "spss_mlp_set_params", "spss_mlp_get_params",
    "spss_mlp_predict", "spss_mlp_gradients", "spss_mlp_set_data", "spss_mlp_train", "spss_mlp_reset_optimizer",
    "spss_mlp_adam_steps",
    "spss_frame", "spss_frame_append", "spss_summary", "spss_summary_append", "spss_read_last_frame",
    "spss_load_last_frame", "spss_dopri5_integrate",
]


class SpssParams(C.Structure):
    _fields_ = [
        ("model", C.c_int), ("n", C.c_int),
        ("v0", C.c_double), ("xi", C.c_double), ("sigma", C.c_double), ("rho", C.c_double),
        ("alpha", C.c_double), ("D_phi", C.c_double), ("xi_r", C.c_double), ("D_r", C.c_double),
        ("x_size", C.c_double), ("y_size", C.c_double), ("verlet_skin", C.c_double),
        ("verlet_mode", C.c_int), ("natfreq_local", C.c_int), ("precision", C.c_int),
        ("cell_subdivision", C.c_int), ("seed", C.c_uint64),
    ]


class SpssReplica(C.Structure):
    _fields_ = [("n", C.c_int), ("xi", C.c_double), ("sigma", C.c_double), ("D_phi", C.c_double),
                ("seed", C.c_uint64)]


class SpssBatchParams(C.Structure):
    _fields_ = [("model", C.c_int), ("v0", C.c_double), ("rho", C.c_double), ("alpha", C.c_double),
                ("cell_subdivision", C.c_int), ("natfreq_local", C.c_int)]


class SpssError(RuntimeError):
    pass


_LIB = None


def load(path=None):
    """Load libspss_hip.so; raises ImportError when it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
    lib = C.CDLL(path)
    vp, i, d, l = C.c_void_p, C.c_int, C.c_double, C.c_long
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    sig = {
        "spss_params_default": (i, [C.POINTER(SpssParams)]),
        "spss_create": (i, [C.POINTER(SpssParams), C.POINTER(vp)]),
        "spss_destroy": (i, [vp]),
        "spss_last_error": (C.c_char_p, []),
        "spss_abi_version": (i, []),
        "spss_set_stream": (i, [vp, vp]),
        "spss_synchronize": (i, [vp]),
        "spss_set_state": (i, [vp, vp]),
        "spss_get_state": (i, [vp, vp]),
        "spss_set_state_device": (i, [vp, vp]),
        "spss_state_device": (vp, [vp]),
        "spss_set_natural_frequencies": (i, [vp, vp, i]),
        "spss_set_higher_order": (i, [vp, i]),
        "spss_get_complementary_force": (i, [vp, vp]),
        "spss_evaluate_rhs": (i, [vp, i, vp, vp, vp, d, d]),
        "spss_evaluate_rhs_device": (i, [vp, i, vp, vp, vp, d, d]),
        "spss_rk4_step": (i, [vp, i, d]),
        "spss_euler_step": (i, [vp, i, d, i, vp]),
        "spss_srk_step": (i, [vp, i, d, i, vp]),
        "spss_run": (i, [vp, i, i, l, d, i]),
        "spss_prepare_run": (i, [vp, i, i, d, i]),
        "spss_run_stats": (i, [vp, C.POINTER(l), C.POINTER(l), C.POINTER(l), i]),
        "spss_get_step": (i, [vp, C.POINTER(l)]),
        "spss_set_step": (i, [vp, l]),
        "spss_calculate_max_displacement": (i, [vp, vp]),
        "spss_verlet_status": (i, [vp, ip, dp, ip]),
        "spss_get_verlet_lists": (i, [vp, vp, vp, l, C.POINTER(l)]),
        "spss_neighbour_pairs": (i, [vp, i, vp, vp, l, C.POINTER(l)]),
        "spss_order_parameter": (i, [vp, i, dp, dp]),
        "spss_apply_periodic_boundary_conditions": (i, [vp]),
        "spss_averaging_reset": (i, [vp]),
        "spss_averaging_accumulate": (i, [vp]),
        "spss_averaging_get": (i, [vp, vp, C.POINTER(l)]),
        "spss_kernel_timer_start": (i, [vp]),
        "spss_kernel_timer_stop": (i, [vp, C.POINTER(d), C.POINTER(l)]),
        "spss_device_copy_bandwidth": (i, [l, i, C.POINTER(d)]),
        "spss_philox4x32_10": (i, [vp, vp, vp]),
        "spss_philox_normals": (i, [C.c_uint64, C.c_uint64, C.c_uint32, l, l, vp, vp]),
        "spss_autotune_cells": (i, [vp, i, d, i, vp, i, C.POINTER(i)]),
        "spss_get_cell_subdivision": (i, [vp, C.POINTER(i)]),
        "spss_set_cell_layout": (i, [vp, i, i]),
        "spss_get_cell_layout": (i, [vp, C.POINTER(i), C.POINTER(i)]),
        "spss_autotune_all_pairs": (i, [vp, i, vp, i, C.POINTER(i)]),
        "spss_get_all_pairs_config": (i, [vp, C.POINTER(i), C.POINTER(i)]),
        "spss_set_all_pairs_config": (i, [vp, i]),
        "spss_batch_create": (i, [C.POINTER(SpssBatchParams), i, vp, C.POINTER(vp)]),
        "spss_batch_destroy": (i, [vp]),
        "spss_batch_set_state": (i, [vp, vp]),
        "spss_batch_get_state": (i, [vp, vp]),
        "spss_batch_set_natural_frequencies": (i, [vp, vp]),
        "spss_batch_reset": (i, [vp, l]),
        "spss_batch_run": (i, [vp, l, d, l, l, i]),
        "spss_batch_order_stats": (i, [vp, vp, vp, vp, vp, vp, C.POINTER(l)]),
        "spss_batch_get_tile_widths": (i, [vp, vp]),
        "spss_batch_set_tile_widths": (i, [vp, vp]),
        "spss_batch_autotune": (i, [vp, d, i, vp, i, C.POINTER(i)]),
        "spss_batch_autotune_cells": (i, [vp, d, i, vp, i, C.POINTER(i)]),
        "spss_batch_get_cell_subdivision": (i, [vp, C.POINTER(i)]),
        "spss_batch_set_cell_layout": (i, [vp, i, i]),
        "spss_batch_get_cell_layout": (i, [vp, C.POINTER(i), C.POINTER(i)]),
        "spss_mlp_create": (i, [C.POINTER(vp), vp, i, i, C.c_uint64]),
        "spss_mlp_destroy": (i, [vp]),
        "spss_mlp_param_count": (i, [vp, C.POINTER(l)]),
        "spss_mlp_set_params": (i, [vp, vp, l]),
        "spss_mlp_get_params": (i, [vp, vp, l]),
        "spss_mlp_predict": (i, [vp, vp, l, vp]),
        "spss_mlp_gradients": (i, [vp, vp, vp, i, vp, C.POINTER(C.c_float)]),
        "spss_mlp_set_data": (i, [vp, vp, vp, l]),
        "spss_mlp_train": (i, [vp, i, i, C.c_float, C.c_float, C.c_float, C.c_float, vp]),
        "spss_mlp_reset_optimizer": (i, [vp]),
        "spss_mlp_adam_steps": (i, [vp, C.POINTER(l)]),
        "spss_frame": (i, [vp, d, vp]),
        "spss_frame_append": (i, [vp, C.c_char_p, d]),
        "spss_summary": (i, [vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "spss_summary_append": (i, [vp, C.c_char_p, d]),
        "spss_read_last_frame": (i, [C.c_char_p, i, i, C.POINTER(C.c_float), vp]),
        "spss_load_last_frame": (i, [vp, C.c_char_p, C.POINTER(d)]),
        "spss_dopri5_integrate": (i, [vp, i, i, d, d, d, d, d, OBSERVER_FN, vp, i, C.POINTER(l), C.POINTER(l)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(rc):
    if rc != SPSS_OK:
        raise SpssError(f"spss error {rc}: {load().spss_last_error().decode()}")
    return rc


def ptr(a):
    """data pointer of a C-contiguous numpy array (None passes NULL)"""
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"], "need a C-contiguous array"
    return a.ctypes.data


def copy_bandwidth(nbytes=1 << 30, repeats=20):
    """attainable HBM GB/s of the library's plain copy kernel"""
    v = C.c_double()
    check(load().spss_device_copy_bandwidth(int(nbytes), int(repeats), C.byref(v)))
    return v.value


def philox4x32_10(ctr, key):
    """one Philox4x32-10 block (host), as the kernels compute it"""
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    check(load().spss_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data))
    return out


def read_last_frame(path, s, n):
    """(t, state) of the last complete frame of a trajectory file (float time, s*n floats)"""
    t = C.c_float()
    st = np.zeros(s * n)
    check(load().spss_read_last_frame(os.fsencode(path), int(s), int(n), C.byref(t), ptr(st)))
    return t.value, st


def philox_normals(seed, step, n, block=0, first=0):
    """(z0, z1): the standard normals the kernels draw for particles first..first+n-1"""
    z0 = np.zeros(n)
    z1 = np.zeros(n)
    check(load().spss_philox_normals(int(seed), int(step), int(block), int(first), int(n), ptr(z0), ptr(z1)))
    return z0, z1
