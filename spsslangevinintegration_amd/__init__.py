"""MI355X (gfx950) right-hand-side hot path for self-propelled particles with
Kuramoto-type phase coupling.

The compute lives in libspss_hip.so (hand-written HIP kernels behind the C ABI
of include/spss_hip.h).  This package is thin ctypes plumbing for tests and
benchmarks; importing it fails if the library has not been built.
"""
from . import _capi  # noqa: F401
from ._capi import (CELL_LAYOUTS, DOPRI5_ADAPTIVE, DOPRI5_CONST, METHOD_ALL_PAIRS, METHOD_AUTO,  # noqa: F401
                    METHOD_GLOBAL_FIELD, METHOD_LINKED_LIST,
                    METHOD_LINKED_LIST_VERLET, METHOD_VERLET, NOISE_DEVICE, NOISE_HOST, NOISE_NONE,
                    NOISE_PHILOX, PRECISION_FP32C, PRECISION_FP64, STEPPER_EULER, STEPPER_RK4, STEPPER_SRK,
                    VERLET_COMPAT, VERLET_EXACT, SpssError)

_capi.load()

from .system import ParticleSystem  # noqa: E402,F401
from .batch import ReplicaBatch  # noqa: E402,F401
from .screen import MLP, SweepScreen  # noqa: E402,F401
