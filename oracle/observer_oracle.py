"""Restatement of the reference's trajectory/summary output and restart (numpy).

TEST INFRASTRUCTURE ONLY: the checker of spss_frame*, spss_summary* and spss_*_last_frame in
tests/ (the product never imports it).

* frame: BinaryObserverFirstOrderPtr::SaveSystemState (Observers/BinaryObserverFirstOrderPtr.cpp:
  108-121; serial twin BinaryObserverFirstOrder.cpp:101-122): copy the FP64 state to float
  (RealOutput = float, Definitions.hpp:18); wrap x and y with the float overload of
  ApplyPeriodicBoundaryConditions (BoundaryConfiguration/PeriodicBoundaryConditions.cpp:137-145:
  v -= floor(v * rsize) * size, evaluated in double on the float value and rounded back by the
  float -= double); wrap phi with phi -= floor(phi / 2 pi) * 2 pi the same way; write the float
  time then the s*n floats.  Model 6 writes the raw float copy
  (BinaryObserverSecondOrderPtr.cpp:108-113).
* summary: SaveSummaryStatistics (BinaryObserverFirstOrderPtr.cpp:126-146; model 6 phi in slot 4,
  BinaryObserverSecondOrderPtr.cpp:175-193): complex<float> accumulation of
  ((float) cos phi_i, (float) sin phi_i) in particle order, divided by (float) n; |Z| and arg Z
  in float; the line "t \\t |Z| \\t arg Z" of operator<< (%g of the float's double value).
* restart: SimulationEngineFirstOrderPtr::InitializeSystemStateFromPreviousSolution
  (ControlEngines/SimulationEngineFirstOrderPtr.cpp:70-94): t_max = file size / ((1 + s n) * 4),
  the frame t_max - 1, floats widened to double.

Pinned in tests/test_oracle_observer.py (hand-worked wraps including the reference's rounding
to exactly L, idempotence on wrapped frames, a pure-Python float32 loop for the sequential sum,
closed forms of |Z|, file round trips with a truncated last frame).
"""
import numpy as np

TWO_PI = 6.283185307179586476925286766559


def frame(state, s, t, x_size=1.0, y_size=1.0):
    """the 1 + s*n floats of one saved frame"""
    f = np.asarray(state, dtype=np.float64).astype(np.float32).reshape(-1, s)
    out = f.copy()
    if s != 6:
        x = f[:, 0].astype(np.float64)
        y = f[:, 1].astype(np.float64)
        p = f[:, 2].astype(np.float64)
        out[:, 0] = (x - np.floor(x * (1.0 / x_size)) * x_size).astype(np.float32)
        out[:, 1] = (y - np.floor(y * (1.0 / y_size)) * y_size).astype(np.float32)
        out[:, 2] = (p - np.floor(p / TWO_PI) * TWO_PI).astype(np.float32)
    return np.concatenate([np.array([t], dtype=np.float32), out.reshape(-1)])


def summary_sums(state, s):
    """the complex<float> sum of the unit phasors in particle order (before the division)"""
    phi = np.asarray(state, dtype=np.float64).reshape(-1, s)[:, 4 if s == 6 else 2]
    c = np.cos(phi).astype(np.float32)
    sn = np.sin(phi).astype(np.float32)
    # add.accumulate adds element after element in float32: the sequential sum
    re = np.add.accumulate(c, dtype=np.float32)[-1]
    im = np.add.accumulate(sn, dtype=np.float32)[-1]
    return np.float32(re), np.float32(im)


def summary(state, s):
    """(|Z|, arg Z) as float32"""
    re, im = summary_sums(state, s)
    n = np.float32(np.asarray(state).size // s)
    z = np.complex64(complex(re / n, im / n))
    return np.float32(np.abs(z)), np.float32(np.arctan2(np.float32(z.imag), np.float32(z.real)))


def summary_line(t, abs_z, arg_z):
    return "%g\t%g\t%g\n" % (float(np.float32(t)), float(abs_z), float(arg_z))


def write_frames(path, frames):
    with open(path, "ab") as f:
        for fr in frames:
            f.write(np.asarray(fr, dtype=np.float32).tobytes())


def read_last_frame(path, s, n):
    raw = np.fromfile(path, dtype=np.uint8)
    per = (1 + s * n) * 4
    t_max = raw.size // per
    if t_max < 1:
        raise ValueError("no complete frame")
    fr = raw[(t_max - 1) * per:t_max * per].view(np.float32)
    return float(fr[0]), fr[1:].astype(np.float64)
