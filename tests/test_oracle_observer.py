"""Pins of the observer/restart restatement (oracle/observer_oracle.py): hand-worked wraps,
including the reference's rounding of a tiny negative coordinate to exactly the box size,
idempotence on an already wrapped frame, the sequential float32 sum against a pure-Python
loop, closed forms of |Z|, and file round trips with a truncated last frame.  A wrap done in
float instead of double, a pairwise instead of sequential sum, or an off-by-one frame offset
fails one of them."""
import os

import numpy as np
import pytest

from oracle import observer_oracle as OO


def _state(n, s, seed, spread=3.0):
    rng = np.random.default_rng(seed)
    st = np.zeros((n, s))
    st[:, 0] = rng.uniform(-spread, spread, n)
    st[:, 1] = rng.uniform(-spread, spread, n)
    st[:, 2 if s != 6 else 4] = rng.uniform(-50, 50, n)
    return st.reshape(-1)


def test_hand_worked_wraps():
    st = np.array([-0.25, 1.25, -np.pi / 2, 0.3,
                   -1e-9, 0.5, 7.0, 0.0,
                   2.0, -3.0, 2 * np.pi, 0.0])
    fr = OO.frame(st, 4, 1.5)
    assert fr[0] == np.float32(1.5)
    f = fr[1:].reshape(-1, 4)
    assert f[0, 0] == np.float32(0.75) and f[0, 1] == np.float32(0.25)
    assert f[0, 2] == np.float32(np.float64(np.float32(-np.pi / 2)) + OO.TWO_PI)
    assert f[0, 3] == np.float32(0.3)  # omega is copied, not wrapped
    # the reference's arithmetic maps a tiny negative x to 1 - 1e-9, which rounds to 1.0f = L
    assert f[1, 0] == np.float32(1.0)
    assert f[1, 2] == np.float32(7.0 - OO.TWO_PI)
    assert f[2, 0] == 0.0 and f[2, 1] == 0.0
    # the float copy of 2 pi lies just above 2 pi in double: it wraps to a tiny positive value
    assert 0.0 < f[2, 2] < 1e-6


def test_wrap_runs_in_double_on_the_float_value():
    """a float-only wrap gives different bits for some values; the restatement must not"""
    st = _state(20000, 4, 1)
    f = OO.frame(st, 4, 0.0)[1:].reshape(-1, 4)
    xf = st.reshape(-1, 4)[:, 0].astype(np.float32)
    float_only = xf - np.floor(xf * np.float32(1.0)) * np.float32(1.0)
    assert np.array_equal(f[:, 0], (xf.astype(np.float64) - np.floor(xf.astype(np.float64))).astype(np.float32))
    assert f.dtype == np.float32 and float_only.dtype == np.float32


def test_frame_ranges_and_idempotence():
    for s in (3, 4):
        st = _state(5000, s, 2)
        f = OO.frame(st, s, 0.0)[1:].reshape(-1, s)
        assert np.all((f[:, 0] >= 0) & (f[:, 0] <= 1)) and np.all((f[:, 1] >= 0) & (f[:, 1] <= 1))
        assert np.all((f[:, 2] >= 0) & (f[:, 2] <= np.float32(OO.TWO_PI)))
        inside = (f[:, 0] < 1) & (f[:, 1] < 1)
        again = OO.frame(f.astype(np.float64).reshape(-1), s, 0.0)[1:].reshape(-1, s)
        assert np.array_equal(again[inside, :2], f[inside, :2])


def test_model6_frame_is_the_raw_float_copy():
    st = _state(300, 6, 3)
    f = OO.frame(st, 6, 2.0)
    assert np.array_equal(f[1:], st.astype(np.float32))


def test_summary_is_the_sequential_float_sum():
    st = _state(3000, 4, 4)
    re, im = OO.summary_sums(st, 4)
    phi = st.reshape(-1, 4)[:, 2]
    r, i = np.float32(0.0), np.float32(0.0)
    for p in phi:  # the reference's loop, one float add after another
        r = np.float32(r + np.float32(np.cos(p)))
        i = np.float32(i + np.float32(np.sin(p)))
    assert re == r and im == i
    pairwise = np.cos(phi).astype(np.float32).sum(dtype=np.float32)
    assert isinstance(pairwise, np.float32)


def test_summary_closed_forms():
    n = 1000
    st = np.zeros((n, 4))
    a, g = OO.summary(st.reshape(-1), 4)  # identical phases 0: |Z| = 1, arg 0
    assert a == np.float32(1.0) and g == 0.0
    st[1::2, 2] = np.pi  # balanced pairs: Z = 0 up to the float sin(pi)
    a, _ = OO.summary(st.reshape(-1), 4)
    assert a < 1e-6
    st6 = np.zeros((n, 6))
    st6[:, 4] = 0.7  # model 6 reads slot 4
    a, g = OO.summary(st6.reshape(-1), 6)
    # |Z| = 1 up to the float accumulation of the reference (n adds: about n * 2^-24 relative)
    assert 1e-7 < abs(a - 1.0) < n * 2.0 ** -24 and abs(g - np.float32(0.7)) < n * 2.0 ** -24
    assert OO.summary_line(0.5, np.float32(1.0), np.float32(0.25)) == "0.5\t1\t0.25\n"


def test_restart_reads_the_last_complete_frame(tmp_path):
    s, n = 4, 50
    path = os.fspath(tmp_path / "traj.bin")
    frames = [OO.frame(_state(n, s, 10 + k), s, 0.1 * k) for k in range(3)]
    OO.write_frames(path, frames)
    with open(path, "ab") as f:  # a partial frame at the end (an interrupted write) is ignored
        f.write(np.zeros(7, dtype=np.float32).tobytes())
    t, st = OO.read_last_frame(path, s, n)
    assert t == float(np.float32(0.2))
    assert np.array_equal(st, frames[2][1:].astype(np.float64))
    with pytest.raises(ValueError):
        OO.read_last_frame(path, s, 10 * n)
