"""The reference's output files and restart on the device (spss_frame*, spss_summary*,
spss_*_last_frame) against the restatement in oracle/observer_oracle.py.

Frames are elementwise float conversions and wraps: bit-exact.  The summary's float sums are
accumulated in particle order exactly as the reference does; the only possible difference is a
phasor component that the device's and the host's double cos/sin round to neighbouring floats,
so |Z| and arg Z are checked to 4 float ulps (and counted when bit-exact).
"""
import os

import numpy as np
import pytest

from oracle import observer_oracle as OO

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spss():
    import spsslangevinintegration_amd as m
    return m


def _state(n, s, seed, spread=3.0):
    rng = np.random.default_rng(seed)
    st = np.zeros((n, s))
    st[:, 0] = rng.uniform(-spread, spread, n)
    st[:, 1] = rng.uniform(-spread, spread, n)
    slot = 4 if s == 6 else 2
    st[:, slot] = rng.uniform(-80, 80, n)
    if s >= 4:
        st[:, 3] = rng.normal(0, 0.3, n)
    st[0, 0], st[1, 1] = -1e-9, 1.0 - 1e-17  # the rounding-to-L corner of the reference's wrap
    return st.reshape(-1)


@pytest.mark.parametrize("model", [3, 4, 6])
def test_frame_bitexact(spss, model):
    n = 20000
    print(f"[gpu] frame model={model} n={n}", flush=True)
    st = _state(n, model, 10 + model)
    g = spss.ParticleSystem(model=model, n=n, rho=0.05)
    g.set_state(st)
    got = g.frame(12.345)
    want = OO.frame(st, model, 12.345)
    assert got.dtype == np.float32 and got.size == 1 + model * n
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("model", [4, 6])
def test_summary_matches_the_sequential_float_sum(spss, model):
    n = 50000
    print(f"[gpu] summary model={model} n={n}", flush=True)
    st = _state(n, model, 20 + model)
    g = spss.ParticleSystem(model=model, n=n, rho=0.05)
    g.set_state(st)
    a, arg = g.summary()
    wa, warg = OO.summary(st, model)
    print(f"  |Z| {a!r} vs {wa!r}, arg {arg!r} vs {warg!r}, bit-exact: {a == wa and arg == warg}")
    assert abs(a - wa) <= 4 * np.spacing(np.float32(wa))
    assert abs(arg - warg) <= 4 * np.spacing(np.float32(abs(warg)))


def test_files_and_restart_round_trip(spss, tmp_path):
    n, model = 5000, 4
    traj = os.fspath(tmp_path / "traj.bin")
    summ = os.fspath(tmp_path / "summary.txt")
    st = _state(n, model, 31)
    g = spss.ParticleSystem(model=model, n=n, v0=0.01, xi=1.0, sigma=2.0, rho=0.05, D_phi=0.05,
                            precision=spss.PRECISION_FP32C, seed=5)
    g.set_state(st)
    states = []
    for k in range(3):
        g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 20, 0.01, spss.NOISE_PHILOX)  # resident order
        g.append_frame(traj, 0.2 * (k + 1))
        g.append_summary(summ, 0.2 * (k + 1))
        states.append(g.get_state().copy())
    assert os.path.getsize(traj) == 3 * (1 + model * n) * 4
    t, last = spss._capi.read_last_frame(traj, model, n)
    want = OO.frame(states[-1], model, 0.6)
    assert t == want[0]
    np.testing.assert_array_equal(last, want[1:].astype(np.float64))
    lines = open(summ).read().splitlines()
    assert len(lines) == 3
    for k, line in enumerate(lines):
        t_s, a_s, g_s = line.split("\t")
        wa, warg = OO.summary(states[k], model)
        assert float(t_s) == float("%g" % np.float32(0.2 * (k + 1)))
        assert abs(float(a_s) - wa) <= 1e-5 and abs(float(g_s) - warg) <= 1e-5
    # restart: a new system takes the last frame as its state
    h = spss.ParticleSystem(model=model, n=n, v0=0.01, xi=1.0, sigma=2.0, rho=0.05, D_phi=0.05)
    assert h.load_last_frame(traj) == pytest.approx(float(np.float32(0.6)))
    np.testing.assert_array_equal(h.get_state(), want[1:].astype(np.float64))


def test_restart_errors(spss, tmp_path):
    path = os.fspath(tmp_path / "short.bin")
    np.zeros(5, dtype=np.float32).tofile(path)
    with pytest.raises(spss.SpssError):
        spss._capi.read_last_frame(path, 4, 100)
    g = spss.ParticleSystem(model=4, n=100, rho=0.1)
    with pytest.raises(spss.SpssError):
        g.load_last_frame(os.fspath(tmp_path / "missing.bin"))
