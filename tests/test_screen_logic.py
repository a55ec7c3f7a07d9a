"""Host logic of the sweep screen (window around the predicted crossing, widening, fallback to
the plain sweep) with a stand-in network and a stand-in simulator: no GPU needed."""
import numpy as np
import pytest

S = pytest.importorskip("spsslangevinintegration_amd.screen")


class _Net:
    def __init__(self, shift):
        self.shift = shift

    def predict(self, x):  # x[:, 0] is sigma mapped to [-1, 1] over [0, 4]
        sig = (x[:, 0].astype(np.float64) + 1.0) * 2.0
        return np.stack([_curve(sig - self.shift), np.full_like(sig, 0.01)], axis=1).astype(np.float32)


def _curve(sig, xc=1.3):
    return 1.0 / (1.0 + np.exp(-(sig - xc) / 0.1))


class _Screen(S.SweepScreen):
    def __init__(self, shift, truth):
        super().__init__(net=_Net(shift))
        self.truth = truth

    def simulate(self, pts):
        self.simulations += pts.size
        return np.stack([self.truth(pts["sigma"]), np.zeros(pts.size)], axis=1)


def _line(G=40):
    return S.points(np.linspace(0, 3, G), 0.05, 0.5, 0.01, 0.03, 4096, np.arange(G, dtype=np.uint64))


@pytest.mark.parametrize("shift", [0.0, 0.3, -0.5, 1.2])
def test_crossing_equals_plain_sweep(shift):
    line = _line()
    sc = _Screen(shift, _curve)
    x_plain, _ = sc.plain_sweep(line)
    n_plain = sc.simulations
    sc.simulations = 0
    x, idx, R, err = sc.screened_sweep(line, window=6)
    assert abs(x - x_plain) < 1e-12
    assert np.all(np.diff(idx) == 1)  # one contiguous window
    np.testing.assert_array_equal(R, _curve(line["sigma"][idx]))
    if shift == 0.0:
        assert sc.simulations == 6 and err < 1e-6
    assert sc.simulations < n_plain


def test_no_crossing_widens_to_the_end_of_the_line():
    """mean |Z| grows with the coupling: no crossing up to the largest sigma means none at all"""
    line = _line()
    sc = _Screen(0.0, lambda s: 0.1 + 0 * s)
    x, idx, _, _ = sc.screened_sweep(line, window=6)
    assert x is None and idx[-1] == line.size - 1 and sc.simulations == idx.size


def test_large_prediction_error_falls_back_to_the_plain_sweep():
    line = _line()
    truth = lambda s: _curve(s, 2.0) * 0.6 + 0.2  # noqa: E731
    sc = _Screen(0.0, truth)
    x, idx, R, err = sc.screened_sweep(line, window=6, max_error=0.1)
    assert idx.size == line.size and err > 0.1
    assert abs(x - S.SweepScreen.crossing(line["sigma"], truth(line["sigma"]), 0.5)) < 1e-12
    sc2 = _Screen(0.0, truth)  # without the bound it only widens to the bracket
    x2, idx2, _, _ = sc2.screened_sweep(line, window=6)
    assert abs(x2 - x) < 1e-12 and idx2.size < line.size


def test_crossing_interpolates_and_handles_the_edges():
    sig = np.array([0.0, 1.0, 2.0])
    assert S.SweepScreen.crossing(sig, np.array([0.0, 0.25, 0.75]), 0.5) == 1.5
    assert S.SweepScreen.crossing(sig, np.array([0.6, 0.7, 0.8]), 0.5) == 0.0
    assert S.SweepScreen.crossing(sig, np.array([0.1, 0.2, 0.3]), 0.5) is None


def test_features_map_the_ranges_to_unit_interval():
    sc = _Screen(0.0, _curve)
    p = S.points([0.0, 4.0], [0.0, 0.2], [0.0, 1.0], [0.0, 0.02], [0.0, 0.05], [2048, 65536], [1, 2])
    f = sc.features(p)
    np.testing.assert_allclose(f[0], -1.0)
    np.testing.assert_allclose(f[1], 1.0)


def test_initial_state_depends_only_on_the_point():
    p = S.points(1.0, 0.1, 0.5, 0.01, 0.03, 2048, 123, coherence0=0.5)[0]
    a, wa = S.initial_state(p, 4)
    b, wb = S.initial_state(p, 4)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(wa, wb)
    st = a.reshape(-1, 4)
    assert np.all(st[:, 2] < np.pi + 1e-12) and np.all(st[:, 3] == 0.0)
    R = abs(np.mean(np.exp(1j * st[:, 2])))
    assert abs(R - 2 / np.pi) < 0.05  # phases U[0, pi): |<e^{i phi}>| = 2/pi


def test_screen_simulates_where_the_slope_is_steep():
    line = _line(80)
    sc = _Screen(0.0, _curve)
    out = sc.screen(line, slope_min=0.5, spread_max=0.05)
    sig = line["sigma"][out["selected"]]
    slope = _curve(line["sigma"]) * (1 - _curve(line["sigma"])) / 0.1  # exact derivative of the curve
    want = np.nonzero(slope > 0.5)[0]
    assert abs(out["selected"].size - want.size) <= 2 and np.all(np.abs(sig - 1.3) < 0.6)
    assert sc.simulations == out["selected"].size
    np.testing.assert_allclose(out["error"][:, 0], 0.0, atol=1e-6)
    sc2 = _Screen(0.0, _curve)
    out2 = sc2.screen(line, slope_min=0.5, spread_max=0.005)  # every predicted spread 0.01 > 0.005
    assert out2["selected"].size == line.size


def test_random_points_cover_the_ranges():
    p = S.random_points(2000, np.random.default_rng(1))
    assert p["sigma"].min() >= 0 and p["sigma"].max() <= 3 and set(np.unique(p["rho"])) == {0.02, 0.03, 0.04}
    assert p["N"].min() >= 2048 and p["N"].max() <= 8192 and len(np.unique(p["seed"])) == 2000
