"""Cell-size autotuner of the FP32C path (spss_autotune_cells).

The probes must leave no trace: state, step counter and complementary force are restored, and
the system then steps exactly (bit for bit) like a system created with the chosen cell layout
(subdivision and x split).  Every layout gives the same neighbour sets (pair lists equal to the
FP64 ones).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spss():
    import spsslangevinintegration_amd as m
    return m


def _make(spss, n, rho, sub, seed=21):
    rng = np.random.default_rng(seed)
    st = np.zeros((n, 4))
    st[:, 0], st[:, 1] = rng.random(n), rng.random(n)
    st[:, 2] = rng.random(n) * 2 * np.pi
    st[:, 3] = rng.normal(0, 0.3, n)
    g = spss.ParticleSystem(model=4, n=n, v0=0.01, xi=1.0, sigma=2.0, rho=rho, alpha=0.1, D_phi=0.05,
                            natfreq_local=1, seed=seed, precision=spss.PRECISION_FP32C, cell_subdivision=sub)
    w = rng.normal(0, 0.2, n)
    g.set_natural_frequencies(w)
    g.set_state(st.reshape(-1))
    return g


@pytest.mark.parametrize("stepper_name", ["EULER", "SRK"])
def test_autotune_restores_and_matches_a_fresh_system(spss, stepper_name):
    stepper = getattr(spss, "STEPPER_" + stepper_name)
    n, rho, dt = 65536, 0.02, 0.002
    g = _make(spss, n, rho, 1)
    g.run(stepper, spss.METHOD_LINKED_LIST, 5, dt, spss.NOISE_PHILOX)
    before = g.get_state().copy()
    step = g.step
    table, k = g.autotune_cells(stepper, dt, probe_steps=5)
    kc, xs = g.cell_layout
    print(f"{stepper_name}: ms/step for layouts {spss.CELL_LAYOUTS}: {np.round(table, 4)}; chosen {(kc, xs)}")
    assert k == kc and g.cell_subdivision == k
    assert np.all(table > 0)
    assert (kc, xs) == spss.CELL_LAYOUTS[int(np.argmin(table))]
    np.testing.assert_array_equal(g.get_state(), before)
    assert g.step == step
    # from here on it steps like a system built with that layout from the same state
    h = _make(spss, n, rho, k)
    h.set_cell_layout(k, xs)
    h.set_state(before)
    h.step = step
    for x in (g, h):
        x.run(stepper, spss.METHOD_LINKED_LIST, 20, dt, spss.NOISE_PHILOX)
    np.testing.assert_array_equal(g.get_state(), h.get_state())


def test_neighbour_sets_do_not_depend_on_the_subdivision(spss):
    n, rho = 20000, 0.03
    rng = np.random.default_rng(5)
    st = np.zeros((n, 4))
    st[:, 0], st[:, 1] = rng.random(n), rng.random(n)
    pairs = []
    for k, xs in spss.CELL_LAYOUTS:
        g = _make(spss, n, rho, k)
        g.set_cell_layout(k, xs)
        g.set_state(st.reshape(-1))
        pairs.append(g.neighbour_pairs(spss.METHOD_LINKED_LIST, st.reshape(-1)))
    assert pairs[0].shape[0] > 10 * n
    for p in pairs[1:]:
        np.testing.assert_array_equal(p, pairs[0])


def test_autotune_refuses_the_fp64_path(spss):
    g = spss.ParticleSystem(model=4, n=1000, rho=0.05)
    with pytest.raises(spss.SpssError):
        g.autotune_cells(spss.STEPPER_EULER, 0.01)
