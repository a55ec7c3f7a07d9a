"""Fused FP32C steps on the resident, cell-ordered copy of the state.

A fused Euler/SRK step (FP32C precision, linked-list method, models 3 and 4) keeps the
state, the particle ids and the natural frequencies in cell order on the device between
steps and re-sorts them after every step from the moves the pair kernel reports.  These
tests check that this is invisible to the caller:

* trajectories track the FP64 HIP path (itself pinned to the oracle) step by step with the
  same Philox draws, including steps where particles jump farther than one cell;
* reading the state (get_state, order parameter, averaging) does not perturb the run:
  bit-identical to an uninterrupted run;
* writes to the state or the natural frequencies between steps (set_state,
  set_natural_frequencies, apply_periodic_boundary_conditions), an RHS evaluation and a
  switch of stepper are all honoured.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# FP32 pair sums with the hardware sine: ~1e-6 relative on sigma*A/n per step, which
# reaches the phase through omega and the positions through v0 * dt * phase
TOL_OMEGA = 1e-6
TOL_PHI = 2e-7
TOL_POS_PER_V0 = 2e-8


@pytest.fixture(scope="module")
def spss():
    import spsslangevinintegration_amd as m
    return m


def _state(n, model, seed):
    rng = np.random.default_rng(seed)
    st = np.zeros((n, model))
    st[:, 0], st[:, 1] = rng.random(n), rng.random(n)
    st[:, 2] = rng.random(n) * 2 * np.pi
    if model == 4:
        st[:, 3] = rng.normal(0, 0.3, n)
    return st.reshape(-1)


def _pair(spss, model, n, rho, sub, v0, seed=9, natf=None):
    kw = dict(model=model, n=n, v0=v0, xi=1.0, sigma=2.0, rho=rho, alpha=0.2, D_phi=0.05, natfreq_local=1,
              seed=seed)
    g64 = spss.ParticleSystem(**kw)
    g32 = spss.ParticleSystem(precision=spss.PRECISION_FP32C, cell_subdivision=sub, **kw)
    st = _state(n, model, seed)
    for g in (g64, g32):
        if natf is not None:
            g.set_natural_frequencies(natf)
        g.set_state(st)
    return g64, g32


def _close(spss, model, a, b, what, v0):
    s = model
    d = np.abs(a.reshape(-1, s) - b.reshape(-1, s))
    print(f"  {what}: max |d| pos {d[:, :2].max():.2e} phase {d[:, 2].max():.2e} "
          f"omega {d[:, 3].max() if model == 4 else 0.0:.2e}", flush=True)
    assert d[:, :2].max() <= TOL_POS_PER_V0 * v0, (what, "position", d[:, :2].max())
    assert d[:, 2].max() <= TOL_PHI, (what, "phase", d[:, 2].max())
    if model == 4:
        assert d[:, 3].max() <= TOL_OMEGA, (what, "omega", d[:, 3].max())


RESIDENT_CASES = [
    # model, n, rho, subdivision, v0, stepper
    (4, 20000, 0.02, 1, 0.01, "euler"),
    (4, 20000, 0.02, 2, 3.0, "euler"),   # v0*dt = 0.015 > cell side 0.01: jumps past the 3x3 block
    (4, 20000, 0.02, 1, 0.01, "srk"),
    (4, 20000, 0.02, 3, 2.5, "srk"),
    (3, 20000, 0.02, 1, 0.01, "euler"),
    (4, 2000, 0.3, 1, 0.5, "euler"),     # 3 x 3 cells
    (4, 20000, 0.02, 1, 1.0, "euler"),   # v0*dt = cell side / 4: many near moves, all eight directions
    (4, 20000, 0.1, 1, 2.0, "srk"),      # ~200 per cell: cells span several waves, 10% move per step
    (4, 20001, 0.02, 2, 3.0, "euler"),   # far jumps with N not a multiple of 4 (the far pass's tail)
]


@pytest.mark.parametrize("model,n,rho,sub,v0,stepper", RESIDENT_CASES)
def test_resident_steps_track_fp64(spss, model, n, rho, sub, v0, stepper):
    print(f"[gpu] resident model={model} n={n} rho={rho} sub={sub} v0={v0} {stepper}", flush=True)
    dt = 0.005
    natf = np.random.default_rng(3).normal(0, 0.2, n)
    g64, g32 = _pair(spss, model, n, rho, sub, v0, natf=natf)
    st_id = spss.STEPPER_SRK if stepper == "srk" else spss.STEPPER_EULER
    for chunk in range(3):
        for g in (g64, g32):
            g.run(st_id, spss.METHOD_LINKED_LIST, 10, dt, spss.NOISE_PHILOX)
        _close(spss, model, g32.get_state(), g64.get_state(), f"after {10 * (chunk + 1)} steps", v0)
    if stepper == "srk":
        c64, c32 = g64.complementary_force(), g32.complementary_force()
        assert np.max(np.abs(c64 - c32)) <= 2e-5 * max(1.0, np.max(np.abs(c64)))
    assert g32.step == 30


def test_resident_clustered_rows_track_fp64(spss):
    """dense clusters in one row of cells: cells far above one workgroup's particles (dense-cell
    chunks of the overflow pass), stages that overflow, and a row that needs more spans than
    the per-step row tiles provide (the last spans widened past the particle bound)"""
    n, rho, v0, dt = 20000, 0.02, 0.01, 0.005
    print(f"[gpu] resident clustered rows n={n} rho={rho}", flush=True)
    rng = np.random.default_rng(17)
    st = _state(n, 4, 17).reshape(-1, 4)
    for q, x0 in enumerate((0.1, 0.4, 0.7)):  # three clusters of 2500 in the cell row [0.40, 0.42)
        sl = slice(q * 2500, (q + 1) * 2500)
        st[sl, 0] = x0 + 0.03 * rng.random(2500)
        st[sl, 1] = 0.401 + 0.018 * rng.random(2500)
    natf = rng.normal(0, 0.2, n)
    g64, g32 = _pair(spss, 4, n, rho, 1, v0, natf=natf)
    for g in (g64, g32):
        g.set_state(st.reshape(-1))
    for chunk in range(2):
        for g in (g64, g32):
            g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 10, dt, spss.NOISE_PHILOX)
        _close(spss, 4, g32.get_state(), g64.get_state(), f"after {10 * (chunk + 1)} steps", v0)


def test_resident_reads_do_not_perturb(spss):
    """get_state / order parameter / averaging between steps leave the run bit-identical."""
    n, rho = 30000, 0.015
    kw = dict(model=4, n=n, v0=0.02, xi=1.0, sigma=2.0, rho=rho, D_phi=0.05, seed=21,
              precision=spss.PRECISION_FP32C)
    st = _state(n, 4, 4)
    a = spss.ParticleSystem(**kw)
    b = spss.ParticleSystem(**kw)
    for g in (a, b):
        g.set_state(st)
    a.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 20, 0.005, spss.NOISE_PHILOX)
    for k in range(4):
        b.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 5, 0.005, spss.NOISE_PHILOX)
        b.get_state()
        b.order_parameter(2)
        b.averaging_accumulate()
    assert np.array_equal(a.get_state(), b.get_state())
    # averaging of 4 snapshots matches the snapshots themselves
    c = spss.ParticleSystem(**kw)
    c.set_state(st)
    want = np.zeros(4 * n)
    for k in range(4):
        c.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 5, 0.005, spss.NOISE_PHILOX)
        want += c.get_state()
    s, cnt = b.averaging_get()
    assert cnt == 4
    assert np.max(np.abs(s - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))


def test_resident_honours_writes_between_steps(spss):
    n, rho, dt = 20000, 0.02, 0.005
    rng = np.random.default_rng(5)
    natf = rng.normal(0, 0.2, n)
    v0 = 0.05
    g64, g32 = _pair(spss, 4, n, rho, 1, v0, natf=natf)

    def both(fn):
        for g in (g64, g32):
            fn(g)

    both(lambda g: g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 8, dt, spss.NOISE_PHILOX))
    # periodic wrap between steps (applied to the resident copy)
    both(lambda g: g.apply_periodic_boundary_conditions())
    s32 = g32.get_state()
    assert 0.0 <= s32[0::4].min() and s32[0::4].max() < 1.0
    _close(spss, 4, s32, g64.get_state(), "after wrap", v0)
    both(lambda g: g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 8, dt, spss.NOISE_PHILOX))
    _close(spss, 4, g32.get_state(), g64.get_state(), "steps after wrap", v0)
    # new natural frequencies
    natf2 = rng.normal(0.5, 0.3, n)
    both(lambda g: g.set_natural_frequencies(natf2))
    both(lambda g: g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 8, dt, spss.NOISE_PHILOX))
    _close(spss, 4, g32.get_state(), g64.get_state(), "after new natural frequencies", v0)
    # stepper switch Euler -> SRK -> Euler
    both(lambda g: g.run(spss.STEPPER_SRK, spss.METHOD_LINKED_LIST, 6, dt, spss.NOISE_PHILOX))
    both(lambda g: g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 6, dt, spss.NOISE_PHILOX))
    _close(spss, 4, g32.get_state(), g64.get_state(), "after stepper switch", v0)
    # an RHS evaluation of another state between steps
    other = _state(n, 4, 99)
    k = np.zeros(4 * n)
    both(lambda g: g.evaluate_rhs(spss.METHOD_LINKED_LIST, other, np.zeros(4 * n), k, 0.0, dt))
    both(lambda g: g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 6, dt, spss.NOISE_PHILOX))
    _close(spss, 4, g32.get_state(), g64.get_state(), "after an RHS evaluation", v0)
    # a new state
    st2 = _state(n, 4, 77)
    both(lambda g: g.set_state(st2))
    both(lambda g: g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 6, dt, spss.NOISE_PHILOX))
    _close(spss, 4, g32.get_state(), g64.get_state(), "after set_state", v0)
    # host-injected noise through the single-step entry point
    z = rng.normal(0, 1, n)
    both(lambda g: g.euler_step(spss.METHOD_LINKED_LIST, dt, spss.NOISE_HOST, z))
    _close(spss, 4, g32.get_state(), g64.get_state(), "after a host-noise step", v0)


@pytest.mark.parametrize("stepper_name", ["EULER", "SRK"])
def test_graph_replayed_run_equals_eager_steps(spss, stepper_name):
    """spss_run replays the bulk of a run from a HIP graph (Philox step read from a device
    counter); the result must equal one-step calls bit for bit, across calls whose resident
    buffer parity differs (the graph is captured again)"""
    stepper = getattr(spss, "STEPPER_" + stepper_name)
    n, rho, dt = 40000, 0.02, 0.003
    print(f"[gpu] graph vs eager {stepper_name} n={n}", flush=True)
    st = _state(n, 4, 77)
    natf = np.random.default_rng(78).normal(0, 0.1, n)
    a = _pair(spss, 4, n, rho, 1, 0.02, natf=natf)[1]
    b = _pair(spss, 4, n, rho, 1, 0.02, natf=natf)[1]
    for g in (a, b):
        g.set_state(st)
    a.run(stepper, spss.METHOD_LINKED_LIST, 57, dt, spss.NOISE_PHILOX)
    for _ in range(57):
        b.run(stepper, spss.METHOD_LINKED_LIST, 1, dt, spss.NOISE_PHILOX)
    assert a.step == b.step == 57
    np.testing.assert_array_equal(a.get_state(), b.get_state())
    a.run(stepper, spss.METHOD_LINKED_LIST, 34, dt, spss.NOISE_PHILOX)
    for _ in range(34):
        b.run(stepper, spss.METHOD_LINKED_LIST, 1, dt, spss.NOISE_PHILOX)
    np.testing.assert_array_equal(a.get_state(), b.get_state())
    if stepper_name == "SRK":
        np.testing.assert_array_equal(a.complementary_force(), b.complementary_force())
