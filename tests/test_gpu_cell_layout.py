"""FP32C cell grids split in x (spss_set_cell_layout): cells 1/(k*xsplit*nc) wide, 1/(k*nc) high.

A particle's candidates are then 2k+1 rows of 2k*xsplit+1 narrow cells, fewer than the reference's
3x3 square cells but holding every pair within rho.  Checked here:

* neighbour sets equal the brute-force FP64 ones (and the oracle's linked-list pairs) bit for bit,
  also with dense clusters (the overflow pass and its dense-cell chunks) and positions outside
  the box;
* the RHS stays within the FP32C tolerance of the oracle, and resident Euler/SRK steps track the
  FP64 HIP path with the same Philox draws, including moves past the 3x3 block of narrow cells;
* a layout set between steps keeps the state: the run continues bit-identically to a system
  created with that layout from the same state and step;
* batch replicas on a split grid step bit-identically to single systems of the same layout.
"""
import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.gpu

TOL_POS = 1e-12
TOL_COUPLING = 2e-6
TOL_OMEGA = 1e-6
TOL_PHI = 2e-7
TOL_POS_PER_V0 = 2e-8


@pytest.fixture(scope="module")
def spss():
    import spsslangevinintegration_amd as m
    return m


def _state(n, seed, model=4, spread=1.0, cluster=False):
    rng = np.random.default_rng(seed)
    st = np.zeros((n, model))
    if cluster:  # a third of the particles in a few dense spots
        m = n // 3
        centres = rng.random((4, 2))
        c = centres[rng.integers(0, 4, m)]
        st[:m, :2] = c + 0.01 * rng.random((m, 2))
        st[m:, 0], st[m:, 1] = rng.random(n - m), rng.random(n - m)
    else:
        st[:, 0] = rng.random(n) * spread - (spread - 1) / 2
        st[:, 1] = rng.random(n) * spread - (spread - 1) / 2
    st[:, 2] = rng.random(n) * 2 * np.pi * 3
    if model == 4:
        st[:, 3] = rng.normal(0, 0.5, n)
    return st.reshape(-1)


def _system(spss, k, xs, **kw):
    g = spss.ParticleSystem(precision=spss.PRECISION_FP32C, cell_subdivision=k, **kw)
    g.set_cell_layout(k, xs)
    assert g.cell_layout == (k, xs) and g.cell_subdivision == k
    return g


PAIR_CASES = [
    # n, rho, subdivision, x split, spread, cluster
    (65536, 0.004, 1, 2, 1.0, False),
    (65536, 0.004, 1, 4, 1.0, False),
    (20000, 0.02, 2, 2, 1.3, False),
    (20000, 0.02, 3, 4, 1.0, False),
    (12000, 0.03, 1, 4, 1.0, True),
    (12000, 0.03, 2, 2, 1.0, True),
    (300, 0.3, 1, 2, 1.0, False),
    (5000, 0.3, 3, 4, 1.4, False),
    (262144, 0.004, 1, 4, 1.05, False),
]


@pytest.mark.parametrize("n,rho,k,xs,spread,cluster", PAIR_CASES)
def test_split_cell_pairs_bitexact(spss, n, rho, k, xs, spread, cluster):
    print(f"[gpu] split-cell pairs n={n} rho={rho} k={k} xsplit={xs} cluster={cluster}", flush=True)
    st = _state(n, seed=n + 7 * xs, spread=spread, cluster=cluster)
    g = _system(spss, k, xs, model=4, n=n, rho=rho)
    got = g.neighbour_pairs(spss.METHOD_LINKED_LIST, st)
    if n <= 65536:  # O(n^2)
        assert np.array_equal(got, pyoracle.pairs_bruteforce(st, n, 4, rho))
    o = pyoracle.Oracle(model=4, n=n, rho=rho)
    assert np.array_equal(got, o.pairs_linked_list(st))


def test_split_cell_layout_rejects_bad_splits(spss):
    g = spss.ParticleSystem(model=4, n=1000, rho=0.05, precision=spss.PRECISION_FP32C)
    for k, xs in ((1, 3), (1, 0), (4, 1)):
        with pytest.raises(spss.SpssError):
            g.set_cell_layout(k, xs)
    assert g.cell_layout == (1, 1)
    f = spss.ParticleSystem(model=4, n=1000, rho=0.05)
    with pytest.raises(spss.SpssError):
        f.set_cell_layout(1, 2)  # FP64: no FP32C cell path


@pytest.mark.parametrize("model,n,rho,k,xs,alpha,cluster", [
    (4, 30000, 0.01, 1, 4, 0.2, False), (3, 20000, 0.02, 1, 2, 0.1, False), (4, 12000, 0.03, 1, 4, 0.3, True)])
def test_split_cell_rhs_vs_oracle(spss, model, n, rho, k, xs, alpha, cluster):
    import oracle
    print(f"[gpu] split-cell rhs model={model} n={n} rho={rho} k={k} xsplit={xs}", flush=True)
    kw = dict(model=model, n=n, v0=0.02, xi=1.5, sigma=2.0, rho=rho, alpha=alpha, natfreq_local=1)
    st = _state(n, seed=n + 1, model=model, cluster=cluster)
    natf = np.random.default_rng(2).normal(0, 0.2, n)
    o = oracle.Oracle(**kw)
    g = _system(spss, k, xs, **kw)
    o.set_natural_frequencies(natf)
    g.set_natural_frequencies(natf)
    kp = np.zeros(model * n)
    ko = np.zeros(model * n)
    kg = np.zeros(model * n)
    o.evaluate_rhs(3, st, kp, ko, 0.0, 0.01)
    g.evaluate_rhs(3, st, kp, kg, 0.0, 0.01)
    for slot in range(model - 1):
        assert np.max(np.abs(kg[slot::model] - ko[slot::model])) <= TOL_POS, slot
    c = model - 1
    err = np.max(np.abs(kg[c::model] - ko[c::model]))
    print(f"  coupling slot max |gpu - oracle| = {err:.3e}", flush=True)
    assert err <= TOL_COUPLING * max(1.0, kw["sigma"])


def _close(model, a, b, what, v0):
    d = np.abs(a.reshape(-1, model) - b.reshape(-1, model))
    print(f"  {what}: max |d| pos {d[:, :2].max():.2e} phase {d[:, 2].max():.2e} "
          f"omega {d[:, 3].max() if model == 4 else 0.0:.2e}", flush=True)
    assert d[:, :2].max() <= TOL_POS_PER_V0 * v0, (what, "position", d[:, :2].max())
    assert d[:, 2].max() <= TOL_PHI, (what, "phase", d[:, 2].max())
    if model == 4:
        assert d[:, 3].max() <= TOL_OMEGA, (what, "omega", d[:, 3].max())


RESIDENT_CASES = [
    # model, n, rho, subdivision, x split, v0, stepper
    (4, 20000, 0.02, 1, 4, 0.01, "euler"),
    (4, 20000, 0.02, 1, 4, 1.0, "euler"),   # v0*dt = one narrow cell: all eight directions, some farther
    (4, 20000, 0.02, 1, 2, 0.01, "srk"),
    (4, 20001, 0.02, 2, 4, 3.0, "euler"),   # jumps far past the 3x3 block of narrow cells
    (3, 20000, 0.02, 1, 2, 0.01, "euler"),
    (4, 20000, 0.1, 1, 4, 2.0, "srk"),      # ~200 per square cell
]


@pytest.mark.parametrize("model,n,rho,k,xs,v0,stepper", RESIDENT_CASES)
def test_split_cell_resident_steps_track_fp64(spss, model, n, rho, k, xs, v0, stepper):
    print(f"[gpu] split-cell resident model={model} n={n} rho={rho} k={k} xsplit={xs} v0={v0} {stepper}",
          flush=True)
    dt = 0.005
    kw = dict(model=model, n=n, v0=v0, xi=1.0, sigma=2.0, rho=rho, alpha=0.2, D_phi=0.05, natfreq_local=1,
              seed=9)
    natf = np.random.default_rng(3).normal(0, 0.2, n)
    st = _state(n, 9, model=model)
    g64 = spss.ParticleSystem(**kw)
    g32 = _system(spss, k, xs, **kw)
    for g in (g64, g32):
        g.set_natural_frequencies(natf)
        g.set_state(st)
    sid = spss.STEPPER_SRK if stepper == "srk" else spss.STEPPER_EULER
    for chunk in range(3):
        for g in (g64, g32):
            g.run(sid, spss.METHOD_LINKED_LIST, 10, dt, spss.NOISE_PHILOX)
        _close(model, g32.get_state(), g64.get_state(), f"after {10 * (chunk + 1)} steps", v0)


def test_split_cell_clustered_rows_track_fp64(spss):
    """dense clusters in one row of square cells, on a grid split 4 ways in x: cells above one
    workgroup's particles (dense-cell chunks of 2kx+1 narrow columns) and stages that overflow"""
    n, rho, v0, dt = 20000, 0.02, 0.01, 0.005
    print(f"[gpu] split-cell clustered rows n={n} rho={rho}", flush=True)
    rng = np.random.default_rng(17)
    st = _state(n, 17).reshape(-1, 4)
    for q, x0 in enumerate((0.1, 0.4, 0.7)):
        sl = slice(q * 2500, (q + 1) * 2500)
        st[sl, 0] = x0 + 0.03 * rng.random(2500)
        st[sl, 1] = 0.401 + 0.018 * rng.random(2500)
    kw = dict(model=4, n=n, v0=v0, xi=1.0, sigma=2.0, rho=rho, alpha=0.2, D_phi=0.05, natfreq_local=1, seed=9)
    natf = rng.normal(0, 0.2, n)
    g64 = spss.ParticleSystem(**kw)
    g32 = _system(spss, 1, 4, **kw)
    for g in (g64, g32):
        g.set_natural_frequencies(natf)
        g.set_state(st.reshape(-1))
    for chunk in range(2):
        for g in (g64, g32):
            g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 10, dt, spss.NOISE_PHILOX)
        _close(4, g32.get_state(), g64.get_state(), f"after {10 * (chunk + 1)} steps", v0)


@pytest.mark.parametrize("graph", [False, True])
def test_layout_set_between_steps_continues_like_a_fresh_system(spss, graph):
    n, rho, dt = 65536, 0.01, 0.002
    print(f"[gpu] layout switch graph={graph}", flush=True)
    kw = dict(model=4, n=n, v0=0.01, xi=1.0, sigma=2.0, rho=rho, alpha=0.1, D_phi=0.05, natfreq_local=1, seed=5)
    st = _state(n, 5)
    natf = np.random.default_rng(6).normal(0, 0.2, n)
    a = spss.ParticleSystem(precision=spss.PRECISION_FP32C, **kw)
    a.set_natural_frequencies(natf)
    a.set_state(st)
    if graph:
        a.prepare_run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, dt, spss.NOISE_PHILOX)
    a.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 12, dt, spss.NOISE_PHILOX)
    mid = a.get_state().copy()
    a.set_cell_layout(1, 4)
    np.testing.assert_array_equal(a.get_state(), mid)
    b = _system(spss, 1, 4, **kw)
    b.set_natural_frequencies(natf)
    b.set_state(mid)
    b.step = a.step
    for g in (a, b):
        if graph:
            g.prepare_run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, dt, spss.NOISE_PHILOX)
        g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 24, dt, spss.NOISE_PHILOX)
    np.testing.assert_array_equal(a.get_state(), b.get_state())


def test_batch_on_split_cells_equals_single_systems(spss):
    model, R, K, dt, v0, rho = 4, 4, 20, 0.005, 0.01, 0.03
    print("[gpu] batch on split cells", flush=True)
    rng = np.random.default_rng(31)
    ns = rng.integers(6000, 20000, R)
    xi, sig, Dp = np.ones(R), rng.uniform(0, 4, R), rng.uniform(0, 0.2, R)
    seeds = 0x5EED0004 + np.arange(R)
    states = [_state(int(ns[r]), 40 + r) for r in range(R)]
    nats = [rng.normal(0, 0.2, int(ns[r])) for r in range(R)]
    b = spss.ReplicaBatch(model, ns, xi, sig, Dp, seeds, v0=v0, rho=rho)
    b.set_cell_layout(1, 4)
    assert b.cell_layout == (1, 4)
    b.set_natural_frequencies(np.concatenate(nats))
    b.set_state(np.concatenate(states))
    b.run(K, dt, observe_every=0, graph=True)
    got = b.get_state()
    b.close()
    off = 0
    for r in range(R):
        g = _system(spss, 1, 4, model=model, n=int(ns[r]), v0=v0, xi=xi[r], sigma=sig[r], rho=rho, D_phi=Dp[r],
                    natfreq_local=1, seed=int(seeds[r]))
        g.set_natural_frequencies(nats[r])
        g.set_state(states[r])
        g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, K, dt, spss.NOISE_PHILOX)
        m = model * int(ns[r])
        np.testing.assert_array_equal(got[off:off + m], g.get_state())
        off += m
        g.close()
