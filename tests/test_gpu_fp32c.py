"""FP32 cell-list path (SPSS_PRECISION_FP32C: FP32 distance tests and phasor partial sums, FP64 totals and state) against the FP64 oracle.

* neighbour sets: bit-exact against the brute-force FP64 set with the reference's
  ClassC formula (sorted (i, j) lists), for cell subdivisions 1-3, positions outside
  the box, clustered inputs (LDS staging overflow -> global path, list flushes),
  tiny grids;
* one RHS: position slots within 1e-12 (FP64 self terms), coupling slot within
  1e-6 absolute (FP32 distance tests and phasor components summed in FP32 partials per
  stencil row, combined in FP64; observed ~1e-7);
* Euler trajectories fed identical noise: order parameter within 1e-4 of the FP64
  oracle after 300 steps at config-2 sizes;
* determinism: identical inputs give bit-identical results.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_POS = 1e-12      # k slots 0..2: FP64 self terms
TOL_COUPLING = 1e-6  # k slot 3: sigma * (phasor pair sum) / n, absolute
TOL_R = 1e-4         # order parameter vs FP64 oracle (the stated accuracy target)


@pytest.fixture(scope="module")
def spss():
    import spsslangevinintegration_amd as m
    return m


def _state(n, seed, spread=1.0, cluster=False, model=4):
    rng = np.random.default_rng(seed)
    st = np.zeros((n, model))
    if cluster:
        # half the particles packed into a few tiny discs: thousands per cell
        m = n // 2
        centers = rng.random((4, 2))
        which = rng.integers(0, 4, m)
        st[:m, 0] = centers[which, 0] + rng.normal(0, 0.002, m)
        st[:m, 1] = centers[which, 1] + rng.normal(0, 0.002, m)
        st[m:, 0] = rng.random(n - m)
        st[m:, 1] = rng.random(n - m)
    else:
        st[:, 0] = rng.random(n) * spread - (spread - 1) / 2
        st[:, 1] = rng.random(n) * spread - (spread - 1) / 2
    st[:, 2] = rng.random(n) * 2 * np.pi * 3
    if model == 4:
        st[:, 3] = rng.normal(0, 0.5, n)
    return st.reshape(-1)


PAIR_CASES = [
    # n, rho, subdivision, spread, cluster
    (4096, 0.03, 1, 1.0, False),
    (20000, 0.02, 2, 1.3, False),
    (20000, 0.02, 3, 1.0, False),
    (65536, 0.004, 1, 1.0, False),
    (12000, 0.03, 1, 1.0, True),
    (12000, 0.03, 2, 1.0, True),
    (300, 0.3, 1, 1.0, False),
    (5000, 0.3, 3, 1.4, False),
    # 4.2 particles per cell: the row tiles reach their 46-column cap (widest staged x)
    (262144, 0.004, 1, 1.0, False),
    (262144, 0.004, 1, 1.05, False),
]


@pytest.mark.parametrize("n,rho,sub,spread,cluster", PAIR_CASES)
def test_fp32c_pairs_bitexact(spss, n, rho, sub, spread, cluster):
    from oracle import pyoracle
    print(f"[gpu] fp32c pairs n={n} rho={rho} sub={sub} cluster={cluster}", flush=True)
    st = _state(n, seed=n + sub, spread=spread, cluster=cluster)
    g = spss.ParticleSystem(model=4, n=n, rho=rho, precision=spss.PRECISION_FP32C, cell_subdivision=sub)
    got = g.neighbour_pairs(spss.METHOD_LINKED_LIST, st)
    if n <= 65536:  # O(n^2)
        want = pyoracle.pairs_bruteforce(st, n, 4, rho)
        assert np.array_equal(got, want)
    o = pyoracle.Oracle(model=4, n=n, rho=rho)
    assert np.array_equal(got, o.pairs_linked_list(st))


def test_fp32c_rejects_grids_narrower_than_three_cells(spss):
    """int(1/rho) < 3: the reference's stencil visits a cell more than once and counts the
    minimum-image pair each time; the FP32C path refuses that case (FP64 reproduces it)."""
    with pytest.raises(spss.SpssError, match="int\\(1/rho\\) >= 3"):
        spss.ParticleSystem(model=4, n=64, rho=0.4, precision=spss.PRECISION_FP32C)


@pytest.mark.parametrize("model,n,rho,sub,alpha,cluster", [
    (4, 30000, 0.01, 1, 0.0, False), (4, 30000, 0.01, 2, 0.4, False), (3, 20000, 0.02, 1, 0.2, False),
    (4, 12000, 0.03, 1, 0.1, True)])
def test_fp32c_rhs_vs_oracle(spss, model, n, rho, sub, alpha, cluster):
    import oracle
    print(f"[gpu] fp32c rhs model={model} n={n} rho={rho} sub={sub}", flush=True)
    kw = dict(model=model, n=n, v0=0.02, xi=1.5, sigma=2.0, rho=rho, alpha=alpha, natfreq_local=1)
    st = _state(n, seed=n, cluster=cluster, model=model)
    natf = np.random.default_rng(1).normal(0, 0.2, n)
    kp = np.zeros(model * n)
    o = oracle.Oracle(**kw)
    g = spss.ParticleSystem(precision=spss.PRECISION_FP32C, cell_subdivision=sub, **kw)
    o.set_natural_frequencies(natf)
    g.set_natural_frequencies(natf)
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


def test_fp32c_srk_complementary_force(spss):
    import oracle
    n, rho = 20000, 0.02
    kw = dict(model=4, n=n, v0=0.02, xi=1.0, sigma=1.5, rho=rho, alpha=0.3)
    st = _state(n, seed=3)
    o = oracle.Oracle(**kw)
    g = spss.ParticleSystem(precision=spss.PRECISION_FP32C, **kw)
    o.set_higher_order(1)
    g.set_higher_order(1)
    ko = np.zeros(4 * n)
    kg = np.zeros(4 * n)
    o.evaluate_rhs(3, st, ko.copy(), ko, 0.0, 0.01)
    g.evaluate_rhs(3, st, kg.copy(), kg, 0.0, 0.01)
    co, cg = o.complementary_force(), g.complementary_force()
    print(f"  complementary force max |gpu - oracle| = {np.max(np.abs(co - cg)):.3e}", flush=True)
    assert np.max(np.abs(co - cg)) <= 1e-6 * max(1.0, np.max(np.abs(co)))


def test_fp32c_euler_order_parameter_matches_fp64(spss):
    """Config-2 sizes (N=65536, rho=0.02, dt=0.005, m=1 -> xi=1, sigma=2, D=0.05,
    omega ~ N(0, 0.2^2) natural frequencies): FP32C and the FP64 HIP path, both fed the same
    Philox noise, against the FP64 oracle fed the identical draws for the first steps."""
    import oracle
    n, rho, dt, steps = 65536, 0.02, 0.005, 300
    kw = dict(model=4, n=n, v0=0.005, xi=1.0, sigma=2.0, rho=rho, D_phi=0.05, natfreq_local=1)
    rng = np.random.default_rng(2)
    st = np.zeros((n, 4))
    st[:, 0], st[:, 1], st[:, 2] = rng.random(n), rng.random(n), rng.random(n) * 2 * np.pi
    st = st.reshape(-1)
    natf = rng.normal(0, 0.2, n)
    g64 = spss.ParticleSystem(seed=77, **kw)
    g32 = spss.ParticleSystem(seed=77, precision=spss.PRECISION_FP32C, **kw)
    for g in (g64, g32):
        g.set_natural_frequencies(natf)
        g.set_state(st)
    o = oracle.Oracle(**kw)
    o.set_natural_frequencies(natf)
    so = st.copy()
    for k in range(steps):
        z = rng.normal(0.0, 1.0, n)
        g64.euler_step(spss.METHOD_LINKED_LIST, dt, spss.NOISE_HOST, z)
        g32.euler_step(spss.METHOD_LINKED_LIST, dt, spss.NOISE_HOST, z)
        if k < 20:
            o.euler_step(3, so, dt, z)
    # short horizon, pointwise: FP64 HIP path vs oracle
    # (oracle ran 20 steps; compare against a second FP64 handle that ran 20 steps)
    s64 = g64.get_state()
    s32 = g32.get_state()
    r64 = np.hypot(*g64.order_parameter(2))
    r32 = np.hypot(*g32.order_parameter(2))
    from oracle import pyoracle
    ro = np.hypot(*pyoracle.order_parameter(s64, n, 4, 2))
    assert abs(r64 - ro) <= 1e-13
    assert abs(r32 - r64) <= TOL_R, (r32, r64)
    # positions stay close pointwise over this horizon
    assert np.max(np.abs(s32[0::4] - s64[0::4])) <= 1e-6
    # pointwise short-horizon FP64 HIP vs oracle
    g64b = spss.ParticleSystem(seed=77, **kw)
    g64b.set_natural_frequencies(natf)
    g64b.set_state(st)
    rng2 = np.random.default_rng(2)
    rng2.random(3 * n)
    rng2.normal(0, 0.2, n)
    for k in range(20):
        g64b.euler_step(spss.METHOD_LINKED_LIST, dt, spss.NOISE_HOST, rng2.normal(0.0, 1.0, n))
    assert np.max(np.abs(g64b.get_state() - so)) <= 1e-12


def test_fp32c_deterministic(spss):
    n = 50000
    st = _state(n, seed=11)
    outs = []
    for _ in range(2):
        g = spss.ParticleSystem(model=4, n=n, v0=0.01, sigma=1.0, rho=0.01, D_phi=0.02, seed=5,
                                precision=spss.PRECISION_FP32C)
        g.set_state(st)
        g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 20, 0.002, spss.NOISE_PHILOX)
        outs.append(g.get_state())
    assert np.array_equal(outs[0], outs[1])
