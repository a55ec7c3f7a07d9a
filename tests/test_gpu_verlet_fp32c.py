"""Verlet lists in FP32C precision (exact mode: r_max = rho + skin, candidates from cells of side
>= r_max), the large configuration's neighbour method, against the FP64 oracle.

* lists: after the first rebuild every particle's list equals the oracle's exact-mode list
  (EvaluateInteractionsWithLinkedListAndVerletNeighborList's build, ParticleSystemFirstOrderPtr.cpp:
  616-815) as a set, bit for bit, also with positions outside the box;
* the large configuration's coefficients (m = 0.5, sigma = 1.5, D = 0.02 -> xi = 2, sigma = 3,
  D_phi = 0.08; Lorentzian(0, 0.05) natural frequencies; skin 0.2 rho; SRK, dt = 0.002) at a
  reduced N with ~52 neighbours: 20 Philox-driven SRK steps against the oracle fed the same draws
  (positions 1e-12, phases 1e-9, omega 1e-7), and the rebuild flag and accumulated displacement
  of CalculateMaxDisplacement (:836-861, exact mode) step for step over 180 steps with several
  rebuilds (v0 = 0.05: about one rebuild every 45 steps);
* graph-replayed steps (spss_run) equal eager steps bit for bit, rebuilds decided on the device;
* the Verlet and the cell-list FP32C paths take the same neighbours: their states agree to the
  FP32 summation order after 50 steps;
* dense cells (a tile with more particles than a workgroup, tiles whose stencil cells overflow the
  LDS stage, lists past the width) against the oracle over 12 SRK steps;
* at N = 1,048,576, rho = 0.004: FP32C Verlet SRK against the FP64 path's exact-mode Verlet SRK fed
  the same Philox draws, |R_FP32C - R_FP64| <= 1e-4 over 5000 steps, the FP64 path pinned to the
  oracle for the first two steps.
"""
import time

import numpy as np
import pytest

import oracle
from oracle.pyoracle import METHOD_LINKED_LIST, METHOD_LINKED_LIST_VERLET, VERLET_EXACT
from spsslangevinintegration_amd import _capi

pytestmark = pytest.mark.gpu

M, SIGMA, D = 0.5, 1.5, 0.02


@pytest.fixture(scope="module")
def spss():
    import spsslangevinintegration_amd as m
    return m


def _config3(n, rho, v0=0.005, seed=3):
    rng = np.random.default_rng(seed)
    st = np.zeros((n, 4))
    st[:, 0], st[:, 1] = rng.random(n), rng.random(n)
    st[:, 2] = rng.random(n) * 2 * np.pi
    w = np.clip(0.05 * np.tan(np.pi * (rng.random(n) - 0.5)), -10.0, 10.0) / M
    kw = dict(model=4, n=n, v0=v0, xi=1.0 / M, sigma=SIGMA / M, rho=rho, D_phi=D / M ** 2, natfreq_local=1,
              verlet_mode=VERLET_EXACT, verlet_skin=0.2 * rho)
    return kw, st.reshape(-1), w


def _pairs_from_lists(lists):
    i = np.repeat(np.arange(len(lists)), [len(l) for l in lists])
    j = np.concatenate(lists) if lists else np.zeros(0, np.int64)
    order = np.lexsort((j, i))
    return np.stack([i[order], j[order]], axis=1)


@pytest.mark.parametrize("n,rho,spread", [(8192, 0.045, 1.0), (65536, 0.016, 1.1), (20000, 0.02, 1.0)])
def test_fp32c_verlet_lists_bitexact(spss, n, rho, spread):
    print(f"[gpu] fp32c verlet lists n={n} rho={rho} spread={spread}", flush=True)
    kw, st, w = _config3(n, rho, seed=n)
    st = st.reshape(-1, 4)
    st[:, :2] = st[:, :2] * spread - (spread - 1) / 2
    st = st.reshape(-1)
    g = spss.ParticleSystem(precision=spss.PRECISION_FP32C, **kw)
    g.set_natural_frequencies(w)
    g.set_state(st)
    g.euler_step(spss.METHOD_LINKED_LIST_VERLET, 0.002, spss.NOISE_NONE)  # enters, builds, steps
    got = _pairs_from_lists(g.verlet_lists())
    o = oracle.Oracle(**kw)
    o.set_natural_frequencies(w)
    so = st.copy()
    o.euler_step(METHOD_LINKED_LIST_VERLET, so, 0.002, np.zeros(n))
    oi, oj = o.verlet_lists()
    want = np.stack([oi, oj], axis=1)[np.lexsort((oj, oi))]
    print(f"  {len(got)} list entries on the GPU, {len(want)} in the oracle", flush=True)
    assert np.array_equal(got, want)
    d = np.abs(g.get_state() - so).reshape(-1, 4)
    assert d[:, :3].max() <= 1e-12  # the step moves x, y, phi with the old phase and omega
    assert d[:, 3].max() <= 1e-8  # omega: the FP32C coupling (~1e-7) times dt


def test_fp32c_verlet_srk_config3_vs_oracle(spss):
    n = 16384
    rho = float(np.sqrt(52.7 / (np.pi * n)))
    kw, st, w = _config3(n, rho)
    seed, dt, steps = 0x5EED0003, 0.002, 20
    print(f"[gpu] fp32c verlet SRK config-3 coefficients n={n} rho={rho:.4f}", flush=True)
    g = spss.ParticleSystem(precision=spss.PRECISION_FP32C, seed=seed, **kw)
    g.set_natural_frequencies(w)
    g.set_state(st)
    o = oracle.Oracle(**kw)
    o.set_natural_frequencies(w)
    so = st.copy()
    g.run(spss.STEPPER_SRK, spss.METHOD_LINKED_LIST_VERLET, steps, dt, spss.NOISE_PHILOX)
    for k in range(steps):
        z0, z1 = _capi.philox_normals(seed, k, n)
        o.srk_step(METHOD_LINKED_LIST_VERLET, so, dt, np.stack([z0, z1], axis=1).reshape(-1))
    d = np.abs(g.get_state() - so).reshape(-1, 4)
    print(f"  max |gpu - oracle| pos {d[:, :2].max():.2e} phase {d[:, 2].max():.2e} omega {d[:, 3].max():.2e}",
          flush=True)
    assert d[:, :2].max() <= 1e-12 and d[:, 2].max() <= 1e-9 and d[:, 3].max() <= 1e-7
    co = o.complementary_force()
    cg = g.complementary_force()
    assert np.max(np.abs(cg - co)) <= 1e-6 * max(1.0, np.max(np.abs(co)))


def test_fp32c_verlet_rebuild_schedule_matches_oracle(spss):
    n = 8192
    kw, st, w = _config3(n, 0.045, v0=0.05, seed=5)
    seed, dt = 0x5EED0003, 0.002
    print("[gpu] fp32c verlet rebuild schedule", flush=True)
    g = spss.ParticleSystem(precision=spss.PRECISION_FP32C, seed=seed, **kw)
    g.set_natural_frequencies(w)
    g.set_state(st)
    o = oracle.Oracle(**kw)
    o.set_natural_frequencies(w)
    so = st.copy()
    rebuilds = 0
    for k in range(180):
        g.srk_step(spss.METHOD_LINKED_LIST_VERLET, dt, spss.NOISE_PHILOX)
        z0, z1 = _capi.philox_normals(seed, k, n)
        o.srk_step(METHOD_LINKED_LIST_VERLET, so, dt, np.stack([z0, z1], axis=1).reshape(-1))
        su, acc, uses = g.verlet_status()
        assert uses and su == o.should_update_lists, k
        assert abs(acc - o.accumulated_displacement) <= 1e-14, k
        rebuilds += int(su)
    print(f"  {rebuilds} rebuilds in 180 steps, same flags and accumulated displacement as the oracle", flush=True)
    assert rebuilds >= 3
    assert np.max(np.abs(g.get_state() - so)[0::4]) <= 1e-10


@pytest.mark.parametrize("stepper", ["euler", "srk"])
def test_fp32c_verlet_graph_replay_equals_eager(spss, stepper):
    n = 20000
    kw, st, w = _config3(n, 0.02, v0=0.05, seed=7)
    sid = spss.STEPPER_SRK if stepper == "srk" else spss.STEPPER_EULER
    out = []
    for graph in (True, False):
        g = spss.ParticleSystem(precision=spss.PRECISION_FP32C, seed=11, **kw)
        g.set_natural_frequencies(w)
        g.set_state(st)
        if graph:
            g.run(sid, spss.METHOD_LINKED_LIST_VERLET, 37, 0.002, spss.NOISE_PHILOX)
            eager, replayed, _ = g.run_stats()
            assert replayed == 36 and eager == 1
        else:
            for _ in range(37):
                if stepper == "srk":
                    g.srk_step(spss.METHOD_LINKED_LIST_VERLET, 0.002, spss.NOISE_PHILOX)
                else:
                    g.euler_step(spss.METHOD_LINKED_LIST_VERLET, 0.002, spss.NOISE_PHILOX)
        out.append((g.get_state(), g.verlet_status()))
        g.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


def test_fp32c_verlet_lists_beyond_the_width_walk_their_cells(spss, monkeypatch):
    """lists longer than the width are not stored: those particles walk the 3 x 3 cells of the last
    rebuild (a superset of their list, in the same order); with the width capped at 24 most
    particles take that path and the run agrees with the uncapped one to the FP32 summation order"""
    n = 20000
    kw, st, w = _config3(n, 0.03, v0=0.05, seed=13)
    res = []
    for cap in (None, "24"):
        if cap:
            monkeypatch.setenv("SPSS_VERLET_WIDTH_CAP", cap)
        g = spss.ParticleSystem(precision=spss.PRECISION_FP32C, seed=3, **kw)
        g.set_natural_frequencies(w)
        g.set_state(st)
        g.run(spss.STEPPER_SRK, spss.METHOD_LINKED_LIST_VERLET, 60, 0.002, spss.NOISE_PHILOX)
        res.append((g.get_state(), g.verlet_status()))
        if cap:
            with pytest.raises(spss.SpssError, match="outgrew"):
                g.verlet_lists()
        g.close()
    d = np.abs(res[0][0] - res[1][0]).reshape(-1, 4)
    print(f"  capped width vs full lists after 60 SRK steps: phase {d[:, 2].max():.2e} omega {d[:, 3].max():.2e}",
          flush=True)
    assert d[:, 3].max() <= 1e-7 and d[:, 2].max() <= 1e-9
    assert res[0][1][0] == res[1][1][0]  # same rebuild flag


def test_fp32c_verlet_dense_cells_vs_oracle(spss):
    """three adjacent Verlet cells holding 400 particles each among a uniform background: the dense
    cells' tile holds more particles than a workgroup (the rest are stepped apart), the tiles around
    them overflow the LDS stage (their lists are built and walked through global memory) and the
    dense particles' lists outgrow the width (they walk their cells); 12 Philox SRK steps against the
    oracle fed the same draws, to the large configuration's tolerances"""
    rho = 0.02
    ncv = int(1.0 / (1.2 * rho))
    h = 1.0 / ncv
    rng = np.random.default_rng(77)
    parts = [rng.random((12000, 2))]
    for cx, cy in ((20, 20), (21, 20), (20, 21)):
        parts.append(np.stack([(cx + 0.02 + 0.96 * rng.random(400)) * h, (cy + 0.02 + 0.96 * rng.random(400)) * h], 1))
    xy = np.concatenate(parts)
    n = len(xy)
    kw, st, w = _config3(n, rho, seed=17)
    st = st.reshape(-1, 4)
    st[:, :2] = xy
    st = st.reshape(-1)
    seed, dt, steps = 0x5EED0007, 0.002, 12
    print(f"[gpu] fp32c verlet dense cells n={n} rho={rho}", flush=True)
    g = spss.ParticleSystem(precision=spss.PRECISION_FP32C, seed=seed, **kw)
    g.set_natural_frequencies(w)
    g.set_state(st)
    g.run(spss.STEPPER_SRK, spss.METHOD_LINKED_LIST_VERLET, steps, dt, spss.NOISE_PHILOX)
    flag, acc, overflow = g.verlet_status()
    o = oracle.Oracle(**kw)
    o.set_natural_frequencies(w)
    so = st.copy()
    for k in range(steps):
        z0, z1 = _capi.philox_normals(seed, k, n)
        o.srk_step(METHOD_LINKED_LIST_VERLET, so, dt, np.stack([z0, z1], axis=1).reshape(-1))
    d = np.abs(g.get_state() - so).reshape(-1, 4)
    print(f"  lists past the width {overflow}; max |gpu - oracle| pos {d[:, :2].max():.2e} "
          f"phase {d[:, 2].max():.2e} omega {d[:, 3].max():.2e}", flush=True)
    assert overflow > 0
    assert d[:, :2].max() <= 1e-12 and d[:, 2].max() <= 1e-9 and d[:, 3].max() <= 1e-7
    g.close()


def test_fp32c_verlet_and_cell_list_take_the_same_neighbours(spss):
    n = 30000
    kw, st, w = _config3(n, 0.015, seed=9)
    res = []
    for method in (spss.METHOD_LINKED_LIST_VERLET, spss.METHOD_LINKED_LIST):
        g = spss.ParticleSystem(precision=spss.PRECISION_FP32C, seed=21, **kw)
        g.set_natural_frequencies(w)
        g.set_state(st)
        g.run(spss.STEPPER_SRK, method, 50, 0.002, spss.NOISE_PHILOX)
        res.append(g.get_state())
        g.close()
    d = np.abs(res[0] - res[1]).reshape(-1, 4)
    print(f"  verlet vs cell list after 50 SRK steps: phase {d[:, 2].max():.2e} omega {d[:, 3].max():.2e}", flush=True)
    assert d[:, 3].max() <= 1e-7 and d[:, 2].max() <= 1e-9


@pytest.mark.timeout(600)
def test_fp32c_verlet_config3_full_size_order_parameter(spss):
    n, rho, seed, dt = 1 << 20, 0.004, 0x5EED0003, 0.002
    kw, st, w = _config3(n, rho, seed=seed)
    print(f"[gpu] config 3 at N={n}: FP32C Verlet SRK vs FP64 Verlet SRK", flush=True)
    systems = []
    for prec in (spss.PRECISION_FP32C, spss.PRECISION_FP64):
        g = spss.ParticleSystem(precision=prec, seed=seed, **kw)
        g.set_natural_frequencies(w)
        g.set_state(st)
        systems.append(g)
    g32, g64 = systems
    # the FP64 path against the oracle over the first two steps
    g64.run(spss.STEPPER_SRK, spss.METHOD_LINKED_LIST_VERLET, 2, dt, spss.NOISE_PHILOX)
    o = oracle.Oracle(**kw)
    o.set_natural_frequencies(w)
    so = st.copy()
    for k in range(2):
        z0, z1 = _capi.philox_normals(seed, k, n)
        o.srk_step(METHOD_LINKED_LIST_VERLET, so, dt, np.stack([z0, z1], axis=1).reshape(-1))
    d64 = np.max(np.abs(g64.get_state() - so))
    print(f"  FP64 Verlet vs oracle after 2 steps: {d64:.2e}", flush=True)
    assert d64 <= 1e-12
    g32.run(spss.STEPPER_SRK, spss.METHOD_LINKED_LIST_VERLET, 2, dt, spss.NOISE_PHILOX)
    t0 = time.time()
    worst = 0.0
    for done in range(1000, 5001, 1000):
        for g in (g32, g64):
            g.run(spss.STEPPER_SRK, spss.METHOD_LINKED_LIST_VERLET, 1000 if done > 1000 else 998, dt,
                  spss.NOISE_PHILOX)
        r32 = np.hypot(*g32.order_parameter(2))
        r64 = np.hypot(*g64.order_parameter(2))
        worst = max(worst, abs(r32 - r64))
        print(f"  step {done}: R_fp32c {r32:.8f} R_fp64 {r64:.8f} |diff| {abs(r32 - r64):.2e} ({time.time() - t0:.1f} s)",
              flush=True)
        assert abs(r32 - r64) <= 1e-4
    su32, _, _ = g32.verlet_status()
    su64, _, _ = g64.verlet_status()
    print(f"  worst |dR| {worst:.2e}", flush=True)
    for g in systems:
        g.close()
