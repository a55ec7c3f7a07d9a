"""Batches of independent replicas (spss_batch_*, ReplicaBatch) against single systems.

Replica r of a batch must step exactly like a single FP32C system with the same parameters,
natural frequencies and Philox seed run through spss_run: bit-identical states, whatever the
tile widths (default, set by hand, or picked by the autotuner) and whether the steps run
eagerly or as graph replays.  The per-replica order-parameter readings equal the order
parameter of the replica's state (FP32 phasors summed in FP64: within 1e-6) and their time
averages follow the readings of a single system.
"""
import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spss():
    import spsslangevinintegration_amd as m
    return m


def _replicas(model, R, seed):
    rng = np.random.default_rng(seed)
    ns = rng.integers(2048, 9000, R)
    m = rng.choice([0.5, 1.0], R)
    sigma = rng.uniform(0, 4, R)
    D = rng.uniform(0, 0.2, R)
    if model == 4:
        xi, sig, Dp = 1.0 / m, sigma / m, D / m ** 2
    else:
        xi, sig, Dp = np.zeros(R), sigma, D
    seeds = 0x5EED0004 + np.arange(R)
    states, nats = [], []
    for r in range(R):
        st = np.zeros((ns[r], model))
        st[:, 0], st[:, 1] = rng.random(ns[r]), rng.random(ns[r])
        st[:, 2] = rng.random(ns[r]) * 2 * np.pi
        states.append(st.reshape(-1))
        w = rng.normal(0, 0.2, ns[r])
        nats.append(w / m[r] if model == 4 else w)
    return ns, xi, sig, Dp, seeds, states, nats


def _singles(spss, model, ns, xi, sig, Dp, seeds, states, nats, K, dt, v0, rho, k=1, xs=1):
    out = []
    for r in range(len(ns)):
        g = spss.ParticleSystem(model=model, n=int(ns[r]), v0=v0, xi=xi[r], sigma=sig[r], rho=rho, D_phi=Dp[r],
                                natfreq_local=1, precision=spss.PRECISION_FP32C, seed=int(seeds[r]),
                                cell_subdivision=k)
        if xs != 1:
            g.set_cell_layout(k, xs)
        g.set_natural_frequencies(nats[r])
        g.set_state(states[r])
        g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, K, dt, spss.NOISE_PHILOX)
        out.append(g.get_state())
        g.close()
    return out


@pytest.mark.parametrize("model,graph", [(4, True), (4, False), (3, True)])
def test_batch_equals_single_systems(spss, model, graph):
    R, K, dt, v0, rho = 6, 24, 0.005, 0.01, 0.03
    print(f"[gpu] batch model={model} graph={graph}", flush=True)
    ns, xi, sig, Dp, seeds, states, nats = _replicas(model, R, 10 + model)
    b = spss.ReplicaBatch(model, ns, xi, sig, Dp, seeds, v0=v0, rho=rho)
    b.set_natural_frequencies(np.concatenate(nats))
    b.set_state(np.concatenate(states))
    b.run(K, dt, observe_every=4, observe_from=8, graph=graph)
    got = b.get_state()
    want = _singles(spss, model, ns, xi, sig, Dp, seeds, states, nats, K, dt, v0, rho)
    for r in range(R):
        assert np.array_equal(b.replica_state(got, r), want[r]), r
    mean, var, cnt, z, step = b.order_stats()
    assert step == K
    # readings after steps 4, 8, ..., 24; averaged from step 8 on (5 readings)
    assert np.all(cnt == 5)
    for r in range(R):
        re, im = pyoracle.order_parameter(want[r], int(ns[r]), model, 2)
        assert abs(z[r] - complex(re, im)) <= 1e-6
    assert np.all(mean >= 0) and np.all(mean <= 1) and np.all(var >= 0)


def test_batch_time_averages_follow_readings(spss):
    model, R, dt, v0, rho = 4, 3, 0.005, 0.01, 0.03
    ns, xi, sig, Dp, seeds, states, nats = _replicas(model, R, 21)
    b = spss.ReplicaBatch(model, ns, xi, sig, Dp, seeds, v0=v0, rho=rho)
    b.set_natural_frequencies(np.concatenate(nats))
    b.set_state(np.concatenate(states))
    b.run(30, dt, observe_every=3, observe_from=9, graph=True)
    mean, var, cnt, z, step = b.order_stats()
    for r in range(R):
        g = spss.ParticleSystem(model=model, n=int(ns[r]), v0=v0, xi=xi[r], sigma=sig[r], rho=rho, D_phi=Dp[r],
                                natfreq_local=1, precision=spss.PRECISION_FP32C, seed=int(seeds[r]))
        g.set_natural_frequencies(nats[r])
        g.set_state(states[r])
        rs = []
        for k in range(10):
            g.run(spss.STEPPER_EULER, spss.METHOD_LINKED_LIST, 3, dt, spss.NOISE_PHILOX)
            if 3 * (k + 1) >= 9:
                rs.append(np.hypot(*g.order_parameter(2)))
        rs = np.array(rs)
        assert cnt[r] == len(rs)
        assert abs(mean[r] - rs.mean()) <= 1e-6
        assert abs(var[r] - rs.var()) <= 1e-6


def test_batch_tile_widths_and_autotune_do_not_change_results(spss):
    """Also with a dense spot (600 particles inside one cell of replica 1): tiles wide enough to
    overflow the LDS stage there take the global-memory walk, narrow ones stay staged; both
    visit the same candidates in the same order."""
    model, R, K, dt, v0, rho = 4, 5, 12, 0.005, 0.01, 0.03
    ns, xi, sig, Dp, seeds, states, nats = _replicas(model, R, 31)
    rng = np.random.default_rng(5)
    s1 = states[1].reshape(-1, model)
    s1[:600, 0] = 0.5 + 0.02 * rng.random(600)
    s1[:600, 1] = 0.5 + 0.02 * rng.random(600)
    states[1] = s1.reshape(-1)
    outs = []
    for mode in ("default", "manual", "tuned"):
        b = spss.ReplicaBatch(model, ns, xi, sig, Dp, seeds, v0=v0, rho=rho)
        b.set_natural_frequencies(np.concatenate(nats))
        b.set_state(np.concatenate(states))
        if mode == "manual":
            b.set_tile_widths(np.array([1, 3, 7, 11, 33]))
        elif mode == "tuned":
            table, chosen = b.autotune(dt, probe_steps=6)
            print(f"  autotune ms/step {np.round(table, 4)} chosen {chosen}", flush=True)
            assert np.all(table > 0) and 0 <= chosen < 6
            assert b.order_stats()[4] == 0  # step counter restored
        b.run(K, dt, observe_every=0, graph=True)
        outs.append(b.get_state())
        b.close()
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_batch_cell_autotune_keeps_replicas_equal_to_single_systems(spss):
    """the cell-size autotuner rebuilds the batch at every cell layout, restores state and step
    counter, and leaves a batch whose replicas are bit-identical to single FP32C systems of the
    chosen layout"""
    model, R, K, dt, v0, rho = 4, 4, 12, 0.005, 0.01, 0.03
    print("[gpu] batch cell autotune", flush=True)
    ns, xi, sig, Dp, seeds, states, nats = _replicas(model, R, 77)
    b = spss.ReplicaBatch(model, ns, xi, sig, Dp, seeds, v0=v0, rho=rho)
    b.set_natural_frequencies(np.concatenate(nats))
    b.set_state(np.concatenate(states))
    table, k = b.autotune_cells(dt, probe_steps=4)
    kc, xs = b.cell_layout
    print(f"  ms/step by layout {spss.CELL_LAYOUTS}: {np.round(table, 4)} chosen {(kc, xs)}", flush=True)
    assert k == kc and b.cell_subdivision == k
    assert table[spss.CELL_LAYOUTS.index((kc, xs))] == table[table > 0].min()
    assert b.order_stats()[4] == 0  # step counter restored
    np.testing.assert_array_equal(b.get_state(), np.concatenate(states))
    b.run(K, dt, observe_every=0, graph=True)
    got = b.get_state()
    b.close()
    want = _singles(spss, model, ns, xi, sig, Dp, seeds, states, nats, K, dt, v0, rho, k=kc, xs=xs)
    off = 0
    for r in range(R):
        m = model * int(ns[r])
        np.testing.assert_array_equal(got[off:off + m], want[r])
        off += m


def test_batches_run_side_by_side_equal_sequential_runs(spss):
    """Batches driven from host threads of their own (each batch has its own stream; the library's
    error state is per thread) give the same states and order-parameter readings, bit for bit, as
    the same batches run one after the other: what the ensemble benchmark relies on."""
    from concurrent.futures import ThreadPoolExecutor
    dt, v0, rho, K = 0.005, 0.01, 0.03, 30
    print("[gpu] batches side by side", flush=True)
    specs = [(4, _replicas(4, 4, 41)), (3, _replicas(3, 3, 42)), (4, _replicas(4, 2, 43))]

    def make():
        out = []
        for model, (ns, xi, sig, Dp, seeds, states, nats) in specs:
            b = spss.ReplicaBatch(model, ns, xi, sig, Dp, seeds, v0=v0, rho=rho)
            b.set_natural_frequencies(np.concatenate(nats))
            b.set_state(np.concatenate(states))
            out.append(b)
        return out

    results = []
    for concurrent in (False, True):
        bs = make()
        if concurrent:
            with ThreadPoolExecutor(max_workers=len(bs)) as ex:
                for f in [ex.submit(b.run, K, dt, 3, 9, True) for b in bs]:
                    f.result()
        else:
            for b in bs:
                b.run(K, dt, observe_every=3, observe_from=9, graph=True)
        results.append([(b.get_state(), b.order_stats()[:3]) for b in bs])
        for b in bs:
            b.close()
    for (s0, z0), (s1, z1) in zip(results[0], results[1]):
        np.testing.assert_array_equal(s0, s1)
        for a, c in zip(z0, z1):
            np.testing.assert_array_equal(a, c)
