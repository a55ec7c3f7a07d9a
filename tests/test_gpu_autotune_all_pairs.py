"""All-pairs autotuner (spss_autotune_all_pairs): every workgroup / j-tile shape of the all-pairs
kernel visits j in ascending order (EvaluateInteractionsWithAllPairs, ParticleSystemFirstOrderPtr.cpp:
225-316), so the RHS is bit-identical for all six shapes and the tuned handle still replays the
reference's 1000-step RK4 run within 1e-10 (golden rk4_allpairs_s4)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spss():
    import spsslangevinintegration_amd as m
    return m


@pytest.mark.parametrize("model,n", [(4, 1024), (6, 700), (3, 3001)])
def test_all_pairs_shapes_give_identical_rhs(spss, model, n):
    print(f"[gpu] all-pairs shapes model={model} n={n}", flush=True)
    rng = np.random.default_rng(n)
    st = rng.random(model * n)
    g = spss.ParticleSystem(model=model, n=n, rho=0.1, sigma=1.3, alpha=0.2, v0=0.02)
    g.set_higher_order(model == 4)
    outs = []
    for c in range(6):
        g.set_all_pairs_config(c)
        k = np.zeros(model * n)
        g.evaluate_rhs(spss.METHOD_ALL_PAIRS, st, np.zeros(model * n), k, 0.0, 0.01)
        outs.append((k, g.complementary_force() if model == 4 else None, g.all_pairs_config))
    for k, c, shape in outs[1:]:
        assert np.array_equal(k, outs[0][0]), shape
        if model == 4:
            assert np.array_equal(c, outs[0][1]), shape
    table, chosen = g.autotune_all_pairs(probe_evals=2)
    print(f"  ms per evaluation {np.round(table, 4).tolist()}, chosen {chosen} -> {g.all_pairs_config}", flush=True)
    assert 0 <= chosen < 6 and table[chosen] == table.min()


def test_tuned_handle_replays_the_reference_run(spss):
    from tests import golden_replay, gpu_engine
    print("[gpu] golden rk4_allpairs_s4 (1000 RK4 steps) through a tuned handle", flush=True)
    chosen = []

    def factory(**kw):
        eng = gpu_engine.GpuEngine(**kw)
        eng.sys.set_state(np.random.default_rng(0).random(kw["model"] * kw["n"]))
        table, c = eng.sys.autotune_all_pairs(probe_evals=2)
        eng.sys.set_all_pairs_config(5 if c == 4 else c)  # a shape other than the default {256, 256}
        chosen.append(eng.sys.all_pairs_config)
        return eng

    res = golden_replay.replay("rk4_allpairs_s4", factory)
    print(f"  shape {chosen[0]}", flush=True)
    for step, got, want in res["states"]:
        assert golden_replay.relerr(got, want) <= 1e-10, step
