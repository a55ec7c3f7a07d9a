"""The oracle's threaded linked-list particle loop (the all-core CPU baseline of bench.py) gives
the serial loop's results bit for bit: each particle's sums are formed by one thread in the
serial order, whatever the block partition (Parallelization/ThreadSharedMemory.cpp:20-23)."""
import numpy as np
import pytest

import oracle
from oracle.pyoracle import METHOD_LINKED_LIST


@pytest.mark.parametrize("model,n,threads,higher", [(4, 5000, 3, False), (4, 4096, 8, True), (3, 3001, 5, False),
                                                     (6, 2000, 4, False)])
def test_threaded_linked_list_equals_serial(model, n, threads, higher):
    rng = np.random.default_rng(n)
    st = rng.random((n, model))
    st[:, :2] = rng.random((n, 2)) * 1.2 - 0.1
    st = st.reshape(-1)
    kw = dict(model=model, n=n, v0=0.01, xi=0.7, sigma=1.3, rho=0.05, alpha=0.2, natfreq_local=1)
    outs = []
    for t in (1, threads):
        o = oracle.Oracle(**kw)
        o.set_natural_frequencies(np.linspace(-0.1, 0.1, n))
        o.set_threads(t)
        if higher:
            o.set_higher_order(1)
        kn = np.zeros(model * n)
        o.evaluate_rhs(METHOD_LINKED_LIST, st, np.zeros(model * n), kn, 0.0, 0.01)
        outs.append((kn, o.complementary_force()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])


def test_threaded_euler_steps_equal_serial():
    n = 6000
    rng = np.random.default_rng(3)
    st = np.zeros((n, 4))
    st[:, 0], st[:, 1], st[:, 2] = rng.random(n), rng.random(n), rng.random(n) * 2 * np.pi
    st = st.reshape(-1)
    res = []
    for t in (1, 7):
        o = oracle.Oracle(model=4, n=n, v0=0.01, xi=2.0, sigma=3.0, rho=0.03, D_phi=0.08)
        o.set_threads(t)
        s = st.copy()
        r2 = np.random.default_rng(9)
        for _ in range(3):
            o.euler_step(METHOD_LINKED_LIST, s, 0.002, r2.normal(size=n))
        res.append(s)
    assert np.array_equal(res[0], res[1])
