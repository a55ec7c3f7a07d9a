"""The sweep-screen network's HIP kernels (spss_mlp_*) against the FP64 oracle
(oracle/mlp_oracle.py, pinned against torch float64 and finite differences).

Tolerances (FP32 kernels against FP64): outputs within 2e-5 of max |y|; per-layer gradients
within 1e-4 relative in the 2-norm and the loss within 1e-5 relative; Adam parameters within
1e-2 lr where the gradient is well determined (|g| > 1e-4 max |g|; the update m/sqrt(v) amplifies
the FP32 gradient error where successive gradients nearly cancel), 1e-7 on average, and within
2 lr per step everywhere (Adam moves a parameter by at most about lr per step).
"""
import numpy as np
import pytest

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu

DIMS = (6, 128, 128, 128, 2)


@pytest.fixture(scope="module")
def spss():
    import spsslangevinintegration_amd as m
    return m


def _data(n, seed, dims=DIMS):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (n, dims[0])).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) + X[:, 1] ** 2, X[:, 2] * X[:, 3] - 0.5 * X[:, 4]], axis=1)
    return X, T[:, :dims[-1]].astype(np.float32)


def _layers(g, dims):
    offs, _ = MO.layout(dims)
    out = []
    for l, (w, b) in enumerate(offs):
        out.append(g[w:b + dims[l + 1]])
    return out


def test_initial_parameters_follow_the_linear_default(spss):
    net = spss.MLP(DIMS, max_batch=256, seed=11)
    p = net.params.astype(np.float64)
    for l, (W, b) in enumerate(MO.unpack(p, DIMS)):
        bound = 1.0 / np.sqrt(DIMS[l])
        assert np.all(np.abs(W) <= bound) and np.all(np.abs(b) <= bound)
        assert abs(W.std() - bound / np.sqrt(3)) < 0.1 * bound  # uniform on [-bound, bound]
    assert not np.array_equal(p, spss.MLP(DIMS, max_batch=256, seed=12).params)
    np.testing.assert_array_equal(net.params, spss.MLP(DIMS, max_batch=256, seed=11).params)


@pytest.mark.parametrize("n", [1, 37, 300, 1000])
def test_forward_matches_oracle(spss, n):
    net = spss.MLP(DIMS, max_batch=256, seed=3)  # n > 256 runs in chunks
    X, _ = _data(n, 1)
    Y = net.predict(X)
    _, _, Y64 = MO.forward(net.params.astype(np.float64), DIMS, X.astype(np.float64))
    err = np.max(np.abs(Y - Y64)) / np.max(np.abs(Y64))
    print(f"forward n={n}: max rel err {err:.2e}")
    assert err < 2e-5


@pytest.mark.parametrize("batch", [1, 33, 200, 512])
def test_gradients_match_oracle(spss, batch):
    net = spss.MLP(DIMS, max_batch=512, seed=4)
    X, T = _data(batch, 2)
    loss, g = net.gradients(X, T)
    loss64, g64 = MO.loss_and_grad(net.params.astype(np.float64), DIMS, X.astype(np.float64),
                                   T.astype(np.float64))
    assert abs(loss - loss64) <= 1e-5 * loss64
    for l, (a, b) in enumerate(zip(_layers(g.astype(np.float64), DIMS), _layers(g64, DIMS))):
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"batch={batch} layer {l}: gradient rel err {rel:.2e}")
        assert rel < 1e-4


def test_adam_steps_match_oracle(spss):
    n, steps, lr = 256, 3, 1e-3
    net = spss.MLP(DIMS, max_batch=256, seed=5)
    X, T = _data(n, 3)
    p = net.params.astype(np.float64)
    net.set_data(X, T)
    losses = net.train(steps, n, lr=lr)  # full batch: every step sees all rows
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    X64, T64 = X.astype(np.float64), T.astype(np.float64)
    for t in range(1, steps + 1):
        l64, g = MO.loss_and_grad(p, DIMS, X64, T64)
        assert abs(losses[t - 1] - l64) <= 1e-5 * l64
        gmax = np.max(np.abs(g))
        p, m, v = MO.adam_step(p, g, m, v, t, lr)
    q = net.params.astype(np.float64)
    well = np.abs(g) > 1e-4 * gmax
    err_well = np.max(np.abs(q - p)[well])
    print(f"adam: max |dp| {err_well:.2e} on {well.mean():.3f} of the parameters, mean {np.mean(np.abs(q - p)):.2e}")
    assert err_well < 1e-2 * lr
    assert np.mean(np.abs(q - p)) < 1e-7
    assert np.max(np.abs(q - p)) <= 2 * lr * steps
    assert net.adam_steps == steps


def test_training_fits_a_smooth_function_and_is_reproducible(spss):
    X, T = _data(1024, 6)
    runs = []
    for _ in range(2):
        net = spss.MLP(DIMS, max_batch=256, seed=7)
        net.set_data(X, T)
        losses = net.train(1500, 128, lr=2e-3)
        runs.append((losses, net.params))
    losses = runs[0][0]
    print(f"training loss {losses[:8].mean():.4f} -> {losses[-8:].mean():.6f}")
    assert losses[-8:].mean() < 0.02 * losses[:8].mean()
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    Xh, Th = _data(512, 8)  # held out
    net = spss.MLP(DIMS, max_batch=256, seed=7)
    net.params = runs[0][1]
    held = np.mean((net.predict(Xh) - Th) ** 2)
    assert held < 0.05 * np.mean((Th - Th.mean(axis=0)) ** 2)


def test_bad_arguments_raise(spss):
    net = spss.MLP((3, 8, 1), max_batch=16, seed=1)
    X = np.zeros((17, 3), np.float32)
    with pytest.raises(spss.SpssError):
        net.gradients(X, np.zeros((17, 1), np.float32))  # beyond max_batch
    with pytest.raises(spss.SpssError):
        net.train(1, 8)  # no data
    net.set_data(X[:4], np.zeros((4, 1), np.float32))
    with pytest.raises(spss.SpssError):
        net.train(1, 8)  # batch larger than the data
    with pytest.raises(spss.SpssError):
        net.params = np.zeros(3, np.float32)
    with pytest.raises(spss.SpssError):
        spss.MLP((3,), max_batch=16)
