"""Pins of the FP64 network oracle (oracle/mlp_oracle.py) that the GPU network kernels are
checked against.  Independent references: torch on the CPU in float64 (functional linear and
exact GELU, autograd, torch.optim.Adam) and central finite differences of the loss.  A dropped
bias, a transposed weight, the tanh GELU instead of the erf form, a missing factor 2 in the
loss gradient or a wrong bias correction in Adam fails one of them.
"""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as MO

DIMS = [6, 16, 16, 16, 2]


def _params(dims, seed):
    rng = np.random.default_rng(seed)
    _, n = MO.layout(dims)
    return rng.normal(0, 0.4, n)


def _torch_net(params, dims):
    layers = MO.unpack(params, dims)
    return [(torch.tensor(W, dtype=torch.float64, requires_grad=True),
             torch.tensor(b, dtype=torch.float64, requires_grad=True)) for W, b in layers]


def _torch_forward(tl, X):
    A = torch.tensor(X, dtype=torch.float64)
    for l, (W, b) in enumerate(tl):
        A = torch.nn.functional.linear(A, W, b)
        if l + 1 < len(tl):
            A = torch.nn.functional.gelu(A, approximate="none")
    return A


def test_layout_counts():
    offs, n = MO.layout([6, 128, 128, 128, 2])
    assert n == 6 * 128 + 128 + 2 * (128 * 128 + 128) + 128 * 2 + 2
    assert offs[0] == (0, 768) and offs[1] == (896, 896 + 128 * 128)


def test_gelu_matches_torch_and_its_derivative():
    z = np.linspace(-6, 6, 1001)
    zt = torch.tensor(z, dtype=torch.float64, requires_grad=True)
    g = torch.nn.functional.gelu(zt, approximate="none")
    g.sum().backward()
    np.testing.assert_allclose(MO.gelu(z), g.detach().numpy(), rtol=0, atol=1e-15)
    np.testing.assert_allclose(MO.gelu_grad(z), zt.grad.numpy(), rtol=0, atol=1e-14)
    # special values: gelu(0) = 0, gelu'(0) = 1/2, gelu(z) -> z for large z
    assert MO.gelu(np.array([0.0]))[0] == 0.0 and MO.gelu_grad(np.array([0.0]))[0] == 0.5
    assert abs(MO.gelu(np.array([10.0]))[0] - 10.0) < 1e-15


def test_forward_matches_torch():
    p = _params(DIMS, 1)
    X = np.random.default_rng(2).normal(size=(37, DIMS[0]))
    _, _, Y = MO.forward(p, DIMS, X)
    Yt = _torch_forward(_torch_net(p, DIMS), X).detach().numpy()
    np.testing.assert_allclose(Y, Yt, rtol=1e-13, atol=1e-14)


def test_gradient_matches_torch_autograd():
    p = _params(DIMS, 3)
    rng = np.random.default_rng(4)
    X = rng.normal(size=(29, DIMS[0]))
    T = rng.normal(size=(29, DIMS[-1]))
    loss, g = MO.loss_and_grad(p, DIMS, X, T)
    tl = _torch_net(p, DIMS)
    lt = torch.mean((_torch_forward(tl, X) - torch.tensor(T)) ** 2)
    lt.backward()
    gt = np.concatenate([np.concatenate([W.grad.numpy().reshape(-1), b.grad.numpy()]) for W, b in tl])
    assert abs(loss - lt.item()) <= 1e-14 * max(1.0, abs(loss))
    np.testing.assert_allclose(g, gt, rtol=1e-11, atol=1e-14)


def test_gradient_matches_finite_differences():
    dims = [6, 8, 8, 2]
    p = _params(dims, 5)
    rng = np.random.default_rng(6)
    X = rng.normal(size=(11, dims[0]))
    T = rng.normal(size=(11, dims[-1]))
    _, g = MO.loss_and_grad(p, dims, X, T)
    h = 1e-6
    fd = np.zeros_like(p)
    for k in range(p.size):
        e = np.zeros_like(p)
        e[k] = h
        fd[k] = (MO.loss_and_grad(p + e, dims, X, T)[0] - MO.loss_and_grad(p - e, dims, X, T)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9)


def test_adam_matches_torch():
    rng = np.random.default_rng(7)
    p0 = rng.normal(size=50)
    grads = [rng.normal(size=50) * s for s in (1.0, 0.3, 2.0, 1e-3)]
    pt = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([pt], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for t, g in enumerate(grads, start=1):
        pt.grad = torch.tensor(g, dtype=torch.float64)
        opt.step()
        p, m, v = MO.adam_step(p, g, m, v, t, 3e-3)
        np.testing.assert_allclose(p, pt.detach().numpy(), rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("mistake", ["no_bias", "tanh_gelu", "half_grad"])
def test_planted_mistakes_are_caught(mistake):
    """the checks above distinguish the oracle from plausible wrong versions of it"""
    p = _params(DIMS, 8)
    rng = np.random.default_rng(9)
    X = rng.normal(size=(13, DIMS[0]))
    T = rng.normal(size=(13, DIMS[-1]))
    tl = _torch_net(p, DIMS)
    Yt = _torch_forward(tl, X).detach().numpy()
    if mistake == "no_bias":
        layers = MO.unpack(p, DIMS)
        A = X
        for l, (W, b) in enumerate(layers):
            A = A @ W.T
            A = MO.gelu(A) if l + 1 < len(layers) else A
        assert np.max(np.abs(A - Yt)) > 1e-3
    elif mistake == "tanh_gelu":
        z = np.linspace(-3, 3, 61)
        approx = 0.5 * z * (1 + np.tanh(np.sqrt(2 / np.pi) * (z + 0.044715 * z ** 3)))
        exact = torch.nn.functional.gelu(torch.tensor(z), approximate="none").numpy()
        assert np.max(np.abs(approx - exact)) > 1e-5
        assert np.max(np.abs(MO.gelu(z) - exact)) < 1e-15
    else:
        _, g = MO.loss_and_grad(p, DIMS, X, T)
        lt = torch.mean((_torch_forward(tl, X) - torch.tensor(T)) ** 2)
        lt.backward()
        gt = np.concatenate([np.concatenate([W.grad.numpy().reshape(-1), b.grad.numpy()]) for W, b in tl])
        assert np.max(np.abs(0.5 * g - gt)) > 1e-6
