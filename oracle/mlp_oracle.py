"""FP64 numpy restatement of the sweep-screen network (spss_mlp_* of the C ABI).

TEST INFRASTRUCTURE ONLY: the checker of the network kernels in tests/ (the product never
imports it).  There is no counterpart in the reference; the functions follow published
definitions:

* fully connected layers Z = A W^T + b with W [out x in] row-major, parameters flat, layer
  after layer (W_l then b_l);
* GELU(z) = z Phi(z) = z (1 + erf(z / sqrt 2)) / 2 (Hendrycks & Gimpel 2016, exact form), after
  every hidden layer; linear output;
* loss mean((Y - T)^2) over all outputs of the batch;
* Adam (Kingma & Ba 2015) in torch.optim.Adam's form without weight decay:
  m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
  p -= (lr / (1 - b1^t)) m / (sqrt(v) / sqrt(1 - b2^t) + eps).

Pinned in tests/test_oracle_mlp.py against torch (float64, CPU: functional linear/gelu,
autograd, torch.optim.Adam) and against central finite differences of the loss.
"""
import numpy as np
from scipy.special import erf

_INV_SQRT2 = 1.0 / np.sqrt(2.0)
_INV_SQRT2PI = 1.0 / np.sqrt(2.0 * np.pi)


def layout(dims):
    """[(w_off, b_off)] per layer and the parameter count for widths dims[0] -> ... -> dims[-1]"""
    offs, n = [], 0
    for l in range(len(dims) - 1):
        w = n
        n += dims[l + 1] * dims[l]
        offs.append((w, n))
        n += dims[l + 1]
    return offs, n


def unpack(params, dims):
    offs, _ = layout(dims)
    out = []
    for l, (w, b) in enumerate(offs):
        W = params[w:b].reshape(dims[l + 1], dims[l])
        out.append((W, params[b:b + dims[l + 1]]))
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z * _INV_SQRT2))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z * _INV_SQRT2)) + z * _INV_SQRT2PI * np.exp(-0.5 * z * z)


def forward(params, dims, X):
    """(Zs, As, Y): pre-activations per layer, inputs per layer, output"""
    layers = unpack(np.asarray(params, dtype=np.float64), dims)
    A = np.asarray(X, dtype=np.float64)
    Zs, As = [], []
    for l, (W, b) in enumerate(layers):
        As.append(A)
        Z = A @ W.T + b
        Zs.append(Z)
        A = gelu(Z) if l + 1 < len(layers) else Z
    return Zs, As, A


def loss_and_grad(params, dims, X, T):
    """(loss, flat gradient) of mean((Y - T)^2)"""
    params = np.asarray(params, dtype=np.float64)
    layers = unpack(params, dims)
    Zs, As, Y = forward(params, dims, X)
    T = np.asarray(T, dtype=np.float64)
    diff = Y - T
    loss = float(np.mean(diff * diff))
    D = 2.0 * diff / diff.size
    offs, n = layout(dims)
    g = np.zeros(n)
    for l in range(len(layers) - 1, -1, -1):
        w, b = offs[l]
        g[w:b] = (D.T @ As[l]).reshape(-1)
        g[b:b + dims[l + 1]] = D.sum(axis=0)
        if l > 0:
            D = (D @ layers[l][0]) * gelu_grad(Zs[l - 1])
    return loss, g


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """one Adam update at step t (1-based); returns new (p, m, v)"""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    bc1 = 1.0 - b1 ** t
    bc2 = 1.0 - b2 ** t
    p = p - (lr / bc1) * m / (np.sqrt(v) / np.sqrt(bc2) + eps)
    return p, m, v
