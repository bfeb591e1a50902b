"""CPU checks of the selector-network model (oracle/mlp.py).

The model's gradient is the checker of the HIP backward kernels; here it is
pinned against central finite differences of its own loss, and its forward
pass against a direct per-sample evaluation of the layer equations.
"""
from __future__ import annotations

import numpy as np

from oracle import mlp


def _setup(seed=0, n=9):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, mlp.NPARAM)
    f = rng.normal(0, 2, (n, mlp.IN))
    t = rng.normal(0, 1, (n, mlp.OUT))
    mu = rng.normal(0, 1, mlp.IN)
    sd = rng.uniform(0.5, 2, mlp.IN)
    return p, f, t, mu, sd


def test_gradient_matches_finite_differences():
    p, f, t, mu, sd = _setup()
    _, g = mlp.loss_and_grad(p, f, t, mu, sd)
    rng = np.random.default_rng(1)
    idx = rng.choice(mlp.NPARAM, 300, replace=False)
    # cover every block: W1, b1, W2, b2, W3, b3
    bounds = np.cumsum([0, mlp.H * mlp.IN, mlp.H, mlp.H * mlp.H, mlp.H, mlp.OUT * mlp.H, mlp.OUT])
    idx = np.concatenate([idx, bounds[:-1], bounds[1:] - 1])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = mlp.loss_and_grad(p + e, f, t, mu, sd)
        lm, _ = mlp.loss_and_grad(p - e, f, t, mu, sd)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_forward_matches_layer_equations():
    p, f, t, mu, sd = _setup(3, 4)
    W1, b1, W2, b2, W3, b3 = mlp.unpack(p)
    _, _, _, y = mlp.forward(p, f, mu, sd)
    for r in range(f.shape[0]):
        x = (f[r] - mu) / sd
        h1 = [max(0.0, sum(W1[j, i] * x[i] for i in range(mlp.IN)) + b1[j]) for j in range(mlp.H)]
        h2 = [max(0.0, sum(W2[j, i] * h1[i] for i in range(mlp.H)) + b2[j]) for j in range(mlp.H)]
        yo = [sum(W3[o, i] * h2[i] for i in range(mlp.H)) + b3[o] for o in range(mlp.OUT)]
        assert np.allclose(y[r], yo, rtol=1e-12, atol=1e-12)


def test_planted_mistake_in_gradient_is_caught():
    # a backward pass that forgets the ReLU mask of the first layer
    p, f, t, mu, sd = _setup(5)
    W1, b1, W2, b2, W3, b3 = mlp.unpack(p)
    x, h1, h2, y = mlp.forward(p, f, mu, sd)
    dy = 2.0 * (y - t) / (f.shape[0] * mlp.OUT)
    dh2 = (dy @ W3) * (h2 > 0)
    dh1_wrong = dh2 @ W2
    g_wrong = (dh1_wrong.T @ x).ravel()
    _, g = mlp.loss_and_grad(p, f, t, mu, sd)
    assert not np.allclose(g_wrong, g[:mlp.H * mlp.IN], rtol=1e-3, atol=1e-9)
