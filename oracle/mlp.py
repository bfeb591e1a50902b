"""CPU model of the tile selector's network, in numpy float64.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
training and inference kernels (blis_amd/csrc/mlp.hip), never as the thing
measured. The reference has no counterpart.

Layout of the parameter vector, all row-major:
    W1[H][IN] b1[H] W2[H][H] b2[H] W3[OUT][H] b3[OUT]
Network: x = (f - mu) / sd; h1 = relu(W1 x + b1); h2 = relu(W2 h1 + b2);
y = W3 h2 + b3. Loss: mean over batch and outputs of (y - t)^2.
The gradients below are the textbook back-propagation of that loss; they are
pinned against central finite differences in tests/test_mlp_oracle.py.
"""
from __future__ import annotations

import numpy as np

IN, H, OUT = 16, 64, 42
NPARAM = H * IN + H + H * H + H + OUT * H + OUT


def unpack(p: np.ndarray):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * IN].reshape(H, IN); o += H * IN
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + OUT * H].reshape(OUT, H); o += OUT * H
    b3 = p[o:o + OUT]; o += OUT
    assert o == NPARAM
    return W1, b1, W2, b2, W3, b3


def forward(p, feats, mu, sd):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = (np.asarray(feats, np.float64) - mu) / sd
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ W3.T + b3
    return x, h1, h2, y


def loss_and_grad(p, feats, targets, mu, sd):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x, h1, h2, y = forward(p, feats, mu, sd)
    n = x.shape[0]
    d = y - np.asarray(targets, np.float64)
    loss = float(np.mean(d * d))
    dy = 2.0 * d / (n * OUT)
    gW3 = dy.T @ h2
    gb3 = dy.sum(axis=0)
    dh2 = (dy @ W3) * (h2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(axis=0)
    dh1 = (dh2 @ W2) * (h1 > 0)
    gW1 = dh1.T @ x
    gb1 = dh1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def predict(p, feats, mu, sd) -> np.ndarray:
    return np.argmin(forward(p, feats, mu, sd)[3], axis=1)


def load_weights(path):
    """Read the library's selector file: b'BLAMDMLP', int32 [1, IN, H, OUT],
    mu[IN], sd[IN], params (float32)."""
    raw = open(path, "rb").read()
    assert raw[:8] == b"BLAMDMLP", "not a selector weight file"
    hdr = np.frombuffer(raw[8:24], dtype="<i4")
    assert tuple(hdr) == (1, IN, H, OUT), hdr
    body = np.frombuffer(raw[24:], dtype="<f4")
    mu, sd, p = body[:IN], body[IN:2 * IN], body[2 * IN:]
    assert p.size == NPARAM
    return p.astype(np.float64), mu.astype(np.float64), sd.astype(np.float64)
