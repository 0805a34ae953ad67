"""TEST INFRASTRUCTURE: NumPy restatement of the rank-predictor MLP.

The reference has no learned component (the learned-rank truncation is an
addition of this build), so this oracle follows the network definition of the
HIP kernels directly:

    y = w3 . relu(W2 relu(W1 x + b1) + b2) + b3,  loss = mean((y - t)^2)

with parameters packed as W1 (H x F), b1, W2 (H x H), b2, w3 (H), b3, and the
bias-corrected Adam update.  Pinned by central finite differences of the loss
(tests/test_oracle_mlp.py), not by the code it is compared with.  Computed in
float64 so that it checks the float32 kernels rather than sharing their
rounding.
"""

from __future__ import annotations

import numpy as np


def param_count(F: int, H: int) -> int:
    return H * F + H + H * H + H + H + 1


def unpack(p: np.ndarray, F: int, H: int):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def init_params(F: int, H: int, rng: np.random.Generator) -> np.ndarray:
    """He-style initialisation (FP32 values)."""
    W1 = rng.normal(scale=np.sqrt(2.0 / F), size=(H, F))
    W2 = rng.normal(scale=np.sqrt(2.0 / H), size=(H, H))
    w3 = rng.normal(scale=np.sqrt(1.0 / H), size=H)
    parts = [W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), w3, np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


def forward(p, x, F, H):
    W1, b1, W2, b2, w3, b3 = unpack(p, F, H)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    y = h2 @ w3 + b3
    return h1, h2, y


def loss_and_grad(p, x, t, F, H):
    """MSE loss and its gradient in the packed parameter layout."""
    W1, b1, W2, b2, w3, b3 = unpack(p, F, H)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    B = x.shape[0]
    h1, h2, y = forward(p, x, F, H)
    d = y - t
    loss = float(np.mean(d * d))
    dy = 2.0 * d / B
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    dh2 = np.outer(dy, w3) * (h2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(axis=0)
    dh1 = (dh2 @ W2) * (h1 > 0)
    gW1 = dh1.T @ x
    gb1 = dh1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g


def adam(p, g, m, v, lr, b1, b2, eps, step):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1 = 1 - b1 ** step
    c2 = 1 - b2 ** step
    p = p - lr * (m / c1) / (np.sqrt(v / c2) + eps)
    return p, m, v
