"""FP64 CPU oracle of the tuner's cost model (test infrastructure only).

Same network and flat parameter layout as the HIP kernels in
gpuprefixsums_amd/csrc/mlp.hip:
    z1 = W1 x + b1, a1 = relu(z1); z2 = W2 a1 + b2, a2 = relu(z2); y = W3 a2 + b3
    loss = mean((y - t)^2)
params = [W1 (H x D row-major), b1 (H), W2 (H x H), b2 (H), W3 (H), b3 (1)].
No reference counterpart. Pinned by central finite differences of `loss`
(tests/test_oracle.py), so the analytic gradient here is the checker for the
GPU backward kernels.
"""
from __future__ import annotations

import numpy as np


def param_count(D: int, H: int) -> int:
    return H * D + H + H * H + H + H + 1


def unpack(p: np.ndarray, D: int, H: int):
    p = np.asarray(p, dtype=np.float64)
    i = 0
    W1 = p[i:i + H * D].reshape(H, D); i += H * D
    b1 = p[i:i + H]; i += H
    W2 = p[i:i + H * H].reshape(H, H); i += H * H
    b2 = p[i:i + H]; i += H
    W3 = p[i:i + H]; i += H
    b3 = p[i]
    return W1, b1, W2, b2, W3, b3


def forward(p, x, D: int, H: int) -> np.ndarray:
    W1, b1, W2, b2, W3, b3 = unpack(p, D, H)
    x = np.asarray(x, dtype=np.float64)
    a1 = np.maximum(x @ W1.T + b1, 0.0)
    a2 = np.maximum(a1 @ W2.T + b2, 0.0)
    return a2 @ W3 + b3


def loss(p, x, t, D: int, H: int) -> float:
    y = forward(p, x, D, H)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def loss_grad(p, x, t, D: int, H: int):
    """Returns (loss, flat gradient) by reverse-mode differentiation in FP64."""
    W1, b1, W2, b2, W3, b3 = unpack(p, D, H)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    B = x.shape[0]
    z1 = x @ W1.T + b1
    a1 = np.maximum(z1, 0.0)
    z2 = a1 @ W2.T + b2
    a2 = np.maximum(z2, 0.0)
    y = a2 @ W3 + b3
    d = y - t
    L = float(np.mean(d * d))
    dy = 2.0 * d / B
    gW3 = a2.T @ dy
    gb3 = dy.sum()
    dz2 = np.outer(dy, W3) * (z2 > 0)
    gW2 = dz2.T @ a1
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2 @ W2) * (z1 > 0)
    gW1 = dz1.T @ x
    gb1 = dz1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return L, g


def init_params(D: int, H: int, seed: int) -> np.ndarray:
    """He-uniform weights, zero biases (the tuner's initialisation)."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (H, D)) * np.sqrt(6.0 / D)
    W2 = rng.uniform(-1, 1, (H, H)) * np.sqrt(6.0 / H)
    W3 = rng.uniform(-1, 1, H) * np.sqrt(6.0 / H)
    return np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), W3, [0.0]])
