"""Float64 CPU restatement of the learned-stop MLP (test infrastructure).

Forward, mean binary cross-entropy and its gradient for the parameter layout of
``hnsw_amd.learned`` (normalisation, W1 16x64, b1, W2, b2), with the weights rounded to
bfloat16 exactly as the GPU kernels round them (round to nearest even). Pinned in
tests/test_learned.py by finite differences.
"""

from __future__ import annotations

import numpy as np

F, H = 16, 64
W1, B1 = 32, 32 + F * H
W2, B2 = B1 + H, B1 + 2 * H
SIZE = B2 + 1


def bf16(x: np.ndarray) -> np.ndarray:
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def unpack(p: np.ndarray, round_weights: bool = True):
    q = bf16(p) if round_weights else np.asarray(p, np.float64)
    mean, istd = p[:F].astype(np.float64), p[F:W1].astype(np.float64)
    w1 = q[W1:B1].astype(np.float64).reshape(F, H)
    b1 = q[B1:W2].astype(np.float64)
    w2 = q[W2:B2].astype(np.float64)
    b2 = float(q[B2])
    return mean, istd, w1, b1, w2, b2


def forward(x: np.ndarray, p: np.ndarray, round_weights: bool = True) -> np.ndarray:
    mean, istd, w1, b1, w2, b2 = unpack(p, round_weights)
    xn = (x.astype(np.float64) - mean) * istd
    z = xn @ w1 + b1
    return np.maximum(z, 0.0) @ w2 + b2


def loss_and_grad(x: np.ndarray, y: np.ndarray, p: np.ndarray, pos_weight: float = 1.0,
                  round_weights: bool = True):
    """Mean loss and its gradient w.r.t. the (rounded) weights; normalisation slots get 0."""
    mean, istd, w1, b1, w2, b2 = unpack(p, round_weights)
    xn = (x.astype(np.float64) - mean) * istd
    z = xn @ w1 + b1
    h = np.maximum(z, 0.0)
    o = h @ w2 + b2
    t = y.astype(np.float64)
    sp_pos = np.logaddexp(0.0, -o)  # softplus(-o)
    sp_neg = np.logaddexp(0.0, o)
    loss = np.mean(pos_weight * t * sp_pos + (1 - t) * sp_neg)
    sig = 1.0 / (1.0 + np.exp(-o))
    dout = (pos_weight * t * (sig - 1.0) + (1 - t) * sig) / x.shape[0]
    g = np.zeros(SIZE, dtype=np.float64)
    g[B2] = dout.sum()
    g[W2:B2] = h.T @ dout
    dh = (dout[:, None] * w2[None, :]) * (z > 0)
    g[B1:W2] = dh.sum(axis=0)
    g[W1:B1] = (xn.T @ dh).reshape(-1)
    return loss, g
