"""CPU oracle of the learned scorer's network (TEST INFRASTRUCTURE ONLY:
tests check the HIP training and inference kernels against it; the engine
never imports it).

Restates in float64 numpy, for the network of relaibilitygpu_amd/csrc/
nr_learn.h (no counterpart in the reference):
  * forward     y_hat = b2 + relu(X @ W1) @ w2           (W1[16][64], row 15
                                                          against feature 15 = 1)
  * loss        mean((y_hat - y)^2 / 2)
  * gradient    d loss / d (W1, w2, b2), flattened in that order
  * inference   the BF16 form: W1' = bf16(W1 * |w2|), s = sign(w2),
                y = b2 + sum_h s_h relu(bf16(x) @ W1'[:, h])
Pinned by central finite differences of the loss (gradient) and by the
identity relu(z)|w| = relu(z|w|) (inference form equals the forward pass
up to BF16 rounding).
"""
from __future__ import annotations

import numpy as np

F, H = 16, 64
N_PARAMS = F * H + H + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    return p[: F * H].reshape(F, H), p[F * H: F * H + H], p[-1]


def forward(p, X):
    W1, w2, b2 = unpack(p)
    Z = np.asarray(X, np.float64) @ W1
    return b2 + np.maximum(Z, 0.0) @ w2


def loss_and_grad(p, X, y):
    """Mean loss and gradient over the rows of X."""
    W1, w2, b2 = unpack(p)
    X = np.asarray(X, np.float64)
    Z = X @ W1
    Hh = np.maximum(Z, 0.0)
    r = b2 + Hh @ w2 - np.asarray(y, np.float64)
    n = len(r)
    gW1 = X.T @ ((Z > 0) * np.outer(r, w2)) / n
    gw2 = Hh.T @ r / n
    gb2 = r.sum() / n
    return 0.5 * float(r @ r) / n, np.concatenate([gW1.ravel(), gw2, [gb2]])


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def infer_bf16(p, X):
    """The inference form the matrix-core kernel evaluates (float64 sums)."""
    W1, w2, b2 = unpack(p)
    W1s = bf16((W1 * np.abs(w2)[None, :]).astype(np.float32)).astype(np.float64)
    Xb = bf16(np.asarray(X, np.float32)).astype(np.float64)
    Z = Xb @ W1s
    s = np.where(w2 < 0, -1.0, 1.0)
    return np.float32(b2) + np.maximum(Z, 0.0) @ s
