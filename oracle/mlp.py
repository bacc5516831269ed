"""Oracle of the learned warm start's network: a numpy FP64 restatement of the
multilayer perceptron (ReLU hidden layers, linear output, mean-squared error)
and of its gradients, with optional bf16 rounding at exactly the points where
the GPU kernels store bf16 (inputs, weights, hidden activations and the
backward signals).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  The reference has no
learned part; this module has no reference file to follow.  Its backward pass
is pinned by central finite differences of `loss` in FP64 without rounding
(tests/test_oracle.py), and its bf16 rounding by torch's own float32 ->
bfloat16 conversion.
"""

from __future__ import annotations

import numpy as np


def bf16(x):
    """Rounds to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(x, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _q(x, quant):
    return bf16(x) if quant else np.asarray(x, np.float64)


def forward(Ws, bs, X, quant=True, S=None):
    """Returns (activations H_0..H_{L-1}, output Y).  Ws[l] is [d_{l+1}, d_l];
    S [d_out, d_in] is the optional linear skip (Y += X S^T)."""
    H = [_q(X, quant)]
    L = len(Ws)
    Y = None
    for l in range(L):
        Z = H[-1] @ _q(Ws[l], quant).T + bs[l]
        if l < L - 1:
            H.append(_q(np.maximum(Z, 0.0), quant))
        else:
            Y = Z
    if S is not None:
        Y = Y + H[0] @ _q(S, quant).T
    return H, Y


def loss(Ws, bs, X, T, quant=False, S=None):
    _, Y = forward(Ws, bs, X, quant, S)
    return float(np.mean((Y - T) ** 2))


def backward(Ws, bs, X, T, quant=True, S=None):
    """Gradients of mean((Y - T)^2) over all entries: (loss, dWs, dbs), plus
    the skip gradient dS as a fourth item when S is given."""
    H, Y = forward(Ws, bs, X, quant, S)
    M, N = Y.shape
    D = Y - T
    G = _q(2.0 / (M * N) * D, quant)
    L = len(Ws)
    dWs = [None] * L
    dbs = [None] * L
    dS = G.T @ H[0] if S is not None else None
    for l in range(L - 1, -1, -1):
        dWs[l] = G.T @ H[l]
        dbs[l] = G.sum(axis=0)
        if l > 0:
            G = _q((G @ _q(Ws[l], quant)) * (H[l] > 0), quant)
    if S is not None:
        return float(np.mean(D ** 2)), dWs, dbs, dS
    return float(np.mean(D ** 2)), dWs, dbs


def adam(w, g, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One bias-corrected Adam step (Kingma & Ba 2015, Algorithm 1) with
    decoupled weight decay (Loshchilov & Hutter 2019, Algorithm 2 with a
    schedule multiplier of 1); returns (w, m1, m2)."""
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    w = w - lr * ((m1 / (1 - b1 ** step)) / (np.sqrt(m2 / (1 - b2 ** step)) + eps) + wd * w)
    return w, m1, m2
