"""FP64 numpy oracle of the warm-start network (pflotran_amd/csrc/pfx_mlp.hip).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
forward, backward, Adam and feature kernels, never by the product.  The
warm-start network has no counterpart in the reference; this file restates
the textbook algorithms it uses:

  * dense layers Z = A W^T + b, GELU(z) = z/2 (1 + erf(z / sqrt 2)) on the
    hidden layers, a linear output layer;
  * mean squared error L = mean((Y - T)^2) and its gradients by
    back-propagation;
  * Adam (Kingma & Ba 2015, Algorithm 1) with bias correction;
  * the feature map of pfx_surrogate.h.

With bf16=True each matrix product rounds both operands to BF16 (round to
nearest even) as the GPU kernels do; products and sums are otherwise FP64.
Pinned by central finite differences of the loss (tests/test_mlp_oracle.py),
by Adam's closed form at step 1, and by known BF16 roundings.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

SQRT1_2 = 0.70710678118654752
INV_SQRT_2PI = 0.39894228040143268


def bf16_round(x) -> np.ndarray:
    """FP32 value of x rounded to BF16 (round to nearest even), as FP64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z * SQRT1_2))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z * SQRT1_2)) + z * INV_SQRT_2PI * np.exp(-0.5 * z * z)


def _q(x, bf16):
    return bf16_round(x) if bf16 else np.asarray(x, dtype=np.float64)


def forward(params, X, bf16=False):
    """params: [(W[n_out][n_in], b[n_out])]; returns (Y, cache)."""
    A = np.asarray(X, dtype=np.float64)
    acts, pre = [A], []
    L = len(params)
    for l, (W, b) in enumerate(params):
        Z = _q(A, bf16) @ _q(W, bf16).T + np.asarray(b, dtype=np.float64)
        pre.append(Z)
        A = Z if l == L - 1 else gelu(Z)
        acts.append(A)
    return A, (acts, pre)


def loss_and_grads(params, X, T, bf16=False):
    """Mean squared error over every output entry and its gradients
    [(dW, db)] by back-propagation."""
    Y, (acts, pre) = forward(params, X, bf16)
    T = np.asarray(T, dtype=np.float64)
    n = Y.size
    loss = float(np.sum((Y - T) ** 2) / n)
    dZ = 2.0 * (Y - T) / n
    grads = [None] * len(params)
    for l in range(len(params) - 1, -1, -1):
        W = params[l][0]
        dW = _q(dZ, bf16).T @ _q(acts[l], bf16)
        db = dZ.sum(axis=0)
        grads[l] = (dW, db)
        if l > 0:
            dA = _q(dZ, bf16) @ _q(W, bf16)
            dZ = dA * gelu_grad(pre[l - 1])
    return loss, grads


def adam_step(w, g, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update at step t (1-based); returns (w, m, v)."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mhat = m / (1.0 - beta1 ** t)
    vhat = v / (1.0 - beta2 ** t)
    return w - lr * mhat / (np.sqrt(vhat) + eps), m, v


def slog(v):
    v = np.asarray(v, dtype=np.float64)
    return np.sign(v) * np.log10(1.0 + np.abs(v) * 1e12)


def features(total, pri_molal, temperature, volfrac, nfeat, N, K):
    """Raw feature rows (slots of pfx_surrogate.h) of cell arrays: slog of
    the totals, ln of the previous free ions, temperature, volume fractions."""
    n = total.shape[0]
    X = np.zeros((n, nfeat))
    X[:, :N] = slog(total[:, :N])
    X[:, N:2 * N] = np.log(np.clip(pri_molal[:, :N], 1e-300, 1e300))
    X[:, 2 * N] = 25.0 if temperature is None else temperature
    if K > 0 and volfrac is not None:
        X[:, 2 * N + 1:2 * N + 1 + K] = volfrac[:, :K]
    return X
