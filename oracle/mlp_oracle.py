"""ORACLE — test infrastructure only.

CPU restatement of the per-column surrogate network of thor_mlp.h in numpy
FP64: n_in -> n_hidden GELU -> n_hidden GELU -> n_out, mean-squared loss and
its gradient. GELU in its tanh form, x sigma(2 c (x + 0.044715 x^3)) with
c = sqrt(2 / pi) (= 0.5 x (1 + tanh(c (x + 0.044715 x^3)))). With bf16=True every tensor the GPU keeps in BF16 (inputs,
weight copies, hidden activations, backward operands) is rounded to BF16 at
the same point (round to nearest even), and the arithmetic is exact FP64, so
the GPU kernels (FP32 accumulation) must agree to FP32 summation error. With
bf16=False it is a plain FP64 network, which a finite-difference check pins.

Only tests/ load this module.
"""
from __future__ import annotations

import numpy as np

__all__ = ["bf16", "split", "forward", "loss_and_grad"]


def bf16(x: np.ndarray) -> np.ndarray:
    """Rounds to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def split(p: np.ndarray, n_in: int, nh: int, n_out: int):
    """Flat parameters -> W1 [n_in, nh], b1, W2 [nh, nh], b2, W3 [nh, n_out], b3."""
    p = np.asarray(p, dtype=np.float64)
    sizes = [n_in * nh, nh, nh * nh, nh, nh * n_out, n_out]
    shapes = [(n_in, nh), (nh,), (nh, nh), (nh,), (nh, n_out), (n_out,)]
    out, o = [], 0
    for n, sh in zip(sizes, shapes):
        out.append(p[o:o + n].reshape(sh))
        o += n
    return out


_C2 = 2.0 * np.sqrt(2.0 / np.pi)  # 1.5957691216057308
_A = 0.044715


def _gelu(z):
    u = _C2 * z * (1.0 + _A * z * z)
    return z / (1.0 + np.exp(-u))


def _gelu_grad(z):
    u = _C2 * z * (1.0 + _A * z * z)
    s = 1.0 / (1.0 + np.exp(-u))
    return s + z * s * (1.0 - s) * _C2 * (1.0 + 3.0 * _A * z * z)


def forward(p, X, dims, bf=True):
    n_in, nh, n_out = dims
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = split(p, n_in, nh, n_out)
    X0 = r(X)
    Z1 = X0 @ r(W1) + b1
    H1 = r(_gelu(Z1))
    Z2 = H1 @ r(W2) + b2
    H2 = r(_gelu(Z2))
    Y = H2 @ r(W3) + b3
    return Y, (X0, Z1, H1, Z2, H2)


def loss_and_grad(p, X, T, dims, bf=True):
    """Mean squared error over all outputs and its gradient (flat layout)."""
    n_in, nh, n_out = dims
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = split(p, n_in, nh, n_out)
    Y, (X0, Z1, H1, Z2, H2) = forward(p, X, dims, bf)
    n = Y.size
    d = Y - T
    loss = float((d * d).sum() / n)
    dY = (2.0 / n) * d
    if bf:
        dY = np.asarray(dY, dtype=np.float32).astype(np.float64)  # the GPU keeps dY in FP32 ...
    dYb = r(dY)                                                    # ... and its BF16 copy for the products
    gW3 = H2.T @ dYb
    gb3 = dY.sum(axis=0)
    dZ2 = (dYb @ r(W3).T) * _gelu_grad(Z2)
    dZ2b = r(dZ2)
    gW2 = H1.T @ dZ2b
    gb2 = dZ2.sum(axis=0)
    dZ1 = (dZ2b @ r(W2).T) * _gelu_grad(Z1)
    dZ1b = r(dZ1)
    gW1 = X0.T @ dZ1b
    gb1 = dZ1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
