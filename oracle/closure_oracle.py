"""CPU oracle for the learned closure's network: a multilayer perceptron with BF16 operands and FP32
accumulation, its backward pass, the mean-squared-error loss and Adam.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it; the product path (``scattersim_amd.closure``) runs the HIP kernels and never
falls back here.

There is no counterpart in the reference (it has no learned component).  The arithmetic mirrors the device
kernels: GEMM operands are rounded to BF16 (nearest, ties to even) and accumulated in FP32 (here in
float64, which differs from the device's FP32 sums by rounding only); activations are the tanh form of
GELU; hidden activations and back-propagated gradients are rounded to BF16 before they enter a GEMM.

Pins (``tests/test_closure_cpu.py``): ``bf16`` against hand-worked bit patterns (ties to even,
overflow to inf, NaN kept), ``gelu`` and its derivative against a central finite difference, and the
whole backward pass against central finite differences of the loss in float64 with rounding switched
off.
"""
from __future__ import annotations

import numpy as np


def bf16(x, enabled: bool = True) -> np.ndarray:
    """Round to BF16 (nearest, ties to even) and return float32; float64 unchanged when not enabled."""
    if not enabled:
        return np.array(x, np.float64)
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    nan = special & ((u & 0xFFFF) != 0)
    r = np.where(special, (u >> 16) | np.where(nan, np.uint64(0x40), np.uint64(0)), r)
    return (r.astype(np.uint32) << 16).view(np.float32)


def gelu(z):
    """tanh form of GELU and its derivative."""
    z = np.asarray(z, np.float64)
    c, a = 0.7978845608028654, 0.044715
    t = np.tanh(c * (z + a * z ** 3))
    return 0.5 * z * (1 + t), 0.5 * (1 + t) + 0.5 * z * (1 - t * t) * c * (1 + 3 * a * z * z)


def forward(params, X, rounding: bool = True):
    """params: [(W (out x in), b (out))]; GELU after every layer but the last.  Returns (Y, cache)."""
    H = bf16(X, rounding).astype(np.float64)
    cache = []
    for i, (W, b) in enumerate(params):
        Wb = bf16(W, rounding).astype(np.float64)
        Z = H @ Wb.T + b
        last = i == len(params) - 1
        cache.append((H, Wb, Z, last))
        if last:
            H = Z
        else:
            g, _ = gelu(Z)
            H = bf16(g, rounding).astype(np.float64)
    return H, cache


def mse(Y, T):
    d = np.asarray(Y, np.float64) - np.asarray(T, np.float64)
    return float(np.mean(d * d)), 2.0 * d / d.size


def backward(cache, dY, rounding: bool = True):
    """Gradients [(dW, db)] of the loss whose gradient at the output is dY."""
    grads = [None] * len(cache)
    dZ = np.asarray(dY, np.float64)
    for i in range(len(cache) - 1, -1, -1):
        H, Wb, Z, last = cache[i]
        if not last:
            _, dg = gelu(Z)
            dZ = dZ * dg
        db = dZ.sum(axis=0)
        dZr = bf16(dZ, rounding).astype(np.float64)
        dW = dZr.T @ H
        grads[i] = (dW, db)
        dZ = dZr @ Wb
    return grads


def adam(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return w, m, v
