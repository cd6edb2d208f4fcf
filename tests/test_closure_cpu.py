"""CPU pins of the learned-closure oracle (oracle/closure_oracle.py) and of the closure's host-side helpers."""
import numpy as np
import pytest

from oracle import closure_oracle as co
from scattersim_amd import closure as cl


def test_bf16_rounding_bit_patterns():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -2.5, 3.0e38, np.inf, np.nan,
                  1e-40], np.float32)
    r = co.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                          # tie -> even (mantissa ...0)
    assert r[2] == 1.0 + 4 * 2 ** -8            # tie -> even (rounds up to ...0)
    assert r[3] == 1.0 + 2 * 2 ** -8            # above the tie -> up
    assert r[4] == -2.5
    assert np.isinf(co.bf16(np.float32(3.4e38)))
    assert np.isinf(r[6]) and np.isnan(r[7])
    bits = cl._bf16_bits(x)
    assert np.array_equal((bits.astype(np.uint32) << 16).view(np.float32)[~np.isnan(x)], r[~np.isnan(x)])


def test_gelu_derivative_matches_finite_difference():
    z = np.linspace(-4, 4, 41)
    _, d = co.gelu(z)
    h = 1e-6
    fd = (co.gelu(z + h)[0] - co.gelu(z - h)[0]) / (2 * h)
    assert np.max(np.abs(d - fd)) < 1e-8
    assert abs(co.gelu(0.0)[0]) == 0 and abs(co.gelu(10.0)[0] - 10.0) < 1e-12


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    dims = [7, 5, 6, 3]
    params = [(rng.standard_normal((dims[i + 1], dims[i])) * 0.5, rng.standard_normal(dims[i + 1]) * 0.1)
              for i in range(3)]
    X = rng.standard_normal((4, 7))
    T = rng.standard_normal((4, 3))
    Y, cache = co.forward(params, X, rounding=False)
    loss, dY = co.mse(Y, T)
    grads = co.backward(cache, dY, rounding=False)
    h = 1e-6
    for li in range(3):
        for which in (0, 1):
            arr = params[li][which]
            for idx in [(0,) * arr.ndim, tuple(s - 1 for s in arr.shape)]:
                old = arr[idx]
                arr[idx] = old + h
                lp = co.mse(co.forward(params, X, rounding=False)[0], T)[0]
                arr[idx] = old - h
                lm = co.mse(co.forward(params, X, rounding=False)[0], T)[0]
                arr[idx] = old
                assert abs((lp - lm) / (2 * h) - grads[li][which][idx]) < 1e-7


def test_adam_first_step_moves_by_lr():
    w, m, v = co.adam(np.array([1.0, -2.0]), np.array([0.3, -5.0]), np.zeros(2), np.zeros(2), 1, 0.01)
    assert np.allclose(w, [1.0 - 0.01, -2.0 + 0.01], atol=1e-9)


def test_order_embedding_keeps_the_zero_order_central():
    # orders of a 3x3 list inside a 7x7 list: the centre goes to the centre
    idx = cl.order_map(3, 7)
    assert idx[4] == 3 + 3 * 7
    x = np.arange(6 * 9) + 1.0
    y = cl.place(x, 3, 7)
    assert np.array_equal(cl.take(y, 7, 3), x)
    assert np.count_nonzero(y) == 54
