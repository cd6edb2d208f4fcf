"""The no-allocation zgesv entry point (cwx_zgesv_batched_dev_noalloc) against the allocating entry points
it shares its kernels with (cwx_zgetrf_batched_dev + cwx_zgetrs_batched_dev, and cwx_zgesv per matrix):
bit for bit on two calls in a row, the second smaller, and refused beyond its scratch limit."""
import ctypes as C

import numpy as np
import pytest

import scattersim_amd as sa
from scattersim_amd import _capi

pytestmark = pytest.mark.gpu


def random_system(rng, n, nrhs, batch):
    A = rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))
    A += 4.0 * np.sqrt(n) * np.eye(n)[None]  # comfortably nonsingular
    B = rng.standard_normal((batch, nrhs, n)) + 1j * rng.standard_normal((batch, nrhs, n))
    # column-major storage per matrix: element (r, c) at r + n c, i.e. the C-order (c, r) array
    return np.ascontiguousarray(A.transpose(0, 2, 1)), np.ascontiguousarray(B)


def noalloc(ctx, At, Bt):
    batch, n, _ = At.shape
    nrhs = Bt.shape[1]
    dA = sa.DeviceArray.from_numpy(ctx, At)
    dB = sa.DeviceArray.from_numpy(ctx, Bt)
    info = np.zeros(batch, np.int32)
    st = _capi.lib().cwx_zgesv_batched_dev_noalloc(ctx.h, n, nrhs, batch, C.c_void_p(dA.ptr), n * n,
                                                   C.c_void_p(dB.ptr), n * nrhs, info.ctypes.data_as(C.c_void_p))
    _capi.check(st)
    return dA.numpy(), dB.numpy(), info


def allocating(ctx, At, Bt):
    batch, n, _ = At.shape
    nrhs = Bt.shape[1]
    L = _capi.lib()
    dA = sa.DeviceArray.from_numpy(ctx, At)
    dB = sa.DeviceArray.from_numpy(ctx, Bt)
    dP = sa.DeviceArray(ctx, (batch, n), np.int32)
    info = np.zeros(batch, np.int32)
    _capi.check(L.cwx_zgetrf_batched_dev(ctx.h, n, batch, C.c_void_p(dA.ptr), n * n, C.c_void_p(dP.ptr), n,
                                         info.ctypes.data_as(C.c_void_p)))
    _capi.check(L.cwx_zgetrs_batched_dev(ctx.h, n, nrhs, batch, C.c_void_p(dA.ptr), n * n, C.c_void_p(dP.ptr), n,
                                         C.c_void_p(dB.ptr), n * nrhs))
    ctx.synchronize()
    return dA.numpy(), dB.numpy()


@pytest.mark.parametrize("shape", [((300, 3, 5), (100, 1, 3)), ((64, 2, 40), (33, 4, 2))])
def test_noalloc_matches_allocating_entry_points(ctx, shape):
    rng = np.random.default_rng(17)
    for n, nrhs, batch in shape:                       # two calls in a row, the second smaller
        At, Bt = random_system(rng, n, nrhs, batch)
        LUa, Xa, info = noalloc(ctx, At, Bt)
        LUb, Xb = allocating(ctx, At, Bt)
        assert not np.any(info)
        assert np.array_equal(LUa, LUb) and np.array_equal(Xa, Xb)
        for b in range(batch):
            A = At[b].T                                  # the matrix itself
            X = Xa[b].T                                  # n x nrhs
            assert np.max(np.abs(A @ X - Bt[b].T)) / np.max(np.abs(Bt[b])) < 1e-12
            Xh = sa.zgesv(ctx, A, Bt[b].T.copy())        # the host entry point (cwx_zgesv), one matrix
            assert np.max(np.abs(Xh - X)) / np.max(np.abs(Xh)) < 1e-13


def test_noalloc_refuses_beyond_scratch(ctx):
    n = 1024
    batch = _capi_max_pivots() // n + 1
    dA = sa.DeviceArray(ctx, (1,), np.complex128)
    st = _capi.lib().cwx_zgesv_batched_dev_noalloc(ctx.h, n, 1, batch, C.c_void_p(dA.ptr), n * n,
                                                   C.c_void_p(dA.ptr), n, None)
    assert st == _capi.CWX_ERR_ARG
    assert "scratch" in _capi.lib().cwx_last_error().decode()


def _capi_max_pivots():
    return ((4 << 20) - 65536 * 4) // 4
