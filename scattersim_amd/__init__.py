"""scattersim_amd: MI355X-native coupled-wave solver (libcwx) with a thin Python host layer.

The compute path is the HIP library ``lib/libcwx.so`` (C ABI in ``include/cwx.h``); this module only
marshals numpy arrays and device pointers into it.  There is no CPU fallback: every call raises when
the library or a GPU is missing.

Python mirrors of the reference's host interfaces:
  * :func:`scattervolume`    the scattervolume solve (ScatterVolume.cpp:506-1019)
  * :class:`Solution`        solution vectors and the CoupledWaveStructure records
  * :func:`zgeev`, :func:`zgesv`, :func:`zinverse`, :func:`zgemm`   third_Lapack.h:38-85
  * :meth:`Context.field_init_f32` / :meth:`Context.field_evaluate`  gpuEvaluator.h:14-17
  * :func:`cw_read_unpacked` CoupledWaveStructure::load + cw_allocate/cw_unpack
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _capi
from ._capi import check, lib

__all__ = ["Context", "DeviceArray", "Solution", "scattervolume", "zgeev", "zgesv", "zinverse", "zgemm",
           "cw_read_unpacked", "read_cw2", "solve_fields", "solve_batch_mixed", "scatterlayer", "layer_stack",
           "layer_write_cw", "beam_polar", "beam_montecarlo", "WAVE_F32", "WAVE_F64"]

WAVE_F32 = np.dtype([("E0", np.complex64, 3), ("k", np.complex64, 3)])   # cpuEvaluator.h:6-10
WAVE_F64 = np.dtype([("E0", np.complex128, 3), ("k", np.complex128, 3)])


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One HIP device and stream (cwx_create)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().cwx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            lib().cwx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_stream(self, stream_handle: Optional[int]):
        check(lib().cwx_set_stream(self.h, C.c_void_p(stream_handle) if stream_handle else None))

    def synchronize(self):
        check(lib().cwx_synchronize(self.h))

    # ---- field evaluation (gpuEvaluator.h:14-17) ----
    def field_init_f32(self, W: np.ndarray, z_layers, begin, end):
        W = np.ascontiguousarray(W, dtype=WAVE_F32)
        z = np.ascontiguousarray(z_layers, dtype=np.float32)
        b = np.ascontiguousarray(begin, dtype=np.int32)
        e = np.ascontiguousarray(end, dtype=np.int32)
        check(lib().cwx_field_init_f32(self.h, _ptr(W), len(W), _ptr(z), len(z), _ptr(b), _ptr(e)))

    def field_init_f64(self, W: np.ndarray, z_layers, begin, end):
        W = np.ascontiguousarray(W, dtype=WAVE_F64)
        z = np.ascontiguousarray(z_layers, dtype=np.float32)
        b = np.ascontiguousarray(begin, dtype=np.int32)
        e = np.ascontiguousarray(end, dtype=np.int32)
        check(lib().cwx_field_init_f64(self.h, _ptr(W), len(W), _ptr(z), len(z), _ptr(b), _ptr(e)))

    def field_init_solution(self, sol: "Solution"):
        check(lib().cwx_field_init_solution(self.h, sol.h))

    def field_evaluate(self, x_start, y_start, z_start, x, y, z, d, N, tol_mode: int = 0):
        """gpu_cw_evaluate: returns (E_xy, E_xz, E_yz), each (N, N, 3) complex64.  tol_mode 1 takes the
        region test of cpu_cw_evaluate_* (z >= z_l, cpuEvaluator.cpp:294-407) instead of z >= z_l - 0.001."""
        out = [np.zeros((N, N, 3), np.complex64) for _ in range(3)]
        check(lib().cwx_field_evaluate_rule(self.h, _ptr(out[0]), _ptr(out[1]), _ptr(out[2]), x_start, y_start,
                                            z_start, x, y, z, d, N, int(tol_mode)))
        return tuple(out)

    def field_solution_grid_dev(self, sol: "Solution", d_out: int, out_f64: bool, tol_mode: int, x0, y0, z0, dx,
                                dy, dz, nx, ny, nz):
        """Separable evaluation of a solution on a grid into device memory ([nz][ny][nx][3] complex)."""
        check(lib().cwx_field_solution_grid_dev(self.h, sol.h, C.c_void_p(d_out), int(out_f64), tol_mode, x0, y0,
                                                z0, dx, dy, dz, nx, ny, nz))

    def field_grid_dev(self, d_out: int, out_f64: bool, tol_mode: int, x0, y0, z0, dx, dy, dz, nx, ny, nz):
        """Evaluate a grid into device memory at address d_out ([nz][ny][nx][3] complex)."""
        check(lib().cwx_field_grid_dev(self.h, C.c_void_p(d_out), int(out_f64), tol_mode, x0, y0, z0, dx, dy,
                                       dz, nx, ny, nz))


class DeviceArray:
    """A device allocation of the context's device holding a numpy-shaped array (cwx_dev_alloc)."""

    def __init__(self, ctx: Context, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(shape) if not isinstance(shape, int) else (shape,)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().cwx_dev_alloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value or 0

    @classmethod
    def from_numpy(cls, ctx: Context, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.shape, a.dtype)
        if d.nbytes:
            check(lib().cwx_memcpy(ctx.h, C.c_void_p(d.ptr), _ptr(a), d.nbytes, 0))
        return d

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().cwx_memcpy(self.ctx.h, _ptr(out), C.c_void_p(self.ptr), self.nbytes, 1))
        return out

    def free(self):
        if self.ptr:
            lib().cwx_dev_free(self.ctx.h, C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _params(lam=1.0, direction=(0, 0, 1), ex=(0, 0), ey=(1, 0), ez=(0, 0), n=(1, 1), kappa=0.0,
            size=(10, 10, 2), z=0.0, coef: Sequence[int] = (), psf=False, external=False, npz=False,
            no_cutoff=False, eig_mode=0) -> _capi.cwx_volume_params:
    p = _capi.cwx_volume_params()
    lib().cwx_volume_params_default(C.byref(p))
    p.lambda_ = lam
    p.direction[:] = [float(v) for v in direction]
    p.ex[:] = [float(v) for v in ex]
    p.ey[:] = [float(v) for v in ey]
    p.ez[:] = [float(v) for v in ez]
    if len(n) != 2:
        raise ValueError("n takes the two half-space indices")
    p.n[:] = [float(v) for v in n]
    p.kappa = float(kappa[0] if isinstance(kappa, (tuple, list)) else kappa)
    p.size[:] = [float(v) for v in size]
    p.z = float(z)
    coef = list(coef)
    if len(coef) > 2:
        raise ValueError("coef takes 0, 1 or 2 values")
    p.ncoef = len(coef)
    for i, c in enumerate(coef):
        p.coef[i] = int(c)
    p.psf, p.external, p.npz, p.no_cutoff = int(psf), int(external), int(npz), int(no_cutoff)
    p.eig_mode = int(eig_mode)
    return p


class Solution:
    """Device-resident result of one solve (cwx_solution)."""

    def __init__(self, ctx: Context, handle: C.c_void_p):
        self.ctx = ctx
        self.h = handle
        info = _capi.cwx_solution_info()
        check(lib().cwx_solution_get_info(self.h, C.byref(info)))
        self.M0, self.M1, self.MF, self.L = info.M0, info.M1, info.MF, info.L
        self.cutoff, self.status, self.k, self.nrec = bool(info.cutoff), info.info, info.k, info.nrec
        self.nlists = info.nlists

    def free(self):
        if self.h:
            lib().cwx_solution_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def _get(self, what: int, dtype=np.complex128):
        n = lib().cwx_solution_size(self.h, what)
        out = np.zeros(max(n, 0), dtype=dtype)
        if n > 0:
            check(lib().cwx_solution_copy(self.ctx.h, self.h, what, _ptr(out)))
        return out

    def device_ptr(self, what: int) -> int:
        return lib().cwx_solution_device_ptr(self.h, what) or 0

    @property
    def x(self):
        return self._get(_capi.SOL_X)

    @property
    def gammas(self):
        return self._get(_capi.SOL_GAMMA).reshape(self.L, 4 * self.MF)

    @property
    def betas(self):
        return self._get(_capi.SOL_BETA).reshape(-1, 4 * self.MF)

    @property
    def Q_check_hat(self):
        """(L, 3MF, 4MF) arrays [Q_check | Q_hat] (column-major storage reshaped)."""
        a = self._get(_capi.SOL_QCH)
        return a.reshape(-1, 4 * self.MF, 3 * self.MF).transpose(0, 2, 1)

    @property
    def z(self):
        return self._get(_capi.SOL_Z, np.float64)

    def setup_array(self, what: int):
        return self._get(what)

    @property
    def records(self):
        return self._get(_capi.SOL_RECORDS).reshape(-1, 6)

    def lists(self):
        off = np.zeros(self.nlists, np.int64)
        cnt = np.zeros(self.nlists, np.int64)
        check(lib().cwx_solution_lists(self.h, _ptr(off), _ptr(cnt)))
        return off, cnt

    def write_cw(self, path: str):
        check(lib().cwx_solution_write_cw(self.ctx.h, self.h, path.encode()))

    def write_cw2(self, path: str):
        """Compact heterogeneous .cw: external waves plus the mode data of every layer (cwx_solution_write_cw2)."""
        check(lib().cwx_solution_write_cw2(self.ctx.h, self.h, path.encode()))


def read_cw2(ctx: Context, path: str) -> Solution:
    """Rebuild a device solution from a compact .cw file written by ``Solution.write_cw2``."""
    h = C.c_void_p()
    check(lib().cwx_solution_read_cw2(ctx.h, path.encode(), C.byref(h)))
    return Solution(ctx, h)


def scattervolume(ctx: Context, sample: np.ndarray, **opts) -> Solution:
    """Solve a sample of complex refractive indices, shape (L, s1, s2) or (s1, s2), C order.

    Keyword options mirror the scattervolume command line (ScatterVolume.cpp:510-530): lam,
    direction, ex, ey, ez, n, kappa, size, z, coef, psf, external, npz; plus no_cutoff and eig_mode.
    ``sample`` may also be a device pointer (int) with ``shape=(L, s1, s2)`` given.
    """
    shape = opts.pop("shape", None)
    p = _params(**opts)
    h = C.c_void_p()
    if isinstance(sample, int):
        L, s1, s2 = shape
        st = lib().cwx_solve_volume_dev(ctx.h, C.byref(p), C.c_void_p(sample), L, s1, s2, C.byref(h))
    else:
        a = np.asarray(sample, dtype=np.complex128)
        if a.ndim == 2:
            a = a[None]
        if a.ndim != 3:
            raise ValueError("sample must be 2-D or 3-D")
        a = np.ascontiguousarray(a)
        st = lib().cwx_solve_volume(ctx.h, C.byref(p), _ptr(a), a.shape[0], a.shape[1], a.shape[2], C.byref(h))
    if st != _capi.CWX_OK:
        msg = lib().cwx_last_error().decode()
        if h:
            lib().cwx_solution_free(h)
        raise _capi.CwxError(st, msg)
    return Solution(ctx, h)


def layer_stack(n=(1.0, 1.4, 1.4, 1.0), kappa=(0.05,), z=(-5.0, 0.0, 5.0)):
    """The stack of scatterlayer's --n / --kappa / --z lists (ScatterLayer.cpp:327-344): L is the longest
    of len(n), len(kappa) + 1 and len(z) + 1; n is padded with its last value, kappa with zeros and z in
    steps of 10.  Returns (L, n[L], kappa[L-1], z[L-1])."""
    n, kappa, z = [float(v) for v in n], [float(v) for v in kappa], [float(v) for v in z]
    L = max(len(n), len(kappa) + 1, len(z) + 1)
    n += [n[-1]] * (L - len(n))
    kappa = (kappa + [0.0] * L)[:L - 1]
    while len(z) + 1 < L:
        z.append(z[-1] + 10.0)
    return L, np.array(n), np.array(kappa), np.array(z[:L - 1])


def _waves(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.complex128)
    if a.ndim != 2 or a.shape[1] != 6:
        raise ValueError("plane waves are (P, 6) records: E0.x, E0.y, E0.z, k.x, k.y, k.z")
    return a


def scatterlayer(ctx: Context, incident, n=(1.0, 1.4, 1.4, 1.0), kappa=(0.05,), z=(-5.0, 0.0, 5.0), lam=1.0):
    """Solve a stack of homogeneous layers for every incident wave (cwx_layer_solve).

    ``incident`` is (P, 6) complex records or a device pointer (int) with ``P`` given as ``(ptr, P)``.
    Returns (reflected, transmitted), each (P, L-1, 6): the reflected wave of layer l and the transmitted
    wave of layer l+1 at boundary l, wound to z[l] (ScatterLayer.cpp:390-420)."""
    L, nn, kk, zz = layer_stack(n, kappa, z)
    a = _waves(incident)
    P = a.shape[0]
    R = np.zeros((P, L - 1, 6), np.complex128)
    T = np.zeros((P, L - 1, 6), np.complex128)
    check(lib().cwx_layer_solve(ctx.h, L, _ptr(nn), _ptr(kk), _ptr(zz), float(lam), P, _ptr(a), _ptr(R), _ptr(T)))
    return R, T


def layer_write_cw(path: str, z, incident, reflected, transmitted):
    """The .cw file scatterlayer writes (cwx_layer_write_cw)."""
    z = np.ascontiguousarray(z, np.float64)
    a = _waves(incident)
    R = np.ascontiguousarray(reflected, np.complex128)
    T = np.ascontiguousarray(transmitted, np.complex128)
    check(lib().cwx_layer_write_cw(path.encode(), len(z) + 1, _ptr(z), a.shape[0], _ptr(a), _ptr(R), _ptr(T)))


def beam_polar(alpha: float, k, E, N0: int, N1: int, beta: float = 0.0) -> np.ndarray:
    """Focused-beam sampling of tira's SolidAnglePolar (cwx_beam_polar): (N1*N0, 6) records."""
    kv = np.ascontiguousarray(k, np.float64)
    Ev = np.ascontiguousarray(E, np.complex128)
    out = np.zeros((N0 * N1, 6), np.complex128)
    check(lib().cwx_beam_polar(float(alpha), float(beta), _ptr(kv), _ptr(Ev), int(N0), int(N1), _ptr(out)))
    return out


def beam_montecarlo(alpha: float, k, E, N: int, beta: float = 0.0, seed: int = 0) -> np.ndarray:
    """Monte-Carlo focused-beam sampling (cwx_beam_montecarlo): (N, 6) records."""
    kv = np.ascontiguousarray(k, np.float64)
    Ev = np.ascontiguousarray(E, np.complex128)
    out = np.zeros((N, 6), np.complex128)
    check(lib().cwx_beam_montecarlo(float(alpha), float(beta), _ptr(kv), _ptr(Ev), int(N), int(seed), _ptr(out)))
    return out


def solve_batch_mixed(ctx: Context, samples, params, max_bytes: int = 0):
    """Solve independent samples whose Fourier orders may differ (cwx_solve_batch_mixed_dev).

    ``samples`` is a list of (L_i, s1, s2) complex arrays on one grid, ``params`` one option dict per
    instance.  Returns (xs, info, orders): xs[i] is instance i's [R; T] (6 M0_i M1_i complex128)."""
    count = len(params)
    pa = (_capi.cwx_volume_params * count)()
    for i, o in enumerate(params):
        pa[i] = _params(**o)
    arrs = [np.asarray(s_, dtype=np.complex128) for s_ in samples]
    arrs = [a[None] if a.ndim == 2 else a for a in arrs]
    s1, s2 = arrs[0].shape[1:]
    if any(a.shape[1:] != (s1, s2) for a in arrs) or len(arrs) != count:
        raise ValueError("one sample per instance, all on one grid")
    Ls = np.array([a.shape[0] for a in arrs], np.int32)
    off = np.zeros(count + 1, np.int64)
    M0 = np.zeros(count, np.int32)
    M1 = np.zeros(count, np.int32)
    check(lib().cwx_solve_batch_mixed_dev(None, count, pa, _ptr(Ls), None, int(s1), int(s2), int(max_bytes), None,
                                          _ptr(off), None, _ptr(M0), _ptr(M1)))
    dS = DeviceArray.from_numpy(ctx, np.ascontiguousarray(np.concatenate(arrs, axis=0)))
    dX = DeviceArray(ctx, (int(off[-1]),), np.complex128)
    info = np.zeros(count, np.int32)
    st = lib().cwx_solve_batch_mixed_dev(ctx.h, count, pa, _ptr(Ls), C.c_void_p(dS.ptr), int(s1), int(s2),
                                         int(max_bytes), C.c_void_p(dX.ptr), _ptr(off), _ptr(info), None, None)
    if st < 0:
        raise _capi.CwxError(st, lib().cwx_last_error().decode())
    x = dX.numpy()
    dS.free()
    dX.free()
    return [x[off[i]:off[i + 1]] for i in range(count)], info, list(zip(M0.tolist(), M1.tolist()))


def solve_fields(ctx: Context, sample, EF, **opts):
    """Solve one sample for K incident fields at once (cwx_solve_fields_dev).

    ``EF`` is (K, 3*MF) complex: per field the Ex, Ey, Ez Fourier coefficients of the incident orders, in
    the order of ``Solution.setup_array(SOL_EF)``.  ``sample`` is an (L, s1, s2) array (copied to the
    device) or a device pointer with ``shape=(L, s1, s2)``.  Returns (K, 6*MF) [R; T] per field.
    ``EF=None`` returns the orders (M0, M1) instead."""
    shape = opts.pop("shape", None)
    p = _params(**opts)
    if isinstance(sample, int):
        L, s1, s2 = shape
        ptr, keep = sample, None
    else:
        a = np.asarray(sample, dtype=np.complex128)
        if a.ndim == 2:
            a = a[None]
        L, s1, s2 = a.shape
        keep = DeviceArray.from_numpy(ctx, np.ascontiguousarray(a))
        ptr = keep.ptr
    M0, M1 = C.c_int(), C.c_int()
    check(lib().cwx_solve_fields_dev(None, C.byref(p), None, L, s1, s2, 0, None, None, C.byref(M0), C.byref(M1)))
    if EF is None:
        return M0.value, M1.value
    E = np.ascontiguousarray(EF, np.complex128)
    if E.ndim == 1:
        E = E[None]
    if E.ndim != 2 or E.shape[1] != 3 * M0.value * M1.value:
        raise ValueError(f"EF must be (K, 3*MF) with 3*MF = {3 * M0.value * M1.value}")
    K = E.shape[0]
    dE = DeviceArray.from_numpy(ctx, E)
    dx = DeviceArray(ctx, (K, 2 * E.shape[1]), np.complex128)
    check(lib().cwx_solve_fields_dev(ctx.h, C.byref(p), C.c_void_p(ptr), L, s1, s2, K, C.c_void_p(dE.ptr),
                                     C.c_void_p(dx.ptr), None, None))
    out = dx.numpy()
    del keep
    return out


def batch_orders(params, Ls, s1: int, s2: int):
    """Fourier orders (M0, M1) a batch of these instances resolves to (no device work)."""
    count = len(params)
    pa = (_capi.cwx_volume_params * count)()
    for i, o in enumerate(params):
        pa[i] = _params(**o)
    Ls = np.ascontiguousarray(Ls, np.int32)
    M0, M1 = C.c_int(), C.c_int()
    check(lib().cwx_solve_batch_dev(None, count, pa, _ptr(Ls), None, int(s1), int(s2), 0, None, None, C.byref(M0),
                                    C.byref(M1)))
    return M0.value, M1.value


def solve_batch(ctx: Context, samples, params, max_bytes: int = 0, shape=None, d_x: Optional[int] = None):
    """Solve independent samples together (cwx_solve_batch_dev).

    ``params`` is one option dict per instance (the keywords of ``scattervolume``); every instance must
    resolve to the same Fourier orders and eigen mode.  ``samples`` is a list of (L_i, s1, s2) complex
    arrays, or a device pointer (int) to the layers of all instances back to back with
    ``shape=(Ls, s1, s2)``.  Returns (x, info, (M0, M1)): x is (count, 6MF) complex128 [R; T] per
    instance.  With ``d_x`` (a device pointer to count x 6MF complex128) the result stays on the device
    and x is None.
    """
    count = len(params)
    pa = (_capi.cwx_volume_params * count)()
    for i, o in enumerate(params):
        pa[i] = _params(**o)
    own = None
    if isinstance(samples, int):
        Ls, s1, s2 = shape
        Ls = np.ascontiguousarray(Ls, np.int32)
        ptr = samples
    else:
        arrs = [np.asarray(s, dtype=np.complex128) for s in samples]
        arrs = [a[None] if a.ndim == 2 else a for a in arrs]
        s1, s2 = arrs[0].shape[1:]
        if any(a.shape[1:] != (s1, s2) for a in arrs):
            raise ValueError("all samples must share the grid")
        Ls = np.array([a.shape[0] for a in arrs], np.int32)
        own = DeviceArray.from_numpy(ctx, np.ascontiguousarray(np.concatenate(arrs, axis=0)))
        ptr = own.ptr
    if len(Ls) != count:
        raise ValueError("one layer count per instance is needed")
    M0, M1 = C.c_int(), C.c_int()
    check(lib().cwx_solve_batch_dev(ctx.h, count, pa, _ptr(Ls), None, int(s1), int(s2), int(max_bytes), None, None,
                                    C.byref(M0), C.byref(M1)))
    MF = M0.value * M1.value
    out = None
    if d_x is None:
        out = DeviceArray(ctx, (count, 6 * MF), np.complex128)
        d_x = out.ptr
    info = np.zeros(count, np.int32)
    st = lib().cwx_solve_batch_dev(ctx.h, count, pa, _ptr(Ls), C.c_void_p(ptr), int(s1), int(s2), int(max_bytes),
                                   C.c_void_p(d_x), _ptr(info), None, None)
    if st < 0:
        raise _capi.CwxError(st, lib().cwx_last_error().decode())
    x = out.numpy() if out is not None else None
    if own is not None:
        own.free()
    if out is not None:
        out.free()
    return x, info, (M0.value, M1.value)


def solve_batch_graph(ctx: Context, d_samples: int, params, Ls, s1: int, s2: int, chunk: int, d_x: int):
    """cwx_solve_batch_graph_dev: the batched solve with each chunk of `chunk` instances replayed from a
    captured HIP graph.  Device pointers in and out; returns (info, graphs captured by this call)."""
    count = len(params)
    pa = (_capi.cwx_volume_params * count)()
    for i, o in enumerate(params):
        pa[i] = _params(**o)
    Ls = np.ascontiguousarray(Ls, np.int32)
    info = np.zeros(count, np.int32)
    cap = C.c_int()
    st = lib().cwx_solve_batch_graph_dev(ctx.h, count, pa, _ptr(Ls), C.c_void_p(d_samples), int(s1), int(s2),
                                         int(chunk), C.c_void_p(d_x), _ptr(info), C.byref(cap))
    if st < 0:
        raise _capi.CwxError(st, lib().cwx_last_error().decode())
    return info, cap.value


def zgeev(ctx: Context, A: np.ndarray):
    """MKL_eigensolve contract: returns (w, VR); A is not modified."""
    n = A.shape[0]
    a = np.asfortranarray(A, dtype=np.complex128).copy(order="F")
    w = np.zeros(n, np.complex128)
    V = np.zeros((n, n), np.complex128, order="F")
    check(lib().cwx_zgeev(ctx.h, n, _ptr(a), n, _ptr(w), _ptr(V), n))
    return w, V


def zgesv(ctx: Context, A: np.ndarray, B: np.ndarray):
    """MKL_linearsolve contract: returns X with A X = B."""
    n = A.shape[0]
    a = np.asfortranarray(A, dtype=np.complex128).copy(order="F")
    b2 = B.reshape(n, -1)
    b = np.asfortranarray(b2, dtype=np.complex128).copy(order="F")
    check(lib().cwx_zgesv(ctx.h, n, b.shape[1], _ptr(a), n, None, _ptr(b), n))
    return b.reshape(B.shape)


def zinverse(ctx: Context, A: np.ndarray):
    n = A.shape[0]
    a = np.asfortranarray(A, dtype=np.complex128)
    out = np.zeros((n, n), np.complex128, order="F")
    check(lib().cwx_zinverse(ctx.h, n, _ptr(a), n, _ptr(out), n))
    return out


def zgemm(ctx: Context, A: np.ndarray, B: np.ndarray, alpha=1.0):
    m, k = A.shape
    n = B.shape[1]
    a = np.asfortranarray(A, dtype=np.complex128)
    b = np.asfortranarray(B, dtype=np.complex128)
    c = np.zeros((m, n), np.complex128, order="F")
    al = _capi.cwx_complex(complex(alpha).real, complex(alpha).imag)
    check(lib().cwx_zgemm(ctx.h, m, n, k, al, _ptr(a), m, _ptr(b), k, _capi.cwx_complex(0, 0), _ptr(c), m))
    return c


def cw_read_unpacked(path: str):
    """Read a .cw file: returns (W [WAVE_F64], z_layers float64, begin, end)."""
    nw = C.c_int64()
    nl = C.c_int()
    check(lib().cwx_cw_read_unpacked(path.encode(), C.byref(nw), C.byref(nl), None, None, None, None))
    W = np.zeros(nw.value, WAVE_F64)
    z = np.zeros(nl.value, np.float64)
    b = np.zeros(nl.value + 1, np.int32)
    e = np.zeros(nl.value + 1, np.int32)
    check(lib().cwx_cw_read_unpacked(path.encode(), C.byref(nw), C.byref(nl), _ptr(W), _ptr(z), _ptr(b), _ptr(e)))
    return W, z, b, e
