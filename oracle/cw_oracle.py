"""CPU FP64 oracle for the scattervolume coupled-wave solve and the plane-wave field evaluation.

TEST INFRASTRUCTURE ONLY.  This module is the checker the HIP path is compared against.  Only the
test-suite (``tests/``), ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product path (``scattersim_amd``) never does, and fails loudly when its HIP library is
missing instead of falling back here.

It restates, in numpy, the algorithm of the reference files below (all paths relative to the
reference checkout):

  * ``FourierWave.h:18``           PI = 3.141592653 (truncated pi, used for k and up/wq)
  * ``FourierWave.h:32-39``        orthogonalize (normalise without conjugation, project E _|_ d)
  * ``FourierWave.h:82-145``       fftShift2d (including its parity quirk for even-row/odd-column input)
  * ``FourierWave.h:152-195``      fftw_fft2 (unnormalised forward DFT / N1N2, shift, centred crop)
  * ``FourierWave.h:283-301``      volume::reformat (each slice read as a column-major map)
  * ``FourierWave.h:319-334``      UpWq_Cal (order series, up/wq, Sx/Sy meshes, principal-branch Sz)
  * ``FourierWave.h:336-397``      phi (circulant-wrapped Fourier convolution blocks of Phi)
  * ``ScatterVolume.cpp:154-168``  InitLayer_n / InitLayer_z
  * ``ScatterVolume.cpp:171-230``  sorter / Eigen_Sort (absolute-eps comparator, interleave)
  * ``ScatterVolume.cpp:235-337``  EigenDecompositionD (zgeev, Gd/Gc phases, Gc chain)
  * ``ScatterVolume.cpp:339-395``  MatTransfer (f1, f2, f3)
  * ``ScatterVolume.cpp:398-466``  SetGaussianConstraints / SetBoundaryConditions
  * ``ScatterVolume.cpp:469-493``  mat2waves (extra n on the external kz)
  * ``ScatterVolume.cpp:574-651``  option post-processing and Fourier-order selection / cut-off
  * ``ScatterVolume.cpp:660-685``  incident-field Fourier coefficients EF and the flat Sx/Sy/Sz
  * ``ScatterVolume.cpp:739-889``  zgesv solve, beta recurrence, Q_check/Q_hat and the Pz (Ez) rows
  * ``ScatterVolume.cpp:907-987``  CoupledWaveStructure assembly (wind to each boundary)
  * ``third_Lapack.h:38-85``       zgeev('N','V') / zgesv / inverse-by-zgesv / zgemm
  * ``CoupledWaveStructure.h:63-125``  .cw container
  * ``cpuEvaluator.cpp:34-102``    cw_allocate / cw_unpack (region map, double -> float)
  * ``cpuEvaluator.cpp:294-407``   cpu_cw_evaluate_xy/yz/xz
  * ``gpuEvaluator.cu:28-68``      evaluate (region test z >= z_l - 0.001, FP32 sum)
  * ``ScatterView.cpp:354-376``    EvaluateVectorSlices (slice geometry)

Dense LAPACK goes through numpy (zgeev / zgesv from the bundled LAPACK); ``set_lapack('torch')``
routes zgeev and zgesv through torch's CPU build (Intel MKL, the library family the reference links).

Pins: every function here is checked in ``tests/test_oracle_pins.py`` against the reference-produced
``x64/a.cw`` known answer, closed forms (uniform layers, Fresnel), a brute-force DFT, an independent
Maxwell formulation for one- and two-dimensional gratings, eigen residuals and energy balance.
Parts with no such check are listed in ``UNPINNED``.
"""
from __future__ import annotations

import functools
import math
import struct
from dataclasses import dataclass, field

import numpy as np

PI_TRUNC = 3.141592653          # FourierWave.h:18
DBL_EPS = np.finfo(np.float64).eps

# Parts of the path that no golden vector, closed form or invariant pins (see tests/test_oracle_pins.py).
UNPINNED = (
    "fftshift2d on even-row/odd-column grids: checked only against a 2x3 case worked by hand from "
    "FourierWave.h:92-135",
    "the mixed x/y ordering of Sx/Sy/Sz/EF against Phi's rows (ScatterVolume.cpp:677-685) for M0>1 and M1>1: "
    "an independent Maxwell formulation agrees once these tables are flattened x-fastest, so everything "
    "but the ordering itself is pinned",
    "the transposed Pz index (ScatterVolume.cpp:872,875) for M0>1 and M1>1: it changes only interior Ez",
    "eigen-sort order of eigenvalues whose imaginary parts differ by less than DBL_EPS",
)


_LAPACK = "numpy"


def set_lapack(backend: str) -> None:
    """Select the LAPACK used for zgeev/zgesv: 'numpy' (bundled LAPACK) or 'torch' (MKL on CPU)."""
    global _LAPACK
    if backend not in ("numpy", "torch"):
        raise ValueError(backend)
    _LAPACK = backend


def zgeev(A: np.ndarray):
    """Right eigenpairs, LAPACK zgeev('N','V') contract (third_Lapack.h:38-45)."""
    if _LAPACK == "torch":
        import torch
        w, v = torch.linalg.eig(torch.from_numpy(np.ascontiguousarray(A)))
        return w.numpy(), v.numpy()
    return np.linalg.eig(A)


def zinv(A: np.ndarray) -> np.ndarray:
    """Inverse through zgesv with an identity right-hand side (third_Lapack.h:56-64)."""
    if _LAPACK == "torch":
        import torch
        return torch.linalg.inv(torch.from_numpy(np.ascontiguousarray(A))).numpy()
    return np.linalg.inv(A)


def zgesv(A: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Solve A x = b (third_Lapack.h:47-54)."""
    if _LAPACK == "torch":
        import torch
        return torch.linalg.solve(torch.from_numpy(np.ascontiguousarray(A)),
                                  torch.from_numpy(np.ascontiguousarray(b))).numpy()
    return np.linalg.solve(A, b)


# ----------------------------------------------------------------------------------------------------
# small helpers
# ----------------------------------------------------------------------------------------------------

def orthogonalize(E, d):
    """FourierWave.h:32-39: E /= sqrt(Ex^2+Ey^2+Ez^2) (no conjugation), then E = d x (E x d)."""
    E = np.asarray(E, dtype=np.complex128)
    d = np.asarray(d, dtype=np.float64).astype(np.complex128)
    mag = np.sqrt(E[0] ** 2 + E[1] ** 2 + E[2] ** 2)
    E = E / mag
    s = np.cross(E, d)
    return np.cross(d, s)


def fftshift2d(mat: np.ndarray) -> np.ndarray:
    """FourierWave.h:82-145, index for index, including the row-parity key used for both axes."""
    m, n = mat.shape
    if m % 2 == 1:
        p, q = m // 2, n // 2
    else:
        p, q = -(-m // 2), -(-n // 2)
    indx = np.empty(m, dtype=np.int64)
    indy = np.empty(n, dtype=np.int64)
    if m % 2 == 1:
        for i in range(0, m - p - 1):
            indx[i] = (m - p) + i
        for i in range(m - p - 1, m):
            indx[i] = i - (m - p - 1)
    else:
        for i in range(0, m - p):
            indx[i] = p + i
        for i in range(m - p, m):
            indx[i] = i - (m - p)
    if n % 2 == 1:
        for i in range(0, n - q - 1):
            indy[i] = (n - q) + i
        for i in range(n - q - 1, n):
            indy[i] = i - (n - q - 1)
    else:
        for i in range(0, n - q):
            indy[i] = q + i
        for i in range(n - q, n):
            indy[i] = i - (n - q)
    return mat[np.ix_(indx, indy)]


def fft2_crop(A: np.ndarray, M1: int, M2: int) -> np.ndarray:
    """FourierWave.h:152-195: forward DFT / (N1 N2), fftShift2d, block(N1/2-M1/2, N2/2-M2/2, M1, M2)."""
    N1, N2 = A.shape
    B = np.fft.fft2(A) / N1 / N2
    S = fftshift2d(B)
    r0 = N1 // 2 - M1 // 2
    c0 = N2 // 2 - M2 // 2
    return S[r0:r0 + M1, c0:c0 + M2]


def dft_index_map(N1: int, N2: int, M1: int, M2: int):
    """Unshifted DFT (row, col) frequency indices that fft2_crop returns, as two index vectors.

    Output entry (r, c) of fft2_crop equals DFT[rows[r], cols[c]] / (N1 N2).  Used by the pruned-DFT
    Phi builder: only these M1 x M2 coefficients are ever needed.
    """
    probe_r = np.arange(N1).reshape(N1, 1) * np.ones((1, N2), dtype=np.int64)
    probe_c = np.ones((N1, 1), dtype=np.int64) * np.arange(N2).reshape(1, N2)
    r0 = N1 // 2 - M1 // 2
    c0 = N2 // 2 - M2 // 2
    rows = fftshift2d(probe_r)[r0:r0 + M1, c0]
    cols = fftshift2d(probe_c)[r0, c0:c0 + M2]
    return rows.astype(np.int64), cols.astype(np.int64)


def reformat(data: np.ndarray) -> list:
    """FourierWave.h:283-301: each slice of the C-order buffer is read as a column-major (s1 x s2) map.

    A 2-D input becomes a single layer.  For square slices this is data[i].T.
    """
    data = np.asarray(data, dtype=np.complex128)
    if data.ndim == 2:
        data = data.reshape((1,) + data.shape)
    if data.ndim != 3:
        raise ValueError("sample must be 2-D or 3-D")
    L, s1, s2 = data.shape
    flat = np.ascontiguousarray(data).reshape(L, s1 * s2)
    return [flat[i].reshape(s2, s1).T.copy() for i in range(L)]


def order_series(M: int) -> np.ndarray:
    """FourierWave.h:320-321 / A19: linspace(-floor(M/2), floor((M-1)/2), M)."""
    return np.arange(M, dtype=np.float64) - (M // 2)


# ----------------------------------------------------------------------------------------------------
# solver
# ----------------------------------------------------------------------------------------------------

@dataclass
class VolumeParams:
    """Command-line options of scattervolume (ScatterVolume.cpp:510-530) with their defaults."""
    lam: float = 1.0
    direction: tuple = (0.0, 0.0, 1.0)
    ex: tuple = (0.0, 0.0)
    ey: tuple = (1.0, 0.0)
    ez: tuple = (0.0, 0.0)
    n: tuple = (1.0, 1.0)
    kappa: tuple = (0.0,)
    size: tuple = (10.0, 10.0, 2.0)
    z: float = 0.0
    coef: tuple = ()
    psf: bool = False
    external: bool = False
    npz: bool = False
    # bypass the silent Fourier-order cut-off (ScatterVolume.cpp:613-642); the CLI keeps it
    no_cutoff: bool = False


def select_orders(p: VolumeParams, num_pixels, dir_n):
    """ScatterVolume.cpp:599-651 (A14).  ``dir_n`` is the normalised direction already times n0."""
    coef = list(p.coef)
    if len(coef) == 0:
        M0, M1 = int(num_pixels[2]), int(num_pixels[1])
    elif len(coef) == 1:
        M0 = M1 = int(math.sqrt(coef[0]))
    elif len(coef) == 2:
        M0, M1 = int(coef[0]), int(coef[1])
    else:
        raise ValueError("--coef takes 0, 1 or 2 values")
    msgs = []
    in_n, size, lam = p.n, p.size, p.lam
    if not p.no_cutoff:
        chk = ((M0 // 2) / size[0] * lam + dir_n[0]) ** 2 + ((M1 // 2) / size[1] * lam + dir_n[1]) ** 2
        if chk >= in_n[0] ** 2:
            msgs.append("cut-off")
            if M0 != 1 and M1 != 1:
                if ((M0 // 2) / size[0] * lam + dir_n[0]) > 0.71 * in_n[0]:
                    M0 = int(0.707 * in_n[0] * 2 * size[0] / lam) - 1
                if ((M1 // 2) / size[1] * lam + dir_n[1]) > 0.71 * in_n[0]:
                    M1 = int(0.707 * in_n[1] * 2 * size[1] / lam) - 1
            if M0 == 1:
                if ((M1 // 2) / size[1] * lam + abs(dir_n[1])) >= in_n[0]:
                    M1 = int(in_n[1] * 2 * size[1] / lam - 1e-10)
                    if M1 % 2 == 0:
                        M1 -= 1
            if M1 == 1:
                if ((M0 // 2) / size[1] * lam + abs(dir_n[0])) >= in_n[0]:
                    M0 = int(in_n[0] * 2 * size[0] / lam - 1e-10)
                    if M0 % 2 == 0:
                        M0 -= 1
    if M0 > num_pixels[2]:
        M0 = int(num_pixels[2])
    if M1 > num_pixels[1]:
        M1 = int(num_pixels[1])
    return M0, M1, msgs


@dataclass
class Setup:
    """Everything the solve needs before the per-layer eigenproblems (host-side precomputation)."""
    M0: int
    M1: int
    MF: int
    L: int
    k: float
    dir_n: np.ndarray
    e: np.ndarray
    z: np.ndarray            # ZL boundary positions
    ni: np.ndarray           # complex indices of the two half-spaces
    up: np.ndarray           # M0
    wq: np.ndarray           # M1
    Sx: np.ndarray           # MF, flat column-major (y fastest)
    Sy: np.ndarray
    Sz0: np.ndarray
    Sz1: np.ndarray
    EF: np.ndarray           # 3MF
    layers: list             # L matrices (s1 x s2) of refractive index
    params: VolumeParams = field(default=None)


def setup(sample: np.ndarray, p: VolumeParams) -> Setup:
    """ScatterVolume.cpp:574-685 up to (but excluding) CalculateD."""
    if len(p.n) != 2:
        raise ValueError("--n takes the two half-space indices")
    d = np.asarray(p.direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    e = orthogonalize([complex(*p.ex), complex(*p.ey), complex(*p.ez)], d)
    dir_n = d * p.n[0]
    k = 2 * PI_TRUNC / p.lam
    ni = np.array([complex(p.n[0], 0.0), complex(p.n[1], p.kappa[0])])
    layers = reformat(sample)
    L = len(layers)
    s1, s2 = layers[0].shape
    num_pixels = (L, s1, s2)
    ZL = L + 1
    z = np.empty(ZL)
    z[0] = p.z - p.size[2] / 2.0
    for i in range(1, ZL):
        z[i] = z[i - 1] + p.size[2] / L
    M0, M1, _ = select_orders(p, num_pixels, dir_n)
    MF = M0 * M1
    ps, qs = order_series(M0), order_series(M1)
    up = 2 * PI_TRUNC * ps / p.size[0] + dir_n[0] * k
    wq = 2 * PI_TRUNC * qs / p.size[1] + dir_n[1] * k
    sx, sy = up / k, wq / k
    meshS0 = np.broadcast_to(sx.reshape(1, M0), (M1, M0)).astype(np.complex128)
    meshS1 = np.broadcast_to(sy.reshape(M1, 1), (M1, M0)).astype(np.complex128)
    # n^2 - Sx^2 - Sy^2 in real arithmetic with a +0 imaginary part (A22: principal branch, decaying)
    sz0 = np.sqrt(ni[0] ** 2 - meshS0 ** 2 - meshS1 ** 2)
    sz1 = np.sqrt(ni[1] ** 2 - meshS0 ** 2 - meshS1 ** 2)
    if p.psf:
        Ef = [e[c] * np.ones((M0, M1), dtype=np.complex128) for c in range(3)]
    else:
        Ef = [fft2_crop(e[c] * np.ones((s1, s2), dtype=np.complex128), M1, M0) for c in range(3)]
    EF = np.concatenate([f.ravel(order="F") for f in Ef])
    return Setup(M0=M0, M1=M1, MF=MF, L=L, k=k, dir_n=dir_n, e=e, z=z, ni=ni, up=up, wq=wq,
                 Sx=meshS0.ravel(order="F").copy(), Sy=meshS1.ravel(order="F").copy(),
                 Sz0=sz0.ravel(order="F").copy(), Sz1=sz1.ravel(order="F").copy(), EF=EF,
                 layers=layers, params=p)


def layer_fourier(layer: np.ndarray, M0: int, M1: int):
    """FourierWave.h:341,343: Nf = fft2_crop(n^2), Nif = fft2_crop(n^-2), each (M1 x M0)."""
    Nf = fft2_crop(layer ** 2, M1, M0)
    Nif = fft2_crop((1.0 / layer) ** 2, M1, M0)
    return Nf, Nif


def circulant_index(M0: int, M1: int):
    """FourierWave.h:366-372 (A3): (row, col) into the M1 x M0 coefficient block for entry (p, p')."""
    qi = np.repeat(np.arange(M1), M0)        # row p = qi*M0 + pi
    pi = np.tile(np.arange(M0), M1)
    indR = (qi[:, None] - qi[None, :]) % M1
    indC = (pi[:, None] - pi[None, :]) % M0
    return (indR + M1 // 2) % M1, (indC + M0 // 2) % M0


def phi_matrix(Nf: np.ndarray, Nif: np.ndarray, up: np.ndarray, wq: np.ndarray, k: float) -> np.ndarray:
    """FourierWave.h:336-397: dense Phi (4MF x 4MF) = [[0, P], [Q, 0]], rows p = qi*M0 + pi."""
    M1, M0 = Nf.shape
    MF = M0 * M1
    R, C = circulant_index(M0, M1)
    A0 = Nf[R, C]
    NI = Nif[R, C]
    upr = np.tile(up, M1).astype(np.complex128)          # up of row / column p
    wqr = np.repeat(wq, M0).astype(np.complex128)
    kinv = 1.0 / k ** 2
    A1 = NI * upr[None, :]
    A2 = NI * wqr[None, :]
    phi = np.zeros((4 * MF, 4 * MF), dtype=np.complex128)
    r = np.arange(MF)
    phi[0:MF, 2 * MF:3 * MF] = (upr * kinv)[:, None] * A2
    phi[0:MF, 3 * MF:4 * MF] = -(upr * kinv)[:, None] * A1
    phi[r, 3 * MF + r] += 1
    phi[MF:2 * MF, 2 * MF:3 * MF] = (wqr * kinv)[:, None] * A2
    phi[MF:2 * MF, 3 * MF:4 * MF] = -(wqr * kinv)[:, None] * A1
    phi[MF + r, 2 * MF + r] += -1
    phi[2 * MF:3 * MF, MF:2 * MF] = -A0
    phi[2 * MF + r, r] += -upr * wqr * kinv
    phi[2 * MF + r, MF + r] += upr * upr * kinv
    phi[3 * MF:4 * MF, 0:MF] = A0
    phi[3 * MF + r, r] += -wqr * wqr * kinv
    phi[3 * MF + r, MF + r] += upr * wqr * kinv
    return phi


def _sorter_cmp(a: complex, b: complex) -> int:
    """ScatterVolume.cpp:177-184 as a three-way comparison (absolute-eps, A9)."""
    if abs(a.imag - b.imag) > DBL_EPS:
        return -1 if a.imag < b.imag else 1
    if abs(a.real - b.real) > DBL_EPS:
        return -1 if a.real < b.real else 1
    if a.imag < b.imag:
        return -1
    if b.imag < a.imag:
        return 1
    return 0


def eigen_sort(w: np.ndarray, V: np.ndarray):
    """ScatterVolume.cpp:190-230: sort ascending, column 2i = i-th largest, 2i+1 = i-th smallest."""
    n = len(w)
    idx = sorted(range(n), key=functools.cmp_to_key(lambda i, j: _sorter_cmp(w[i], w[j])))
    order = np.empty(n, dtype=np.int64)
    for i in range(n // 2):
        order[2 * i] = idx[n - 1 - i]
        order[2 * i + 1] = idx[i]
    return w[order], V[:, order]


def gd_gc(gamma: np.ndarray, V: np.ndarray, k: float, dz: float):
    """ScatterVolume.cpp:297-321: even (down) columns get e^{ik g dz} in Gc, odd (up) e^{-ik g dz} in Gd."""
    Ci = np.exp(1j * k * gamma * dz)
    Di = np.exp(1j * k * gamma * (-dz))
    odd = (np.arange(len(gamma)) % 2) == 1
    Gd = V * np.where(odd, Di, 1.0)[None, :]
    Gc = V * np.where(odd, 1.0, Ci)[None, :]
    return Gd, Gc


def mat_transfer(s: Setup):
    """ScatterVolume.cpp:339-395: f1, f2, f3 (4MF x 3MF)."""
    MF = s.MF
    phase = np.exp(1j * s.k * s.z[0] * s.Sz0)
    f1 = np.zeros((4 * MF, 3 * MF), dtype=np.complex128)
    f2 = np.zeros_like(f1)
    f3 = np.zeros_like(f1)
    r = np.arange(MF)
    f1[r, r] = phase
    f1[MF + r, MF + r] = phase
    f1[2 * MF + r, MF + r] = -phase * s.Sz0
    f1[2 * MF + r, 2 * MF + r] = phase * s.Sy
    f1[3 * MF + r, r] = phase * s.Sz0
    f1[3 * MF + r, 2 * MF + r] = -phase * s.Sx
    f2[r, r] = 1
    f2[MF + r, MF + r] = 1
    f2[2 * MF + r, MF + r] = s.Sz0
    f2[2 * MF + r, 2 * MF + r] = s.Sy
    f2[3 * MF + r, r] = -s.Sz0
    f2[3 * MF + r, 2 * MF + r] = -s.Sx
    f3[r, r] = -1
    f3[MF + r, MF + r] = -1
    f3[2 * MF + r, MF + r] = s.Sz1
    f3[2 * MF + r, 2 * MF + r] = -s.Sy
    f3[3 * MF + r, r] = -s.Sz1
    f3[3 * MF + r, 2 * MF + r] = s.Sx
    return f1, f2, f3


def gauss_rows(s: Setup) -> np.ndarray:
    """ScatterVolume.cpp:126-131, 398-414: rows 0..2MF of A (k.E = 0 for reflected and transmitted)."""
    MF = s.MF
    G = np.zeros((2 * MF, 6 * MF), dtype=np.complex128)
    m = np.arange(MF)
    G[m, m] = s.Sx
    G[m, MF + m] = s.Sy
    G[m, 2 * MF + m] = -s.Sz0
    G[MF + m, 3 * MF + m] = s.Sx
    G[MF + m, 4 * MF + m] = s.Sy
    G[MF + m, 5 * MF + m] = s.Sz1
    return G


@dataclass
class Solution:
    setup: Setup
    x: np.ndarray                # 6MF: [Rx, Ry, Rz, Tx, Ty, Tz]
    gammas: list                 # per layer, sorted/interleaved eigenvalues (4MF)
    vectors: list                # per layer, eigenvectors in the same order (4MF x 4MF)
    betas: list                  # per layer (4MF)
    Q_check: list                # per layer (3MF x 2MF)
    Q_hat: list
    Nifs: list                   # per layer (M1 x M0)
    A: np.ndarray = None         # boundary system (6MF x 6MF), kept for residual checks
    b: np.ndarray = None


def eig_layers(s: Setup):
    """Per-layer Phi, zgeev and sort (ScatterVolume.cpp:241-279, FourierWave.h:303-315)."""
    gammas, vectors, Nifs, phis = [], [], [], []
    for layer in s.layers:
        Nf, Nif = layer_fourier(layer, s.M0, s.M1)
        Nifs.append(Nif)
        Phi = phi_matrix(Nf, Nif, s.up, s.wq, s.k)
        w, V = zgeev(Phi)
        g, Vs = eigen_sort(w, V)
        gammas.append(g)
        vectors.append(Vs)
        phis.append(Phi)
    return gammas, vectors, Nifs, phis


def eig_layers_reduced(s: Setup):
    """The reduced eigenproblem the HIP path's eigen mode 1 solves instead of zgeev of the 4MF Phi (no
    counterpart in the reference; used to time the CPU side of that mode and to check it).  Phi =
    [[0, P], [Q, 0]] squares to diag(PQ, QP), so each eigenpair (mu, a) of PQ gives gamma = +-sqrt(mu)
    with vectors [a; +-Q a / gamma]; they are sorted and interleaved like zgeev's (ScatterVolume.cpp:190-230).
    Invariant outputs (x, fields) equal the full form's wherever no gamma is near 0."""
    gammas, vectors, Nifs = [], [], []
    MF = s.MF
    for layer in s.layers:
        Nf, Nif = layer_fourier(layer, s.M0, s.M1)
        Nifs.append(Nif)
        Phi = phi_matrix(Nf, Nif, s.up, s.wq, s.k)
        P, Q = Phi[0:2 * MF, 2 * MF:], Phi[2 * MF:, 0:2 * MF]
        mu, a = zgeev(P @ Q)
        g = np.sqrt(mu.astype(np.complex128))
        Qa = Q @ a
        w = np.concatenate([g, -g])
        V = np.concatenate([np.concatenate([a, a], axis=1),
                            np.concatenate([Qa / g[None, :], -Qa / g[None, :]], axis=1)], axis=0)
        V = V / np.linalg.norm(V, axis=0)[None, :]
        gs, Vs = eigen_sort(w, V)
        gammas.append(gs)
        vectors.append(Vs)
    return gammas, vectors, Nifs


def solve(sample: np.ndarray, p: VolumeParams, eig=None, EF=None) -> Solution:
    """The scattervolume solve, ScatterVolume.cpp:574-889.

    ``eig`` may replace the per-layer eigen step: a callable (Setup) -> (gammas, vectors, Nifs);
    the default is :func:`eig_layers`.  ``EF`` (3MF) replaces the incident field's Fourier
    coefficients of ScatterVolume.cpp:660-674.
    """
    s = setup(sample, p)
    if EF is not None:
        s.EF = np.asarray(EF, np.complex128).copy()
    MF, L, k, z = s.MF, s.L, s.k, s.z
    ZL = L + 1
    if eig is None:
        gammas, vectors, Nifs, _ = eig_layers(s)
    else:
        gammas, vectors, Nifs = eig(s)
    GD, GC = [], []
    Gd_static = Gc_static = None
    for i in range(L):
        dz = (z[ZL - 1] - z[0]) if L == 1 else (z[i + 1] - z[i])
        Gd, Gc = gd_gc(gammas[i], vectors[i], k, dz)
        GD.append(Gd)
        GC.append(Gc)
        if i == 0:
            Gd_static, Gc_static = Gd, Gc
        else:
            Gc_static = (Gc @ zinv(Gd)) @ Gc_static
    f1, f2, f3 = mat_transfer(s)
    A = np.zeros((6 * MF, 6 * MF), dtype=np.complex128)
    b = np.zeros(6 * MF, dtype=np.complex128)
    A[0:2 * MF] = gauss_rows(s)
    A[2 * MF:, 0:3 * MF] = f2
    A[2 * MF:, 3 * MF:] = (Gd_static @ zinv(Gc_static)) @ f3
    b[2 * MF:] = -(f1 @ s.EF)
    x = zgesv(A.copy(), b.copy())
    betas, Q_check, Q_hat = [], [], []
    if not p.external:
        beta = None
        for i in range(L):
            Ginv = zinv(GD[i])
            if i == 0:
                beta = (Ginv @ f1) @ s.EF + (Ginv @ f2) @ x[0:3 * MF]
            else:
                beta = (Ginv @ GC[i - 1]) @ beta
            betas.append(beta)
        ps, qs = order_series(s.M0), order_series(s.M1)
        WQ = 2.0 * qs * math.pi / p.size[1]            # ScatterVolume.cpp:813-814 uses M_PI
        UP = 2.0 * ps * math.pi / p.size[0]
        upz = np.tile(UP + s.dir_n[0] * k, s.M1)       # up of column p' = qj*M0 + pj
        wqz = np.repeat(WQ + s.dir_n[1] * k, s.M0)
        Cz = pz_matrix(s.M0, s.M1)
        for i in range(L):
            V = vectors[i]
            Q1 = V[:, 0::2] * betas[i][0::2][None, :]
            Q2 = V[:, 1::2] * betas[i][1::2][None, :]
            Qc = np.zeros((3 * MF, 2 * MF), dtype=np.complex128)
            Qh = np.zeros_like(Qc)
            Qc[0:2 * MF] = Q1[0:2 * MF]
            Qh[0:2 * MF] = Q2[0:2 * MF]
            if not p.npz:
                Cn = Nifs[i].ravel(order="F")[Cz]
                I1, I2 = Q1[2 * MF:3 * MF], Q2[2 * MF:3 * MF]
                J1, J2 = Q1[3 * MF:4 * MF], Q2[3 * MF:4 * MF]
                Qc[2 * MF:] = (-1.0 / k) * (Cn @ (upz[:, None] * J1 - wqz[:, None] * I1))
                Qh[2 * MF:] = (-1.0 / k) * (Cn @ (upz[:, None] * J2 - wqz[:, None] * I2))
            Q_check.append(Qc)
            Q_hat.append(Qh)
    return Solution(setup=s, x=x, gammas=gammas, vectors=vectors, betas=betas,
                    Q_check=Q_check, Q_hat=Q_hat, Nifs=Nifs, A=A, b=b)


def pz_matrix(M0: int, M1: int) -> np.ndarray:
    """ScatterVolume.cpp:864-875 (A7): linear index into the column-major M1 x M0 Nif for (p, p')."""
    qi = np.repeat(np.arange(M1), M0)
    pi = np.tile(np.arange(M0), M1)
    indR = (qi[:, None] - qi[None, :] + M1) % M1
    indC = (pi[:, None] - pi[None, :] + M0) % M0
    return ((indR + M1 // 2) % M1) * M0 + (indC + M0 // 2) % M0


# ----------------------------------------------------------------------------------------------------
# .cw assembly, container and field evaluation
# ----------------------------------------------------------------------------------------------------

@dataclass
class CW:
    """CoupledWaveStructure (CoupledWaveStructure.h:48-61): records are (E0x,E0y,E0z,kx,ky,kz)."""
    Pi: np.ndarray                       # (nPi, 6) complex128
    z: list = field(default_factory=list)
    Pr: list = field(default_factory=list)   # per boundary (n, 6)
    Pt: list = field(default_factory=list)


def _wave(kx, ky, kz, E):
    return np.stack([E[0], E[1], E[2], kx, ky, kz], axis=-1).astype(np.complex128)


def _wind(rec: np.ndarray, zb: float) -> np.ndarray:
    """planewave::wind(0, 0, -z_l): E0 *= exp(i k . (0, 0, -z_l)) (A20)."""
    out = rec.copy()
    ph = np.exp(1j * rec[:, 5] * (-zb))
    out[:, 0:3] *= ph[:, None]
    return out


def build_cw(sol: Solution) -> CW:
    """ScatterVolume.cpp:907-987: incident, reflected, transmitted and interior waves per boundary."""
    s, p = sol.setup, sol.setup.params
    MF, L, k, z = s.MF, s.L, s.k, s.z
    ZL = L + 1
    x = sol.x
    pidx = np.arange(MF)                         # loop kk<M1, j<M0 gives p = kk*M0 + j = 0..MF-1
    kx, ky = s.Sx * k, s.Sy * k
    Pi = _wave(kx, ky, s.Sz0 * k, [s.EF[pidx], s.EF[MF + pidx], s.EF[2 * MF + pidx]])
    r = _wave(kx, ky, -s.Sz0 * k * p.n[0], [x[pidx], x[MF + pidx], x[2 * MF + pidx]])
    t = _wave(kx, ky, s.Sz1 * k * p.n[1], [x[3 * MF + pidx], x[4 * MF + pidx], x[5 * MF + pidx]])
    cw = CW(Pi=Pi, z=[0.0] * ZL, Pr=[np.zeros((0, 6), np.complex128) for _ in range(ZL)],
            Pt=[np.zeros((0, 6), np.complex128) for _ in range(ZL)])

    def modes(layer, parity, Q):
        # for each p, 2MF waves (Sx(p)k, Sy(p)k, gamma[2jj+parity] k; Q[p,jj], Q[MF+p,jj], Q[2MF+p,jj])
        g = sol.gammas[layer][parity::2] * k
        KX = np.repeat(kx, 2 * MF)
        KY = np.repeat(ky, 2 * MF)
        KZ = np.tile(g, MF)
        E = [Q[c * MF:(c + 1) * MF, :].reshape(-1) for c in range(3)]
        return _wave(KX, KY, KZ, E)

    for l in range(ZL):
        if l == 0:
            cw.z[l] = z[l]
            cw.Pr[l] = _wind(r, z[l])
            if not p.external:
                cw.Pt[l] = _wind(modes(l, 0, sol.Q_check[l]), z[l])
        elif l == ZL - 1:
            cw.z[l] = z[l]
            cw.Pt[l] = _wind(t, z[l])
            if not p.external:
                cw.Pr[l] = _wind(modes(l - 1, 1, sol.Q_hat[l - 1]), z[l])
        else:
            if not p.external:
                cw.Pr[l] = _wind(modes(l - 1, 1, sol.Q_hat[l - 1]), z[l])
                cw.Pt[l] = _wind(modes(l, 0, sol.Q_check[l]), z[l])
                cw.z[l] = z[l]
    return cw


CW_MAGIC_V2 = 0     # leading is_volume byte of the 96-byte-record layout (ScatterView.py:82-83)


def cw_write(cw: CW, path: str) -> None:
    """Container of CoupledWaveStructure.h:63-92 behind an is_volume byte, 96-byte records (A17)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<B", CW_MAGIC_V2))
        f.write(struct.pack("<QQ", 8, len(cw.Pi)))
        f.write(np.ascontiguousarray(cw.Pi, dtype=np.complex128).tobytes())
        f.write(struct.pack("<Q", len(cw.z)))
        for l in range(len(cw.z)):
            f.write(struct.pack("<dQ", cw.z[l], len(cw.Pr[l])))
            f.write(np.ascontiguousarray(cw.Pr[l], dtype=np.complex128).tobytes())
            f.write(struct.pack("<Q", len(cw.Pt[l])))
            f.write(np.ascontiguousarray(cw.Pt[l], dtype=np.complex128).tobytes())


def cw_read(path: str) -> CW:
    """Reads the 96-byte-record layout, or the legacy 80-byte layout (CoupledWaveStructure.h:94-125).

    Legacy records hold (E0.x, E0.y, k.x, k.y, k.z); E0.z is returned as 0.
    """
    buf = open(path, "rb").read()
    if buf[0] in (0, 1) and struct.unpack_from("<Q", buf, 1)[0] in (4, 8):
        off, rec_c = 1, 6
    else:
        off, rec_c = 0, 5
    prec, = struct.unpack_from("<Q", buf, off)
    if prec != 8:
        raise ValueError("only double-precision .cw files are supported")
    off += 8

    def records(off):
        n, = struct.unpack_from("<Q", buf, off)
        off += 8
        a = np.frombuffer(buf, dtype=np.complex128, count=n * rec_c, offset=off).reshape(n, rec_c)
        off += n * rec_c * 16
        if rec_c == 5:
            a = np.concatenate([a[:, 0:2], np.zeros((n, 1), np.complex128), a[:, 2:5]], axis=1)
        return a.copy(), off

    Pi, off = records(off)
    nb, = struct.unpack_from("<Q", buf, off)
    off += 8
    cw = CW(Pi=Pi)
    for _ in range(nb):
        zb, = struct.unpack_from("<d", buf, off)
        off += 8
        pr, off = records(off)
        pt, off = records(off)
        cw.z.append(zb)
        cw.Pr.append(pr)
        cw.Pt.append(pt)
    return cw


def unpack(cw: CW, dtype=np.complex64):
    """cpuEvaluator.cpp:34-102: flat wave table (Pi, then Pr/Pt per boundary) and region ranges."""
    parts = [cw.Pi]
    begin, end = [0], []
    total = len(cw.Pi)
    for l in range(len(cw.z)):
        total += len(cw.Pr[l])
        end.append(total)
        begin.append(total)
        total += len(cw.Pt[l])
        parts.extend([cw.Pr[l], cw.Pt[l]])
    end.append(total)
    W = np.concatenate(parts).astype(dtype)
    zl = np.asarray(cw.z, dtype=np.float32)
    return W, zl, np.asarray(begin, np.int64), np.asarray(end, np.int64)


def region_of(zv, z_layers, tol: float):
    """Region index: last boundary with z >= z_l - tol (gpuEvaluator.cu:32-37; tol=0 on the CPU)."""
    zv = np.asarray(zv, dtype=np.float64)
    l = np.zeros(zv.shape, dtype=np.int64)
    for li in range(len(z_layers)):
        l = np.where(zv >= np.float64(z_layers[li]) - tol, li + 1, l)
    return l


def field_points(W, z_layers, begin, end, pts: np.ndarray, tol: float = 0.001) -> np.ndarray:
    """FP64 sum E(r) = sum_w E0_w exp(i k_w . r) over the region of each point (gpuEvaluator.cu:28-68).

    ``pts`` is (n, 3) float; the result is (n, 3) complex128.  Coordinates and wave data are taken
    exactly as stored (float32 after unpack), the arithmetic is double.
    """
    pts = np.asarray(pts)
    W = np.asarray(W).astype(np.complex128)
    reg = region_of(pts[:, 2].astype(np.float32), z_layers, tol)
    out = np.zeros((len(pts), 3), dtype=np.complex128)
    P = pts.astype(np.float64)
    for l in np.unique(reg):
        sel = np.nonzero(reg == l)[0]
        w = W[begin[l]:end[l]]
        if len(w) == 0:
            continue
        for c0 in range(0, len(sel), 4096):
            s = sel[c0:c0 + 4096]
            kr = P[s] @ w[:, 3:6].T            # (n, nw)
            ph = np.exp(1j * kr)
            out[s] = ph @ w[:, 0:3]
    return out


def grid_coord(origin, step, i) -> np.ndarray:
    """float32 coordinate origin + step * i rounded once, as a fused multiply-add computes it
    (gpuEvaluator.cu:78-79 under contraction).  The exact value is formed in extended precision."""
    o = np.longdouble(np.float32(origin))
    d = np.longdouble(np.float32(step))
    return (o + d * np.asarray(i, dtype=np.longdouble)).astype(np.float32)


def slice_points(axis: int, N: int, extent: float, center, plane_position):
    """ScatterView.cpp:354-376 with gpuEvaluator.cu:70-129: points of the xy / xz / yz slice.

    Returns (N*N, 3) float32 in the output order of the (N, N, 3) slice (row index = second axis).
    """
    d = np.float32(extent / (N - 1))
    start = np.asarray(center, np.float32) - np.float32(extent / 2)
    i = np.arange(N)
    out = np.zeros((N, N, 3), dtype=np.float32)
    if axis == 2:      # xy: E_xy[yi, xi] at (x_start + d xi, y_start + d yi, z)
        out[..., 0] = grid_coord(start[0], d, i)[None, :]
        out[..., 1] = grid_coord(start[1], d, i)[:, None]
        out[..., 2] = np.float32(plane_position[2])
    elif axis == 1:    # xz: E_xz[zi, xi] at (x_start + d xi, y, z_start + d zi)
        out[..., 0] = grid_coord(start[0], d, i)[None, :]
        out[..., 1] = np.float32(plane_position[1])
        out[..., 2] = grid_coord(start[2], d, i)[:, None]
    else:              # yz: E_yz[zi, yi] at (x, y_start + d yi, z_start + d zi)
        out[..., 0] = np.float32(plane_position[0])
        out[..., 1] = grid_coord(start[1], d, i)[None, :]
        out[..., 2] = grid_coord(start[2], d, i)[:, None]
    return out.reshape(N * N, 3)


def grid_points(x0, y0, z0, dx, dy, dz, nx, ny, nz) -> np.ndarray:
    """(nz*ny*nx, 3) float32 points of a regular grid in [iz][iy][ix] order."""
    X = grid_coord(x0, dx, np.arange(nx))
    Y = grid_coord(y0, dy, np.arange(ny))
    Z = grid_coord(z0, dz, np.arange(nz))
    P = np.zeros((nz, ny, nx, 3), np.float32)
    P[..., 0] = X[None, None, :]
    P[..., 1] = Y[None, :, None]
    P[..., 2] = Z[:, None, None]
    return P.reshape(-1, 3)


def far_field_flux(sol: Solution):
    """Power carried along z by incident, reflected and transmitted orders (for energy balance).

    For a wave with amplitude E and normalised kz/k = s, the z-flux is proportional to Re(s) |E|^2;
    returns (incident, reflected, transmitted) summed over orders, using the same kz as the .cw
    (ScatterVolume.cpp:474-487, including the extra n factor of A13).
    """
    s = sol.setup
    MF, p = s.MF, s.params
    R = sol.x[0:3 * MF].reshape(3, MF)
    T = sol.x[3 * MF:6 * MF].reshape(3, MF)
    E = s.EF.reshape(3, MF)
    inc = np.sum(np.real(s.Sz0) * np.sum(np.abs(E) ** 2, axis=0))
    ref = np.sum(np.real(s.Sz0) * np.sum(np.abs(R) ** 2, axis=0))
    tra = np.sum(np.real(s.Sz1) * np.sum(np.abs(T) ** 2, axis=0))
    return inc, ref, tra
