"""GPU parity of the dense complex linear algebra (third_Lapack.h:38-85 seam) against numpy LAPACK."""
import numpy as np
import pytest

import scattersim_amd as sa
from oracle import cw_oracle as o

pytestmark = pytest.mark.gpu


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (5, 7, 3), (64, 64, 64), (100, 37, 129), (130, 65, 17)])
def test_zgemm(ctx, m, n, k):
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    A, B = crand(rng, m, k), crand(rng, k, n)
    C = sa.zgemm(ctx, A, B, alpha=0.5 - 0.25j)
    ref = (0.5 - 0.25j) * (A @ B)
    assert np.max(np.abs(C - ref)) <= 1e-13 * np.max(np.abs(ref)) * np.sqrt(k)


@pytest.mark.parametrize("n,nrhs,kind", [(1, 1, "getrf"), (37, 5, "getrf"), (1350, 1, "getrf"), (900, 7, "getrf"),
                                         (37, 3, "general"), (1350, 1, "general")])
def test_getrs_row_interchanges(ctx, n, nrhs, kind):
    """zgetrs with identity factors is the row interchange alone (zlaswp): B <- P B for (t, ipiv[t]) in order,
    with zgetrf's pivots (ipiv[t] >= t) and with arbitrary ones, against numpy applying the swaps."""
    rng = np.random.default_rng(n + nrhs)
    B = crand(rng, nrhs, n)  # column-major n x nrhs
    if kind == "getrf":
        ipiv = np.array([rng.integers(t, n) for t in range(n)], np.int32)
    else:
        ipiv = rng.integers(0, n, n).astype(np.int32)
    ref = B.T.copy()
    for t in range(n):
        p = ipiv[t]
        ref[[t, p]] = ref[[p, t]]
    LU = np.eye(n, dtype=np.complex128)
    dLU = sa.DeviceArray.from_numpy(ctx, LU)
    dP = sa.DeviceArray.from_numpy(ctx, ipiv)
    dB = sa.DeviceArray.from_numpy(ctx, B)
    sa.check(sa.lib().cwx_zgetrs_batched_dev(ctx.h, n, nrhs, 1, dLU.ptr, n * n, dP.ptr, n, dB.ptr, n * nrhs))
    X = dB.numpy().T
    assert np.array_equal(X, ref)
    for d in (dLU, dP, dB):
        d.free()


def op(X, t):
    return X if t == 0 else (X.T if t == 1 else X.conj().T)


@pytest.mark.parametrize("m,n,k,ta,tb,batch", [(1, 1, 1, 0, 0, 1), (100, 37, 129, 0, 0, 2), (65, 130, 17, 1, 2, 1),
                                               (484, 484, 484, 0, 0, 3), (200, 96, 300, 2, 1, 2)])
def test_zgemm_fp32_matrix_cores(ctx, m, n, k, ta, tb, batch):
    """cwx_zgemm_split32_batched_dev against the FP64 product: elementwise within 16 * 2^-24 of |A| |B|
    (FP32 accumulation of at most 12 matrix-core steps per 16 values of k, FP64 across them), for the split
    form (3 terms) and the rounded form (1 term).  Both errors are printed: the split form is no more accurate
    once k passes a few dozen (1.6e-8 against 1.5e-8 at k = 484), since the FP32 rounding of the hi hi
    partial sums, not the operands' rounding, sets the error."""
    rng = np.random.default_rng(m + 3 * n + 7 * k + ta + tb)
    shA = (k, m) if ta else (m, k)
    shB = (n, k) if tb else (k, n)
    A, B, C0 = crand(rng, batch, *shA), crand(rng, batch, *shB), crand(rng, batch, m, n)
    alpha, beta = 0.75 + 0.5j, -0.5 + 0.25j
    lda, ldb = shA[0], shB[0]
    dA = sa.DeviceArray.from_numpy(ctx, np.ascontiguousarray(A.transpose(0, 2, 1)))
    dB = sa.DeviceArray.from_numpy(ctx, np.ascontiguousarray(B.transpose(0, 2, 1)))
    errs = {}
    for terms in (3, 1):
        dC = sa.DeviceArray.from_numpy(ctx, np.ascontiguousarray(C0.transpose(0, 2, 1)))
        sa.check(sa.lib().cwx_zgemm_split32_batched_dev(
            ctx.h, ta, tb, m, n, k, sa._capi.cwx_complex(alpha.real, alpha.imag), dA.ptr, lda, A[0].size, dB.ptr, ldb,
            B[0].size, sa._capi.cwx_complex(beta.real, beta.imag), dC.ptr, m, m * n, batch, terms))
        C = dC.numpy().transpose(0, 2, 1)
        dC.free()
        worst = 0.0
        for i in range(batch):
            oa, ob = op(A[i], ta), op(B[i], tb)
            ref = alpha * (oa @ ob) + beta * C0[i]
            # componentwise scale: |alpha| (|Re A| + |Im A|)(|Re B| + |Im B|)
            absA = np.abs(oa.real) + np.abs(oa.imag)
            absB = np.abs(ob.real) + np.abs(ob.imag)
            scale = abs(alpha) * (absA @ absB)
            worst = max(worst, float(np.max(np.abs(C[i] - ref) / scale)))
        errs[terms] = worst
    print(f"{m}x{n}x{k} op({ta},{tb}) batch {batch}: split {errs[3]:.2e}, rounded {errs[1]:.2e} of |A||B|", flush=True)
    assert errs[3] < 16 * 2.0 ** -24 and errs[1] < 16 * 2.0 ** -24
    dA.free()
    dB.free()


@pytest.mark.parametrize("n,nrhs", [(1, 1), (2, 3), (31, 1), (32, 4), (33, 2), (100, 1), (300, 7), (1100, 3), (1100, 1),
                                    (2000, 1)])
def test_zgesv(ctx, n, nrhs):
    rng = np.random.default_rng(n * 7 + nrhs)
    A = crand(rng, n, n) + 2 * np.eye(n)
    B = crand(rng, n, nrhs)
    X = sa.zgesv(ctx, A, B)
    res = np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X))
    assert res < 1e-14 * max(10, n)
    assert np.max(np.abs(X - np.linalg.solve(A, B))) <= 1e-10 * np.max(np.abs(X))


@pytest.mark.parametrize("n,batch", [(40, 3), (200, 2), (450, 4), (675, 1), (900, 1), (1100, 1)])
@pytest.mark.parametrize("panel", [0, 1, 2])
def test_zgetrf_pivots_and_factors_match_lapack(ctx, n, batch, panel):
    """zgetrf's pivots equal LAPACK's (scipy lu_factor: izamax on |re| + |im|, first of equals) and its packed
    factors agree to rounding, with each panel kernel (lu_panel 2: whole rows per lane, 1: a column pair per
    wavefront, both with 32- or 16-column panels by height; 0: rows per lane pair), across the panel
    heights they choose between."""
    import scipy.linalg as sl
    rng = np.random.default_rng(n + batch)
    A = np.stack([crand(rng, n, n) for _ in range(batch)])
    A[0][:, 3] = 1.0  # a column of equal entries: the first of them is the pivot
    dA = sa.DeviceArray.from_numpy(ctx, np.ascontiguousarray(A.transpose(0, 2, 1)))
    dP = sa.DeviceArray(ctx, (batch, n), np.int32)
    L = sa.lib()
    try:
        sa.check(L.cwx_set_param(b"lu_panel", panel))
        sa.check(L.cwx_zgetrf_batched_dev(ctx.h, n, batch, dA.ptr, n * n, dP.ptr, n, None))
    finally:
        sa.check(L.cwx_set_param(b"lu_panel", sa._capi.CWX_PARAM_TABLE))
    LU = dA.numpy().transpose(0, 2, 1)
    piv = dP.numpy()
    for k in range(batch):
        lu, pv = sl.lu_factor(A[k])
        assert np.array_equal(piv[k], pv), k
        assert np.max(np.abs(LU[k] - lu)) <= 1e-12 * np.max(np.abs(lu)) * np.sqrt(n), k
    dA.free()
    dP.free()


def test_zgesv_needs_pivoting(ctx):
    # zero leading pivot: solvable only with row interchanges
    A = np.array([[0, 1, 2], [1, 0, 3], [4, -3, 8]], dtype=np.complex128)
    b = np.array([1, 2, 3], dtype=np.complex128)
    x = sa.zgesv(ctx, A, b)
    assert np.allclose(A @ x, b, atol=1e-14)


def test_zinverse(ctx):
    rng = np.random.default_rng(3)
    A = crand(rng, 50, 50) + 3 * np.eye(50)
    Ai = sa.zinverse(ctx, A)
    assert np.max(np.abs(A @ Ai - np.eye(50))) < 1e-12


def check_eig(A, w, V, tol_res=None):
    n = A.shape[0]
    nA = np.linalg.norm(A)
    res = np.linalg.norm(A @ V - V * w[None, :]) / (max(nA, 1e-300) * np.sqrt(n))
    assert res < (tol_res or 1e-14 * max(10, n)), res
    # zgeev normalisation: unit 2-norm, largest component real and non-negative
    assert np.allclose(np.linalg.norm(V, axis=0), 1.0, atol=1e-12)
    k = np.argmax(np.abs(V) ** 2, axis=0)
    vk = V[k, np.arange(n)]
    assert np.all(vk.imag == 0.0)
    # eigenvalue multiset against LAPACK
    wr = np.linalg.eigvals(A)
    for x in w:
        assert np.min(np.abs(wr - x)) <= 1e-10 * max(nA, 1.0)
    for x in wr:
        assert np.min(np.abs(w - x)) <= 1e-10 * max(nA, 1.0)


@pytest.mark.parametrize("n", [1, 2, 3, 10, 33, 50, 98, 196, 450])
def test_zgeev_random(ctx, n):
    rng = np.random.default_rng(n)
    A = crand(rng, n, n)
    w, V = sa.zgeev(ctx, A)
    check_eig(A, w, V)


def test_zgeev_structured(ctx):
    # already triangular, diagonal with repeated values, and a Jordan-free degenerate case
    T = np.triu(np.arange(1, 26).reshape(5, 5)).astype(np.complex128)
    w, V = sa.zgeev(ctx, T)
    check_eig(T, w, V)
    D = np.diag([2, 2, 1j, 1j, -3]).astype(np.complex128)
    w, V = sa.zgeev(ctx, D)
    check_eig(D, w, V)
    assert np.linalg.matrix_rank(V) == 5


def test_zgeev_repeated_clusters(ctx):
    # a normal matrix with 12 eigenvalues of multiplicity 25 each (n = 300): the deflation window sees
    # whole clusters converge together, and the Schur basis inside a cluster is arbitrary
    rng = np.random.default_rng(4)
    n = 300
    vals = np.repeat(crand(rng, 12), 25)
    Q, _ = np.linalg.qr(crand(rng, n, n))
    A = (Q * vals[None, :]) @ Q.conj().T
    w, V = sa.zgeev(ctx, A)
    check_eig(A, w, V, tol_res=1e-12)
    assert np.max(np.abs(np.sort_complex(w) - np.sort_complex(vals))) < 1e-10


def test_zgeev_uniform_layer_phi(ctx):
    # Phi of a uniform layer: eigenvalues +-sqrt(n^2 - Sx^2 - Sy^2), each doubly degenerate
    p = o.VolumeParams(lam=0.5, size=(4, 4, 1), coef=(3, 3), n=(1.0, 1.0))
    s = o.setup(np.full((1, 8, 8), 1.4 + 0j), p)
    Nf, Nif = o.layer_fourier(s.layers[0], s.M0, s.M1)
    Phi = o.phi_matrix(Nf, Nif, s.up, s.wq, s.k)
    w, V = sa.zgeev(ctx, Phi)
    check_eig(Phi, w, V)
    assert np.linalg.matrix_rank(V, tol=1e-8) == Phi.shape[0]
    sx = np.tile(s.up / s.k, s.M1)
    sy = np.repeat(s.wq / s.k, s.M0)
    g = np.sqrt((1.4 ** 2 - sx ** 2 - sy ** 2).astype(np.complex128))
    expect = np.sort_complex(np.concatenate([g, g, -g, -g]))
    assert np.max(np.abs(np.sort_complex(w) - expect)) < 1e-12


def test_zgeev_degenerate_layer_phi_large(ctx):
    # Phi of a uniform layer at 9x9 orders (n = 324): every eigenvalue is exactly doubly degenerate, so whole
    # pairs meet in the deflation windows and the reordering swaps equal diagonal entries
    p = o.VolumeParams(lam=0.5, size=(4, 4, 1), coef=(9, 9), n=(1.0, 1.0))
    s = o.setup(np.full((1, 16, 16), 1.4 + 0.01j), p)
    Nf, Nif = o.layer_fourier(s.layers[0], s.M0, s.M1)
    Phi = o.phi_matrix(Nf, Nif, s.up, s.wq, s.k)
    w, V = sa.zgeev(ctx, Phi)
    check_eig(Phi, w, V)
    assert np.linalg.matrix_rank(V, tol=1e-8) == Phi.shape[0]


def test_zgeev_split_hessenberg(ctx):
    # a block upper-triangular matrix (three diagonal blocks of 60 coupled only upwards) keeps exact zeros on
    # its Hessenberg subdiagonal, so the active block splits while deflation windows and sweeps run
    rng = np.random.default_rng(9)
    n = 180
    A = np.triu(crand(rng, n, n))
    for b0 in (0, 60, 120):
        A[b0:b0 + 60, b0:b0 + 60] = crand(rng, 60, 60)
    w, V = sa.zgeev(ctx, A)
    check_eig(A, w, V)


def test_batched_dev_entry_points(ctx):
    rng = np.random.default_rng(11)
    n, b = 40, 5
    A = crand(rng, b, n, n)
    lib = sa.lib()
    # column-major per matrix: store transposes in row-major
    dA = sa.DeviceArray.from_numpy(ctx, np.ascontiguousarray(A.transpose(0, 2, 1)))
    dW = sa.DeviceArray(ctx, (b, n), np.complex128)
    dV = sa.DeviceArray(ctx, (b, n, n), np.complex128)
    info = np.zeros(b, np.int32)
    sa.check(lib.cwx_zgeev_batched_dev(ctx.h, n, b, dA.ptr, n * n, dW.ptr, n, dV.ptr, n * n, info.ctypes.data))
    W = dW.numpy()
    V = dV.numpy().transpose(0, 2, 1)
    for i in range(b):
        check_eig(A[i], W[i], V[i])
    # LU + solve on the device
    M = crand(rng, b, n, n) + 3 * np.eye(n)
    B = crand(rng, b, n, 3)
    dM = sa.DeviceArray.from_numpy(ctx, np.ascontiguousarray(M.transpose(0, 2, 1)))
    dB = sa.DeviceArray.from_numpy(ctx, np.ascontiguousarray(B.transpose(0, 2, 1)))
    dP = sa.DeviceArray(ctx, (b, n), np.int32)
    sa.check(lib.cwx_zgetrf_batched_dev(ctx.h, n, b, dM.ptr, n * n, dP.ptr, n, info.ctypes.data))
    sa.check(lib.cwx_zgetrs_batched_dev(ctx.h, n, 3, b, dM.ptr, n * n, dP.ptr, n, dB.ptr, n * 3))
    X = dB.numpy().transpose(0, 2, 1)
    for i in range(b):
        assert np.max(np.abs(M[i] @ X[i] - B[i])) < 1e-12
    # GEMM with a conjugate-transposed operand
    G = crand(rng, b, n, n)
    dG = sa.DeviceArray.from_numpy(ctx, np.ascontiguousarray(G.transpose(0, 2, 1)))
    dC = sa.DeviceArray(ctx, (b, n, n), np.complex128)
    sa.check(lib.cwx_zgemm_batched_dev(ctx.h, 2, 0, n, n, n, sa._capi.cwx_complex(1, 0), dG.ptr, n, n * n, dG.ptr, n,
                                       n * n, sa._capi.cwx_complex(0, 0), dC.ptr, n, n * n, b))
    Cm = dC.numpy().transpose(0, 2, 1)
    for i in range(b):
        assert np.max(np.abs(Cm[i] - G[i].conj().T @ G[i])) < 1e-12 * np.abs(Cm[i]).max()


def test_zgeev_large_batch_matches_small_batch(ctx):
    # A batch of at least two matrices per CU runs the QR with two smaller workgroups per CU; the
    # arithmetic of every matrix is the same either way, so results are bit-identical to solving
    # the matrices a few at a time.  n = 100 takes the deflation windows and multishift sweeps (active block >= 48).
    rng = np.random.default_rng(21)
    n, b = 100, 640
    A = crand(rng, b, n, n)
    lib = sa.lib()

    def run(Ab):
        k = Ab.shape[0]
        dA = sa.DeviceArray.from_numpy(ctx, np.ascontiguousarray(Ab.transpose(0, 2, 1)))
        dW = sa.DeviceArray(ctx, (k, n), np.complex128)
        dV = sa.DeviceArray(ctx, (k, n, n), np.complex128)
        info = np.zeros(k, np.int32)
        sa.check(lib.cwx_zgeev_batched_dev(ctx.h, n, k, dA.ptr, n * n, dW.ptr, n, dV.ptr, n * n,
                                           info.ctypes.data))
        out = dW.numpy(), dV.numpy().transpose(0, 2, 1), info
        for d in (dA, dW, dV):
            d.free()
        return out

    W, V, info = run(A)
    assert np.all(info == 0)
    picks = [0, 1, 317, b - 1]
    Ws, Vs, infos = run(A[picks])
    assert np.all(infos == 0)
    for i, p in enumerate(picks):
        assert np.array_equal(W[p], Ws[i]) and np.array_equal(V[p], Vs[i])
        check_eig(A[p], W[p], V[p])


@pytest.mark.parametrize("n,batch", [(450, 4), (196, 1), (900, 2)])
def test_zgeev_teams_match_one_workgroup(ctx, n, batch):
    """The Hessenberg panels and the QR iteration with helper workgroups per matrix (the default when the
    matrices do not fill the device) give exactly the eigenvalues and eigenvectors of one workgroup per
    matrix: every row of the panels' products is summed in the same order, and every row of the QR
    receives the same rotations in the same order."""
    import ctypes as C
    from scattersim_amd import _capi
    rng = np.random.default_rng(n + batch)
    A = np.stack([crand(rng, n, n) for _ in range(batch)])
    out = []
    try:
        for team in (0, -1):
            _capi.check(_capi.lib().cwx_set_param(b"qr_team", team))
            _capi.check(_capi.lib().cwx_set_param(b"lahr2_team", team))
            dA = sa.DeviceArray.from_numpy(ctx, np.ascontiguousarray(A.transpose(0, 2, 1)))
            dW = sa.DeviceArray(ctx, (batch, n), np.complex128)
            dV = sa.DeviceArray(ctx, (batch, n, n), np.complex128)
            info = np.zeros(batch, np.int32)
            _capi.check(_capi.lib().cwx_zgeev_batched_dev(ctx.h, n, batch, C.c_void_p(dA.ptr), n * n,
                                                          C.c_void_p(dW.ptr), n, C.c_void_p(dV.ptr), n * n,
                                                          info.ctypes.data_as(C.c_void_p)))
            out.append((dW.numpy(), dV.numpy()))
    finally:
        _capi.lib().cwx_set_param(b"qr_team", -1)
        _capi.lib().cwx_set_param(b"lahr2_team", -1)
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    for b in range(batch):
        check_eig(A[b], out[1][0][b], out[1][1][b].T)
