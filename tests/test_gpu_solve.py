"""GPU parity of the scattervolume solve (ScatterVolume.cpp:574-987) against the FP64 oracle.

Compared quantities are the ones every correct solve shares (eigenvector scaling and the basis of
degenerate modes are free, A10 of the survey): the boundary solution x = [R; T], the sorted eigenvalue
multiset per layer, and the fields the .cw records produce at points inside every region.  Bar:
1e-9 relative for the FP64 path.
"""
import os
import tempfile

import numpy as np
import pytest

import scattersim_amd as sa
from oracle import cw_oracle as o
from tests.helpers import cw_from_records, disks_sample, fields, golden, probe_points, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-9


def gpu_solve(ctx, sample, p: o.VolumeParams, eig_mode=0):
    return sa.scattervolume(ctx, sample, lam=p.lam, direction=p.direction, ex=p.ex, ey=p.ey, ez=p.ez, n=p.n,
                            kappa=p.kappa, size=p.size, z=p.z, coef=p.coef, psf=p.psf, external=p.external,
                            npz=p.npz, no_cutoff=p.no_cutoff, eig_mode=eig_mode)


def compare(ctx, sample, p, eig_mode=0, tol=TOL, check_gamma=True):
    ref = o.solve(sample, p)
    sol = gpu_solve(ctx, sample, p, eig_mode)
    s = ref.setup
    assert (sol.M0, sol.M1, sol.L) == (s.M0, s.M1, s.L)
    assert rel_err(sol.x, ref.x) < tol
    if check_gamma:
        G = sol.gammas
        for l in range(s.L):
            assert rel_err(np.sort_complex(G[l]), np.sort_complex(ref.gammas[l])) < tol
    # fields of the .cw records inside every region
    off, cnt = sol.lists()
    mine = cw_from_records(sol.records, cnt, sol.z)
    theirs = o.build_cw(ref)
    pts = probe_points(s.z, p.size)
    assert rel_err(fields(mine, pts), fields(theirs, pts)) < tol
    return sol, ref


@pytest.mark.parametrize("eig_mode", [0, 1])
def test_acw_known_answer(ctx, eig_mode):
    """layers.npy through the M = 1 path against the reference-produced x64/a.cw (TE waves)."""
    g = golden("acw_te.npz")
    sample = golden("samples.npz")["layers"]
    Pi, Pr, Pt, z = g["Pi"], g["Pr"], g["Pt"], g["z"]
    worst = 0.0
    for j in range(0, len(Pi), 9):
        kv = Pi[j, 3:6].real
        km = np.linalg.norm(kv)
        sol = sa.scattervolume(ctx, sample, lam=2 * o.PI_TRUNC / km, direction=tuple(kv / km), ex=(0, 0), ey=(1, 0),
                               n=(1.0, 1.0), size=(10, 10, 6), z=0.0, eig_mode=eig_mode)
        off, cnt = sol.lists()
        mine = cw_from_records(sol.records, cnt, sol.z)
        ref = o.CW(Pi=Pi[j:j + 1], z=list(z), Pr=[Pr[j:j + 1, l] for l in range(3)],
                   Pt=[Pt[j:j + 1, l] for l in range(3)])
        pts = np.array([[0.3, 0.0, zz] for zz in (-4.0, -2.0, -0.5, 0.5, 2.5, 4.0)])
        Em = fields(mine, pts)[:, 1] * Pi[j, 1]
        Ea = fields(ref, pts)[:, 1]
        worst = max(worst, rel_err(Em, Ea))
    assert worst < 1e-12


@pytest.mark.parametrize("coef", [(1, 1), (3, 3), (5, 5), (4, 3), (5, 1)])
def test_circle_single_layer(ctx, coef):
    sample = golden("samples.npz")["circle"]
    p = o.VolumeParams(lam=0.6, size=(4.0, 4.0, 1.0), coef=coef, ex=(1, 0), ey=(0, 0), n=(1.0, 1.2), kappa=(0.01,))
    compare(ctx, sample, p)


@pytest.mark.parametrize("eig_mode", [0, 1])
def test_three_layer_oblique(ctx, eig_mode):
    sample = disks_sample(16, 3, 4.0, 4, 1.0, lambda r: 1.4 + 0.02j * r.random(), seed=5)
    p = o.VolumeParams(lam=0.7, direction=(0.3, 0.1, 1.0), ex=(1, 0.5), ey=(0.2, 0), size=(4.0, 4.0, 1.5),
                       coef=(5, 5), n=(1.0, 1.33))
    compare(ctx, sample, p, eig_mode=eig_mode)


def test_reduced_chain_block_form_matches_order_4mf_chain(ctx):
    """The reduced mode's transfer chain in block form (LUs of the order-2MF eigenvectors and Q, T_g assembled
    from their products) against the same solve through LUs of the order-4MF Gd (chain_block 0, the
    reference's formulation, ScatterVolume.cpp:327-334) and against the oracle: eight layers at oblique
    incidence, interior betas included (the fields of the .cw records inside every layer)."""
    sample = disks_sample(16, 8, 4.0, 4, 1.0, lambda r: 1.3 + 0.3 * r.random() + 0.02j * r.random(), seed=21)
    p = o.VolumeParams(lam=0.6, direction=(0.25, -0.1, 1.0), ex=(1, 0.3), ey=(0.1, 0.2), size=(4.0, 4.0, 2.0),
                       coef=(7, 7), n=(1.0, 1.2))
    L = sa.lib()
    try:
        sa.check(L.cwx_set_param(b"chain_block", 0))
        s4, _ = compare(ctx, sample, p, eig_mode=1)
        x4 = s4.x.copy()
        sa.check(L.cwx_set_param(b"chain_block", 1))
        blk, _ = compare(ctx, sample, p, eig_mode=1)
    finally:
        sa.check(L.cwx_set_param(b"chain_block", sa._capi.CWX_PARAM_TABLE))
    assert rel_err(blk.x, x4) < 1e-11


@pytest.mark.parametrize("eig_mode", [0, 1])
def test_solves_by_inverse_diagonal_blocks_match_substitution(ctx, eig_mode):
    """The transfer chain's solves (reduced mode) and S^-1 f3 (both modes) with the factors' 64 x 64
    diagonal blocks applied as products by their inverses (getrs_inv 1), against substitution in 32 x 32
    diagonal blocks (getrs_inv 0) and against the oracle.  15x15-order problems are not needed: 7x7 orders
    give 98- and 196-row factors, whose last diagonal blocks are cut by the matrix's end."""
    sample = disks_sample(16, 6, 4.0, 4, 1.0, lambda r: 1.3 + 0.3 * r.random() + 0.02j * r.random(), seed=23)
    p = o.VolumeParams(lam=0.6, direction=(0.2, 0.1, 1.0), ex=(1, 0.2), ey=(0.1, 0.3), size=(4.0, 4.0, 2.0),
                       coef=(7, 7), n=(1.0, 1.2))
    L = sa.lib()
    try:
        sa.check(L.cwx_set_param(b"getrs_inv", 0))
        s0, _ = compare(ctx, sample, p, eig_mode=eig_mode)
        x0 = s0.x.copy()
        sa.check(L.cwx_set_param(b"getrs_inv", 1))
        s1, _ = compare(ctx, sample, p, eig_mode=eig_mode)
    finally:
        sa.check(L.cwx_set_param(b"getrs_inv", sa._capi.CWX_PARAM_TABLE))
    assert rel_err(s1.x, x0) < 1e-11


def test_buffer_cache_does_not_change_results(ctx):
    """Solves with the library's device-buffer cache off (alloc_cache 0: every buffer returned at once) and on
    (released buffers reused by the next solve, after one device synchronisation) give the same bits,
    including a solve whose buffers come back from the cache while the previous solution is still held."""
    sample = disks_sample(16, 4, 4.0, 4, 1.0, lambda r: 1.3 + 0.2 * r.random() + 0.01j * r.random(), seed=29)
    p = o.VolumeParams(lam=0.6, direction=(0.1, 0.2, 1.0), size=(4.0, 4.0, 1.5), coef=(5, 5), n=(1.0, 1.2))
    L = sa.lib()
    try:
        sa.check(L.cwx_set_param(b"alloc_cache", 0))
        x0 = gpu_solve(ctx, sample, p, eig_mode=1).x.copy()
        sa.check(L.cwx_set_param(b"alloc_cache", 1))
        a = gpu_solve(ctx, sample, p, eig_mode=1)
        xa = a.x.copy()
        a.free()
        b = gpu_solve(ctx, sample, p, eig_mode=1)  # takes the blocks a released
        c = gpu_solve(ctx, sample, p, eig_mode=1)  # while b is alive
        assert np.array_equal(xa, x0) and np.array_equal(b.x, x0) and np.array_equal(c.x, x0)
    finally:
        sa.check(L.cwx_set_param(b"alloc_cache", sa._capi.CWX_PARAM_TABLE))


def test_external_and_npz(ctx):
    sample = disks_sample(12, 2, 3.0, 3, 1.0, 1.5, seed=9)
    p = o.VolumeParams(lam=0.8, size=(3.0, 3.0, 1.0), coef=(3, 3), external=True)
    ref = o.solve(sample, p)
    sol = gpu_solve(ctx, sample, p)
    assert rel_err(sol.x, ref.x) < TOL
    off, cnt = sol.lists()
    assert list(cnt) == [len(ref.setup.EF) // 3] + [9, 0, 0, 0, 0, 9]
    p2 = o.VolumeParams(lam=0.8, size=(3.0, 3.0, 1.0), coef=(3, 3), npz=True)
    sol2 = gpu_solve(ctx, sample, p2)
    Q = sol2.Q_check_hat
    assert np.all(Q[:, 18:27, :] == 0)


def test_cw_file_roundtrip(ctx):
    sample = disks_sample(8, 1, 2.0, 2, 1.0, 1.3, seed=1)
    p = o.VolumeParams(lam=0.9, size=(2.0, 2.0, 0.5), coef=(3, 3))
    sol = gpu_solve(ctx, sample, p)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "c.cw")
        sol.write_cw(path)
        cw = o.cw_read(path)
        W, zl, b, e = sa.cw_read_unpacked(path)
    recs = sol.records
    flat = np.concatenate([cw.Pi] + [x for pair in zip(cw.Pr, cw.Pt) for x in pair])
    assert np.array_equal(flat, recs)
    assert np.array_equal(W["E0"], recs[:, 0:3]) and np.array_equal(W["k"], recs[:, 3:6])
    assert np.allclose(zl, sol.z)


def test_reduced_mode_falls_back_to_full_phi_at_zero_gamma(ctx):
    """A uniform n = 1 layer between n = 1.2 half-spaces at lambda = 0.5 um in a 4 um cell with 17 x 1
    orders: orders +-8 have Sx = 1 = n inside the layer, so gamma = 0 there and the reduced form's pair
    gamma = +-sqrt(mu) is defective.  The layer must go through the full Phi (the reference's zgeev of
    it), so eigen mode 1 gives exactly mode 0's solution, which matches the oracle.  With a few pixels of
    n = 1.3 at lambda = 0.49 um nothing grazes and mode 1 keeps the reduced path: its x then differs from
    mode 0 in the last bits."""
    sample = np.ones((1, 32, 32), np.complex128)
    base = dict(size=(4.0, 4.0, 1.0), coef=(17, 1), ex=(0, 0), ey=(1, 0), n=(1.2, 1.2), no_cutoff=True)
    p = o.VolumeParams(lam=0.5, **base)
    ref = o.solve(sample, p)
    assert np.min(np.abs(ref.gammas[0])) == 0.0
    x0 = gpu_solve(ctx, sample, p, 0).x
    x1 = gpu_solve(ctx, sample, p, 1).x
    assert np.all(np.isfinite(x1))
    assert np.array_equal(x0, x1)
    assert rel_err(x0, ref.x) < TOL
    q = o.VolumeParams(lam=0.49, **base)
    het = sample.copy()
    het[0, 10:20, 5:9] = 1.3
    refq = o.solve(het, q)
    y0 = gpu_solve(ctx, het, q, 0).x
    y1 = gpu_solve(ctx, het, q, 1).x
    assert not np.array_equal(y0, y1)
    assert rel_err(y1, refq.x) < TOL and rel_err(y0, refq.x) < TOL
