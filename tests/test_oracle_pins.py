"""Pins of the CPU oracle against data and mathematics it does not share code with.

* the reference-produced x64/a.cw (scatterlayer, homogeneous stack) through the M = 1 path;
* brute-force DFT against fft2_crop, and a hand-derived case of fftShift2d's parity quirk;
* closed forms: uniform-layer eigenvalues, an index-matched slab (no reflection), the order cut-off;
* an independent Maxwell formulation (reduced 2MF eigenproblem, one global continuity system with both
  half-spaces parametrised by (Ex, Ey), Fourier coefficients from numpy's FFT of the permittivity and
  the wrapped-difference truncation rule): the oracle itself for gratings (M1 = 1), and for crossed
  2-D structures the oracle with its Sx/Sy/Sz/EF flattened x-fastest -- the one place where the
  reference pairs orders differently from Phi (ScatterVolume.cpp:682-685 vs FourierWave.h:359-378);
* energy balance for lossless stacks; eigen residuals; .cw container round trips.
"""
import os
import tempfile

import numpy as np
import pytest

from oracle import cw_oracle as o
from tests.helpers import golden, rel_err


# ----------------------------------------------------------------------------------------------
def test_acw_known_answer_all_te_waves():
    g = golden("acw_te.npz")
    sample = golden("samples.npz")["layers"]
    Pi, Pr, Pt, z = g["Pi"], g["Pr"], g["Pt"], g["z"]
    assert len(Pi) == 63 and list(z) == [-3.0, 0.0, 3.0]
    worst = 0.0
    for j in range(len(Pi)):
        kv = Pi[j, 3:6].real
        km = np.linalg.norm(kv)
        p = o.VolumeParams(lam=2 * o.PI_TRUNC / km, direction=tuple(kv / km), ex=(0, 0), ey=(1, 0),
                           size=(10., 10., 6.), z=0.0)
        sol = o.solve(sample, p)
        mine = o.build_cw(sol)
        ref = o.CW(Pi=Pi[j:j + 1], z=list(z), Pr=[Pr[j:j + 1, l] for l in range(3)],
                   Pt=[Pt[j:j + 1, l] for l in range(3)])
        pts = np.array([[0.3, 0.0, zz] for zz in (-4.0, -3.5, -2.0, -0.5, 0.5, 2.5, 3.5, 5.0)])
        W, zl, b, e = o.unpack(mine, np.complex128)
        Wa, zla, ba, ea = o.unpack(ref, np.complex128)
        Em = o.field_points(W, zl, b, e, pts, tol=0.0)[:, 1] * Pi[j, 1]
        Ea = o.field_points(Wa, zla, ba, ea, pts, tol=0.0)[:, 1]
        worst = max(worst, rel_err(Em, Ea))
    assert worst < 1e-12


def test_a_cw_fixture_is_continuous():
    """Ey and the Hx-proxy dEy/dz of the fixture are continuous at every boundary (pins wind/regions)."""
    g = golden("acw_te.npz")
    Pi, Pr, Pt, z = g["Pi"], g["Pr"], g["Pt"], g["z"]
    for j in range(0, len(Pi), 7):
        cw = o.CW(Pi=Pi[j:j + 1], z=list(z), Pr=[Pr[j:j + 1, l] for l in range(3)],
                  Pt=[Pt[j:j + 1, l] for l in range(3)])
        W, zl, b, e = o.unpack(cw, np.complex128)
        for zb in z:
            above = o.field_points(W, zl, b, e, np.array([[0.2, 0.0, zb - 1e-7]]), tol=0.0)
            below = o.field_points(W, zl, b, e, np.array([[0.2, 0.0, zb + 1e-7]]), tol=0.0)
            assert abs(above[0, 1] - below[0, 1]) <= 1e-5 * abs(Pi[j, 1])


# ----------------------------------------------------------------------------------------------
@pytest.mark.parametrize("N1,N2,M1,M2", [(4, 4, 3, 3), (5, 7, 3, 5), (6, 6, 5, 4), (9, 5, 1, 5), (8, 8, 8, 8)])
def test_fft2_crop_matches_brute_force_dft(N1, N2, M1, M2):
    rng = np.random.default_rng(N1 * 31 + N2)
    A = rng.standard_normal((N1, N2)) + 1j * rng.standard_normal((N1, N2))
    got = o.fft2_crop(A, M1, M2)
    i = np.arange(N1)[:, None]
    j = np.arange(N2)[None, :]
    for r in range(M1):
        for c in range(M2):
            u = r - M1 // 2          # centred harmonic indices when both sizes share parity
            v = c - M2 // 2
            ref = np.sum(A * np.exp(-2j * np.pi * (u * i / N1 + v * j / N2))) / (N1 * N2)
            assert abs(got[r, c] - ref) < 1e-13


def test_fftshift_quirk_even_rows_odd_cols():
    """FourierWave.h:92-101,122-135: the column shift keys on the row parity; 2 x 3 by hand."""
    a = np.array([[1, 2, 3], [4, 5, 6]])
    assert np.array_equal(o.fftshift2d(a), np.array([[4, 5, 6], [1, 2, 3]]))
    b = np.arange(15).reshape(3, 5)
    assert np.array_equal(o.fftshift2d(b), np.fft.fftshift(b))
    c = np.arange(24).reshape(4, 6)
    assert np.array_equal(o.fftshift2d(c), np.fft.fftshift(c))


def test_reformat_reads_slices_column_major():
    data = np.arange(2 * 3 * 4).reshape(2, 3, 4).astype(np.complex128)
    L = o.reformat(data)
    # column-major (3 x 4) map of each C-order slice buffer
    assert np.array_equal(L[1], data[1].ravel().reshape(4, 3).T)
    sq = np.arange(9).reshape(1, 3, 3).astype(np.complex128)
    assert np.array_equal(o.reformat(sq)[0], sq[0].T)


# ----------------------------------------------------------------------------------------------
def test_uniform_layer_eigenvalues_closed_form():
    p = o.VolumeParams(lam=0.55, size=(3.0, 2.5, 1.0), coef=(5, 3), direction=(0.2, -0.1, 1.0), n=(1.0, 1.0))
    s = o.setup(np.full((1, 10, 10), 1.45 + 0.01j), p)
    gam, V, _, phis = o.eig_layers(s)
    sx = np.tile(s.up / s.k, s.M1)
    sy = np.repeat(s.wq / s.k, s.M0)
    g = np.sqrt(((1.45 + 0.01j) ** 2 - sx ** 2 - sy ** 2).astype(np.complex128))
    expect = np.sort_complex(np.concatenate([g, g, -g, -g]))
    assert np.max(np.abs(np.sort_complex(gam[0]) - expect)) < 1e-12
    # even columns (down) have Im >= 0 for this lossy layer
    assert np.all(gam[0][0::2].imag > 0)
    Phi = phis[0]
    res = np.linalg.norm(Phi @ V[0] - V[0] * gam[0][None, :]) / np.linalg.norm(Phi)
    assert res < 1e-13


def test_index_matched_slab_transmits_unchanged():
    """n0 = n1 = n(layer): no reflection; T at z_last equals the incident field propagated there."""
    p = o.VolumeParams(lam=0.7, size=(2.0, 2.0, 1.3), coef=(3, 3), n=(1.25, 1.25), direction=(0.1, 0.2, 1.0),
                       ex=(1, 0), ey=(0.3, 0.2))
    sol = o.solve(np.full((2, 6, 6), 1.25 + 0j), p)
    s, MF = sol.setup, sol.setup.MF
    assert np.max(np.abs(sol.x[:3 * MF])) < 1e-12
    Tz = sol.x[3 * MF:].reshape(3, MF)
    Einc = s.EF.reshape(3, MF) * np.exp(1j * s.k * s.Sz0 * s.z[-1])[None, :]
    assert np.max(np.abs(Tz - Einc)) < 1e-12


def test_order_cutoff_known_answers():
    # a 21 x 21 request at 4 um and lambda 0.5 is silently cut to 10 x 10 (ScatterVolume.cpp:613-624)
    p = o.VolumeParams(lam=0.5, size=(4.0, 4.0, 1.0), coef=(21, 21))
    assert o.select_orders(p, (1, 128, 128), np.array([0.0, 0.0, 1.0]))[:2] == (10, 10)
    p = o.VolumeParams(lam=0.5, size=(4.0, 4.0, 1.0), coef=(21, 21), no_cutoff=True)
    assert o.select_orders(p, (1, 128, 128), np.array([0.0, 0.0, 1.0]))[:2] == (21, 21)
    # one value = M x M with M = floor(sqrt(v)); clamp to the pixel counts
    p = o.VolumeParams(lam=2.0, size=(40.0, 40.0, 1.0), coef=(50,))
    assert o.select_orders(p, (1, 5, 9), np.array([0.0, 0.0, 1.0]))[:2] == (7, 5)


# ----------------------------------------------------------------------------------------------
# independent Maxwell formulation (Laurent rule, wrapped-difference truncation, x-fastest orders)
# ----------------------------------------------------------------------------------------------
def _wrapped_conv(coef2d, M0, M1):
    """C[(q,p), (q',p')] = c[wrap(q-q'), wrap(p-p')]; coef2d = full 2-D DFT / N (rows: y harmonics)."""
    N1, N2 = coef2d.shape
    q = np.repeat(np.arange(M1), M0)
    p = np.tile(np.arange(M0), M1)
    dq = (q[:, None] - q[None, :] + M1 // 2) % M1 - M1 // 2
    dp = (p[:, None] - p[None, :] + M0 // 2) % M0 - M0 // 2
    return coef2d[dq % N1, dp % N2]


def independent_rcwa(layers_yx, period, thick, k, sdir, E0, n0, n1, M0, M1):
    """Reflected / transmitted (Ex, Ey, Ez) per order (x-fastest) and fluxes, from one global system.

    Maxwell's curl equations in Fourier space with Kx, Ky diagonal, [eps] and [1/eps] convolution
    matrices (Laurent's rule); layer modes from the 2MF eigenproblem of P Q; the half-spaces carry
    plane waves parametrised by (Ex, Ey) with Ez from k.E = 0; continuity of Ex, Ey, Hx, Hy at every
    boundary.  Layer l spans thick[l]; boundary 0 sits at z = 0 where the incident wave has amplitude E0.
    """
    MF = M0 * M1
    px = np.arange(M0) - M0 // 2
    qy = np.arange(M1) - M1 // 2
    sx = np.tile(sdir[0] * n0 + 2 * o.PI_TRUNC * px / period[0] / k, M1)
    sy = np.repeat(sdir[1] * n0 + 2 * o.PI_TRUNC * qy / period[1] / k, M0)
    Kx, Ky = np.diag(sx).astype(complex), np.diag(sy).astype(complex)
    I = np.eye(MF)
    sz0 = np.sqrt((n0 ** 2 - sx ** 2 - sy ** 2).astype(complex))
    sz1 = np.sqrt((n1 ** 2 - sx ** 2 - sy ** 2).astype(complex))

    def wave(a, b, c, Ex, Ey):
        Ez = -(a * Ex + b * Ey) / c
        return Ez, b * Ez - c * Ey, c * Ex - a * Ez

    modes = []
    for eps in layers_yx:
        E = _wrapped_conv(np.fft.fft2(eps ** 2) / eps.size, M0, M1)
        Nn = _wrapped_conv(np.fft.fft2(1.0 / eps ** 2) / eps.size, M0, M1)
        P = np.block([[Kx @ Nn @ Ky, I - Kx @ Nn @ Kx], [Ky @ Nn @ Ky - I, -Ky @ Nn @ Kx]])
        Q = np.block([[-Kx @ Ky, Kx @ Kx - E], [E - Ky @ Ky, Ky @ Kx]])
        mu, w = np.linalg.eig(P @ Q)
        g = np.sqrt(mu)
        g = np.where((g.imag < 0) | ((np.abs(g.imag) < 1e-14) & (g.real < 0)), -g, g)
        h = (Q @ w) / g[None, :]
        modes.append((g, np.vstack([w, h]), np.vstack([w, -h])))
    L, M = len(layers_yx), MF
    A = np.zeros((4 * M * (L + 1), 4 * M + 4 * M * L), complex)
    rhs = np.zeros(4 * M * (L + 1), complex)
    c0 = (M1 // 2) * M0 + M0 // 2
    Ez0, Hx0, Hy0 = wave(sx[c0], sy[c0], sz0[c0], E0[0], E0[1])
    inc = np.zeros(4 * M, complex)
    inc[[c0, M + c0, 2 * M + c0, 3 * M + c0]] = [E0[0], E0[1], Hx0, Hy0]
    il = lambda l: 4 * M + 4 * M * l  # noqa: E731
    for bnd in range(L + 1):
        rows = slice(4 * M * bnd, 4 * M * (bnd + 1))
        if bnd == 0:
            rhs[rows] = -inc
        for j in range(M):
            for pol, (ex, ey) in enumerate(((1, 0), (0, 1))):
                if bnd == 0:
                    _, Hx, Hy = wave(sx[j], sy[j], -sz0[j], ex, ey)
                    A[4 * M * bnd + np.array([j, M + j, 2 * M + j, 3 * M + j]), pol * M + j] += [ex, ey, Hx, Hy]
                if bnd == L:
                    _, Hx, Hy = wave(sx[j], sy[j], sz1[j], ex, ey)
                    A[4 * M * bnd + np.array([j, M + j, 2 * M + j, 3 * M + j]), 2 * M + pol * M + j] -= [ex, ey, Hx, Hy]
        if bnd >= 1:
            g, vp, vm = modes[bnd - 1]
            A[rows, il(bnd - 1):il(bnd - 1) + 2 * M] += vp * np.exp(1j * k * g * thick[bnd - 1])[None, :]
            A[rows, il(bnd - 1) + 2 * M:il(bnd - 1) + 4 * M] += vm
        if bnd < L:
            g, vp, vm = modes[bnd]
            A[rows, il(bnd):il(bnd) + 2 * M] -= vp
            A[rows, il(bnd) + 2 * M:il(bnd) + 4 * M] -= vm * np.exp(1j * k * g * thick[bnd])[None, :]
    x = np.linalg.solve(A, rhs)
    R = np.array([x[:M], x[M:2 * M], np.zeros(M, complex)])
    T = np.array([x[2 * M:3 * M], x[3 * M:4 * M], np.zeros(M, complex)])
    R[2] = (sx * R[0] + sy * R[1]) / sz0
    T[2] = -(sx * T[0] + sy * T[1]) / sz1
    fi = sz0[c0].real * (abs(E0[0]) ** 2 + abs(E0[1]) ** 2 + abs(Ez0) ** 2)
    fr = np.sum(sz0.real * np.sum(np.abs(R) ** 2, axis=0))
    ft = np.sum(sz1.real * np.sum(np.abs(T) ** 2, axis=0))
    return R, T, (fi, fr, ft)


def _x_fastest(a, M0, M1):
    """Column-major (y-fastest) flat (M1 x M0) array -> x-fastest flat order."""
    return a.reshape(M0, M1).T.ravel()


def oracle_consistent_order(sample, p):
    """The oracle with Sx/Sy/Sz/EF flattened x-fastest like Phi's rows (undoing the mixed ordering)."""
    orig = o.setup

    def setup_x_fastest(smp, prm):
        s = orig(smp, prm)
        for name in ("Sx", "Sy", "Sz0", "Sz1"):
            setattr(s, name, _x_fastest(getattr(s, name), s.M0, s.M1).copy())
        s.EF = np.concatenate([_x_fastest(s.EF[c * s.MF:(c + 1) * s.MF], s.M0, s.M1) for c in range(3)])
        return s

    o.setup = setup_x_fastest
    try:
        return o.solve(sample, p)
    finally:
        o.setup = orig


def stack_case(seed, L, M0, M1, lossy, sdir, pol, N=16, lam=0.73, period=(2.0, 2.0), d=0.4):
    rng = np.random.default_rng(seed)
    layers = []
    sample = np.zeros((L, N, N), complex)
    yy, xx = np.mgrid[0:N, 0:N]
    for l in range(L):
        eps = np.full((N, N), 1.0 + 0j)
        for _ in range(3):
            cx, cy, r = rng.uniform(0, N), rng.uniform(0, N), rng.uniform(1.5, 4)
            eps[((xx - cx + N / 2) % N - N / 2) ** 2 + (((yy - cy + N / 2) % N - N / 2) ** 2) * (M1 > 1) < r * r] = \
                rng.uniform(1.2, 1.6) + (0.03j if lossy else 0)
        if M1 == 1:
            eps = np.repeat(eps[:1], N, axis=0)     # y-invariant grating
        layers.append(eps)                           # eps[iy, ix]
        sample[l] = eps.T                            # scattervolume reads each slice transposed (reformat)
    sdir = np.asarray(sdir, float) / np.linalg.norm(sdir)
    p = o.VolumeParams(lam=lam, size=(period[0], period[1], d * L), coef=(M0, M1), direction=tuple(sdir),
                       ex=(pol[0], 0), ey=(pol[1], 0), n=(1.0, 1.0), z=d * L / 2, no_cutoff=True)
    return sample, p, layers, [d] * L, sdir


CASES = [
    (0, 1, 7, 1, False, (0, 0, 1), (0, 1)),        # grating, TE, normal incidence
    (1, 2, 9, 1, False, (0, 0, 1), (1, 0)),        # grating, TM
    (2, 3, 5, 1, True, (0.3, 0, 1), (1, 1)),       # grating, oblique, mixed polarisation, lossy
    (3, 2, 7, 1, False, (0.2, 0.25, 1), (1, 0.3)),  # grating, conical incidence couples TE and TM
    (4, 2, 5, 5, False, (0, 0, 1), (1, 0)),        # 2-D crossed structure
    (5, 2, 5, 3, True, (0.1, 0.2, 1), (0.4, 1)),   # 2-D, M0 != M1, oblique, lossy
]


@pytest.mark.parametrize("seed,L,M0,M1,lossy,sdir,pol", CASES)
def test_independent_maxwell_formulation(seed, L, M0, M1, lossy, sdir, pol):
    """M1 = 1: the oracle itself.  M1 > 1: the oracle with the orders of Sx/Sy/Sz/EF flattened
    x-fastest (the reference pairs Phi's x-fastest rows with y-fastest Sx/Sy, ScatterVolume.cpp:682-685;
    everything else is pinned)."""
    sample, p, layers, thick, sdir = stack_case(seed, L, M0, M1, lossy, sdir, pol)
    sol = o.solve(sample, p) if M1 == 1 else oracle_consistent_order(sample, p)
    s = sol.setup
    assert (s.M0, s.M1) == (M0, M1) and s.z[0] == 0.0
    R, T, flux = independent_rcwa(layers, p.size[:2], thick, s.k, sdir, s.e, 1.0, 1.0, M0, M1)
    MF = M0 * M1
    assert rel_err(sol.x[:3 * MF].reshape(3, MF), R) < 1e-10
    assert rel_err(sol.x[3 * MF:].reshape(3, MF), T) < 1e-10
    if not lossy:
        assert abs(flux[1] + flux[2] - flux[0]) < 1e-10 * flux[0]


@pytest.mark.parametrize("M", [3, 5])
def test_energy_balance_lossless_2d(M):
    """Lossless stack with odd orders (a Hermitian truncation): reflected + transmitted = incident,
    with the reference's own order pairing."""
    sample, p, *_ = stack_case(11, 2, M, M, False, (0.1, 0.05, 1), (1, 0.2))
    sol = o.solve(sample, p)
    fi, fr, ft = o.far_field_flux(sol)
    assert abs(fr + ft - fi) < 1e-10 * fi


def test_boundary_system_residual_and_gauss_constraints():
    rng = np.random.default_rng(4)
    sample = 1.0 + 0.4 * rng.random((3, 8, 8)) + 0.01j * rng.random((3, 8, 8))
    p = o.VolumeParams(lam=0.9, size=(2.0, 2.0, 0.9), coef=(3, 3), direction=(0.1, 0, 1), n=(1.0, 1.3),
                       kappa=(0.02,))
    sol = o.solve(sample, p)
    r = np.linalg.norm(sol.A @ sol.x - sol.b) / (np.linalg.norm(sol.A) * np.linalg.norm(sol.x))
    assert r < 1e-14
    s, MF = sol.setup, sol.setup.MF
    R = sol.x[:3 * MF].reshape(3, MF)
    T = sol.x[3 * MF:].reshape(3, MF)
    assert np.max(np.abs(s.Sx * R[0] + s.Sy * R[1] - s.Sz0 * R[2])) < 1e-13
    assert np.max(np.abs(s.Sx * T[0] + s.Sy * T[1] + s.Sz1 * T[2])) < 1e-13


def test_orthogonalize():
    E = o.orthogonalize([0, 1, 0], [0, 0, 1])
    assert np.allclose(E, [0, 1, 0])
    d = np.array([0.6, 0, 0.8])
    E = o.orthogonalize([1, 0, 0], d)
    assert abs(np.dot(E, d)) < 1e-15
    # normalisation without conjugation: (1, 1j, 0) has zero "length"
    with np.errstate(all="ignore"):
        E = o.orthogonalize([1, 1j, 0], [0, 0, 1])
    assert not np.all(np.isfinite(E))


# ----------------------------------------------------------------------------------------------
def test_cw_roundtrip_and_legacy_reader():
    sample = golden("samples.npz")["npy3"]
    p = o.VolumeParams(lam=0.6, size=(4.0, 4.0, 1.0), coef=(3, 3))
    sol = o.solve(sample, p)
    cw = o.build_cw(sol)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "x.cw")
        o.cw_write(cw, path)
        back = o.cw_read(path)
    assert np.array_equal(back.Pi, cw.Pi)
    for a, b in zip(back.Pr + back.Pt, cw.Pr + cw.Pt):
        assert np.array_equal(a, b)
    assert back.z == cw.z
    nrec = len(cw.Pi) + sum(len(a) for a in cw.Pr + cw.Pt)
    MF = sol.setup.MF
    assert nrec == MF * (3 + 4 * MF * sol.setup.L)


def test_reduced_eigen_form_matches_full_phi():
    """The oracle's reduced-form eigen step (P Q instead of the 4MF Phi, used to time the CPU side of the
    reduced mode) gives the full form's boundary solution and fields on a three-layer oblique stack."""
    from tests.helpers import disks_sample, fields, probe_points, rel_err
    s = disks_sample(16, 3, 4.0, 4, 1.0, lambda r: 1.4 + 0.02j * r.random(), seed=5)
    p = o.VolumeParams(lam=0.6, size=(4.0, 4.0, 1.0), coef=(5, 5), direction=(0.2, 0.1, 1.0))
    a = o.solve(s, p)
    b = o.solve(s, p, eig=o.eig_layers_reduced)
    assert rel_err(b.x, a.x) < 1e-12
    pts = probe_points(a.setup.z, p.size)
    assert rel_err(fields(o.build_cw(b), pts), fields(o.build_cw(a), pts)) < 1e-11
