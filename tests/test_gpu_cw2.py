"""Compact heterogeneous .cw (docs/cw_file_format.pptx slide 1, read by ScatterView.py:68-178).

The file is parsed here in the order of ScatterView.py's ``coupledwave.load`` and its content is checked
against what the format means, not against the writer's own loops:

- the external lists are the solve's incident, reflected and transmitted records (bitwise);
- the mode data give back the field inside every layer: the per-order coefficients of Ex, Ey and the two
  magnetic blocks from beta, gamma and gg as cpuEvaluator.cpp:171-205 forms them, Ez from ft as
  cpuEvaluator.cpp:215-231 does, summed over the orders against the plane-wave sum of the full .cw records
  (FP64, 1e-9);
- ft equals the oracle's n^-2 coefficients, and for a square sample solved at full order ScatterView.py's
  ``getVolume`` (171-178) turns it back into the sample;
- a solution rebuilt from the file holds the same boundary solution, mode matrices and records, and renders
  the same field grid.
"""
import os
import struct
import tempfile

import numpy as np
import pytest

import scattersim_amd as sa
from oracle import cw_oracle as o
from tests.helpers import cw_from_records, disks_sample, fields, probe_points, rel_err

pytestmark = pytest.mark.gpu


def read_compact(path):
    """The layout ScatterView.py:76-164 reads; whatever follows the layers is returned as ``tail``."""
    buf = open(path, "rb").read()
    off = 0

    def take(fmt):
        nonlocal off
        v = struct.unpack_from("<" + fmt, buf, off)
        off += struct.calcsize("<" + fmt)
        return v

    def cplx(n):
        nonlocal off
        a = np.frombuffer(buf, np.complex128, n, off).copy()
        off += 16 * n
        return a

    def waves():
        (n,) = take("Q")
        return cplx(6 * n).reshape(n, 6)

    (is_volume,) = take("?")
    (precision,) = take("Q")
    Pi = waves()
    (nb,) = take("Q")
    z, R, T = [], [], []
    for _ in range(nb):
        z.append(take("d")[0])
        R.append(waves())
        T.append(waves())
    out = dict(is_volume=is_volume, precision=precision, Pi=Pi, z=z, R=R, T=T)
    if is_volume:
        M0, M1 = take("ii")
        D = take("ddd")
        (L,) = take("I")
        n4 = 4 * M0 * M1
        layers = []
        for _ in range(L):
            beta, gamma = cplx(n4), cplx(n4)
            gg = cplx(n4 * n4).reshape(n4, n4)
            ft = cplx(M0 * M1)
            layers.append(dict(beta=beta, gamma=gamma, gg=gg, ft=ft))
        out.update(M=(M0, M1), D=D, layers=layers)
    out["tail"] = buf[off:]
    out["tail_offset"] = off
    return out


def layer_field(f, l, setup, size, pts):
    """Field at points inside layer l from the compact file's mode data (cpuEvaluator.cpp:150-231).

    Down modes (even) carry their phase from the layer's top boundary, up modes (odd) from its bottom one.
    Component n of a block is the order at (n % M1, n // M1) of the (M1 x M0) order grid, as the inverse
    transform of cpuEvaluator.cpp:241-243 maps it; its transverse wave vector is the records' (Sx, Sy) k.
    """
    M0, M1 = f["M"]
    MF = M0 * M1
    k = setup.k
    lay = f["layers"][l]
    zt, zb = f["z"][l], f["z"][l + 1]
    even = (np.arange(4 * MF) % 2) == 0
    UP = 2 * np.pi * (np.arange(M0) - M0 // 2) / size[0] + setup.dir_n[0] * k
    WQ = 2 * np.pi * (np.arange(M1) - M1 // 2) / size[1] + setup.dir_n[1] * k
    qi, pi = np.divmod(np.arange(MF), M0)                      # Ez row qi*M0 + pi, column qj*M0 + pj
    R = (((qi[:, None] - qi[None, :]) % M1) + M1 // 2) % M1
    Cc = (((pi[:, None] - pi[None, :]) % M0) + M0 // 2) % M0
    Cn = lay["ft"][R * M0 + Cc]
    upc = UP[pi][None, :]
    wqc = WQ[qi][None, :]
    kx, ky = (setup.Sx * k).real, (setup.Sy * k).real
    out = []
    for x, y, zz in pts:
        ph = np.where(even, np.exp(1j * k * lay["gamma"] * (zz - zt)), np.exp(1j * k * lay["gamma"] * (zz - zb)))
        c = lay["gg"] @ (lay["beta"] * ph)
        I, J = c[2 * MF:3 * MF], c[3 * MF:]
        ez = (-1.0 / k) * (Cn @ (upc[0] * J - wqc[0] * I))
        e = np.exp(1j * (kx * x + ky * y))
        out.append([c[:MF] @ e, c[MF:2 * MF] @ e, ez @ e])
    return np.array(out)


CASES = {
    # orders differ along x and y, oblique incidence, two layers of an absorbing non-square grid
    "oblique_5x3": (lambda: disks_sample(12, 2, 3.0, 3, 1.0, 1.4 + 0.01j, seed=3)[:, :, :10],
                    o.VolumeParams(lam=0.8, direction=(0.2, -0.1, 1.0), ex=(1, 0), ey=(0.3, 0.1),
                                   size=(3.0, 2.5, 1.0), coef=(5, 3), n=(1.0, 1.2))),
    "three_layers_5x5": (lambda: disks_sample(10, 3, 3.0, 3, 1.0, lambda r: 1.45 + 0.02j * r.random(), seed=8),
                         o.VolumeParams(lam=0.7, direction=(0.1, 0.25, 1.0), ex=(1, 0.5), ey=(0.2, 0),
                                        size=(3.0, 3.0, 1.2), coef=(5, 5), n=(1.0, 1.33))),
}


def solve(ctx, sample, p, eig_mode=0):
    return sa.scattervolume(ctx, sample, lam=p.lam, direction=p.direction, ex=p.ex, ey=p.ey, ez=p.ez, n=p.n,
                            kappa=p.kappa, size=p.size, z=p.z, coef=p.coef, psf=p.psf, external=p.external,
                            npz=p.npz, no_cutoff=p.no_cutoff, eig_mode=eig_mode)


@pytest.mark.parametrize("case", sorted(CASES))
def test_compact_file_layout_and_mode_fields(ctx, case):
    make, p = CASES[case]
    sample = make()
    sol = solve(ctx, sample, p)
    ref = o.solve(sample, p)
    s = ref.setup
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "c.cw")
        sol.write_cw2(path)
        f = read_compact(path)
    L, MF = sol.L, sol.MF
    assert f["is_volume"] is True and f["precision"] == 8
    assert f["M"] == (sol.M0, sol.M1) and tuple(f["D"]) == tuple(p.size) and len(f["layers"]) == L
    # external lists bitwise, interior boundaries empty
    rec = sol.records
    off, cnt = sol.lists()
    seg = lambda i: rec[off[i]:off[i] + cnt[i]]
    assert np.array_equal(f["Pi"], seg(0))
    assert np.array_equal(np.array(f["z"]), sol.z)
    assert np.array_equal(f["R"][0], seg(1)) and len(f["T"][0]) == 0
    for b in range(1, L):
        assert len(f["R"][b]) == 0 and len(f["T"][b]) == 0
    assert len(f["R"][L]) == 0 and np.array_equal(f["T"][L], seg(2 + 2 * L))
    assert len(f["R"][0]) == MF and len(f["T"][L]) == MF
    # per-layer data: beta and gamma as the solution holds them, ft against the oracle
    G, B = sol.gammas, sol.betas
    for l in range(L):
        assert np.array_equal(f["layers"][l]["gamma"], G[l])
        assert np.array_equal(f["layers"][l]["beta"], B[l])
        assert rel_err(f["layers"][l]["ft"], ref.Nifs[l].ravel(order="F")) < 1e-12
    # the mode data reproduce the field of the full record set inside every layer
    full = cw_from_records(rec, cnt, sol.z)
    pts = probe_points(sol.z, p.size, seed=2)
    Er = fields(full, pts)
    for l in range(L):
        sel = (pts[:, 2] > sol.z[l]) & (pts[:, 2] < sol.z[l + 1])
        assert sel.sum() >= 2
        Em = layer_field(f, l, s, p.size, pts[sel])
        assert rel_err(Em, Er[sel]) < 1e-9, (l, rel_err(Em, Er[sel]))
    # the trailer follows the layers; the reference reader never looks at it
    assert f["tail"][:8] == b"CWXTRL01"


def test_compact_mode_columns_normalised_like_zgeev_in_both_modes(ctx):
    """gg's columns are the solution's mode vectors.  zgeev (third_Lapack.h:38-45) returns each eigenvector
    of Phi with unit 2-norm and its largest component real and positive; the reduced mode forms
    v+- = [a; +-Q a / gamma] from P Q's eigenvectors and must rescale each to the same convention.  For a
    simple eigenvalue that fixes the vector, so the reduced mode's columns equal the full mode's."""
    make, p = CASES["three_layers_5x5"]
    sample = make()
    gg = {}
    for mode in (0, 1):
        sol = solve(ctx, sample, p, mode)
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "c.cw")
            sol.write_cw2(path)
            f = read_compact(path)
        gg[mode] = [(lay["gg"], lay["gamma"]) for lay in f["layers"]]
        sol.free()
    worst_norm = worst_imag = worst_diff = 0.0
    compared = 0
    for (g0, gam0), (g1, gam1) in zip(gg[0], gg[1]):
        assert np.allclose(gam0, gam1, rtol=1e-10, atol=1e-12)
        for g in (g0, g1):
            worst_norm = max(worst_norm, float(np.max(np.abs(np.linalg.norm(g, axis=0) - 1.0))))
            big = g[np.argmax(np.abs(g), axis=0), np.arange(g.shape[1])]
            assert np.all(big.real > 0)
            worst_imag = max(worst_imag, float(np.max(np.abs(big.imag))))
        gap = np.abs(gam0[:, None] - gam0[None, :]) + np.eye(len(gam0)) * 1e9
        simple = np.min(gap, axis=1) > 1e-6 * np.max(np.abs(gam0))
        assert simple.sum() > len(gam0) // 2
        worst_diff = max(worst_diff, float(np.max(np.abs(g1[:, simple] - g0[:, simple]))))
        compared += int(simple.sum())
    print(f"|norm - 1| {worst_norm:.1e}, imag of largest {worst_imag:.1e}, reduced vs full on {compared} simple "
          f"modes {worst_diff:.1e}", flush=True)
    assert worst_norm < 1e-12 and worst_imag < 1e-12 and worst_diff < 1e-8


def test_compact_file_ft_inverts_to_sample(ctx):
    """ScatterView.py:171-178 (getVolume) on a square sample solved at full order gives the sample back."""
    rng = np.random.default_rng(11)
    sample = 1.3 + 0.2 * rng.random((2, 7, 7)) + 0.01j * rng.random((2, 7, 7))
    sol = sa.scattervolume(ctx, sample, lam=0.9, size=(2.0, 2.0, 0.6), coef=(7, 7), no_cutoff=True,
                           ex=(1, 0), ey=(0, 0))
    assert (sol.M0, sol.M1) == (7, 7)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "c.cw")
        sol.write_cw2(path)
        f = read_compact(path)
    M0, M1 = f["M"]
    for l, lay in enumerate(f["layers"]):
        img = np.fft.ifft2(np.fft.ifftshift(lay["ft"].reshape(M1, M0))) * M0 * M1
        assert rel_err(1.0 / np.sqrt(img), sample[l]) < 1e-12


@pytest.mark.parametrize("npz", [False, True])
def test_compact_file_read_back(ctx, npz):
    make, p = CASES["oblique_5x3"]
    sample = make()
    sol = sa.scattervolume(ctx, sample, lam=p.lam, direction=p.direction, ex=p.ex, ey=p.ey, size=p.size,
                           coef=p.coef, n=p.n, npz=npz)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "c.cw")
        sol.write_cw2(path)
        back = sa.read_cw2(ctx, path)
    assert (back.M0, back.M1, back.L, back.nrec, back.cutoff) == (sol.M0, sol.M1, sol.L, sol.nrec, sol.cutoff)
    assert np.array_equal(back.x, sol.x)
    assert np.array_equal(back.records, sol.records)
    from scattersim_amd import _capi
    for what in (_capi.SOL_GD, _capi.SOL_GC, _capi.SOL_NIF, _capi.SOL_EF):
        assert np.array_equal(back.setup_array(what), sol.setup_array(what)), what
    assert np.array_equal(back.Q_check_hat, sol.Q_check_hat)
    g = (-0.3, 0.1, -1.2, 0.11, 0.13, 0.1, 27, 19, 25)
    a = sa.DeviceArray(ctx, (g[8], g[7], g[6], 3), np.complex128)
    b = sa.DeviceArray(ctx, (g[8], g[7], g[6], 3), np.complex128)
    ctx.field_solution_grid_dev(sol, a.ptr, True, 0, *g)
    ctx.field_solution_grid_dev(back, b.ptr, True, 0, *g)
    ctx.synchronize()
    assert np.array_equal(a.numpy(), b.numpy())


def test_compact_file_without_trailer_is_refused(ctx):
    """A file in the bare reference layout lacks the wavelength and incident field: reading it fails cleanly."""
    make, p = CASES["oblique_5x3"]
    sol = solve(ctx, make(), p)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "c.cw")
        sol.write_cw2(path)
        f = read_compact(path)
        bare = os.path.join(d, "bare.cw")
        with open(path, "rb") as src, open(bare, "wb") as dst:
            dst.write(src.read()[:f["tail_offset"]])
        with pytest.raises(sa._capi.CwxError) as e:
            sa.read_cw2(ctx, bare)
        assert e.value.code == sa._capi.CWX_ERR_IO
        cut = os.path.join(d, "cut.cw")
        with open(path, "rb") as src, open(cut, "wb") as dst:
            dst.write(src.read()[:f["tail_offset"] // 2])
        with pytest.raises(sa._capi.CwxError) as e:
            sa.read_cw2(ctx, cut)
        assert e.value.code == sa._capi.CWX_ERR_IO
        full = os.path.join(d, "full.cw")
        sol.write_cw(full)
        with pytest.raises(sa._capi.CwxError):
            sa.read_cw2(ctx, full)
    # an --external solve keeps no interior amplitudes: nothing compact to write
    ext = sa.scattervolume(ctx, make(), lam=p.lam, size=p.size, coef=p.coef, external=True)
    with tempfile.TemporaryDirectory() as d:
        with pytest.raises(sa._capi.CwxError) as e:
            ext.write_cw2(os.path.join(d, "e.cw"))
        assert e.value.code == sa._capi.CWX_ERR_ARG


def test_cli_compact_file_renders_like_full_file(tmp_path):
    """scattervolume --compact, then scatterview / scatterviewsample --nogui on it, against the same slices of
    the full file (ScatterView.cpp:354-376, ScatterViewSample.cpp:354-376 slice geometry)."""
    import subprocess
    bin_dir = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "scattersim_amd", "bin")

    def run(tool, *args):
        r = subprocess.run([os.path.join(bin_dir, tool), *map(str, args)], capture_output=True, text=True,
                           timeout=120)
        return r

    sample = disks_sample(12, 2, 3.0, 3, 1.0, 1.4 + 0.01j, seed=21)
    np.save(tmp_path / "s.npy", sample)
    common = ["--sample", tmp_path / "s.npy", "--lambda", 0.7, "--size", 3, 3, 1, "--coef", 5, 5,
              "--direction", 0.1, 0, 1]
    r = run("scattervolume", *common, "--output", tmp_path / "full.cw")
    assert r.returncode == 0, r.stderr
    r = run("scattervolume", *common, "--compact", "--output", tmp_path / "compact.cw")
    assert r.returncode == 0, r.stderr
    assert os.path.getsize(tmp_path / "compact.cw") < os.path.getsize(tmp_path / "full.cw")
    view = ["--nogui", "--size", 4, "--resolution", 5, "--axis", 1, "--center", 0.3, 0.2, 0.017]
    r = run("scatterview", *view, "--input", tmp_path / "full.cw", "--output", tmp_path / "full.npy")
    assert r.returncode == 1 and "takes" in r.stdout, r.stderr
    r = run("scatterview", *view, "--input", tmp_path / "compact.cw", "--output", tmp_path / "compact.npy")
    assert r.returncode == 1 and "takes" in r.stdout, r.stderr
    a, b = np.load(tmp_path / "compact.npy"), np.load(tmp_path / "full.npy")
    assert a.shape == b.shape == (32, 32, 3)
    # the full file is evaluated from FP32 waves (cw_unpack), the compact one from the FP64 records
    assert rel_err(a, b) < 1e-4
    # scatterviewsample: the slices of all three axes from the compact file through the separable evaluator,
    # and from the full file through the wave table, against scatterview on the full file
    for axis in (0, 1, 2):
        geo = ["--nogui", "--resolution", 5, "--axis", axis, "--slice", 0.21, "--center", 0.3, 0.2, 0.017]
        r = run("scatterview", *geo, "--size", 4, "--input", tmp_path / "full.cw", "--output", tmp_path / "v.npy")
        assert r.returncode == 1, r.stderr
        r = run("scatterviewsample", *geo, "--extent", 4, "--input", tmp_path / "compact.cw", "--output",
                tmp_path / "sc.npy")
        assert r.returncode == 1 and "Visualization working." in r.stdout, r.stderr
        r = run("scatterviewsample", *geo, "--extent", 4, "--input", tmp_path / "full.cw", "--output",
                tmp_path / "sf.npy")
        assert r.returncode == 1, r.stderr
        v, sc, sf = (np.load(tmp_path / n) for n in ("v.npy", "sc.npy", "sf.npy"))
        assert sc.shape == (32, 32, 3) and sc.dtype == np.complex64
        assert np.array_equal(sf, v)
        assert rel_err(sc, v) < 1e-4, (axis, rel_err(sc, v))


def test_cli_scatterviewsample_slices_match_oracle(tmp_path):
    """scatterviewsample --nogui on a compact file: the slices of all three axes, inside and outside the
    sample, against the oracle's FP64 sum over its own .cw records at the same points (ScatterViewSample.cpp
    :354-376 geometry, gpuEvaluator.cu:34 region rule).  complex64 output: 2e-6 of the slice maximum."""
    import subprocess
    bin_dir = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "scattersim_amd", "bin")
    sample = disks_sample(16, 3, 3.0, 3, 1.0, lambda r: 1.4 + 0.02j * r.random(), seed=33)
    np.save(tmp_path / "s.npy", sample)
    p = o.VolumeParams(lam=0.7, size=(3.0, 3.0, 1.2), coef=(5, 5), direction=(0.1, 0.0, 1.0))
    r = subprocess.run([os.path.join(bin_dir, "scattervolume"), "--sample", str(tmp_path / "s.npy"), "--lambda", "0.7",
                        "--size", "3", "3", "1.2", "--coef", "5", "5", "--direction", "0.1", "0", "1", "--compact",
                        "--output", str(tmp_path / "c.cw")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ref = o.build_cw(o.solve(sample, p))
    W, zl, b, e = o.unpack(ref, np.complex128)
    N, extent, center, sl = 32, 4.0, (0.3, 0.2, 0.017), 0.21
    for axis in (0, 1, 2):
        r = subprocess.run([os.path.join(bin_dir, "scatterviewsample"), "--nogui", "--resolution", "5", "--axis",
                            str(axis), "--slice", str(sl), "--center", *map(str, center), "--extent", str(extent),
                            "--input", str(tmp_path / "c.cw"), "--output", str(tmp_path / "sc.npy")],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 1, r.stderr
        got = np.load(tmp_path / "sc.npy").reshape(-1, 3)
        plane = [0.0, 0.0, 0.0]
        plane[axis] = sl
        pts = o.slice_points(axis, N, extent, center, plane)
        E = o.field_points(W, zl, b, e, pts, tol=0.001)
        err = rel_err(got, E)
        print(f"axis {axis}: {err:.2e}", flush=True)
        assert err < 2e-6
