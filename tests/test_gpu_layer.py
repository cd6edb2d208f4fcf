"""GPU parity of the scatterlayer solve (cwx_layer_solve, ScatterLayer.cpp:110-420) against the FP64 layer
oracle and x64/a.cw.

Stacks of up to 10 layers run the one-wavefront-per-wave kernel, taller ones the batched-LU path; both are
compared with the oracle on random stacks and mixed polarisations (1e-10 relative: the reference's QR, the
oracle's LU and the GPU's LU agree to rounding on these well-conditioned systems).  The a.cw TE incidences
and the scatterlayer tool's default beam are compared with the reference's own output file.
"""
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

import scattersim_amd as sa
from oracle import cw_oracle as o
from oracle import layer_oracle as lo
from tests.helpers import golden, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "scattersim_amd", "bin")
TOL = 1e-10


def random_waves(rng, P, n0, lam, spread=1.2):
    k = 2 * math.pi / lam * n0
    th = spread * rng.random(P)
    ph = 2 * math.pi * rng.random(P)
    kv = k * np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    E = rng.standard_normal((P, 3)) + 1j * rng.standard_normal((P, 3))
    E -= kv * (np.sum(kv * E, 1) / np.sum(kv * kv, 1))[:, None]
    return np.concatenate([E, kv.astype(complex)], 1)


@pytest.mark.parametrize("L", [2, 3, 4, 7, 10, 12, 20])
def test_layer_solve_matches_oracle(ctx, L):
    rng = np.random.default_rng(L)
    n = (1.0 + 0.5 * rng.random(),) + tuple(1.1 + rng.random(L - 1))
    kappa = tuple(0.05 * rng.random(L - 1))
    z = tuple(np.cumsum(0.1 + 0.5 * rng.random(L - 1)) - 0.7)
    lam = 0.5 + rng.random()
    inc = random_waves(rng, 37, n[0], lam)
    R, T = sa.scatterlayer(ctx, inc, n=n, kappa=kappa, z=z, lam=lam)
    Ro, To = lo.solve(lo.LayerStack(n=n, kappa=kappa, z=z, lam=lam), inc)
    assert R.shape == (37, L - 1, 6)
    assert rel_err(R, Ro) < TOL and rel_err(T, To) < TOL


def test_layer_solve_acw_te_known_answer(ctx):
    g = golden("acw_te.npz")
    R, T = sa.scatterlayer(ctx, g["Pi"], n=(1.0, 1.4, 1.4, 1.0), kappa=(0.05,), z=(-3.0, 0.0, 3.0), lam=1.0)
    drop = lambda a: np.concatenate([a[..., 0:2], a[..., 3:6]], -1)
    assert rel_err(drop(R), drop(g["Pr"])) < 1e-12
    assert rel_err(drop(T), drop(g["Pt"])) < 1e-12


def test_layer_solve_many_waves_and_empty(ctx):
    rng = np.random.default_rng(7)
    inc = random_waves(rng, 20000, 1.0, 0.9)
    R, T = sa.scatterlayer(ctx, inc, n=(1.0, 1.5, 1.2), kappa=(0.01, 0.0), z=(0.0, 0.8), lam=0.9)
    pick = rng.choice(20000, 40, replace=False)
    Ro, To = lo.solve(lo.LayerStack(n=(1.0, 1.5, 1.2), kappa=(0.01, 0.0), z=(0.0, 0.8), lam=0.9), inc[pick])
    assert rel_err(R[pick], Ro) < TOL and rel_err(T[pick], To) < TOL
    R0, T0 = sa.scatterlayer(ctx, np.zeros((0, 6), complex), n=(1.0, 1.5), kappa=(0.0,), z=(0.0,))
    assert R0.shape == (0, 1, 6)


def test_cli_default_beam_against_acw(tmp_path):
    """scatterlayer with a.cw's stack and the default 64 x 64 polar beam: the incident list equals a.cw's,
    its TE incidences give a.cw's records, and every incidence agrees with the oracle."""
    r = subprocess.run([os.path.join(BIN, "scatterlayer"), "--z", "-3", "0", "3", "--output", str(tmp_path / "a.cw"),
                        "--output_sample", str(tmp_path / "s.npy")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "samples:" in r.stdout and "64 x 64 = 4096" in r.stdout
    cw = o.cw_read(str(tmp_path / "a.cw"))
    beam = golden("acw_beam.npz")["Pi"]
    mine = np.stack([cw.Pi[:, 0].real, cw.Pi[:, 1].real, cw.Pi[:, 3].real, cw.Pi[:, 4].real, cw.Pi[:, 5].real], 1)
    assert np.max(np.abs(mine - beam)) < 1e-14
    assert list(cw.z) == [-3.0, 0.0, 3.0]
    g = golden("acw_te.npz")
    drop = lambda a: np.concatenate([a[..., 0:2], a[..., 3:6]], -1)
    for l in range(3):
        assert rel_err(drop(cw.Pr[l][g["index"]]), drop(g["Pr"][:, l])) < 1e-12
        assert rel_err(drop(cw.Pt[l][g["index"]]), drop(g["Pt"][:, l])) < 1e-12
    pick = np.arange(0, 4096, 97)
    Ro, To = lo.solve(lo.LayerStack(n=(1.0, 1.4, 1.4, 1.0), kappa=(0.05,), z=(-3.0, 0.0, 3.0), lam=1.0), cw.Pi[pick])
    for l in range(3):
        assert rel_err(cw.Pr[l][pick], Ro[:, l]) < TOL and rel_err(cw.Pt[l][pick], To[:, l]) < TOL
    # the sample file holds the layer indices repeated once per incident wave (ScatterLayer.cpp:118-125, 426-434)
    s = np.load(tmp_path / "s.npy")
    assert s.shape == (4 * 4096 - 2, 1, 1) and s[0, 0, 0] == 1.4 + 0.05j and s[1, 0, 0] == 1.4


def test_cli_single_wave_renders(tmp_path):
    """alpha = 0: one plane wave; scatterview renders the file and matches the oracle's field."""
    r = subprocess.run([os.path.join(BIN, "scatterlayer"), "--alpha", "0", "--direction", "0.2", "0", "1",
                        "--ex", "1", "0", "--ey", "0", "0", "--n", "1", "1.5", "1.2", "--kappa", "0.02", "0",
                        "--z", "-0.5", "0.5", "--lambda", "0.7", "--output", str(tmp_path / "b.cw"),
                        "--output_sample", str(tmp_path / "s.npy")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    s = np.load(tmp_path / "s.npy")
    assert s.shape == (1, 1, 1) and s[0, 0, 0] == 1.5 + 0.02j
    r = subprocess.run([os.path.join(BIN, "scatterview"), "--nogui", "--input", str(tmp_path / "b.cw"), "--size", "3",
                        "--resolution", "5", "--output", str(tmp_path / "v.npy")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 1, r.stderr
    v = np.load(tmp_path / "v.npy")
    cw = o.cw_read(str(tmp_path / "b.cw"))
    W, zl, b, e = o.unpack(cw, np.complex128)
    pts = o.slice_points(1, 32, 3.0, (0, 0, 0), (0, 0, 0))
    ref = o.field_points(W, zl, b, e, pts, tol=0.001).reshape(32, 32, 3)
    assert rel_err(v, ref) < 1e-5


def test_layer_solve_psf_te_known_answer(ctx):
    """One interface (L = 2): the TE waves of data/psf.cw, scatterplane's default focus."""
    g = golden("psf.npz")
    Pi = g["Pi"][g["index"]]
    inc = np.zeros((len(Pi), 6), np.complex128)
    inc[:, 0:2] = Pi[:, 0:2]
    inc[:, 3:6] = Pi[:, 2:5]
    R, T = sa.scatterlayer(ctx, inc, n=(1.0, 1.4), kappa=(0.05,), z=(0.0,), lam=1.0)
    drop = lambda a: np.concatenate([a[..., 0:2], a[..., 3:6]], -1)
    assert rel_err(drop(R[:, 0]), drop(g["Pr"])) < 1e-12
    assert rel_err(drop(T[:, 0]), drop(g["Pt"])) < 1e-12
