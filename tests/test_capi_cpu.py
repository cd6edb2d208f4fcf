"""CPU-side checks of the drop-in boundary: library symbols, error paths without a GPU, the .cw reader,
the command-line tools' option handling, and the oracle's C field restatement."""
import ctypes as C
import os
import re
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import scattersim_amd as sa
from scattersim_amd import _capi
from oracle import cw_oracle as o
from tests.helpers import golden, rel_err

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "scattersim_amd", "bin")


def declared_functions():
    text = open(os.path.join(ROOT, "include", "cwx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cwx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = sa.lib()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    # and the Python binding declares each of them
    assert set(names) == set(_capi.SIGNATURES)


def test_params_struct_layout_matches_header():
    p = _capi.cwx_volume_params()
    sa.lib().cwx_volume_params_default(C.byref(p))
    assert p.lambda_ == 1.0 and list(p.direction) == [0, 0, 1] and list(p.ey) == [1, 0]
    assert list(p.n) == [1, 1] and list(p.size) == [10, 10, 2] and p.ncoef == 0 and p.eig_mode == 0
    assert C.sizeof(_capi.cwx_volume_params) == 8 * (1 + 3 + 2 * 4 + 1 + 3 + 1) + 4 * 8


def test_errors_without_device_or_with_bad_arguments():
    lib = sa.lib()
    h = C.c_void_p()
    st = lib.cwx_create(0, C.byref(h))
    if st == _capi.CWX_OK:
        lib.cwx_destroy(h)
        pytest.skip("a GPU is present")
    assert st == _capi.CWX_ERR_DEVICE
    assert b"device" in lib.cwx_last_error()
    assert lib.cwx_synchronize(None) == _capi.CWX_ERR_ARG
    assert lib.cwx_zgeev(None, 3, None, 3, None, None, 3) == _capi.CWX_ERR_ARG
    with pytest.raises(RuntimeError):
        sa.Context(0)


def test_cw_reader_on_reference_file_layout():
    """The C reader parses legacy 80-byte records (the reference's own a.cw layout) and ours."""
    g = golden("acw_te.npz")
    Pi, Pr, Pt, z = g["Pi"], g["Pr"], g["Pt"], g["z"]
    with tempfile.TemporaryDirectory() as d:
        legacy = os.path.join(d, "legacy.cw")
        with open(legacy, "wb") as f:
            f.write(np.uint64(8).tobytes() + np.uint64(len(Pi)).tobytes())
            f.write(np.ascontiguousarray(Pi[:, [0, 1, 3, 4, 5]]).tobytes())
            f.write(np.uint64(3).tobytes())
            for l in range(3):
                f.write(np.float64(z[l]).tobytes())
                for lst in (Pr[:, l], Pt[:, l]):
                    f.write(np.uint64(len(lst)).tobytes() + np.ascontiguousarray(lst[:, [0, 1, 3, 4, 5]]).tobytes())
        W, zl, b, e = sa.cw_read_unpacked(legacy)
        cw = o.CW(Pi=Pi, z=list(z), Pr=[Pr[:, l] for l in range(3)], Pt=[Pt[:, l] for l in range(3)])
        W2, zl2, b2, e2 = o.unpack(cw, np.complex128)
        assert np.array_equal(W["E0"], W2[:, 0:3]) and np.array_equal(W["k"], W2[:, 3:6])
        assert np.array_equal(b, b2) and np.array_equal(e, e2) and np.allclose(zl, zl2)
        ours = os.path.join(d, "ours.cw")
        o.cw_write(cw, ours)
        W3, *_ = sa.cw_read_unpacked(ours)
        assert np.array_equal(W3["E0"], W2[:, 0:3])
        # truncated file -> IO error
        with open(ours, "rb") as f:
            data = f.read()
        with open(ours, "wb") as f:
            f.write(data[:-40])
        with pytest.raises(_capi.CwxError) as ei:
            sa.cw_read_unpacked(ours)
        assert ei.value.code == _capi.CWX_ERR_IO


def run(tool, *args):
    return subprocess.run([os.path.join(BIN, tool), *args], capture_output=True, text=True, timeout=60)


def test_cli_help_and_errors():
    r = run("scattervolume", "--help")
    assert r.returncode == 1 and "--lambda" in r.stdout and "--coef" in r.stdout
    r = run("scatterthinvolume", "--help")
    assert r.returncode == 1
    r = run("scattervolume", "--lambda", "0.5")
    assert r.returncode == 1 and "--sample" in r.stderr
    r = run("scatterview", "--help")
    assert r.returncode == 1 and "--nogui" in r.stdout
    r = run("scatterview", "c.cw")
    assert r.returncode == 1 and "--nogui" in r.stderr
    r = run("scatterview", "--nogui", "--input", "/nonexistent.cw")
    assert r.returncode == 1 and "not found" in r.stdout
    r = run("scatterviewsample", "--help")
    assert r.returncode == 1 and "--extent" in r.stdout and "--nogui" in r.stdout
    r = run("scatterviewsample", "c.cw")
    assert r.returncode == 1 and "--nogui" in r.stderr
    r = run("scatterviewsample", "--nogui", "--input", "/nonexistent.cw")
    assert r.returncode == 1 and "not found" in r.stdout


def test_cli_negative_values_parse(tmp_path):
    """scattervolume accepts negative numbers as option values (ScatterVolume.cpp:531-535).  Parsing comes
    first: a value list that loses its negative entry fails there (status 1, "take 3 values"), while a
    parsed command reaches device creation, which fails without a GPU (status 2, "no HIP device") and
    solves with one (status 0)."""
    np.save(tmp_path / "s.npy", np.ones((4, 4), np.complex128))
    r = run("scattervolume", "--sample", str(tmp_path / "s.npy"), "--direction", "0", "1", "--output",
            str(tmp_path / "bad.cw"))
    assert r.returncode == 1 and "take 3 values" in r.stderr
    r = run("scattervolume", "--sample", str(tmp_path / "s.npy"), "--direction", "-0.2", "0", "1", "--z", "-1.5",
            "--ex", "-1", "0", "--ey", "0", "-0.5", "--output", str(tmp_path / "o.cw"))
    if r.returncode == 0:
        cw = o.cw_read(str(tmp_path / "o.cw"))
        assert np.isclose(cw.z[0], -1.5 - 1.0) and cw.Pi[0, 3].real < 0
    else:
        assert r.returncode == 2 and "no HIP device" in r.stderr, r.stderr


def test_oracle_c_field_restatement():
    from oracle import build as ob
    lib = ob.load()
    rng = np.random.default_rng(2)
    n = 200
    W = np.zeros((n, 6), np.complex128)
    W[:, :3] = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    W[:, 3:] = 8 * rng.uniform(-1, 1, (n, 3)) + 0.02j * rng.random((n, 3))
    zl = np.array([-0.5, 0.5], np.float32)
    b = np.array([0, 60, 130], np.int32)
    e = np.array([60, 130, 200], np.int32)
    g = (-0.2, 0.1, -1.0, 0.05, 0.04, 0.1, 8, 6, 21)
    out64 = np.zeros((g[8], g[7], g[6], 3), np.complex128)
    lib.cw_eval_grid_f64(W.ctypes.data, zl.ctypes.data, b.ctypes.data, e.ctypes.data, 2, 0, *g, out64.ctypes.data)
    pts = o.grid_points(*g)
    ref = o.field_points(W, zl, b, e, pts, tol=0.001)
    assert rel_err(out64.reshape(-1, 3), ref) < 1e-12
    W32 = W.astype(np.complex64)
    out32 = np.zeros((g[8], g[7], g[6], 3), np.complex64)
    lib.cw_eval_grid_f32(W32.ctypes.data, zl.ctypes.data, b.ctypes.data, e.ctypes.data, 2, 0, *g, out32.ctypes.data)
    ref32 = o.field_points(W32, zl, b, e, pts, tol=0.001)
    assert rel_err(out32.reshape(-1, 3), ref32) < 1e-4


def test_library_order_selection_matches_oracle():
    """The library's Fourier-order selection and silent cut-off (host code, no device work) against the
    oracle's select_orders (ScatterVolume.cpp:599-651) over a seeded sweep of --coef / --size / --lambda /
    --direction / --n and sample shapes, plus the known answers: 21 x 21 at 4 um and lambda 0.5 -> 10 x 10,
    the in_n[1] quirk of the M1 branch, and the clamp to the pixel counts."""
    rng = np.random.default_rng(8)
    cases = [dict(lam=0.5, size=(4.0, 4.0, 1.0), coef=(21, 21)),
             dict(lam=0.5, size=(4.0, 4.0, 1.0), coef=(21, 21), no_cutoff=True),
             dict(lam=1.0, size=(10.0, 10.0, 4.0), coef=(), n=(1.0, 1.5)),
             dict(lam=0.6, size=(3.0, 3.0, 1.0), coef=(49,))]
    shapes = [(1, 128, 128), (1, 128, 128), (100, 1, 100), (1, 5, 9)]
    expect = [(10, 10), (21, 21), None, (7, 5)]
    for _ in range(400):
        th, ph = rng.uniform(0, 1.2), rng.uniform(0, 2 * np.pi)
        c = dict(lam=float(rng.uniform(0.3, 1.5)), size=(float(rng.uniform(1, 12)), float(rng.uniform(1, 12)), 1.0),
                 direction=(float(np.sin(th) * np.cos(ph)), float(np.sin(th) * np.sin(ph)), float(np.cos(th))),
                 n=(float(rng.uniform(1, 1.6)), float(rng.uniform(1, 1.6))))
        k = int(rng.integers(0, 3))
        c["coef"] = () if k == 0 else ((int(rng.integers(1, 40)),) if k == 1 else
                                       (int(rng.integers(1, 25)), int(rng.integers(1, 25))))
        cases.append(c)
        shapes.append((1, int(rng.integers(1, 40)), int(rng.integers(1, 40))))
        expect.append(None)
    for c, shp, ex in zip(cases, shapes, expect):
        p = o.VolumeParams(**c)
        d = np.asarray(p.direction, float)
        M0, M1, _ = o.select_orders(p, shp, d / np.linalg.norm(d) * p.n[0])
        if M0 < 1 or M1 < 1:
            # the heuristic can cut an axis to no orders at all; the reference then runs on MF = 0, the
            # library refuses
            with pytest.raises(_capi.CwxError) as ei:
                sa.batch_orders([c], [shp[0]], shp[1], shp[2])
            assert ei.value.code == _capi.CWX_ERR_ARG and "no orders" in str(ei.value)
            continue
        got = sa.batch_orders([c], [shp[0]], shp[1], shp[2])
        assert got == (M0, M1), (c, shp, got, (M0, M1))
        if ex is not None:
            assert got == ex


def test_tuning_table_and_overrides(tmp_path):
    """The launch-shape table (key a_lo a_hi b_lo b_hi value per line) is read from CWX_TUNING_FILE; the
    first matching rule wins, cwx_set_param overrides a key and CWX_PARAM_TABLE returns to the table.
    Runs in a child process, because the library reads its table once."""
    table = tmp_path / "t.txt"
    table.write_text("# comment\nqr_team 1 500 1 16 2  # two helpers\nqr_team 1 500 17 64 0\n"
                     "qr_team 1 1000 1 100 5\nbatch_chunk_mib 1 1000 1 99999 8192\n")
    code = f"""
import ctypes as C, sys
sys.path.insert(0, {ROOT!r})
from scattersim_amd import _capi
L = _capi.lib()
assert L.cwx_tuning_table().decode() == {str(table)!r}
g = lambda k, a, b: L.cwx_get_param(k.encode(), a, b)
assert g("qr_team", 450, 1) == 2 and g("qr_team", 450, 32) == 0 and g("qr_team", 900, 32) == 5
assert g("qr_team", 2000, 1) == -1 and g("qr_threads", 450, 1) == -1 and g("batch_chunk_mib", 484, 10) == 8192
assert L.cwx_set_param(b"qr_team", 3) == 0 and g("qr_team", 450, 1) == 3
assert L.cwx_set_param(b"qr_team", _capi.CWX_PARAM_TABLE) == 0 and g("qr_team", 450, 1) == 2
assert L.cwx_set_param(b"qr_threads", 300) == _capi.CWX_ERR_ARG
assert L.cwx_set_param(b"no_such_key", 1) == _capi.CWX_ERR_ARG
print("ok")
"""
    env = dict(os.environ, CWX_TUNING_FILE=str(table))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=60)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr
