"""Launch-shape tuner of libcwx: times the library's choices on this GPU and writes the table it loads.

    python -m scattersim_amd.tune [--out scattersim_amd/data/tuning.txt] [--quick]

Keys (see cwx_set_param in include/cwx.h; none of them changes a result, only the time):

* qr_team (a = n, b = batch): helper workgroups per matrix of the QR iteration, for eigenproblems that
  leave most of the device idle.  Searched over 0 .. 8 helpers on random complex matrices.
* qr_threads (a = n, b = batch): 256 or 512 threads per QR workgroup for batches that fill the device.
* batch_chunk_mib (a = 4 M0 M1, b = instances): device memory per chunk of the batched solve, searched on
  a prefix of the batched sweep.

Each rule covers the range of shapes between the neighbouring measured points.  The table is plain text,
one rule per line (key a_lo a_hi b_lo b_hi value), read by the library at the first lookup.
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import time

import numpy as np

from . import DeviceArray, _capi, lib
from . import solve_batch
from . import workloads

PKG = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(PKG, "data", "tuning.txt")


def _zgeev_ms(ctx, A: np.ndarray, reps: int = 2) -> float:
    batch, n, _ = A.shape
    dW = DeviceArray(ctx, (batch, n), np.complex128)
    dV = DeviceArray(ctx, (batch, n, n), np.complex128)
    best = float("inf")
    for _ in range(reps):
        dA = DeviceArray.from_numpy(ctx, A)
        ctx.synchronize()
        t0 = time.perf_counter()
        _capi.check(lib().cwx_zgeev_batched_dev(ctx.h, n, batch, C.c_void_p(dA.ptr), n * n, C.c_void_p(dW.ptr), n,
                                                C.c_void_p(dV.ptr), n * n, None))
        ctx.synchronize()
        best = min(best, (time.perf_counter() - t0) * 1e3)
        dA.free()
    dW.free()
    dV.free()
    return best


def _random(n: int, batch: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))).astype(np.complex128)


def _ranges(points):
    """Inclusive ranges splitting the axis halfway (geometrically) between consecutive measured points."""
    pts = sorted(points)
    cuts = [int(np.sqrt(a * b)) for a, b in zip(pts[:-1], pts[1:])]
    lo = [1] + [c + 1 for c in cuts]
    hi = cuts + [1 << 30]
    return dict(zip(pts, zip(lo, hi)))


def tune_qr_team(ctx, cus: int, sizes, batches, log):
    rules = []
    nr, br = _ranges(sizes), _ranges(batches)
    for n in sizes:
        for batch in batches:
            if batch >= cus:
                continue
            A = _random(n, batch, n + batch)
            cand = [h for h in (0, 1, 2, 3, 4, 6, 8) if h == 0 or (h + 1) * batch <= cus]
            times = {}
            for h in cand:
                _capi.check(lib().cwx_set_param(b"qr_team", h))
                times[h] = _zgeev_ms(ctx, A)
            best = min(times, key=times.get)
            log(f"qr_team n={n} batch={batch}: " + ", ".join(f"{h}: {t:.1f} ms" for h, t in times.items()) +
                f" -> {best}")
            rules.append(("qr_team", *nr[n], *br[batch], best, times))
    _capi.check(lib().cwx_set_param(b"qr_team", _capi.CWX_PARAM_TABLE))
    return rules


def tune_qr_threads(ctx, cus: int, shapes, log):
    rules = []
    for n, batch in shapes:
        A = _random(n, batch, 7 * n + batch)
        times = {}
        for t in (256, 512):
            _capi.check(lib().cwx_set_param(b"qr_threads", t))
            times[t] = _zgeev_ms(ctx, A, reps=1)
        best = min(times, key=times.get)
        log(f"qr_threads n={n} batch={batch}: " + ", ".join(f"{t}: {v:.1f} ms" for t, v in times.items()) +
            f" -> {best}")
        rules.append(("qr_threads", n // 2 + 1, 2 * n, 2 * cus, 1 << 30, best, times))
    _capi.check(lib().cwx_set_param(b"qr_threads", _capi.CWX_PARAM_TABLE))
    return rules


def tune_batch_chunk(ctx, count: int, log):
    packed, Ls, params = workloads.sweep(count)
    for p in params:
        p["eig_mode"] = 1
    N = packed.shape[1]
    d_s = DeviceArray.from_numpy(ctx, packed)
    times = {}
    for mib in (8192, 16384, 32768, 65536):
        _capi.check(lib().cwx_set_param(b"batch_chunk_mib", mib))
        solve_batch(ctx, d_s.ptr, params[:min(count, 64)], shape=(Ls[:min(count, 64)], N, N))
        ctx.synchronize()
        t0 = time.perf_counter()
        solve_batch(ctx, d_s.ptr, params, shape=(Ls, N, N))
        ctx.synchronize()
        times[mib] = (time.perf_counter() - t0) * 1e3
    _capi.check(lib().cwx_set_param(b"batch_chunk_mib", _capi.CWX_PARAM_TABLE))
    d_s.free()
    best = min(times, key=times.get)
    log(f"batch_chunk_mib (sweep, {count} instances): " + ", ".join(f"{k}: {v:.0f} ms" for k, v in times.items()) +
        f" -> {best}")
    n4 = 4 * 11 * 11
    return [("batch_chunk_mib", n4 // 2, 2 * n4, 2, 1 << 30, best, times)]


def main(argv=None):
    from . import Context
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--quick", action="store_true", help="fewer shapes (a smoke run of the tuner)")
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    ctx = Context(0)
    import subprocess
    cus = 256
    try:
        out = subprocess.run(["rocminfo"], capture_output=True, text=True, timeout=60).stdout
        vals = [int(l.split(":")[1]) for l in out.splitlines() if "Compute Unit:" in l]
        cus = max(vals) if vals else cus
    except Exception:
        pass
    sizes = (196, 450, 900) if a.quick else (196, 450, 900, 1764)
    batches = (1, 32) if a.quick else (1, 8, 32, 64)
    rules = tune_qr_team(ctx, cus, sizes, batches, log)
    rules += tune_qr_threads(ctx, cus, ((242, 1024),) if a.quick else ((242, 2048), (450, 1024)), log)
    rules += tune_batch_chunk(ctx, 256 if a.quick else 1024, log)
    lines = ["# libcwx launch-shape table, written by python -m scattersim_amd.tune on this kind of GPU",
             f"# {time.strftime('%Y-%m-%d')}, {cus} compute units; key a_lo a_hi b_lo b_hi value  # measured ms"]
    for key, alo, ahi, blo, bhi, val, times in rules:
        meas = " ".join(f"{k}:{v:.1f}" for k, v in times.items())
        lines.append(f"{key} {alo} {ahi} {blo} {bhi} {val}  # {meas}")
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    with open(a.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    log(f"wrote {a.out}")
    ctx.close()


if __name__ == "__main__":
    main()
