"""Benchmark of the MI355X coupled-wave solve and field evaluation.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config volume32|plain_small] [--eig-mode 0|1]

One step = one scattervolume solve of the configuration's sample (already resident in HBM) through
libcwx: Fourier coefficients, Phi, per-layer eigendecomposition, transfer chain, boundary LU, beta,
Q/Ez rows and the .cw records on the device.  `value` is solves/s on one GPU.  The field of the
last solution on the configuration's grid is timed separately ("field": separable synthesis,
complex64 out, with its error against the direct FP64 plane-wave sum).  The roofline entry covers
the eigensolver stage (cwx_zgeev_batched_dev over all layers), timed with HIP events on the stream
it runs on.  cpu_baseline times the oracle (numpy/MKL FP64 restatement of the reference algorithm)
on a bounded sample, and the oracle's C field sum on a few planes.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0

# HBM bytes per eigensolver-stage launch, (workload, eig_mode) -> bytes: rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE summed over the stage's kernels, each scaled by what the counter reads for a 1 GiB device copy
# (run with --copy-calib 1024).  Counters cannot be read inside a timed run, so the profiled figure is
# carried here and must be refreshed whenever the stage's kernels change.
PROFILED_STAGE_TRAFFIC = {("volume32", 1): 49.6e9}
# HBM bytes of one k_layer_reg<4> launch over the 1024 x 1024 beam, measured and scaled the same way.
PROFILED_LAYER_TRAFFIC = {"beam_1024x1024": 807e6}


def eig_flops(n: int) -> float:
    """Nominal zgeev work: 80 n^3 real flops (a complex multiply-add counts 8)."""
    return 80.0 * n ** 3


def solve_flops(MF: int, L: int, eig_mode: int) -> float:
    """Nominal FP64 flops of one solve in the conventions zgeev = 80n^3, getrf = 8/3 n^3,
    getrs(nrhs) = 8 n^2 nrhs, zgemm(m,n,k) = 8mnk (the reference's own per-layer operations)."""
    n4, n2, n6 = 4 * MF, 2 * MF, 6 * MF
    n_eig = n4 if eig_mode == 0 else n2
    f = L * eig_flops(n_eig)
    f += L * 8 / 3 * n4 ** 3                      # LU of every Gd
    f += (L - 1) * 8 * n4 ** 3                    # T_i = Gd_i^-1 Gc_{i-1}
    f += max(L - 1, 0) * 8 * n4 ** 3              # transfer product
    f += 8 / 3 * n4 ** 3 + 8 * n4 ** 2 * 3 * MF + 8 * n4 * n4 * 3 * MF   # boundary block
    f += 8 / 3 * n6 ** 3
    return f


def cpu_threads() -> int:
    """Host threads of the CPU baselines: the CPU share the box gives one GPU (OMP_NUM_THREADS there),
    else every core."""
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_baseline(sample, opts, layers: int, eig_mode: int):
    """Oracle (FP64 numpy + MKL zgeev/zgesv through torch) on the first `layers` layers, with the eigen
    step of `eig_mode` (0: zgeev of the 4MF Phi as the reference; 1: the reduced P Q form), from the
    sample to the assembled .cw; the time is scaled to all layers."""
    import torch
    from oracle import cw_oracle as o
    o.set_lapack("torch")
    threads = cpu_threads()
    torch.set_num_threads(threads)
    p = o.VolumeParams(lam=opts["lam"], direction=opts["direction"], ex=opts["ex"], ey=opts["ey"],
                       n=opts["n"], size=opts["size"], coef=opts["coef"])
    sub = sample[:layers]
    t0 = time.perf_counter()
    sol = o.solve(sub, p, eig=o.eig_layers_reduced if eig_mode else None)
    o.build_cw(sol)
    t = time.perf_counter() - t0
    L = sample.shape[0]
    return t * L / layers, threads


def plane_z(g, iz: int) -> float:
    """z of grid plane iz as the kernels compute it: dz * iz + z0 rounded once to float32."""
    return float(np.float32(float(np.float32(g["dz"])) * iz + float(np.float32(g["z0"]))))


def field_leg(ctx, sol, g, stream, dev):
    """Separable synthesis of the solution's field on the grid (complex64 out, FP64 inside), timed with
    HIP events on the library's stream; its error against the direct FP64 plane-wave sum (the
    reference's per-point algorithm) on the first planes of the grid."""
    import torch
    args = (g["x0"], g["y0"], g["z0"], g["dx"], g["dy"], g["dz"], g["nx"], g["ny"], g["nz"])
    out = torch.empty((g["nz"], g["ny"], g["nx"], 3), dtype=torch.complex64, device=dev)
    ctx.field_solution_grid_dev(sol, out.data_ptr(), False, 0, *args)
    torch.cuda.synchronize()
    times = []
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ctx.field_solution_grid_dev(sol, out.data_ptr(), False, 0, *args)
        e1.record(stream)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) * 1e-3)
    tf = sum(times) / len(times)
    npts = g["nx"] * g["ny"] * g["nz"]
    # error of the complex64 output against the complex128 form of the same evaluator (FP64 throughout;
    # it matches the oracle's direct FP64 sum over the .cw records to about 1e-13), on every point
    ref = torch.empty((g["nz"], g["ny"], g["nx"], 3), dtype=torch.complex128, device=dev)
    ctx.field_solution_grid_dev(sol, ref.data_ptr(), True, 0, *args)
    torch.cuda.synchronize()
    err, scale = 0.0, ref.abs().max().item()
    for z0 in range(0, g["nz"], 16):
        err = max(err, (out[z0:z0 + 16].to(torch.complex128) - ref[z0:z0 + 16]).abs().max().item())
    err /= scale
    del ref
    bytes_out = npts * 3 * 8
    gbs = bytes_out / tf / 1e9
    # the coefficient pass reads each layer's 3MF x 4MF mode matrix once per group of up to 8 consecutive
    # planes inside that layer (the library's grouping: region by z >= z_l - 0.001, float32 coordinates)
    zl = np.asarray(sol.z, np.float64)
    dz32, z032 = float(np.float32(g["dz"])), float(np.float32(g["z0"]))
    zs = np.array([np.float32(dz32 * i + z032) for i in range(g["nz"])], np.float64)
    region = np.array([int(np.sum(z >= zl.astype(np.float32).astype(np.float64) - 0.001)) for z in zs])
    groups_in_layers = 0
    i = 0
    while i < len(region):
        j = i + 1
        while j < len(region) and j - i < 8 and region[j] == region[i]:
            j += 1
        if 1 <= region[i] <= len(zl) - 1:
            groups_in_layers += 1
        i = j
    bytes_q = groups_in_layers * (3 * sol.MF) * (4 * sol.MF) * 16
    return {"form": "separable exp(i kx x).C(z).exp(i ky y), FP64 inside, complex64 out",
            "gpts_per_s": npts / tf / 1e9, "points": npts, "ms": tf * 1e3,
            "rel_err_vs_fp64": err, "err_reference": "complex128 separable form on every point, max |dE| / max |E|",
            "hbm_write_gbs": gbs, "hbm_frac": gbs / PEAK_HBM_GBS,
            "mode_matrix_read_bytes": bytes_q,
            "hbm_rw_frac": (bytes_out + bytes_q) / tf / 1e9 / PEAK_HBM_GBS}


def field_c3_leg(ctx, stream, dev):
    """The 7 x 7 solution of the plain_small sample evaluated on a 512 x 512 x 128 grid over the cell and
    z in [-1, 1] um (above, through and below the layer), complex64 out."""
    import scattersim_amd as sa
    from scattersim_amd import workloads
    sample, opts, _ = workloads.plain_small()
    sol = sa.scattervolume(ctx, sample, **opts)
    g = workloads.field_c3_grid()
    res = field_leg(ctx, sol, g, stream, dev)
    sol.free()
    res["workload"] = "plain_small solution on 512x512x128 (z in [-1, 1] um)"
    return res


def cpu_field_baseline(sol, g):
    """The oracle's C restatement of the reference's CPU field sum (OpenMP, FP64) on a few planes."""
    from oracle import build as ob
    from oracle import cw_oracle as o
    lib = ob.load()
    rec = sol.records
    off, cnt = sol.lists()
    lists = [rec[a:a + c] for a, c in zip(off, cnt)]
    cw = o.CW(Pi=lists[0], z=list(sol.z), Pr=lists[1::2], Pt=lists[2::2])
    W, zl, b, e = o.unpack(cw, np.complex128)
    W = np.ascontiguousarray(W)
    zl = np.ascontiguousarray(zl, np.float32)
    b = np.ascontiguousarray(b, np.int32)
    e = np.ascontiguousarray(e, np.int32)
    # one plane inside the stack, as many rows as keep the sum near 2e9 wave evaluations
    iz = g["nz"] // 2
    z = plane_z(g, iz)
    l = sum(1 for zz in zl if z >= zz - 0.001)
    nwaves = int(e[l] - b[l])
    ny = max(1, min(g["ny"], int(2e9 / max(nwaves, 1) / g["nx"])))
    out = np.empty((1, ny, g["nx"], 3), np.complex128)
    t0 = time.perf_counter()
    lib.cw_eval_grid_f64(W.ctypes.data, zl.ctypes.data, b.ctypes.data, e.ctypes.data, len(zl), 0,
                         g["x0"], g["y0"], z, g["dx"], g["dy"], g["dz"], g["nx"], ny, 1, out.ctypes.data)
    t = time.perf_counter() - t0
    npts = g["nx"] * ny
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"gpts_per_s": npts / t / 1e9, "cores": threads, "kind": "port",
            "sample": f"{ny} rows of plane {iz} ({npts} points, {nwaves} waves each), direct FP64 sum"}


def batched_leg(ctx, count: int, plain_count: int, dev):
    """The batched sweep (256 wavelengths x 8 angles over 4-24-layer stacks, 11 x 11 orders): the first
    `count` instances solved together by cwx_solve_batch_dev, against the first `plain_count` of them
    solved one at a time through the same entry point (same outputs, so the same work per instance)."""
    import torch
    import scattersim_amd as sa
    from scattersim_amd import workloads
    packed, Ls, params = workloads.sweep(count)
    for p in params:
        p["eig_mode"] = 1
    N = packed.shape[1]
    M0, M1 = sa.batch_orders(params, Ls, N, N)
    n6 = 6 * M0 * M1
    d_s = torch.from_numpy(packed).to(dev)
    d_x = torch.zeros((count, n6), dtype=torch.complex128, device=dev)
    d_xp = torch.zeros((plain_count, n6), dtype=torch.complex128, device=dev)
    npix = N * N
    offs = np.concatenate([[0], np.cumsum(Ls)[:-1]]).astype(np.int64)
    # the plain subset is spread evenly over the sweep (every stack, wavelength and angle band), so that
    # its layers per instance match the sweep's; the rates are also compared per layer
    pick = np.unique(np.round(np.linspace(0, count - 1, plain_count)).astype(np.int64))
    plain_count = len(pick)

    def plain(j):
        i = int(pick[j])
        ptr = d_s.data_ptr() + int(offs[i]) * npix * 16
        sa.solve_batch(ctx, ptr, [params[i]], shape=([int(Ls[i])], N, N), d_x=d_xp.data_ptr() + j * n6 * 16)

    plain(0)                                   # warm-up
    # warm-up over several full chunks, so that the timed call finds its workspaces already grown
    wu = min(count, 512)
    sa.solve_batch(ctx, d_s.data_ptr(), params[:wu], shape=(Ls[:wu], N, N), d_x=d_x.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sa.solve_batch(ctx, d_s.data_ptr(), params, shape=(Ls, N, N), d_x=d_x.data_ptr())
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    t0 = time.perf_counter()
    for i in range(plain_count):
        plain(i)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    same = bool(torch.equal(d_x[torch.from_numpy(pick).to(dev)], d_xp))
    # the same sweep as launch sets captured in HIP graphs: chunks of 128 instances (the 16 stacks in the
    # same order in each) replay the graph captured for the first; the first call captures, the second is
    # timed (replays only)
    chunk = 128 if count % 128 == 0 else count
    d_xg = torch.zeros_like(d_x)
    _, cap = sa.solve_batch_graph(ctx, d_s.data_ptr(), params, Ls, N, N, chunk, d_xg.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ginfo, cap2 = sa.solve_batch_graph(ctx, d_s.data_ptr(), params, Ls, N, N, chunk, d_xg.data_ptr())
    torch.cuda.synchronize()
    tg = time.perf_counter() - t0
    graph = {"solves_per_s": count / tg, "s": tg, "chunk_instances": chunk, "graphs_captured": cap,
             "captured_in_timed_call": cap2, "equals_stream": bool(torch.equal(d_xg, d_x)),
             "over_stream": (count / tg) / (count / tb)}
    del d_xg
    fl = sum(solve_flops(M0 * M1, int(L), 1) for L in Ls)
    lp = int(Ls[pick].sum())
    return {"workload": "sweep: 256 lambda x 8 angles, stacks of 4-24 layers, 11x11 orders, reduced eigen mode",
            "instances": count, "layers": int(Ls.sum()), "batched_solves_per_s": count / tb,
            "batched_s": tb, "plain_solves_per_s": plain_count / tp, "plain_instances": plain_count,
            "plain_subset": "evenly spaced over the sweep", "mean_layers_sweep": float(Ls.mean()),
            "mean_layers_plain": float(Ls[pick].mean()),
            "batched_over_plain": (count / tb) / (plain_count / tp),
            "batched_over_plain_per_layer": (float(Ls.sum()) / tb) / (lp / tp), "batched_equals_plain": same,
            "batched_tflops_nominal": fl / tb / 1e12, "graph": graph}


def layer_flops(L: int) -> float:
    """Algorithmic FP64 flops of one incident wave of a scatterlayer stack: Gaussian elimination of the
    augmented 6L x (6L+1) system (complex multiply-add = 8 flops) and the back-substitution."""
    n = 6 * L
    return 8.0 * (n - 1) * n * (n + 1) / 3.0 + 8.0 * n * (n - 1) / 2.0


def layer_leg(ctx, stream, dev, no_cpu: bool):
    """scatterlayer (cwx_layer_solve_dev) on the a.cw stack (n = 1 | 1.4+0.05i | 1.4 | 1, z = -3 0 3,
    lambda = 1): the tool's default 64 x 64 polar beam and a 1024 x 1024 beam, waves resident in HBM."""
    import math
    import torch
    import scattersim_amd as sa
    lib = sa.lib()
    L, n, kap, z = sa.layer_stack((1.0, 1.4, 1.4, 1.0), (0.05,), (-3.0, 0.0, 3.0))
    out = {"workload": "a.cw stack (4 layers), polar beam alpha 1 about +z, lambda 1"}
    for name, N in (("beam_64x64", 64), ("beam_1024x1024", 1024)):
        inc = sa.beam_polar(1.0, (0.0, 0.0, 2 * math.pi), (0, 1, 0), N, N)
        P = len(inc)
        d_inc = torch.from_numpy(inc).to(dev)
        d_r = torch.zeros((P, L - 1, 6), dtype=torch.complex128, device=dev)
        d_t = torch.zeros_like(d_r)

        def call():
            sa.check(lib.cwx_layer_solve_dev(ctx.h, L, n.ctypes.data, kap.ctypes.data, z.ctypes.data, 1.0, P,
                                             d_inc.data_ptr(), d_r.data_ptr(), d_t.data_ptr()))
        call()
        torch.cuda.synchronize()
        ts = []
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            call()
            e1.record(stream)
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        t = min(ts)
        fl = P * layer_flops(L)
        out[name] = {"waves": P, "ms": t * 1e3, "waves_per_s": P / t,
                     "roofline": {"kernel": "k_layer_reg<4>", "bound": "fp64 vector", "achieved": fl / t / 1e12,
                                  "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                                  "frac": fl / t / 1e12 / PEAK_FP64_TFLOPS, "flops_per_wave": layer_flops(L),
                                  "traffic": PROFILED_LAYER_TRAFFIC.get(name),
                                  "bytes_per_wave": 96 * (1 + 2 * (L - 1))}}
    if not no_cpu:
        from oracle import layer_oracle as lo
        inc = sa.beam_polar(1.0, (0.0, 0.0, 2 * math.pi), (0, 1, 0), 64, 64)[::8]
        t0 = time.perf_counter()
        lo.solve(lo.LayerStack(n=(1.0, 1.4, 1.4, 1.0), kappa=(0.05,), z=(-3.0, 0.0, 3.0), lam=1.0), inc)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": len(inc) / tc, "unit": "waves/s", "cores": 1, "kind": "port",
                               "sample": f"{len(inc)} of the 4096 beam waves through the layer oracle"}
    return out


def chain_gemm_leg(ctx, stream, dev):
    """The transfer chain's complex products (MKL_multiply at ScatterVolume.cpp:330-334, 768-781) at the
    chain shapes of the sweep (n = 4 MF = 484) and of volume32 (n = 900), FP64 against the FP32 matrix
    cores (cwx_zgemm_split32_batched_dev, split operands or operands rounded to FP32); HIP events on the
    library's stream; error = max |C - C_fp64| / max |C_fp64| on random unit-variance operands."""
    import torch
    import scattersim_amd as sa
    lib = sa.lib()
    one, zero = sa._capi.cwx_complex(1, 0), sa._capi.cwx_complex(0, 0)
    out = {}
    for n, batch in ((484, 128), (900, 24)):
        g = torch.Generator(device=dev).manual_seed(n)
        A = torch.randn((batch, n, n), dtype=torch.complex128, device=dev, generator=g)
        B = torch.randn((batch, n, n), dtype=torch.complex128, device=dev, generator=g)
        C = torch.empty_like(A)
        fl = 8.0 * n ** 3 * batch
        row = {"batch": batch, "flops_per_launch": fl}
        ref = None
        for name, terms in (("fp64", 0), ("fp32_split3", 3), ("fp32_rounded", 1)):
            def call():
                if terms == 0:
                    sa.check(lib.cwx_zgemm_batched_dev(ctx.h, 0, 0, n, n, n, one, A.data_ptr(), n, n * n,
                                                       B.data_ptr(), n, n * n, zero, C.data_ptr(), n, n * n, batch))
                else:
                    sa.check(lib.cwx_zgemm_split32_batched_dev(ctx.h, 0, 0, n, n, n, one, A.data_ptr(), n, n * n,
                                                               B.data_ptr(), n, n * n, zero, C.data_ptr(), n, n * n,
                                                               batch, terms))
            call()
            torch.cuda.synchronize()
            ts = []
            for _ in range(3):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                call()
                e1.record(stream)
                torch.cuda.synchronize()
                ts.append(e0.elapsed_time(e1) * 1e-3)
            t = min(ts)
            if ref is None:
                ref = C.clone()
                err = 0.0
            else:
                err = float((C - ref).abs().max() / ref.abs().max())
            row[name] = {"ms": t * 1e3, "tflops": fl / t / 1e12, "rel_err_vs_fp64": err}
        out[f"n{n}"] = row
        del A, B, C, ref
    torch.cuda.synchronize()
    return out


def fields_leg(ctx, d_sample, shape, opts, eig_mode: int, plain_ms: float):
    """All 2MF unit transverse incident fields of the workload's cell in one cwx_solve_fields_dev call (the
    order basis of its reflection and transmission matrices), against one plain solve per field."""
    import torch
    import scattersim_amd as sa
    L, s1, s2 = shape
    M0, M1 = sa.solve_fields(ctx, d_sample.data_ptr(), None, shape=shape, eig_mode=eig_mode, **opts)
    MF = M0 * M1
    K = 2 * MF
    EF = torch.zeros((K, 3 * MF), dtype=torch.complex128, device=d_sample.device)
    idx = torch.arange(MF, device=d_sample.device)
    EF[idx, idx] = 1.0
    EF[MF + idx, MF + idx] = 1.0
    d_x = torch.zeros((K, 6 * MF), dtype=torch.complex128, device=d_sample.device)
    p = sa._params(eig_mode=eig_mode, **opts)
    import ctypes as C
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sa.check(sa.lib().cwx_solve_fields_dev(ctx.h, C.byref(p), C.c_void_p(d_sample.data_ptr()), L, s1, s2, K,
                                           C.c_void_p(EF.data_ptr()), C.c_void_p(d_x.data_ptr()), None, None))
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    return {"workload": "order basis (2MF unit Ex / Ey fields) of the benchmark cell", "fields": K, "s": t,
            "fields_per_s": K / t, "plain_solves_per_s": 1e3 / plain_ms,
            "speedup_over_plain": (K / t) / (1e3 / plain_ms)}


def compact_leg(sol):
    """Full and compact .cw of the benchmark solution: sizes, write times, and the rebuild of a device
    solution from the compact file (host I/O to the node's local disk)."""
    import tempfile
    import scattersim_amd as sa
    with tempfile.TemporaryDirectory() as d:
        full, comp = os.path.join(d, "c.cw"), os.path.join(d, "c2.cw")
        t0 = time.perf_counter()
        sol.write_cw(full)
        t1 = time.perf_counter()
        sol.write_cw2(comp)
        t2 = time.perf_counter()
        back = sa.read_cw2(sol.ctx, comp)
        t3 = time.perf_counter()
        same = bool(np.array_equal(back.x, sol.x))
        back.free()
        return {"full_bytes": os.path.getsize(full), "compact_bytes": os.path.getsize(comp),
                "write_full_s": t1 - t0, "write_compact_s": t2 - t1, "rebuild_s": t3 - t2,
                "rebuilt_x_identical": same}


PEAK_BF16_TFLOPS = 2500.0


def learned_leg(ctx, no_cpu: bool):
    """The learned closure (c4: 7x7 orders plus permittivity coefficients in, 21x21 R/T out) on 256 fresh
    random-disk layers: its error and the low-order solution's error against the 21x21 reference solve, and
    the throughput of low-order solve plus network against that solve.  The network was trained on 20 000
    solver-generated pairs (python -m scattersim_amd.closure) and ships as scattersim_amd/data/closure_c4.npz."""
    import scattersim_amd as sa
    from scattersim_amd import closure as cl
    net = cl.load(os.path.join(ROOT, "scattersim_amd", "data", "closure_c4.npz"))
    cfg = net.cfg
    count = 256
    layers = cl.disk_layers(cfg, count, seed=987654)
    cl.solve_orders(ctx, cfg, layers, cfg.low)                 # warm-up at full size: workspaces sized
    t0 = time.perf_counter()
    x_low = cl.solve_orders(ctx, cfg, layers, cfg.low)
    t_low = time.perf_counter() - t0
    t0 = time.perf_counter()
    x_high = cl.solve_orders(ctx, cfg, layers, cfg.high, chunk=256)
    t_high = time.perf_counter() - t0
    X = cl.features(cfg, layers, x_low)
    net.predict(ctx, X, x_low)                                 # warm-up at full size
    t0 = time.perf_counter()
    pred = net.predict(ctx, X, x_low)
    t_inf = time.perf_counter() - t0
    e_net = cl.rel_err_rows(pred, x_high)
    e_base = cl.rel_err_rows(net.baseline(x_low), x_high)
    # the network's forward pass at a large batch (the library's stream, synchronised around the timed loop)
    d0 = cl._pad4(cfg.d_in)
    dims = [d0, *cfg.hidden, cfg.d_out]
    B = 16384
    w = list(net.weights)
    w[0] = cl._pad_cols(w[0], d0)
    mlp = cl.MLP(ctx, dims, B, weights=w)
    Xn = cl._bf16_bits(cl._pad_cols(np.tile((X - net.x_mean) / net.x_std, (B // count, 1)).astype(np.float32), d0))
    dX = sa.DeviceArray.from_numpy(ctx, Xn)
    mlp.forward(dX, 0, B, keep=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        mlp.forward(dX, 0, B, keep=False)
    ctx.synchronize()
    t_fw = (time.perf_counter() - t0) / reps
    fl = 2.0 * B * sum(dims[i] * dims[i + 1] for i in range(len(dims) - 1))
    out = {"workload": "c4: random-disk layers (128x128, 8 disks, n 1.3-1.7, 4 um cell, 1 um thick, lambda 0.5), "
                       "7x7 -> 21x21 orders", "samples": count,
           "rel_err_closure": float(e_net.mean()), "rel_err_low_order": float(e_base.mean()),
           "high_order_solves_per_s": count / t_high, "low_order_solves_per_s": count / t_low,
           "closure_samples_per_s": count / (t_low + t_inf),
           "closure_over_high_order": (count / (t_low + t_inf)) / (count / t_high),
           "params": int(sum(int(np.prod(w.shape)) for w in net.weights)),
           "forward_roofline": {"kernel": "k_dense_bf16 (3 layers, batch 16384, inference)", "bound": "mfma bf16",
                                "achieved": fl / t_fw / 1e12, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                                "frac": fl / t_fw / 1e12 / PEAK_BF16_TFLOPS, "ms": t_fw * 1e3}}
    if not no_cpu:
        from oracle import closure_oracle as co
        params = [(net.weights[2 * i].astype(np.float64), net.weights[2 * i + 1].astype(np.float64))
                  for i in range(len(dims) - 1)]
        Xc = ((X - net.x_mean) / net.x_std).astype(np.float32)
        t0 = time.perf_counter()
        co.forward(params, Xc)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": count / tc, "unit": "network forwards/s", "kind": "port",
                               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                               "sample": f"{count} forwards through the closure oracle (numpy)"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="volume32")
    ap.add_argument("--eig-mode", type=int, default=1)
    ap.add_argument("--cpu-layers", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-field", action="store_true")
    ap.add_argument("--sweep-count", type=int, default=2048, help="instances of the batched sweep leg (2048: the whole sweep)")
    ap.add_argument("--sweep-plain", type=int, default=24, help="of which solved one at a time")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-next", action="store_true", help="skip the layer-stack, many-field and compact legs")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (a blit of known traffic that scales the "
                         "profiler's FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the solver has no multi-GPU path (layers and instances "
              "would shard across GPUs with no exchange between them)", file=sys.stderr)
        sys.exit(2)

    import torch

    import scattersim_amd as sa
    from scattersim_amd import workloads

    sample, opts, grid = workloads.CONFIGS[args.config]()
    L, s1, s2 = sample.shape
    dev = torch.device("cuda:0")
    d_sample = torch.from_numpy(sample).to(dev)
    stream = torch.cuda.Stream()          # default priority; the library queues all work on it
    ctx = sa.Context(0)
    ctx.set_stream(stream.cuda_stream)
    torch.cuda.synchronize()
    if args.copy_calib > 0:
        src_buf = torch.ones(args.copy_calib << 20, dtype=torch.uint8, device=dev)
        dst_buf = torch.empty_like(src_buf)
        dst_buf.copy_(src_buf)
        torch.cuda.synchronize()
        del src_buf, dst_buf

    def one():
        sol = sa.scattervolume(ctx, d_sample.data_ptr(), shape=(L, s1, s2), eig_mode=args.eig_mode, **opts)
        return sol

    for _ in range(args.warmup):
        one().free()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sol = one()
        if _ != args.steps - 1:
            sol.free()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) / args.steps * 1e3
    MF = sol.MF
    # the same solve with the reference's own eigen formulation (zgeev of the 4MF Phi), the CLI default
    other = 1 - args.eig_mode
    one_other = lambda: sa.scattervolume(ctx, d_sample.data_ptr(), shape=(L, s1, s2), eig_mode=other, **opts)
    one_other().free()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    k_other = max(1, min(args.steps, 3))
    for _ in range(k_other):
        one_other().free()
    torch.cuda.synchronize()
    ms_other = (time.perf_counter() - t0) / k_other * 1e3

    # ---- eigensolver stage: HIP events on the stream it runs on ----
    lib = sa.lib()
    p = sa._params(eig_mode=args.eig_mode, **opts)
    import ctypes as C
    nmat = C.c_int()
    blocks = 1 if args.eig_mode else 0
    sa.check(lib.cwx_phi_build_dev(ctx.h, C.byref(p), None, L, s1, s2, blocks, None, C.byref(nmat)))
    n = nmat.value
    per = 2 if blocks else 1
    phi = torch.zeros((per * L, n, n), dtype=torch.complex128, device=dev)
    sa.check(lib.cwx_phi_build_dev(ctx.h, C.byref(p), C.c_void_p(d_sample.data_ptr()), L, s1, s2, blocks,
                                   C.c_void_p(phi.data_ptr()), C.byref(nmat)))
    if blocks:
        Pm = phi[:L].transpose(1, 2).contiguous()
        Qm = phi[L:].transpose(1, 2).contiguous()
        A0 = torch.matmul(Pm, Qm).transpose(1, 2).contiguous()
    else:
        A0 = phi
    W = torch.zeros((L, n), dtype=torch.complex128, device=dev)
    V = torch.zeros((L, n, n), dtype=torch.complex128, device=dev)
    reps = 2
    times = []
    torch.cuda.synchronize()
    for _ in range(reps):
        A = A0.clone()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        sa.check(lib.cwx_zgeev_batched_dev(ctx.h, n, L, A.data_ptr(), n * n, W.data_ptr(), n, V.data_ptr(), n * n,
                                           None))
        e1.record(stream)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) * 1e-3)
    t_eig = min(times)
    fl = L * eig_flops(n)
    achieved = fl / t_eig / 1e12

    # ---- field evaluation of the solution on the configuration grid ----
    field = None
    field_c3 = None
    if not args.no_field:
        field = field_leg(ctx, sol, grid, stream, dev)
        field_c3 = field_c3_leg(ctx, stream, dev)
    cpu_field = None
    if not args.no_field and not args.no_cpu:
        cpu_field = cpu_field_baseline(sol, grid)
    nxt = None
    if not args.no_next:
        nxt = {"layer": layer_leg(ctx, stream, dev, args.no_cpu),
               "fields": fields_leg(ctx, d_sample, (L, s1, s2), opts, args.eig_mode, ms),
               "compact_cw": compact_leg(sol),
               "chain_gemm": chain_gemm_leg(ctx, stream, dev),
               "learned": learned_leg(ctx, args.no_cpu)}
    sol.free()

    batched = None
    if not args.no_batched:
        batched = batched_leg(ctx, args.sweep_count, min(args.sweep_plain, args.sweep_count), dev)

    cpu = None
    if not args.no_cpu:
        # the headline's eigen mode, and the other one beside it (the reference's full form is the slower)
        lay = {1: min(args.cpu_layers, L), 0: min(max(1, args.cpu_layers // 2), L)}
        t_cpu, threads = cpu_baseline(sample, opts, lay[args.eig_mode], args.eig_mode)
        t_cpu_o, _ = cpu_baseline(sample, opts, lay[other], other)
        name = {0: "full 4MF Phi", 1: "reduced 2MF PQ"}
        cpu = {"value": 1.0 / t_cpu, "unit": "solves/s", "cores": threads, "kind": "port",
               "eig_mode": name[args.eig_mode],
               "sample": f"oracle solve ({name[args.eig_mode]}) of the first {lay[args.eig_mode]} of {L} layers, "
                         f"scaled by {L}/{lay[args.eig_mode]}",
               "other_mode": {"eig_mode": name[other], "value": 1.0 / t_cpu_o, "unit": "solves/s",
                              "sample": f"first {lay[other]} of {L} layers, scaled"},
               "field": cpu_field}

    line = {
        "metric": "sample solves/sec + field Gpts/s at <=1e-6 rel err vs FP64 (plain vs batched)",
        "value": 1e3 / ms,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (f64)",
        "data": "synthetic (seeded)",
        "config": {"workload": args.config, "layers": L, "grid": [s1, s2], "orders": [sol.M0, sol.M1],
                   "eig_n": n, "eig_mode": "full 4MF Phi" if args.eig_mode == 0 else "reduced 2MF PQ",
                   "field_grid": [grid["nx"], grid["ny"], grid["nz"]],
                   "parallelism": "1 GPU"},
        "solve_tflops_nominal": solve_flops(MF, L, args.eig_mode) / (ms * 1e-3) / 1e12,
        "other_eig_mode": {"eig_mode": "full 4MF Phi" if other == 0 else "reduced 2MF PQ",
                           "eig_n": (4 if other == 0 else 2) * MF, "solves_per_s": 1e3 / ms_other,
                           "ms_per_step": ms_other, "steps": k_other,
                           "solve_tflops_nominal": solve_flops(MF, L, other) / (ms_other * 1e-3) / 1e12},
        "field": field,
        "field_c3": field_c3,
        "batched": batched,
        "next": nxt,
        "roofline": {"kernel": "eigensolver stage (gehrd+unghr+hqr+trevc+gemm+normalize)", "bound": "mfma",
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": PROFILED_STAGE_TRAFFIC.get((args.config, args.eig_mode)),
                     "ms_per_launch": t_eig * 1e3, "flops_per_launch": fl},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.close()


if __name__ == "__main__":
    main()
