"""Writes configs.npz: oracle solutions of the benchmark configurations at their full sizes.

    python tests/golden/make_configs.py            # about 3 minutes on 8 cores

Every solve is the oracle's restatement of ScatterVolume.cpp:574-889 with zgeev of the full 4MF x 4MF
property matrix (third_Lapack.h:38-45, ScatterVolume.cpp:255-262) and zgesv / inverses from Intel MKL
through torch's CPU build, the LAPACK family the reference links.  The GPU tests compare the HIP path,
in both eigen modes, with these arrays (tests/test_gpu_configs.py).

Stored per configuration (prefix c0_, c1_, c2_<i>_, c4_<i>_):

* x        the boundary solution [Rx Ry Rz Tx Ty Tz] (6MF)
* gam      the sorted, interleaved eigenvalues of every layer (L x 4MF)
* pts, E   points inside every region (above, two per layer, below) and the FP64 field that the
           solution's .cw records give there (tolerance-free region rule, cpuEvaluator.cpp:294-407)
* condA    2-norm condition number of the boundary system A (ScatterVolume.cpp:417-466, 739)
* condS    condition number of the transfer product Gc_static (ScatterVolume.cpp:330-334) and of the
           Gd_static Gc_static^-1 factor that enters A

plus c0_gpts / c0_gE: the FP64 direct field of the c0 solution at 256 seeded points of the 512 x 512 x 128
grid of the mixed-precision field configuration, and c1_gpts / c1_gE the same for the c1 solution on its
256 x 256 x 64 grid.  The samples are not stored: scattersim_amd.workloads and scattersim_amd.closure
regenerate them from their seeds.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import cw_oracle as o  # noqa: E402
from scattersim_amd import closure, workloads  # noqa: E402

# c2 instances: (stack, wavelength, incidence angle in degrees) at the sweep's extremes, shortest and
# longest stack of the 16 seeded stacks
C2_CASES = (("short", 0.4, 0.0), ("short", 0.8, 35.0), ("long", 0.8, 0.0), ("long", 0.4, 35.0))
C4_LAYERS = 2
C4_SEED = 7


def probe_points(z, cell, seed=0, per_region=6):
    rng = np.random.default_rng(seed)
    zs = [z[0] - 0.7]
    for a, b in zip(z[:-1], z[1:]):
        zs += [a + 0.3 * (b - a), a + 0.77 * (b - a)]
    zs.append(z[-1] + 0.6)
    pts = [[rng.uniform(0, cell[0]), rng.uniform(0, cell[1]), zz] for zz in zs for _ in range(per_region)]
    return np.array(pts)


def fields(sol, pts):
    W, zl, b, e = o.unpack(o.build_cw(sol), np.complex128)
    return o.field_points(W, zl, b, e, pts, tol=0.0)


def grid_sample(grid, count, seed):
    """`count` seeded points of a regular grid (x0, dx, nx, ...), as the float32 coordinates the
    evaluators form (one fused multiply-add per axis)."""
    rng = np.random.default_rng(seed)
    ix = rng.integers(0, grid["nx"], count)
    iy = rng.integers(0, grid["ny"], count)
    iz = rng.integers(0, grid["nz"], count)
    P = np.stack([o.grid_coord(grid["x0"], grid["dx"], ix), o.grid_coord(grid["y0"], grid["dy"], iy),
                  o.grid_coord(grid["z0"], grid["dz"], iz)], axis=1)
    return np.stack([ix, iy, iz], 1), P


def conds(sol):
    s = sol.setup
    MF, L, k, z = s.MF, s.L, s.k, s.z
    ZL = L + 1
    Gd_s = Gc_s = None
    for i in range(L):
        dz = (z[ZL - 1] - z[0]) if L == 1 else (z[i + 1] - z[i])
        Gd, Gc = o.gd_gc(sol.gammas[i], sol.vectors[i], k, dz)
        if i == 0:
            Gd_s, Gc_s = Gd, Gc
        else:
            Gc_s = (Gc @ o.zinv(Gd)) @ Gc_s
    return (float(np.linalg.cond(sol.A)), float(np.linalg.cond(Gc_s)),
            float(np.linalg.cond(Gd_s @ o.zinv(Gc_s))))


def record(out, prefix, sol, pts):
    out[prefix + "x"] = sol.x
    out[prefix + "gam"] = np.stack(sol.gammas)
    out[prefix + "pts"] = pts
    out[prefix + "E"] = fields(sol, pts)
    cA, cS, cB = conds(sol)
    out[prefix + "condA"] = np.array(cA)
    out[prefix + "condS"] = np.array([cS, cB])
    print(f"  {prefix}: cond(A) {cA:.3e}, cond(Gc_static) {cS:.3e}, cond(Gd0 Gc_static^-1) {cB:.3e}", flush=True)


def main():
    o.set_lapack("torch")
    out = {}
    t0 = time.time()
    # c0: plain small
    sample, opts, grid = workloads.plain_small(0)
    p = o.VolumeParams(**opts)
    sol = o.solve(sample, p)
    record(out, "c0_", sol, probe_points(sol.setup.z, p.size))
    idx, P = grid_sample(workloads.field_c3_grid(), 256, 11)
    W, zl, b, e = o.unpack(o.build_cw(sol), np.complex128)
    out["c0_gidx"], out["c0_gpts"], out["c0_gE"] = idx, P, o.field_points(W, zl, b, e, P, tol=0.0)
    print(f"c0 done {time.time() - t0:.1f} s", flush=True)
    # c1: volume32
    sample, opts, grid = workloads.volume32(0)
    p = o.VolumeParams(**opts)
    sol = o.solve(sample, p)
    record(out, "c1_", sol, probe_points(sol.setup.z, p.size))
    idx, P = grid_sample(grid, 256, 12)
    W, zl, b, e = o.unpack(o.build_cw(sol), np.complex128)
    out["c1_gidx"], out["c1_gpts"], out["c1_gE"] = idx, P, o.field_points(W, zl, b, e, P, tol=0.0)
    del sol, W
    print(f"c1 done {time.time() - t0:.1f} s", flush=True)
    # c2: four sweep instances
    Ls = [m.shape[0] for m in workloads.sweep_stacks()]
    stack = {"short": int(np.argmin(Ls)), "long": int(np.argmax(Ls))}
    meta = []
    for i, (which, lam, th) in enumerate(C2_CASES):
        smp, prm = workloads.sweep_instance(stack[which], lam, th)
        p = o.VolumeParams(**prm)
        sol = o.solve(smp, p)
        record(out, f"c2_{i}_", sol, probe_points(sol.setup.z, p.size))
        meta.append([stack[which], lam, th])
        print(f"c2 case {i} (stack {stack[which]}, {smp.shape[0]} layers) done {time.time() - t0:.1f} s", flush=True)
    out["c2_cases"] = np.array(meta)
    # c4: judge layers at 21 x 21 orders (the order cut-off bypassed)
    cfg = closure.ClosureConfig()
    layers = closure.disk_layers(cfg, C4_LAYERS, seed=C4_SEED)
    for i in range(C4_LAYERS):
        prm = cfg.opts(cfg.high)
        prm.pop("eig_mode")
        p = o.VolumeParams(**prm)
        sol = o.solve(layers[i][None], p)
        record(out, f"c4_{i}_", sol, probe_points(sol.setup.z, p.size))
        print(f"c4 layer {i} done {time.time() - t0:.1f} s", flush=True)
    out["c4_seed"] = np.array(C4_SEED)
    np.savez_compressed(os.path.join(HERE, "configs.npz"), **out)
    print(f"wrote configs.npz in {time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
