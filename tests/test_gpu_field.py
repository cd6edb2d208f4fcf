"""GPU parity of the plane-wave field evaluator (gpuEvaluator.cu:28-165) against the FP64 oracle sum."""
import numpy as np
import pytest

import scattersim_amd as sa
from oracle import cw_oracle as o
from tests.helpers import golden, rel_err

pytestmark = pytest.mark.gpu


def random_table(seed, nwaves=(40, 120, 90, 60), z_layers=(-1.0, 0.0, 1.5), kscale=12.0, lossy=True):
    """Random waves per region: Pi+Pr0 | Pt0+Pr1 | Pt1+Pr2 | Pt2, as cw_allocate lays them out."""
    rng = np.random.default_rng(seed)
    n = sum(nwaves)
    W = np.zeros((n, 6), np.complex128)
    W[:, 0:3] = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    W[:, 3:5] = kscale * rng.uniform(-0.7, 0.7, (n, 2))
    W[:, 5] = kscale * rng.uniform(0.3, 1.0, n) * np.where(rng.random(n) < 0.5, 1, -1)
    if lossy:
        W[:, 3:6] += 1j * rng.uniform(0, 0.05, (n, 3))
    b = np.zeros(len(nwaves), np.int64)
    b[1:] = np.cumsum(nwaves)[:-1]
    e = np.cumsum(nwaves)
    return W, np.asarray(z_layers, np.float32), b, e


def to_struct(W, dtype):
    out = np.zeros(len(W), dtype)
    out["E0"] = W[:, 0:3]
    out["k"] = W[:, 3:6]
    return out


grid_points = o.grid_points


@pytest.mark.parametrize("tol_mode", [0, 1])
def test_grid_f64_table(ctx, tol_mode):
    W, zl, b, e = random_table(1)
    ctx.field_init_f64(to_struct(W, sa.WAVE_F64), zl, b, e)
    g = (-0.4, -0.3, -2.0, 0.05, 0.07, 0.125, 17, 9, 33)  # z crosses every boundary, incl. exactly z_l
    out = sa.DeviceArray(ctx, (g[8], g[7], g[6], 3), np.complex128)
    ctx.field_grid_dev(out.ptr, True, tol_mode, *g)
    ctx.synchronize()
    pts = grid_points(*g)
    ref = o.field_points(W, zl, b, e, pts, tol=0.001 if tol_mode == 0 else 0.0)
    got = out.numpy().reshape(-1, 3)
    assert rel_err(got, ref) < 1e-6


def test_grid_f32_table_reference_arithmetic(ctx):
    W, zl, b, e = random_table(2, kscale=6.0)
    W32 = W.astype(np.complex64)
    ctx.field_init_f32(to_struct(W32, sa.WAVE_F32), zl, b, e)
    g = (0.1, 0.2, -2.0, 0.1, 0.1, 0.2, 16, 8, 20)
    out = sa.DeviceArray(ctx, (g[8], g[7], g[6], 3), np.complex64)
    ctx.field_grid_dev(out.ptr, False, 0, *g)
    ctx.synchronize()
    ref = o.field_points(W32, zl, b, e, grid_points(*g), tol=0.001)
    # FP32 phases with |k.r| <~ 30: relative error of a few 1e-6 per term
    assert rel_err(out.numpy().reshape(-1, 3), ref) < 5e-5


def test_slices_match_oracle(ctx):
    W, zl, b, e = random_table(3)
    ctx.field_init_f64(to_struct(W, sa.WAVE_F64), zl, b, e)
    N, extent = 32, 3.0
    center = (0.1, -0.2, 0.05)
    plane = (0.3, -0.1, 0.4)
    d = np.float32(extent / (N - 1))
    start = np.asarray(center, np.float32) - np.float32(extent / 2)
    Exy, Exz, Eyz = ctx.field_evaluate(float(start[0]), float(start[1]), float(start[2]), *plane, float(d), N)
    for axis, got in ((2, Exy), (1, Exz), (0, Eyz)):
        pts = o.slice_points(axis, N, extent, center, plane)
        ref = o.field_points(W, zl, b, e, pts, tol=0.001).reshape(N, N, 3)
        assert rel_err(got, ref) < 1e-6, axis


def test_acw_fixture_field(ctx):
    """Evaluate the TE subset of the reference's a.cw (scatterlayer output) on the GPU."""
    g = golden("acw_te.npz")
    Pi, Pr, Pt, z = g["Pi"], g["Pr"], g["Pt"], g["z"]
    cw = o.CW(Pi=Pi, z=list(z), Pr=[Pr[:, l] for l in range(3)], Pt=[Pt[:, l] for l in range(3)])
    W, zl, b, e = o.unpack(cw, np.complex128)
    ctx.field_init_f64(to_struct(W, sa.WAVE_F64), zl, b, e)
    N = 64
    Exy, Exz, Eyz = ctx.field_evaluate(-2.0, -2.0, -5.0, 0.0, 0.0, 0.0, 10.0 / (N - 1), N)
    pts = o.slice_points(1, N, 10.0, (3.0, 3.0, 0.0), (0.0, 0.0, 0.0))
    ref = o.field_points(W, zl, b, e, pts, tol=0.001).reshape(N, N, 3)
    assert rel_err(Exz, ref) < 1e-6


@pytest.mark.parametrize("external", [False, True])
def test_separable_matches_direct_and_oracle(ctx, external):
    """Order-separable synthesis of a solution's field equals the direct sum over its .cw records."""
    from tests.helpers import disks_sample
    sample = disks_sample(12, 3, 3.0, 3, 1.0, lambda r: 1.45 + 0.02j * r.random(), seed=4)
    opts = dict(lam=0.7, direction=(0.2, 0.1, 1.0), ex=(1, 0), ey=(0.3, 0), size=(3.0, 3.0, 1.2), coef=(5, 3),
                n=(1.0, 1.2), external=external)
    sol = sa.scattervolume(ctx, sample, **opts)
    g = (-0.3, 0.1, -1.4, 0.11, 0.13, 0.1, 37, 21, 29)      # planes above, inside every layer and below
    sep = sa.DeviceArray(ctx, (g[8], g[7], g[6], 3), np.complex128)
    direct = sa.DeviceArray(ctx, (g[8], g[7], g[6], 3), np.complex128)
    ctx.field_solution_grid_dev(sol, sep.ptr, True, 0, *g)
    ctx.field_init_solution(sol)
    ctx.field_grid_dev(direct.ptr, True, 0, *g)
    ctx.synchronize()
    a, b = sep.numpy(), direct.numpy()
    assert rel_err(a, b) < 1e-6           # the direct FP64 kernel reduces phases to FP32 sincos
    p = o.VolumeParams(lam=0.7, direction=(0.2, 0.1, 1.0), ex=(1, 0), ey=(0.3, 0), size=(3.0, 3.0, 1.2),
                       coef=(5, 3), n=(1.0, 1.2), external=external)
    ref = o.build_cw(o.solve(sample, p))
    W, zl, bb, ee = o.unpack(ref, np.complex128)
    pts = grid_points(*g)
    E = o.field_points(W, zl, bb, ee, pts, tol=0.001).reshape(a.shape)
    assert rel_err(a, E) < 1e-9
    sol.free()


@pytest.mark.parametrize("external", [False, True])
def test_separable_complex64_matches_fp64(ctx, external):
    """The complex64 evaluator (FP64 coefficients and phase tables, FP32 matrix-core synthesis) stays
    within 1e-6 of the FP64 direct sum over the .cw records, over planes above, inside every layer and
    below; grid sizes that are not multiples of the 16 x 64 tiles are included."""
    from tests.helpers import disks_sample
    sample = disks_sample(16, 3, 3.0, 3, 1.0, lambda r: 1.45 + 0.02j * r.random(), seed=9)
    opts = dict(lam=0.7, direction=(0.2, 0.1, 1.0), ex=(1, 0), ey=(0.3, 0), size=(3.0, 3.0, 1.2), coef=(7, 5),
                n=(1.0, 1.2), external=external)
    sol = sa.scattervolume(ctx, sample, **opts)
    g = (-0.3, 0.1, -1.4, 0.031, 0.027, 0.05, 117, 75, 61)
    out = sa.DeviceArray(ctx, (g[8], g[7], g[6], 3), np.complex64)
    ref = sa.DeviceArray(ctx, (g[8], g[7], g[6], 3), np.complex128)
    ctx.field_solution_grid_dev(sol, out.ptr, False, 0, *g)
    ctx.field_solution_grid_dev(sol, ref.ptr, True, 0, *g)
    ctx.synchronize()
    a, b = out.numpy().astype(np.complex128), ref.numpy()
    assert rel_err(a, b) < 1e-6
    # and the FP64 separable form against the direct sum (region map and records)
    direct = sa.DeviceArray(ctx, (g[8], g[7], g[6], 3), np.complex128)
    ctx.field_init_solution(sol)
    ctx.field_grid_dev(direct.ptr, True, 0, *g)
    ctx.synchronize()
    assert rel_err(a, direct.numpy()) < 2e-6
    sol.free()
