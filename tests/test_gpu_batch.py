"""GPU batched solve (cwx_solve_batch_dev): many independent samples with uneven layer counts and their
own wavelength, incidence and half-space indices, solved together, against one-at-a-time solves of the
same library (bit-identical: every matrix sees the same arithmetic in either form) and the oracle."""
import numpy as np
import pytest

import scattersim_amd as sa
from oracle import cw_oracle as o
from tests.helpers import disks_sample, rel_err

pytestmark = pytest.mark.gpu


def instances(seed, count, M=(5, 3), grid=12):
    rng = np.random.default_rng(seed)
    samples, params = [], []
    for i in range(count):
        L = int(rng.integers(1, 5))
        lam = float(rng.uniform(0.6, 0.9))
        nd = 1.5 + 0.004 / lam ** 2
        samples.append(disks_sample(grid, L, 3.0, 2, 1.0, lambda r, nd=nd: nd + 0.01j * r.random(), seed=seed * 100 + i))
        th = float(rng.uniform(0.0, 0.5))
        params.append(dict(lam=lam, direction=(np.sin(th), 0.0, np.cos(th)), ex=(1, 0), ey=(0.2, 0.1),
                           size=(3.0, 3.0, 0.4 * L + 0.3), coef=M, n=(1.0, float(rng.uniform(1.0, 1.3))),
                           no_cutoff=True))
    return samples, params


@pytest.mark.parametrize("eig_mode", [0, 1])
def test_batch_equals_single_solves(ctx, eig_mode):
    samples, params = instances(7, 9)
    for p in params:
        p["eig_mode"] = eig_mode
    x, info, (M0, M1) = sa.solve_batch(ctx, samples, params)
    assert (M0, M1) == (5, 3) and x.shape == (9, 6 * 15)
    assert np.all(info == 0)
    for s, p, xb in zip(samples, params, x):
        sol = sa.scattervolume(ctx, s, **p)
        assert np.array_equal(sol.x, xb)
        sol.free()


def test_large_batch_equals_single_solves(ctx):
    # about 1 750 layers at 9x9 orders (162 x 162 eigenproblems): a batch this large takes the launch
    # shapes chosen for many matrices (two QR workgroups per CU, deflation windows, Hessenberg panels),
    # which must leave every instance's arithmetic as in its own solve
    samples, params = instances(11, 700, M=(9, 9), grid=12)
    for p in params:
        p["eig_mode"] = 1
    x, info, _ = sa.solve_batch(ctx, samples, params)
    assert np.all(info == 0)
    for i in (0, 351, 699):
        sol = sa.scattervolume(ctx, samples[i], **params[i])
        assert np.array_equal(sol.x, x[i]), i
        sol.free()


def test_batch_against_oracle_and_chunking(ctx):
    samples, params = instances(3, 6, M=(3, 3), grid=10)
    x_all, info, _ = sa.solve_batch(ctx, samples, params)
    # a tiny workspace budget forces one instance per chunk; results must not change
    x_chunk, info2, _ = sa.solve_batch(ctx, samples, params, max_bytes=1)
    assert np.array_equal(x_all, x_chunk) and np.all(info == 0) and np.all(info2 == 0)
    for s, p, xb in zip(samples, params, x_all):
        op = o.VolumeParams(lam=p["lam"], direction=p["direction"], ex=p["ex"], ey=p["ey"], size=p["size"],
                            coef=p["coef"], n=p["n"], no_cutoff=True)
        ref = o.solve(s, op)
        assert rel_err(xb, ref.x) < 1e-9


def test_batch_device_input_and_errors(ctx):
    samples, params = instances(5, 4)
    packed = np.ascontiguousarray(np.concatenate(samples, axis=0))
    Ls = [s.shape[0] for s in samples]
    d = sa.DeviceArray.from_numpy(ctx, packed)
    x_dev, info, _ = sa.solve_batch(ctx, d.ptr, params, shape=(Ls, packed.shape[1], packed.shape[2]))
    x_host, _, _ = sa.solve_batch(ctx, samples, params)
    assert np.array_equal(x_dev, x_host)
    # instances whose orders differ cannot share a batch
    bad = [dict(params[0]), dict(params[1], coef=(3, 3))]
    with pytest.raises(sa._capi.CwxError):
        sa.solve_batch(ctx, samples[:2], bad)
    d.free()


def test_batch_with_mixed_orders(ctx):
    """cwx_solve_batch_mixed_dev: instances at different orders in one call (grouped by orders), each equal bit
    for bit to its own solve; the cut-off picks the orders of one instance itself."""
    samples, params = instances(11, 7, M=(3, 3), grid=12)
    coefs = [(3, 3), (5, 5), (3, 3), (5, 3), (5, 5), (3, 5), (5, 3)]
    for p, c in zip(params, coefs):
        p["coef"] = c
    params[6]["no_cutoff"] = False
    xs, info, orders = sa.solve_batch_mixed(ctx, samples, params)
    assert np.all(info == 0) and len(xs) == 7
    for s, p, x, (M0, M1) in zip(samples, params, xs, orders):
        sol = sa.scattervolume(ctx, s, **p)
        assert (sol.M0, sol.M1) == (M0, M1) and x.shape == (6 * M0 * M1,)
        assert np.array_equal(sol.x, x)
        sol.free()
    assert orders[1] == (5, 5) and orders[5] == (3, 5)


@pytest.mark.parametrize("eig_mode,count,chunk", [(1, 96, 32), (0, 32, 16)])
def test_graph_replays_equal_stream_solve(ctx, eig_mode, count, chunk):
    """The batched sweep through captured HIP graphs (cwx_solve_batch_graph_dev): chunks of `chunk`
    instances (the sweep's 16 stacks in the same order in every chunk) replay the graph captured for the
    first one.  Every replay equals the stream-ordered batched solve bit for bit, on two calls with other
    work and synchronisations in between (the second call captures nothing)."""
    from scattersim_amd import workloads
    packed, Ls, params = workloads.sweep(count)
    for p in params:
        p["eig_mode"] = eig_mode
    N = packed.shape[1]
    x_ref, info, (M0, M1) = sa.solve_batch(ctx, [packed[a:a + l] for a, l in
                                                 zip(np.concatenate([[0], np.cumsum(Ls)[:-1]]), Ls)], params)
    assert not np.any(info)
    n6 = 6 * M0 * M1
    d_s = sa.DeviceArray.from_numpy(ctx, packed)
    d_x = sa.DeviceArray(ctx, (count, n6), np.complex128)
    for call in range(2):
        ginfo, captured = sa.solve_batch_graph(ctx, d_s.ptr, params, Ls, N, N, chunk, d_x.ptr)
        assert not np.any(ginfo)
        assert captured == (1 if call == 0 else 0)
        xg = d_x.numpy()
        for c0 in range(0, count, chunk):
            assert np.array_equal(xg[c0:c0 + chunk], x_ref[c0:c0 + chunk]), (call, c0)
        # eager work between the calls
        sol = sa.scattervolume(ctx, packed[:Ls[0]], **params[0])
        assert np.array_equal(sol.x, x_ref[0])
        sol.free()
        ctx.synchronize()
        d_x.numpy()[:] = 0
