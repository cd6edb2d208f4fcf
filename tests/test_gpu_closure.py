"""GPU checks of the learned closure: the BF16 GEMM and the network's kernels against the CPU oracle
(oracle/closure_oracle.py), the gradients of a whole network, and what a trained network gains over the
low-order solution on solves it has not seen."""
import numpy as np
import pytest

import scattersim_amd as sa
from scattersim_amd import closure as cl
from oracle import closure_oracle as co

pytestmark = pytest.mark.gpu


def dev(ctx, a):
    return sa.DeviceArray.from_numpy(ctx, np.ascontiguousarray(a))


def bf(a):
    return cl._bf16_bits(np.asarray(a, np.float32))


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (17, 33, 9), (64, 64, 16), (130, 70, 200), (256, 300, 37)])
@pytest.mark.parametrize("layout", ["nn", "nt", "tn"])
def test_bf16_gemm_matches_fp32_accumulation(ctx, M, N, K, layout):
    rng = np.random.default_rng(M * 1000 + N + K)
    A = co.bf16(rng.standard_normal((M, K)))
    B = co.bf16(rng.standard_normal((K, N)))
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    # storage of A and B per layout, with the strides that read them back as (M x K) and (K x N)
    if layout == "nn":
        Ast, sam, sak = A, K, 1
        Bst, sbk, sbn = B, N, 1
    elif layout == "nt":
        Ast, sam, sak = A, K, 1
        Bst, sbk, sbn = np.ascontiguousarray(B.T), 1, K
    else:
        Ast, sam, sak = np.ascontiguousarray(A.T), 1, M
        Bst, sbk, sbn = B, N, 1
    dA, dB, dC = dev(ctx, bf(Ast)), dev(ctx, bf(Bst)), dev(ctx, C0)
    sa.check(sa.lib().cwx_bf16_gemm(ctx.h, M, N, K, cl._p(dA), sam, sak, cl._p(dB), sbk, sbn, 1.5, 0.5, cl._p(dC), N))
    ref = 1.5 * (A.astype(np.float64) @ B.astype(np.float64)) + 0.5 * C0
    got = dC.numpy()
    assert np.max(np.abs(got - ref)) <= 1e-5 * max(1.0, np.max(np.abs(ref))) * np.sqrt(K)


def test_activation_loss_and_adam_kernels(ctx):
    rng = np.random.default_rng(3)
    M, N = 37, 45
    Zacc = rng.standard_normal((M, N)).astype(np.float32) * 2
    b = rng.standard_normal(N).astype(np.float32)
    dZacc, db_ = dev(ctx, Zacc), dev(ctx, b)
    Z = sa.DeviceArray(ctx, (M, N), np.float32)
    H = sa.DeviceArray(ctx, (M, N), np.uint16)
    Hf = sa.DeviceArray(ctx, (M, N), np.float32)
    L = sa.lib()
    sa.check(L.cwx_mlp_act_fwd(ctx.h, M, N, 1, cl._p(dZacc), cl._p(db_), cl._p(Z), cl._p(H), cl._p(Hf)))
    g, dg = co.gelu(Zacc.astype(np.float64) + b)
    assert np.max(np.abs(Hf.numpy() - g)) < 1e-5
    Hb = (H.numpy().astype(np.uint32) << 16).view(np.float32)
    assert np.max(np.abs(Hb - co.bf16(Hf.numpy()))) == 0
    dH = rng.standard_normal((M, N)).astype(np.float32)
    ddH = dev(ctx, dH)
    dZ = sa.DeviceArray(ctx, (M, N), np.float32)
    dbias = sa.DeviceArray(ctx, (N,), np.float32)
    sa.check(L.cwx_mlp_act_bwd(ctx.h, M, N, 1, cl._p(Z), cl._p(ddH), cl._p(dZ), None, cl._p(dbias)))
    assert np.max(np.abs(dZ.numpy() - dH * dg)) < 1e-5
    assert np.max(np.abs(dbias.numpy() - (dH * dg).sum(0))) < 1e-4
    T = rng.standard_normal((M, N)).astype(np.float32)
    dT = dev(ctx, T)
    dY = sa.DeviceArray(ctx, (M, N), np.float32)
    loss = sa.DeviceArray(ctx, (1,), np.float32)
    sa.check(L.cwx_mlp_mse(ctx.h, M, N, cl._p(Hf), cl._p(dT), cl._p(dY), None, cl._p(loss)))
    lref, gref = co.mse(Hf.numpy(), T)
    assert abs(loss.numpy()[0] - lref) < 1e-5 * lref and np.max(np.abs(dY.numpy() - gref)) < 1e-8
    w = rng.standard_normal(100).astype(np.float32)
    gw = rng.standard_normal(100).astype(np.float32)
    dw, dgw = dev(ctx, w), dev(ctx, gw)
    m, v = dev(ctx, np.zeros(100, np.float32)), dev(ctx, np.zeros(100, np.float32))
    wb = sa.DeviceArray(ctx, (100,), np.uint16)
    wr, mr, vr = w.astype(np.float64), np.zeros(100), np.zeros(100)
    for step in (1, 2, 3):
        sa.check(L.cwx_mlp_adam(ctx.h, 100, 1e-2, 0.9, 0.999, 1e-8, step, cl._p(dw), cl._p(dgw), cl._p(m), cl._p(v),
                                cl._p(wb)))
        wr, mr, vr = co.adam(wr, gw.astype(np.float64), mr, vr, step, 1e-2)
    assert np.max(np.abs(dw.numpy() - wr)) < 1e-6
    assert np.array_equal(wb.numpy(), bf(dw.numpy()))


@pytest.mark.parametrize("M,N,K,ldx", [(1, 8, 8, 8), (129, 136, 72, 80), (300, 2700, 256, 256), (257, 256, 928, 936)])
@pytest.mark.parametrize("act", [0, 1])
def test_dense_layer_matches_fp32_accumulation(ctx, M, N, K, ldx, act):
    """The fused dense layer (Z = X W^T + bias, H = act(Z)) at widths that are multiples of 8, which take the
    LDS-staged double-depth kernel: row and column tails, a depth that is not a multiple of 32, a row
    stride wider than K, against FP64 products of the same BF16 values."""
    rng = np.random.default_rng(M + N + K + act)
    X = co.bf16(rng.standard_normal((M, ldx)))
    W = co.bf16(rng.standard_normal((N, K)))
    b = rng.standard_normal(N).astype(np.float32)
    dX, dW, db = dev(ctx, bf(X)), dev(ctx, bf(W)), dev(ctx, b)
    Z = sa.DeviceArray(ctx, (M, N), np.float32)
    H = sa.DeviceArray(ctx, (M, N), np.uint16)
    Hf = sa.DeviceArray(ctx, (M, N), np.float32)
    sa.check(sa.lib().cwx_mlp_dense_fwd(ctx.h, M, N, K, cl._p(dX), ldx, cl._p(dW), K, cl._p(db), act, cl._p(Z),
                                        cl._p(H), cl._p(Hf)))
    zref = X[:, :K].astype(np.float64) @ W.astype(np.float64).T + b
    tol = 1e-5 * max(1.0, np.max(np.abs(zref))) * np.sqrt(K)
    assert np.max(np.abs(Z.numpy() - zref)) <= tol
    href = co.gelu(zref)[0] if act else zref
    assert np.max(np.abs(Hf.numpy() - href)) <= 2 * tol
    assert np.array_equal(H.numpy(), bf(Hf.numpy()))


@pytest.mark.parametrize("dims", [[37, 29, 41, 23], [36, 28, 40, 22], [40, 32, 48, 24]])
def test_network_gradients_match_oracle(ctx, dims):
    """Forward and backward of a three-layer network on the GPU against the oracle: odd widths take the
    strided GEMM and a separate bias/activation pass, widths that are multiples of 4 the fused dense layer
    (multiples of 8: its LDS-staged form)."""
    rng = np.random.default_rng(11)
    B = 19
    net = cl.MLP(ctx, dims, B, seed=5)
    params = [(net.W[i].numpy().astype(np.float64), net.b[i].numpy().astype(np.float64)) for i in range(3)]
    X = rng.standard_normal((B, dims[0])).astype(np.float32)
    T = rng.standard_normal((B, dims[-1])).astype(np.float32)
    dX, dT = dev(ctx, bf(X)), dev(ctx, T)
    Y = net.forward(dX, 0, B).numpy()
    net.mse(dT, 0, B)
    net.backward(dX, 0, B, net.dY)
    Yr, cache = co.forward(params, X)
    lref, dYr = co.mse(Yr, T)
    grads = co.backward(cache, dYr)
    assert np.linalg.norm(Y - Yr) / np.linalg.norm(Yr) < 2e-3
    assert abs(net.loss.numpy()[0] - lref) < 1e-2 * lref
    for i in range(3):
        gW, gb = net.gW[i].numpy(), net.gb[i].numpy()
        assert np.linalg.norm(gW - grads[i][0]) / np.linalg.norm(grads[i][0]) < 2e-2, i
        assert np.linalg.norm(gb - grads[i][1]) / np.linalg.norm(grads[i][1]) < 2e-2, i


def test_trained_closure_beats_low_order_solution(ctx):
    """A small closure (3x3 orders in, 9x9 reference solve, 32x32 grids) trained on 3000 solves: on 1000 held-out
    layers its high-order R/T are closer to the 9x9 solve than the low-order solution alone (what the method
    gains from the network)."""
    cfg = cl.ClosureConfig(grid=32, cell=2.0, thickness=0.5, lam=0.5, disks=4, low=3, high=9, band=7, eps_orders=5,
                           hidden=(256, 256))
    data = cl.make_dataset(ctx, cfg, 4000, seed=1)
    tr = cl.Dataset(X=data.X[:3000], T=data.T[:3000], x_low=data.x_low[:3000], x_high=data.x_high[:3000])
    net = cl.train(ctx, cfg, tr, epochs=60, batch=64, lr=1e-3, seed=0)
    assert net.history[-1] < 0.5 * net.history[0]
    te = slice(3000, 4000)
    pred = net.predict(ctx, data.X[te], data.x_low[te])
    base = net.baseline(data.x_low[te])
    e_net = cl.rel_err_rows(pred, data.x_high[te]).mean()
    e_base = cl.rel_err_rows(base, data.x_high[te]).mean()
    print(f"held-out relative error: low-order {e_base:.4f}, closure {e_net:.4f}")
    assert e_net < 0.85 * e_base
