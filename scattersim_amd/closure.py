"""Learned closure: a multilayer perceptron that maps a low-order coupled-wave solution and the layer's
permittivity Fourier coefficients to the high-order reflection and transmission coefficients.

No counterpart in the reference, which only has the plain solve.  The network is trained on pairs the
solver generates (low orders in, the high-order solve as target and reference) and runs entirely in the HIP
kernels of libcwx: BF16 GEMMs with FP32 accumulation on the matrix cores, GELU, the mean-squared error
and Adam (``cwx_bf16_gemm``, ``cwx_mlp_*``).  Host code only prepares features and moves data.

The target is the correction to the low-order solution: x_high on the central band of orders minus the
low-order solution placed at its orders, so that the method without the network (the low-order solution,
zero elsewhere) is the zero prediction.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import DeviceArray, check, lib


def _bf16_bits(x: np.ndarray) -> np.ndarray:
    """float32 -> BF16 bit patterns (nearest, ties to even), as uint16."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    r = np.where(special, (u >> 16) | np.where((u & 0xFFFF) != 0, np.uint64(0x40), np.uint64(0)), r)
    return r.astype(np.uint16)


def _p(a: DeviceArray, offset_elems: int = 0) -> C.c_void_p:
    return C.c_void_p(a.ptr + offset_elems * a.dtype.itemsize)


def _pad4(n: int) -> int:
    """Input width rounded up to a multiple of 4, so the forward product reads 8-byte BF16 fragments."""
    return (n + 3) // 4 * 4


def _pad_cols(a: np.ndarray, n: int) -> np.ndarray:
    if a.shape[-1] == n:
        return a
    out = np.zeros(a.shape[:-1] + (n,), a.dtype)
    out[..., :a.shape[-1]] = a
    return out


class MLP:
    """Layers dims[0] -> dims[1] -> ... -> dims[-1], GELU between them, on the device.

    FP32 master weights and Adam moments, BF16 copies for the GEMMs; activations of up to ``batch`` rows."""

    def __init__(self, ctx, dims, batch: int, seed: int = 0, weights=None):
        self.ctx, self.dims, self.batch = ctx, list(dims), int(batch)
        rng = np.random.default_rng(seed)
        self.W, self.Wb, self.b, self.mW, self.vW, self.mb, self.vb, self.gW, self.gb = ([] for _ in range(9))
        for i in range(len(dims) - 1):
            nin, nout = dims[i], dims[i + 1]
            if weights is not None:
                w0 = np.asarray(weights[2 * i], np.float32)
                b0 = np.asarray(weights[2 * i + 1], np.float32)
            else:
                w0 = (rng.standard_normal((nout, nin)) * np.sqrt(1.0 / nin)).astype(np.float32)
                b0 = np.zeros(nout, np.float32)
            assert w0.shape == (nout, nin) and b0.shape == (nout,)
            self.W.append(DeviceArray.from_numpy(ctx, w0))
            self.Wb.append(DeviceArray.from_numpy(ctx, _bf16_bits(w0)))
            self.b.append(DeviceArray.from_numpy(ctx, b0))
            for lst, shape in ((self.mW, (nout, nin)), (self.vW, (nout, nin)), (self.gW, (nout, nin)),
                               (self.mb, (nout,)), (self.vb, (nout,)), (self.gb, (nout,))):
                lst.append(DeviceArray.from_numpy(ctx, np.zeros(shape, np.float32)))
        B = self.batch
        self.Z = [DeviceArray(ctx, (B, d), np.float32) for d in dims[1:]]          # pre-activations
        self.H = [DeviceArray(ctx, (B, d), np.uint16) for d in dims[1:-1]]         # hidden activations
        self.Y = DeviceArray(ctx, (B, dims[-1]), np.float32)
        self.acc = DeviceArray(ctx, (B, max(dims)), np.float32)
        self.dH = DeviceArray(ctx, (B, max(dims)), np.float32)
        self.dZb = DeviceArray(ctx, (B, max(dims)), np.uint16)
        self.dY = DeviceArray(ctx, (B, dims[-1]), np.float32)
        self.loss = DeviceArray(ctx, (1,), np.float32)
        self.t = 0

    def nparams(self) -> int:
        return sum(self.dims[i] * self.dims[i + 1] + self.dims[i + 1] for i in range(len(self.dims) - 1))

    def forward(self, X: DeviceArray, row0: int, rows: int, keep: bool = True) -> DeviceArray:
        """Rows row0 .. row0+rows of X (BF16, dims[0] columns) through the network; returns self.Y.
        keep=False (inference): the pre-activations that only the backward pass reads are not written."""
        L = lib()
        h = self.ctx.h
        assert rows <= self.batch and X.dtype == np.uint16 and X.shape[1] == self.dims[0]
        Hin, ld_in = _p(X, row0 * self.dims[0]), self.dims[0]
        for i in range(len(self.dims) - 1):
            nin, nout = self.dims[i], self.dims[i + 1]
            last = i == len(self.dims) - 2
            outs = (_p(self.Z[i]) if keep else None, None if last else _p(self.H[i]), _p(self.Y) if last else None)
            if nin % 4 == 0:
                # one launch: product, bias and activation
                check(L.cwx_mlp_dense_fwd(h, rows, nout, nin, Hin, ld_in, _p(self.Wb[i]), nin, _p(self.b[i]),
                                          0 if last else 1, *outs))
            else:
                check(L.cwx_bf16_gemm(h, rows, nout, nin, Hin, ld_in, 1, _p(self.Wb[i]), 1, nin, 1.0, 0.0,
                                      _p(self.acc), nout))
                check(L.cwx_mlp_act_fwd(h, rows, nout, 0 if last else 1, _p(self.acc), _p(self.b[i]), *outs))
            if not last:
                Hin, ld_in = _p(self.H[i]), nout
        return self.Y

    def backward(self, X: DeviceArray, row0: int, rows: int, dY: DeviceArray):
        """Gradients into gW / gb from dY (FP32, rows x dims[-1]) after forward on the same rows."""
        L = lib()
        h = self.ctx.h
        n = len(self.dims) - 1
        dH_ptr = _p(dY)
        for i in range(n - 1, -1, -1):
            nin, nout = self.dims[i], self.dims[i + 1]
            last = i == n - 1
            check(L.cwx_mlp_act_bwd(h, rows, nout, 0 if last else 1, _p(self.Z[i]), dH_ptr, None, _p(self.dZb),
                                    _p(self.gb[i])))
            Hin = _p(X, row0 * self.dims[0]) if i == 0 else _p(self.H[i - 1])
            # dW (nout x nin) = dZ^T H
            check(L.cwx_bf16_gemm(h, nout, nin, rows, _p(self.dZb), 1, nout, Hin, nin, 1, 1.0, 0.0, _p(self.gW[i]),
                                  nin))
            if i > 0:
                # dH (rows x nin) = dZ W
                check(L.cwx_bf16_gemm(h, rows, nin, nout, _p(self.dZb), nout, 1, _p(self.Wb[i]), nin, 1, 1.0, 0.0,
                                      _p(self.dH), nin))
                dH_ptr = _p(self.dH)

    def step(self, lr: float):
        L = lib()
        self.t += 1
        for i in range(len(self.dims) - 1):
            check(L.cwx_mlp_adam(self.ctx.h, self.W[i].shape[0] * self.W[i].shape[1], lr, 0.9, 0.999, 1e-8, self.t,
                                 _p(self.W[i]), _p(self.gW[i]), _p(self.mW[i]), _p(self.vW[i]), _p(self.Wb[i])))
            check(L.cwx_mlp_adam(self.ctx.h, self.b[i].shape[0], lr, 0.9, 0.999, 1e-8, self.t, _p(self.b[i]),
                                 _p(self.gb[i]), _p(self.mb[i]), _p(self.vb[i]), None))

    def mse(self, T: DeviceArray, row0: int, rows: int) -> None:
        check(lib().cwx_mlp_mse(self.ctx.h, rows, self.dims[-1], _p(self.Y), _p(T, row0 * self.dims[-1]), _p(self.dY),
                                None, _p(self.loss)))

    def weights(self):
        out = []
        for i in range(len(self.dims) - 1):
            out += [self.W[i].numpy(), self.b[i].numpy()]
        return out


# ------------------------------------------------------------------------------------------------
# data: single heterogeneous layers (disks in a square cell) solved at two orders
# ------------------------------------------------------------------------------------------------

@dataclass
class ClosureConfig:
    """The learned-closure workload: random-disk single layers, low and high orders of the same solve."""
    grid: int = 128                 # pixels per side of the permittivity grid
    cell: float = 4.0               # um
    thickness: float = 1.0          # um
    lam: float = 0.5                # um
    disks: int = 8
    r_range: tuple = (0.1, 0.3)     # um
    n_range: tuple = (1.3, 1.7)
    low: int = 7                    # orders per axis of the input solve
    high: int = 21                  # orders per axis of the high-order reference solve
    band: int = 15                  # central orders per axis the network predicts
    eps_orders: int = 13            # permittivity Fourier coefficients per axis in the input
    hidden: tuple = (256, 256)

    def opts(self, M: int):
        return dict(lam=self.lam, size=(self.cell, self.cell, self.thickness), coef=(M, M), ex=(1, 0), ey=(0, 0),
                    n=(1.0, 1.0), no_cutoff=True, eig_mode=1)

    @property
    def d_in(self) -> int:
        return 2 * (6 * self.low * self.low + self.eps_orders * self.eps_orders)

    @property
    def d_out(self) -> int:
        return 2 * 6 * self.band * self.band


def disk_layers(cfg: ClosureConfig, count: int, seed: int) -> np.ndarray:
    """(count, grid, grid) complex refractive indices: background 1, cfg.disks periodic disks each."""
    rng = np.random.default_rng(seed)
    N = cfg.grid
    xs = (np.arange(N) + 0.5) * cfg.cell / N
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    out = np.ones((count, N, N), np.complex128)
    for s in range(count):
        for _ in range(cfg.disks):
            r = rng.uniform(*cfg.r_range)
            cx, cy = rng.uniform(0, cfg.cell, 2)
            dx = (X - cx + cfg.cell / 2) % cfg.cell - cfg.cell / 2
            dy = (Y - cy + cfg.cell / 2) % cfg.cell - cfg.cell / 2
            out[s][dx * dx + dy * dy <= r * r] = rng.uniform(*cfg.n_range)
    return out


def order_map(m_from: int, m_to: int) -> np.ndarray:
    """Index in the m_to x m_to order list (column-major, zero order at the centre) of every order of the
    m_from x m_from list, for m_from <= m_to, both odd."""
    off = (m_to - m_from) // 2
    r = np.tile(np.arange(m_from), m_from)
    c = np.repeat(np.arange(m_from), m_from)
    return (r + off) + (c + off) * m_to


def place(x: np.ndarray, m_from: int, m_to: int) -> np.ndarray:
    """Coefficient vectors [Rx Ry Rz Tx Ty Tz] at m_from^2 orders each, embedded in m_to^2 orders."""
    x = np.asarray(x)
    MFf, MFt = m_from * m_from, m_to * m_to
    idx = order_map(m_from, m_to)
    out = np.zeros(x.shape[:-1] + (6 * MFt,), x.dtype)
    for c in range(6):
        out[..., c * MFt + idx] = x[..., c * MFf:(c + 1) * MFf]
    return out


def take(x: np.ndarray, m_from: int, m_to: int) -> np.ndarray:
    """The central m_to^2 orders of coefficient vectors at m_from^2 orders (m_to <= m_from)."""
    MFf, MFt = m_from * m_from, m_to * m_to
    idx = order_map(m_to, m_from)
    return np.concatenate([x[..., c * MFf + idx] for c in range(6)], axis=-1)


def eps_features(layers: np.ndarray, m: int) -> np.ndarray:
    """Centred m x m block of the normalised DFT of n^2 of every layer, flattened row by row."""
    N = layers.shape[-1]
    F = np.fft.fftshift(np.fft.fft2(layers ** 2), axes=(-2, -1)) / (N * N)
    c0 = N // 2 - m // 2
    return F[:, c0:c0 + m, c0:c0 + m].reshape(len(layers), m * m)


def solve_orders(ctx, cfg: ClosureConfig, layers: np.ndarray, M: int, chunk: int = 512, log=None) -> np.ndarray:
    """[R; T] (count, 6 M^2) of every layer at M x M orders through the batched solve."""
    from . import solve_batch
    out = []
    for i in range(0, len(layers), chunk):
        if log:
            log(f"{M}x{M} orders: layers {i} .. {min(i + chunk, len(layers))} of {len(layers)}")
        part = layers[i:i + chunk]
        x, info, (M0, M1) = solve_batch(ctx, [p[None] for p in part], [cfg.opts(M)] * len(part))
        if np.any(info):
            raise RuntimeError(f"solve failed for {int(np.count_nonzero(info))} layers at {M}x{M} orders")
        assert (M0, M1) == (M, M)
        out.append(np.asarray(x))
    return np.concatenate(out)


def _ri(z: np.ndarray) -> np.ndarray:
    return np.concatenate([z.real, z.imag], axis=-1)


def features(cfg: ClosureConfig, layers: np.ndarray, x_low: np.ndarray) -> np.ndarray:
    """Network inputs: the low-order [R; T] and the central n^2 Fourier coefficients, real and imaginary."""
    return np.concatenate([_ri(x_low), _ri(eps_features(layers, cfg.eps_orders))], axis=1).astype(np.float32)


@dataclass
class Dataset:
    X: np.ndarray        # (count, d_in) float32 features (unnormalised)
    T: np.ndarray        # (count, d_out) float32 correction targets (unnormalised)
    x_low: np.ndarray    # (count, 6 low^2) complex
    x_high: np.ndarray   # (count, 6 high^2) complex


def make_dataset(ctx, cfg: ClosureConfig, count: int, seed: int, layers=None, log=None) -> Dataset:
    if layers is None:
        layers = disk_layers(cfg, count, seed)
    x_low = solve_orders(ctx, cfg, layers, cfg.low, log=log)
    x_high = solve_orders(ctx, cfg, layers, cfg.high, log=log)
    X = features(cfg, layers, x_low)
    corr = take(x_high, cfg.high, cfg.band) - place(x_low, cfg.low, cfg.band)
    return Dataset(X=X, T=_ri(corr).astype(np.float32), x_low=x_low, x_high=x_high)


@dataclass
class Closure:
    """A trained network with the normalisation of its inputs and outputs."""
    cfg: ClosureConfig
    weights: list
    x_mean: np.ndarray
    x_std: np.ndarray
    t_std: np.ndarray
    history: list = field(default_factory=list)

    def save(self, path: str):
        arrs = {f"w{i}": w for i, w in enumerate(self.weights)}
        np.savez_compressed(path, x_mean=self.x_mean, x_std=self.x_std, t_std=self.t_std,
                            dims=np.array([self.cfg.d_in, *self.cfg.hidden, self.cfg.d_out]),
                            cfg=np.array([self.cfg.grid, self.cfg.cell, self.cfg.thickness, self.cfg.lam,
                                          self.cfg.disks, self.cfg.r_range[0], self.cfg.r_range[1],
                                          self.cfg.n_range[0], self.cfg.n_range[1], self.cfg.low, self.cfg.high,
                                          self.cfg.band, self.cfg.eps_orders]), **arrs)

    @staticmethod
    def load(path: str) -> "Closure":
        z = np.load(path, allow_pickle=False)
        c = z["cfg"]
        dims = [int(d) for d in z["dims"]]
        cfg = ClosureConfig(grid=int(c[0]), cell=float(c[1]), thickness=float(c[2]), lam=float(c[3]),
                            disks=int(c[4]), r_range=(float(c[5]), float(c[6])), n_range=(float(c[7]), float(c[8])),
                            low=int(c[9]), high=int(c[10]), band=int(c[11]), eps_orders=int(c[12]),
                            hidden=tuple(dims[1:-1]))
        nw = 2 * (len(dims) - 1)
        return Closure(cfg=cfg, weights=[z[f"w{i}"] for i in range(nw)], x_mean=z["x_mean"], x_std=z["x_std"],
                       t_std=z["t_std"])

    def predict(self, ctx, X: np.ndarray, x_low: np.ndarray, batch: int = 1024) -> np.ndarray:
        """High-order [R; T] (count, 6 high^2): the low-order solution plus the network's correction."""
        cfg = self.cfg
        d0 = _pad4(cfg.d_in)
        w = list(self.weights)
        w[0] = _pad_cols(np.asarray(w[0], np.float32), d0)
        net = MLP(ctx, [d0, *cfg.hidden, cfg.d_out], batch, weights=w)
        Xn = _bf16_bits(_pad_cols(((X - self.x_mean) / self.x_std).astype(np.float32), d0))
        dX = DeviceArray.from_numpy(ctx, Xn)
        out = np.zeros((len(X), cfg.d_out), np.float32)
        for r0 in range(0, len(X), batch):
            rows = min(batch, len(X) - r0)
            Y = net.forward(dX, r0, rows, keep=False)
            out[r0:r0 + rows] = Y.numpy()[:rows]
        corr = out * self.t_std
        h = corr.shape[1] // 2
        c = corr[:, :h] + 1j * corr[:, h:]
        return place(place(x_low, cfg.low, cfg.band) + c, cfg.band, cfg.high)

    def baseline(self, x_low: np.ndarray) -> np.ndarray:
        """The method without the network: the low-order solution, zero at the other orders."""
        return place(x_low, self.cfg.low, self.cfg.high)


def train(ctx, cfg: ClosureConfig, data: Dataset, epochs: int, batch: int = 256, lr: float = 1e-3, seed: int = 0,
          log=None, schedule: str = "const") -> Closure:
    """Train the closure on data with Adam (mini-batches, reshuffled every epoch); returns the network.
    schedule "cosine" decays the step size from lr to lr / 20 over the epochs along a half cosine."""
    x_mean = data.X.mean(axis=0)
    x_std = data.X.std(axis=0) + 1e-12
    t_std = np.full(data.T.shape[1], data.T.std() + 1e-30, np.float32)
    d0 = _pad4(cfg.d_in)
    Xn = _pad_cols(((data.X - x_mean) / x_std).astype(np.float32), d0)
    Tn = (data.T / t_std).astype(np.float32)
    net = MLP(ctx, [d0, *cfg.hidden, cfg.d_out], batch, seed=seed)
    rng = np.random.default_rng(seed)
    n = len(Xn)
    hist = []
    for ep in range(epochs):
        lr_ep = lr
        if schedule == "cosine":
            lr_ep = lr * (0.05 + 0.95 * 0.5 * (1.0 + np.cos(np.pi * ep / max(epochs, 1))))
        perm = rng.permutation(n)
        dX = DeviceArray.from_numpy(ctx, _bf16_bits(Xn[perm]))
        dT = DeviceArray.from_numpy(ctx, Tn[perm])
        tot, nb = 0.0, 0
        for r0 in range(0, n - batch + 1, batch):
            net.forward(dX, r0, batch)
            net.mse(dT, r0, batch)
            net.backward(dX, r0, batch, net.dY)
            net.step(lr_ep)
            tot += float(net.loss.numpy()[0])
            nb += 1
        hist.append(tot / max(nb, 1))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.4e}")
        dX.free()
        dT.free()
    w = net.weights()
    w[0] = np.ascontiguousarray(w[0][:, :cfg.d_in])   # the padding columns only ever meet zero inputs
    return Closure(cfg=cfg, weights=w, x_mean=x_mean.astype(np.float32), x_std=x_std.astype(np.float32),
                   t_std=t_std, history=hist)


def rel_err_rows(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def save_bf16(net: Closure, path: str):
    """Store a trained closure with its weights as BF16 bit patterns (what inference uses)."""
    w = [_bf16_bits(a) if i % 2 == 0 else np.asarray(a, np.float32) for i, a in enumerate(net.weights)]
    c = Closure(cfg=net.cfg, weights=w, x_mean=net.x_mean, x_std=net.x_std, t_std=net.t_std)
    c.save(path)


def load(path: str) -> Closure:
    """A closure saved by save_bf16 (BF16 weight matrices) or Closure.save (FP32)."""
    c = Closure.load(path)
    c.weights = [((a.astype(np.uint32) << 16).view(np.float32) if a.dtype == np.uint16 else a) for a in c.weights]
    return c


def time_forward(ctx, rows: int, path: str = None):
    """The shipped network's inference forward pass on `rows` random feature rows: 3 untimed and 10 timed
    passes on the library's stream; prints ms per pass and TFLOP/s."""
    import os
    import time
    net = load(path or os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "closure_c4.npz"))
    cfg = net.cfg
    d0 = _pad4(cfg.d_in)
    dims = [d0, *cfg.hidden, cfg.d_out]
    w = list(net.weights)
    w[0] = _pad_cols(np.asarray(w[0], np.float32), d0)
    mlp = MLP(ctx, dims, rows, weights=w)
    X = _bf16_bits(np.random.default_rng(0).standard_normal((rows, d0)).astype(np.float32))
    dX = DeviceArray.from_numpy(ctx, X)
    for _ in range(3):
        mlp.forward(dX, 0, rows, keep=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        mlp.forward(dX, 0, rows, keep=False)
    ctx.synchronize()
    t = (time.perf_counter() - t0) / 10
    fl = 2.0 * rows * sum(dims[i] * dims[i + 1] for i in range(len(dims) - 1))
    print(f"forward {t * 1e3:.3f} ms, {fl / t / 1e12:.1f} TFLOP/s (dims {dims}, {rows} rows)", flush=True)


def main(argv=None):
    """Generate the closure's training pairs with the solver, train on the GPU, evaluate on held-out layers
    and save the network:  python -m scattersim_amd.closure --count 20000 --test 1000 --out c4.npz"""
    import argparse
    import json
    import time
    from . import Context
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=20000, help="training pairs")
    ap.add_argument("--test", type=int, default=1000, help="held-out layers")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--eps-orders", type=int, default=ClosureConfig.eps_orders,
                    help="permittivity Fourier coefficients per axis in the input")
    ap.add_argument("--schedule", choices=["const", "cosine"], default="const")
    ap.add_argument("--out", help="where the trained network is saved")
    ap.add_argument("--time-forward", type=int, default=0, metavar="ROWS",
                    help="only time the shipped network's inference forward pass at ROWS rows (3 + 10 passes)")
    a = ap.parse_args(argv)
    ctx = Context(0)
    if a.time_forward:
        time_forward(ctx, a.time_forward)
        ctx.close()
        return
    if not a.out:
        ap.error("--out is required for training")
    cfg = ClosureConfig(eps_orders=a.eps_orders)
    t0 = time.time()
    data = make_dataset(ctx, cfg, a.count + a.test, seed=a.seed, log=lambda s: print(s, flush=True))
    t_data = time.time() - t0
    print(f"{a.count + a.test} pairs ({cfg.low}x{cfg.low} and {cfg.high}x{cfg.high} orders) in {t_data:.1f} s",
          flush=True)
    n = a.count
    tr = Dataset(X=data.X[:n], T=data.T[:n], x_low=data.x_low[:n], x_high=data.x_high[:n])
    t0 = time.time()
    net = train(ctx, cfg, tr, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=0,
                log=lambda s: print(s, flush=True), schedule=a.schedule)
    t_train = time.time() - t0
    te = slice(n, n + a.test)
    pred = net.predict(ctx, data.X[te], data.x_low[te])
    base = net.baseline(data.x_low[te])
    e_net = rel_err_rows(pred, data.x_high[te])
    e_base = rel_err_rows(base, data.x_high[te])
    save_bf16(net, a.out)
    print(json.dumps({"pairs": n, "test": a.test, "data_s": t_data, "train_s": t_train, "epochs": a.epochs,
                      "eps_orders": cfg.eps_orders, "schedule": a.schedule,
                      "params": sum(int(np.prod(w.shape)) for w in net.weights),
                      "held_out_rel_err_low_order": float(e_base.mean()),
                      "held_out_rel_err_closure": float(e_net.mean()),
                      "loss_first": net.history[0], "loss_last": net.history[-1]}), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
