"""Seeded synthetic samples and option sets of the benchmark configurations.

Each sample is a complex128 grid of refractive index in C order, (L, Ny, Nx), as scattervolume reads
it (the solver squares n itself, FourierWave.h:341).
"""
from __future__ import annotations

import numpy as np


def _wrap(d, period):
    return (d + period / 2) % period - period / 2


def disks(N: int, cell: float, ndisk: int, n_bg, n_disk, rng, rmin=0.1, rmax=0.3):
    """One N x N slice: background plus periodically wrapped disks of radius U[rmin, rmax]."""
    xs = (np.arange(N) + 0.5) * cell / N
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    out = np.full((N, N), n_bg, dtype=np.complex128)
    for _ in range(ndisk):
        r = rng.uniform(rmin, rmax)
        cx, cy = rng.uniform(0, cell, 2)
        nd = n_disk(rng) if callable(n_disk) else n_disk
        out[_wrap(X - cx, cell) ** 2 + _wrap(Y - cy, cell) ** 2 <= r * r] = nd
    return out


def plain_small(seed: int = 0):
    """One 128 x 128 layer of 8 disks (r 0.1-0.3 um, n = 1.4) in n = 1.0, 4 x 4 x 1 um cell,
    lambda 0.5 um, normal incidence, x-polarised, 7 x 7 orders (a 196 x 196 eigenproblem)."""
    rng = np.random.default_rng(seed)
    sample = disks(128, 4.0, 8, 1.0, 1.4 + 0j, rng)[None]
    opts = dict(lam=0.5, direction=(0, 0, 1), ex=(1, 0), ey=(0, 0), n=(1.0, 1.0), size=(4.0, 4.0, 1.0),
                coef=(7, 7))
    grid = dict(x0=0.0, y0=0.0, z0=-0.5 + 0.5 / 32, dx=4.0 / 128, dy=4.0 / 128, dz=1.0 / 32, nx=128, ny=128, nz=32)
    return sample, opts, grid


def volume32(seed: int = 0):
    """64 x 64 x 32 voxels sliced into 32 layers: random spheres with n ~ U[1.4, 1.6] and
    kappa ~ U[0, 0.01] in an n = 1.33 background, 8 x 8 x 4 um cell, lambda 0.6 um, 15 x 15 orders
    (900 x 900 eigenproblems), field on a 256 x 256 x 64 grid through the volume."""
    rng = np.random.default_rng(seed)
    L, N = 32, 64
    cell = (8.0, 8.0, 4.0)
    xs = (np.arange(N) + 0.5) * cell[0] / N
    zs = (np.arange(L) + 0.5) * cell[2] / L
    Z, Y, X = np.meshgrid(zs, xs, xs, indexing="ij")
    sample = np.full((L, N, N), 1.33 + 0j, dtype=np.complex128)
    for _ in range(14):
        r = rng.uniform(0.6, 1.4)
        c = rng.uniform(0, 1, 3) * np.asarray(cell)
        n = rng.uniform(1.4, 1.6) + 1j * rng.uniform(0.0, 0.01)
        d2 = _wrap(X - c[0], cell[0]) ** 2 + _wrap(Y - c[1], cell[1]) ** 2 + (Z - c[2]) ** 2
        sample[d2 <= r * r] = n
    opts = dict(lam=0.6, direction=(0, 0, 1), ex=(1, 0), ey=(0, 0), n=(1.33, 1.33), size=cell, coef=(15, 15))
    grid = dict(x0=0.0, y0=0.0, z0=-2.0 + 2.0 / 64, dx=8.0 / 256, dy=8.0 / 256, dz=4.0 / 64, nx=256, ny=256, nz=64)
    return sample, opts, grid


def _stack_masks(L: int, N: int, cell: float, rng):
    """Material masks of one seeded stack: every layer holds 3-6 disks (r 0.5-2 um) or 2-4 rectangles
    (sides 1-5 um), periodically wrapped in a cell x cell square."""
    xs = (np.arange(N) + 0.5) * cell / N
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    masks = np.zeros((L, N, N), dtype=bool)
    for l in range(L):
        if rng.random() < 0.5:
            for _ in range(int(rng.integers(3, 7))):
                r = rng.uniform(0.5, 2.0)
                cx, cy = rng.uniform(0, cell, 2)
                masks[l] |= _wrap(X - cx, cell) ** 2 + _wrap(Y - cy, cell) ** 2 <= r * r
        else:
            for _ in range(int(rng.integers(2, 5))):
                w, h = rng.uniform(1.0, 5.0, 2)
                cx, cy = rng.uniform(0, cell, 2)
                masks[l] |= (np.abs(_wrap(X - cx, cell)) <= w / 2) & (np.abs(_wrap(Y - cy, cell)) <= h / 2)
    return masks


SWEEP_CELL = 12.0


def sweep_stacks(seed: int = 0, nstacks: int = 16, N: int = 64):
    """The seeded stacks (material masks, (L, N, N) bool) of the batched sweep."""
    rng = np.random.default_rng(seed)
    return [_stack_masks(int(rng.integers(4, 25)), N, SWEEP_CELL, rng) for _ in range(nstacks)]


def sweep_instance(stack: int, lam: float, theta_deg: float, seed: int = 0, nstacks: int = 16, N: int = 64):
    """One instance of the sweep's family: stack `stack` of sweep_stacks() at wavelength `lam` (um) and
    incidence `theta_deg` in the xz plane.  Returns (sample (L, N, N) complex128, option dict)."""
    m = sweep_stacks(seed, nstacks, N)[stack]
    th = np.deg2rad(theta_deg)
    sample = np.where(m, 1.5 + 0.004 / lam ** 2, 1.0).astype(np.complex128)
    return sample, dict(lam=float(lam), direction=(float(np.sin(th)), 0.0, float(np.cos(th))), n=(1.0, 1.0),
                        size=(SWEEP_CELL, SWEEP_CELL, 0.1 * m.shape[0]), coef=(11, 11))


def field_c3_grid():
    """The mixed-precision field grid: 512 x 512 x 128 points over the 4 x 4 um cell of plain_small, z in
    [-1, 1] um (plane centres), so that the planes cross the layer and both half-spaces."""
    return dict(x0=0.0, y0=0.0, z0=-1.0 + 1.0 / 128, dx=4.0 / 512, dy=4.0 / 512, dz=2.0 / 128, nx=512, ny=512,
                nz=128)


def sweep(count: int = 2048, seed: int = 0, nstacks: int = 16, N: int = 64):
    """Batched sweep: 256 wavelengths in [0.4, 0.8] um x 8 incidence angles in [0, 35] deg (instance i is
    wavelength i // 8 at angle i % 8; the first `count` instances are returned).  Instance i uses stack
    i % nstacks, one of `nstacks` seeded stacks of 4-24 layers (0.1 um each) of disks or rectangles in
    n = 1.0 made of a Cauchy material n(lambda) = 1.5 + 0.004 / lambda^2; 12 x 12 um cell (11 x 11 orders
    stay below the cut-off at lambda 0.8 and 35 deg), 11 x 11 orders, s-polarised incidence in the xz plane.

    Returns (packed, Ls, params): the layers of all instances back to back, (sum L, N, N) complex128;
    the layer count and the option dict of every instance."""
    cell = SWEEP_CELL
    stacks = sweep_stacks(seed, nstacks, N)
    lams = np.linspace(0.4, 0.8, 256)
    thetas = np.deg2rad(np.linspace(0.0, 35.0, 8))
    Ls = np.array([stacks[i % nstacks].shape[0] for i in range(count)], np.int32)
    packed = np.empty((int(Ls.sum()), N, N), np.complex128)
    params = []
    off = 0
    for i in range(count):
        lam, th = float(lams[i // 8]), float(thetas[i % 8])
        m = stacks[i % nstacks]
        packed[off:off + m.shape[0]] = np.where(m, 1.5 + 0.004 / lam ** 2, 1.0)
        off += m.shape[0]
        params.append(dict(lam=lam, direction=(float(np.sin(th)), 0.0, float(np.cos(th))), n=(1.0, 1.0),
                           size=(cell, cell, 0.1 * m.shape[0]), coef=(11, 11)))
    return packed, Ls, params


CONFIGS = {"volume32": volume32, "plain_small": plain_small}
