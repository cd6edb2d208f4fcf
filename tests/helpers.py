"""Shared helpers of the test-suite: seeded samples and field comparisons through the oracle."""
from __future__ import annotations

import os

import numpy as np

from oracle import cw_oracle as o

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name: str):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def disks_sample(N: int, L: int, cell: float, ndisk: int, n_bg: complex, n_disk, seed: int, rmin=0.1, rmax=0.3):
    """(L, N, N) complex128 grid of refractive index: background plus periodic random disks."""
    rng = np.random.default_rng(seed)
    out = np.full((L, N, N), n_bg, dtype=np.complex128)
    xs = (np.arange(N) + 0.5) * cell / N
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    for l in range(L):
        for _ in range(ndisk):
            r = rng.uniform(rmin, rmax)
            cx, cy = rng.uniform(0, cell, 2)
            dx = (X - cx + cell / 2) % cell - cell / 2
            dy = (Y - cy + cell / 2) % cell - cell / 2
            nd = n_disk(rng) if callable(n_disk) else n_disk
            out[l][dx ** 2 + dy ** 2 <= r ** 2] = nd
    return out


def cw_from_records(records: np.ndarray, counts: np.ndarray, z) -> o.CW:
    """Split a flat record array (Pi, Pr_0, Pt_0, Pr_1, ...) into the oracle's CW structure."""
    lists, off = [], 0
    for c in counts:
        lists.append(records[off:off + c])
        off += c
    return o.CW(Pi=lists[0], z=list(z), Pr=lists[1::2], Pt=lists[2::2])


def probe_points(z, cell, seed=0, per_region=6):
    """Points strictly inside every region (above, each layer, below) of a stack with boundaries z."""
    rng = np.random.default_rng(seed)
    zs = [z[0] - 0.7]
    for a, b in zip(z[:-1], z[1:]):
        zs += [a + 0.3 * (b - a), a + 0.77 * (b - a)]
    zs.append(z[-1] + 0.6)
    pts = []
    for zz in zs:
        for _ in range(per_region):
            pts.append([rng.uniform(0, cell[0]), rng.uniform(0, cell[1]), zz])
    return np.array(pts)


def fields(cw: o.CW, pts: np.ndarray) -> np.ndarray:
    W, zl, b, e = o.unpack(cw, np.complex128)
    return o.field_points(W, zl, b, e, pts, tol=0.0)


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def multiset_err(a, b) -> float:
    """Relative distance of two eigenvalue multisets: each value of ``a`` is matched to one of ``b`` by a
    minimum-cost assignment, so that near-ties in any sort order cannot pair the wrong values."""
    from scipy.optimize import linear_sum_assignment
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    assert a.shape == b.shape
    cost = np.abs(a[:, None] - b[None, :])
    r, c = linear_sum_assignment(cost)
    return float(np.max(cost[r, c]) / max(np.max(np.abs(b)), 1e-300))
