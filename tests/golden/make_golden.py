"""Writes the committed fixtures of this directory from data files shipped with the reference.

    python tests/golden/make_golden.py /path/to/reference

* acw_te.npz  -- the 63 TE incidences of x64/a.cw (scatterlayer output for the stack
                 n = 1 | 1.4+0.05i | 1.4 | 1, boundaries z = -3, 0, 3, lambda = 1): for each incident
                 wave its record and the reflected/transmitted record at every boundary, as
                 (E0.x, E0.y, E0.z = 0, k.x, k.y, k.z) complex128 rows.
* acw_beam.npz -- all 4096 incident waves of x64/a.cw (the focused beam of scatterlayer's defaults,
                 alpha = 1, 64 x 64 polar samples about +z) as real (E0.x, E0.y, k.x, k.y, k.z) rows;
                 every imaginary part in the file is zero.
* psf.npz     -- data/psf.cw (scatterplane's defaults: a 64 x 64 polar beam of half-angle 0.5 about
                 (1, 0, 1), lambda = 1, one interface at z = 0 into 1.4+0.05i): all 4096 incident waves as
                 real (E0.x, E0.y, k.x, k.y, k.z) rows, and for the waves with k.y = 0 and E0.y != 0 (TE)
                 their reflected and transmitted records.
* samples.npz -- the input grids data/layers.npy, data/circle.npy, 3_npy.npy, data/cylinder.npy
                 (100 layers of 1 x 100: the M1 = 1 path with a 100-deep chain), data/triangle.npy
                 (225 x 225, odd) and data/RecBar.npy (stored in Fortran order, kept so).
The reference's records hold only (E0.x, E0.y, k); for ky = 0 TE waves E0.y is the global Ey.
"""
import os
import struct
import sys

import numpy as np


def read_cw80(path):
    buf = open(path, "rb").read()
    off = 0
    prec, = struct.unpack_from("<Q", buf, off)
    assert prec == 8
    off += 8

    def recs(off):
        n, = struct.unpack_from("<Q", buf, off)
        off += 8
        a = np.frombuffer(buf, np.complex128, n * 5, off).reshape(n, 5)
        off += n * 80
        return np.concatenate([a[:, :2], np.zeros((n, 1), np.complex128), a[:, 2:]], axis=1), off

    Pi, off = recs(off)
    nb, = struct.unpack_from("<Q", buf, off)
    off += 8
    z, Pr, Pt = [], [], []
    for _ in range(nb):
        zz, = struct.unpack_from("<d", buf, off)
        off += 8
        r, off = recs(off)
        t, off = recs(off)
        z.append(zz)
        Pr.append(r)
        Pt.append(t)
    assert off == len(buf), "trailing bytes"
    return Pi, np.array(z), Pr, Pt


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    Pi, z, Pr, Pt = read_cw80(os.path.join(ref, "x64", "a.cw"))
    te = np.nonzero((np.abs(Pi[:, 4]) == 0) & (np.abs(Pi[:, 1]) > 0))[0]
    np.savez_compressed(os.path.join(here, "acw_te.npz"), index=te, z=z, Pi=Pi[te],
                        Pr=np.stack([p[te] for p in Pr], 1), Pt=np.stack([p[te] for p in Pt], 1))
    raw = np.concatenate([Pi[:, 0:2], Pi[:, 3:6]], axis=1)
    assert np.all(raw.imag == 0)
    np.savez_compressed(os.path.join(here, "acw_beam.npz"), Pi=raw.real.copy())
    Pi, z, Pr, Pt = read_cw80(os.path.join(ref, "data", "psf.cw"))
    raw = np.concatenate([Pi[:, 0:2], Pi[:, 3:6]], axis=1)
    assert np.all(raw.imag == 0) and len(z) == 1
    te = np.nonzero((np.abs(Pi[:, 4]) == 0) & (np.abs(Pi[:, 1]) > 0))[0]
    np.savez_compressed(os.path.join(here, "psf.npz"), Pi=raw.real.copy(), z=z, index=te, Pr=Pr[0][te],
                        Pt=Pt[0][te])
    np.savez_compressed(os.path.join(here, "samples.npz"),
                        layers=np.load(os.path.join(ref, "data", "layers.npy")),
                        circle=np.load(os.path.join(ref, "data", "circle.npy")),
                        npy3=np.load(os.path.join(ref, "3_npy.npy")),
                        cylinder=np.load(os.path.join(ref, "data", "cylinder.npy")),
                        triangle=np.load(os.path.join(ref, "data", "triangle.npy")),
                        recbar=np.load(os.path.join(ref, "data", "RecBar.npy")))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
