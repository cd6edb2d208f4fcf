"""CPU FP64 oracle for scatterlayer: plane waves through a stack of homogeneous layers.

TEST INFRASTRUCTURE ONLY.  This module is the checker the HIP path is compared against.  Only the
test-suite (``tests/``), ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product path (``scattersim_amd``) never does.

It restates, in numpy, ScatterLayer.cpp (paths relative to the reference checkout):

  * ``ScatterLayer.cpp:327-344``  layer count and the padding of n / kappa / z to L layers
  * ``ScatterLayer.cpp:118-125``  complex indices ri (the first layer real) and boundary positions
  * ``ScatterLayer.cpp:132-146``  InitSz: s = direction * n0, sz_l = sqrt(ri_l^2 - sx^2 - sy^2)
  * ``ScatterLayer.cpp:152-165``  SetBoundaryConditions: layer 0 carries the incident E0, the last layer
                                  has no reflected wave
  * ``ScatterLayer.cpp:167-184``  SetGaussianConstraints (the unit direction, not s, in the x/y terms)
  * ``ScatterLayer.cpp:186-232``  SetBoundaryConstraints (equations 8-11 of each interface, phases from
                                  the layer thicknesses zn / zp)
  * ``ScatterLayer.cpp:240-266``  mat2waves (reflected wave of layer l, transmitted wave of layer l+1)
  * ``ScatterLayer.cpp:390-420``  the per-incidence solve and the .cw lists, each wave wound to its
                                  boundary (planewave::wind(0, 0, -z_l))

Unknowns are ordered as the reference orders them: x[6 l + 3 d + c] for layer l, direction d
(0 transmitted, 1 reflected) and component c.  The reference solves with a column-pivoted
Householder QR; any backward-stable solve agrees to rounding, and numpy's LU is used here.

Pins: ``tests/test_oracle_pins.py`` checks this restatement against the reference-produced
``x64/a.cw`` (every boundary's reflected and transmitted records of a 512-wave subset of its 4096
focused-beam incidences), against the Fresnel coefficients of a single interface and against
the transfer-matrix closed form of a slab.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class LayerStack:
    """The options of ScatterLayer.cpp:284-301 that define the stack (defaults as there)."""
    n: tuple = (1.0, 1.4, 1.4, 1.0)
    kappa: tuple = (0.05,)
    z: tuple = (-5.0, 0.0, 5.0)
    lam: float = 1.0


def expand(st: LayerStack):
    """ScatterLayer.cpp:327-344: L = max(len(n), len(kappa)+1, len(z)+1); pad n with its last value, kappa
    with zeros, z in steps of 10.  Returns (L, ri, z)."""
    n, kappa, z = list(st.n), list(st.kappa), list(st.z)
    L = len(n)
    if len(kappa) + 1 > L:
        L = len(kappa) + 1
    if len(z) + 1 > L:
        L = len(z) + 1
    n = n + [n[-1]] * (L - len(n))
    kappa = (kappa + [0.0] * (L - 1))[:L - 1]
    while len(z) + 1 < L:
        z.append(z[-1] + 10.0)
    ri = [complex(n[0], 0.0)] + [complex(n[l], kappa[l - 1]) for l in range(1, L)]
    return L, np.array(ri, np.complex128), np.array(z[:L - 1], np.float64)


def system(L: int, ri, z, k: float, d, E0):
    """A and b of one incidence: d is the unit propagation direction, E0 the incident amplitude."""
    n0 = ri[0].real
    s = np.asarray(d, np.float64) * n0
    sz = np.sqrt(ri ** 2 - s[0] ** 2 - s[1] ** 2)
    A = np.zeros((6 * L, 6 * L), np.complex128)
    b = np.zeros(6 * L, np.complex128)
    T, R = 0, 1
    col = lambda l, dd, c: 6 * l + 3 * dd + c
    # boundary conditions
    for c in range(3):
        A[c, col(0, T, c)] = 1.0
        A[3 + c, col(L - 1, R, c)] = 1.0
        b[c] = E0[c]
    # Gauss's law per wave (the unit direction in the transverse terms)
    for l in range(L - 1):
        A[6 + l, col(l, R, 0)] = d[0]
        A[6 + l, col(l, R, 1)] = d[1]
        A[6 + l, col(l, R, 2)] = -sz[l]
    start = 6 + L - 2
    for l in range(1, L):
        A[start + l, col(l, T, 0)] = d[0]
        A[start + l, col(l, T, 1)] = d[1]
        A[start + l, col(l, T, 2)] = sz[l]
    # interface constraints (equations 8-11)
    start = 6 + 2 * (L - 1)
    for l in range(L - 1):
        zn = z[0] if l == 0 else z[l] - z[l - 1]
        zp = 0.0 if l == L - 2 else z[l] - z[l + 1]
        en = np.exp(1j * k * sz[l] * zn)
        ep = np.exp(-1j * k * sz[l + 1] * zp)
        r = start + 4 * l
        A[r, col(l, T, 0)] = en
        A[r, col(l, R, 0)] = 1.0
        A[r, col(l + 1, T, 0)] = -1.0
        A[r, col(l + 1, R, 0)] = -ep
        A[r + 1, col(l, T, 1)] = en
        A[r + 1, col(l, R, 1)] = 1.0
        A[r + 1, col(l + 1, T, 1)] = -1.0
        A[r + 1, col(l + 1, R, 1)] = -ep
        A[r + 2, col(l, T, 2)] = d[1] * en
        A[r + 2, col(l, T, 1)] = -sz[l] * en
        A[r + 2, col(l, R, 2)] = d[1]
        A[r + 2, col(l, R, 1)] = sz[l]
        A[r + 2, col(l + 1, T, 2)] = -d[1]
        A[r + 2, col(l + 1, T, 1)] = sz[l + 1]
        A[r + 2, col(l + 1, R, 2)] = -d[1] * ep
        A[r + 2, col(l + 1, R, 1)] = -sz[l + 1] * ep
        A[r + 3, col(l, T, 0)] = sz[l] * en
        A[r + 3, col(l, T, 2)] = -d[0] * en
        A[r + 3, col(l, R, 0)] = -sz[l]
        A[r + 3, col(l, R, 2)] = -d[0]
        A[r + 3, col(l + 1, T, 0)] = -sz[l + 1]
        A[r + 3, col(l + 1, T, 2)] = d[0]
        A[r + 3, col(l + 1, R, 0)] = sz[l + 1] * ep
        A[r + 3, col(l + 1, R, 2)] = d[0] * ep
    return A, b, sz


def solve(st: LayerStack, incident: np.ndarray):
    """incident: (P, 6) records (E0.xyz, k.xyz) with real k.  Returns (Pr, Pt), each (P, L-1, 6): the
    reflected wave of layer l and the transmitted wave of layer l+1 at boundary l, wound to z_l."""
    L, ri, z = expand(st)
    kv = 2.0 * math.pi / st.lam                           # k_vac = k (ScatterLayer.cpp:351-352)
    P = len(incident)
    Pr = np.zeros((P, L - 1, 6), np.complex128)
    Pt = np.zeros((P, L - 1, 6), np.complex128)
    for p in range(P):
        kr = incident[p, 3:6].real
        d = kr / np.linalg.norm(kr)                       # planewave::getDirection
        A, b, sz = system(L, ri, z, kv, d, incident[p, 0:3])
        x = np.linalg.solve(A, b)
        s = d * ri[0].real
        for l in range(L - 1):
            r = np.array([x[6 * l + 3], x[6 * l + 4], x[6 * l + 5], s[0] * kv, s[1] * kv, -sz[l] * kv])
            t = np.array([x[6 * (l + 1)], x[6 * (l + 1) + 1], x[6 * (l + 1) + 2], s[0] * kv, s[1] * kv,
                          sz[l + 1] * kv])
            for out, w in ((Pr, r), (Pt, t)):
                w = w.copy()
                w[0:3] *= np.exp(1j * w[5] * (-z[l]))      # planewave::wind(0, 0, -z_l)
                out[p, l] = w
    return Pr, Pt


def reference_wave(st: LayerStack, direction=(0.0, 0.0, 1.0), ex=(0.0, 0.0), ey=(1.0, 0.0), ez=(0.0, 0.0)):
    """The single incident wave of ScatterLayer.cpp:346-353 (alpha = 0): orthogonalised E, k = n0 k_vac d."""
    d = np.asarray(direction, np.float64)
    d = d / np.linalg.norm(d)
    E = np.array([complex(*ex), complex(*ey), complex(*ez)])
    E = E / np.sqrt(E[0] ** 2 + E[1] ** 2 + E[2] ** 2)
    dc = d.astype(np.complex128)
    E = np.cross(dc, np.cross(E, dc))
    _, ri, _ = expand(st)
    k = d * ri[0].real * 2.0 * math.pi / st.lam
    return np.array([[E[0], E[1], E[2], k[0], k[1], k[2]]], np.complex128)
