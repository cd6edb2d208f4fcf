"""Many incident fields on one sample (cwx_solve_fields_dev): one eigen-decomposition, one transfer chain and one
boundary factorisation for K right-hand sides.  Parity against the FP64 oracle given the same incident Fourier
coefficients (ScatterVolume.cpp:660-674 replaced by the caller's), against the plain and --psf solves, and by
linearity in the incident field."""
import numpy as np
import pytest

import scattersim_amd as sa
from scattersim_amd import _capi
from oracle import cw_oracle as o
from tests.helpers import disks_sample, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-9


def case():
    sample = disks_sample(12, 3, 3.0, 3, 1.0, lambda r: 1.45 + 0.02j * r.random(), seed=31)
    opts = dict(lam=0.7, direction=(0.2, 0.1, 1.0), ex=(1, 0), ey=(0.3, 0), size=(3.0, 3.0, 1.2), coef=(5, 3),
                n=(1.0, 1.3))
    p = o.VolumeParams(lam=0.7, direction=(0.2, 0.1, 1.0), ex=(1, 0), ey=(0.3, 0), size=(3.0, 3.0, 1.2),
                       coef=(5, 3), n=(1.0, 1.3))
    return sample, opts, p


@pytest.mark.parametrize("eig_mode", [0, 1])
def test_single_field_equals_plain_solve(ctx, eig_mode):
    sample, opts, _ = case()
    sol = sa.scattervolume(ctx, sample, eig_mode=eig_mode, **opts)
    EF = sol.setup_array(_capi.SOL_EF)
    x = sa.solve_fields(ctx, sample, EF[None], eig_mode=eig_mode, **opts)
    assert x.shape == (1, 6 * sol.MF)
    assert rel_err(x[0], sol.x) < 1e-13


def test_psf_solve_matches_oracle_and_fields(ctx):
    """--psf (ScatterVolume.cpp:670-674): every order carries the incident E."""
    sample, opts, p = case()
    sol = sa.scattervolume(ctx, sample, psf=True, **opts)
    p.psf = True
    ref = o.solve(sample, p)
    assert rel_err(sol.x, ref.x) < TOL
    EF = sol.setup_array(_capi.SOL_EF)
    assert rel_err(EF, ref.setup.EF) < 1e-15
    x = sa.solve_fields(ctx, sample, EF[None], **opts)
    assert rel_err(x[0], sol.x) < 1e-13


def test_random_fields_match_oracle_and_linearity(ctx):
    sample, opts, p = case()
    M0, M1 = sa.solve_fields(ctx, sample, None, **opts)
    MF = M0 * M1
    rng = np.random.default_rng(5)
    K = 6
    EF = rng.standard_normal((K, 3 * MF)) + 1j * rng.standard_normal((K, 3 * MF))
    EF[2] = 0
    EF[2, MF // 2] = 1.0                         # one order, Ex only
    x = sa.solve_fields(ctx, sample, EF, **opts)
    for k in range(K):
        ref = o.solve(sample, p, EF=EF[k])
        assert rel_err(x[k], ref.x) < TOL, k
    a = rng.standard_normal(K) + 1j * rng.standard_normal(K)
    xs = sa.solve_fields(ctx, sample, (a @ EF)[None], **opts)
    assert rel_err(xs[0], a @ x) < 1e-11


def test_order_basis_of_a_cell(ctx):
    """All 2MF unit transverse fields of a cell at once (the order basis of its reflection and transmission
    matrices): sampled columns agree with the oracle."""
    sample = disks_sample(10, 2, 2.0, 2, 1.0, 1.5, seed=4)
    opts = dict(lam=0.9, size=(2.0, 2.0, 0.8), coef=(3, 3), n=(1.0, 1.0), ex=(1, 0), ey=(0, 0))
    p = o.VolumeParams(lam=0.9, size=(2.0, 2.0, 0.8), coef=(3, 3), n=(1.0, 1.0), ex=(1, 0), ey=(0, 0))
    M0, M1 = sa.solve_fields(ctx, sample, None, **opts)
    MF = M0 * M1
    EF = np.zeros((2 * MF, 3 * MF), complex)
    for m in range(MF):
        EF[m, m] = 1.0                    # Ex of order m
        EF[MF + m, MF + m] = 1.0          # Ey of order m
    x = sa.solve_fields(ctx, sample, EF, **opts)
    for k in (0, MF // 2, MF + MF // 2, 2 * MF - 1):
        ref = o.solve(sample, p, EF=EF[k])
        assert rel_err(x[k], ref.x) < TOL, k
