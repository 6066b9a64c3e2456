"""Pin the filter and VCD oracles: closed-form Ram-Lak taps, explicit O(C^2)
convolution, and -- against the reference's own outputs in tests/golden/
(written by tests/golden/make_golden.py) -- the partition RNG sequence, the
auto-regularization values and the seeded VCD trajectory.  The quadratic-
prior fixed point is checked against an FP64 conjugate-gradient solve."""

import os

import numpy as np
import pytest

from oracle import filters as of
from oracle import parallel as op
from oracle import vcd

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def test_ramp_taps_closed_form():
    C = 33
    h = of.ramp_filter(C)
    n = np.arange(-C + 1, C)
    ref = np.where(n == 0, 0.25, np.where(n % 2 == 1, -1.0 / (np.pi ** 2 * n ** 2), 0.0))
    ref = np.where((n % 2 == 0) & (n != 0), 0.0, ref)
    np.testing.assert_allclose(h, ref, atol=1e-15)


def test_row_filter_is_valid_convolution():
    rng = np.random.default_rng(0)
    s = rng.standard_normal((2, 3, 9))
    taps = rng.standard_normal(17)
    out = of.apply_row_filter(s, taps)
    ref = np.zeros_like(s)
    for c in range(9):
        for k in range(9):
            ref[..., c] += s[..., k] * taps[c - k + 8]
    np.testing.assert_allclose(out, ref, atol=1e-12)


def _load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_partitions_match_reference_golden():
    g = _load('partitions.npz')
    for seed, shape in ((0, (32, 32, 4)), (7, (40, 28, 2))):
        np.random.seed(seed)
        parts = vcd.gen_set_of_pixel_partitions(shape, [1, 2, 4, 8, 16, 32, 64, 128, 128,
                                                         128, 128])
        for k, p in enumerate(parts):
            assert np.array_equal(p, g[f's{seed}_p{k}']), (seed, k)
        assert np.array_equal(np.random.permutation(16), g[f's{seed}_next'])


def test_autoregularization_matches_reference_golden():
    g = _load('autoreg.npz')
    for name in ('a', 'b'):
        sino = g[f'{name}_sino']
        w = np.exp(-np.clip(sino, 0, None) * 10).astype(np.float32)
        rs, _ = op.auto_recon_shape(sino.shape)
        for tag, weights, sharp in (('u', None, 1.0), ('w', w, 0.7)):
            got = vcd.auto_regularization(sino, weights, rs, 1.0, 1.0, 1.0, 30.0, sharp)
            np.testing.assert_allclose(got, g[f'{name}_{tag}'], rtol=1e-7)


@pytest.mark.parametrize('case', ['vcd_qggmrf', 'vcd_weighted_pos'])
def test_vcd_trajectory_matches_reference_golden(case):
    g = _load(f'{case}.npz')
    sino, angles = g['sino'], g['angles']
    rs, _ = op.auto_recon_shape(sino.shape)
    geom = dict(num_rows=rs[0], num_cols=rs[1], num_channels=sino.shape[2],
                delta_det_channel=1.0, det_channel_offset=0.0, delta_voxel=1.0,
                delta_voxel_row=1.0)
    params = {k[6:]: g[k].item() for k in g.files if k.startswith('param_')}
    w = g['weights'] if 'weights' in g.files else None
    np.random.seed(int(g['seed']))
    recon, st = vcd.vcd_recon(vcd.ParallelSystem(angles, geom, rs), sino, params,
                              int(g['iters']), weights=w)
    np.testing.assert_allclose([st['sigma_y'], st['sigma_x']], g['sigma'][:2], rtol=1e-7)
    np.testing.assert_allclose(st['fm_rmse'], g['fm_rmse'], rtol=2e-5)
    np.testing.assert_allclose(st['alpha'], g['alpha'], rtol=2e-4)
    ref = g['recon']
    nrmse = np.sqrt(np.mean((recon - ref) ** 2)) / np.sqrt(np.mean(ref ** 2))
    assert nrmse < 1e-4, nrmse


def test_quadratic_vcd_converges_to_normal_equations():
    """p = q = 2: the MAP estimate solves (A^T A/sy^2 + H_prior) x = A^T y/sy^2,
    H_prior the discrete Laplacian form of the quadratic prior."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    V, R, C = 24, 2, 12
    angles = np.linspace(0, np.pi, V, endpoint=False)
    rs, _ = op.auto_recon_shape((V, R, C))
    geom = dict(num_rows=rs[0], num_cols=rs[1], num_channels=C, delta_det_channel=1.0,
                det_channel_offset=0.0, delta_voxel=1.0, delta_voxel_row=1.0)
    rng = np.random.default_rng(0)
    sysm = vcd.ParallelSystem(angles, geom, rs)
    x_true = np.zeros((rs[0] * rs[1], R))
    x_true[sysm.pix] = rng.random((len(sysm.pix), R))
    y = sysm.from_vec(sysm.A @ x_true[sysm.pix]).astype(np.float32)
    sy, sx = 0.5, 0.3
    params = dict(p=2.0, q=2.0, T=1.0, sigma_y=sy, sigma_x=sx, sigma_prox=sx,
                  auto_regularize_flag=False, granularity=[1, 4], partition_sequence=[1],
                  max_alpha=1.5)
    np.random.seed(0)
    recon, st = vcd.vcd_recon(sysm, y, params, 300)
    # Normal equations over all grid pixels with the prior's quadratic form:
    # loss = sum_pairs b_axis * d^2/(4 sx^2) (rho at p=q=2).
    n = rs[0] * rs[1] * R
    lin = np.arange(n).reshape(rs[0], rs[1], R)
    b_axis = 1.0 / 6.0
    rows, cols, vals = [], [], []
    for axis in range(3):
        a = np.take(lin, np.arange(lin.shape[axis] - 1), axis=axis).ravel()
        bb = np.take(lin, np.arange(1, lin.shape[axis]), axis=axis).ravel()
        k = 2 * b_axis / (4 * sx ** 2)
        rows += [a, bb, a, bb]
        cols += [a, bb, bb, a]
        vals += [np.full(a.size, k)] * 2 + [np.full(a.size, -k)] * 2
    Hp = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                       shape=(n, n))
    # A over the ROR pixels, embedded in the full grid (voxel-major layout).
    A_full = sp.lil_matrix((V * C * R, n))
    Acsr = sysm.A.tocsr()
    Ac = Acsr.tocoo()
    for r_ in range(R):
        A_full[Ac.row * R + r_, sysm.pix[Ac.col] * R + r_] = Ac.data
    A_full = A_full.tocsr()
    # restrict unknowns to ROR voxels (the VCD never updates the others)
    ror = np.zeros((rs[0] * rs[1], R), bool)
    ror[sysm.pix] = True
    keep = np.nonzero(ror.ravel())[0]
    M = (A_full.T @ A_full) / sy ** 2 + Hp
    M = M[keep][:, keep]
    rhs = (A_full.T @ sysm.to_vec(y)) / sy ** 2
    rhs = rhs[keep]
    sol, info = spla.cg(M, rhs, rtol=1e-12, maxiter=5000)
    assert info == 0
    got = recon.reshape(-1)[keep]
    err = np.linalg.norm(got - sol) / np.linalg.norm(sol)
    assert err < 1e-3, err
