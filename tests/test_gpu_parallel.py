"""GPU parity of the parallel-beam HIP projectors against the FP64 oracle.

Bars (north star): projections within 1e-5 relative L2 in FP32; adjoint
identity <Ax, y> = <x, A^T y> to 1e-6 with FP64 inner products; gradients
through the autograd wrapper equal the back projection.  The oracle is the
sparse-matrix restatement in oracle/parallel.py, itself pinned on the CPU by
brute force and closed forms (tests/test_oracle_parallel.py).
"""

import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL_PROJ = 1e-5


def _model(sino_shape, **params):
    import mbirtorch_amd as mb
    angles = np.linspace(0, np.pi, sino_shape[0], endpoint=False)
    m = mb.ParallelBeamModel(sino_shape, angles)
    if params:
        m.set_params(**params)
    return m, angles


def _geom(m):
    rs = m.get_params('recon_shape')
    ddc, dco, dv, vra = m.get_params(['delta_det_channel', 'det_channel_offset',
                                      'delta_voxel', 'voxel_row_aspect'])
    return dict(num_rows=rs[0], num_cols=rs[1], num_channels=m.get_params('sinogram_shape')[2],
                delta_det_channel=ddc, det_channel_offset=dco, delta_voxel=dv,
                delta_voxel_row=vra * dv)


VARIANTS = {
    'square': dict(),
    'det_offset': dict(det_channel_offset=2.37),
    'wide_voxel': dict(delta_voxel=2.0),          # W up to 2.83: 3-tap general path
    'fine_channel': dict(delta_det_channel=0.75),  # non-unit channel: division path
    'row_aspect': dict(voxel_row_aspect=0.7),
}


def _variant_model(name, sino_shape=(45, 8, 64)):
    m, angles = _model(sino_shape)
    p = VARIANTS[name]
    if p:
        m.set_params(**p)
    return m, angles


@pytest.mark.parametrize('variant', list(VARIANTS))
def test_forward_matches_oracle(variant):
    from oracle import parallel as op
    print(f'[gpu] forward parity variant={variant}', flush=True)
    m, angles = _variant_model(variant)
    rs = m.get_params('recon_shape')
    rng = np.random.default_rng(1)
    vol = rng.random(rs).astype(np.float32)
    sino = m.forward_project(vol)
    idx = op.full_indices(rs)
    ref = op.forward(vol.reshape(-1, rs[2])[idx], idx, angles, _geom(m),
                     psf=m.get_psf_radius())
    assert sino.shape == ref.shape
    err = rel_l2(sino, ref)
    print(f'forward {variant}: rel L2 vs FP64 oracle = {err:.2e}')
    assert err < TOL_PROJ


@pytest.mark.parametrize('variant', list(VARIANTS))
@pytest.mark.parametrize('coeff_power', [1, 2])
def test_back_matches_oracle(variant, coeff_power):
    from oracle import parallel as op
    print(f'[gpu] back parity variant={variant} cp={coeff_power}', flush=True)
    m, angles = _variant_model(variant)
    V, R, C = m.get_params('sinogram_shape')
    rs = m.get_params('recon_shape')
    rng = np.random.default_rng(2)
    sino = rng.random((V, R, C)).astype(np.float32)
    idx = op.full_indices(rs)
    got = m.sparse_back_project(sino, idx, coeff_power=coeff_power).cpu().numpy()
    ref = op.back(sino, idx, angles, _geom(m), psf=m.get_psf_radius(),
                  coeff_power=coeff_power)
    err = rel_l2(got, ref)
    print(f'back {variant} cp={coeff_power}: rel L2 = {err:.2e}')
    assert err < TOL_PROJ


@pytest.mark.parametrize('R', [1, 3, 8, 16, 20, 24, 32, 48, 64, 80])
def test_row_counts(R):
    """Row widths below, at and above one wavefront (lane layouts / row tiles)."""
    from oracle import parallel as op
    print(f'[gpu] row count R={R}', flush=True)
    m, angles = _model((24, R, 40))
    rs = m.get_params('recon_shape')
    rng = np.random.default_rng(R)
    vol = rng.standard_normal(rs).astype(np.float32)
    sino = rng.standard_normal((24, R, 40)).astype(np.float32)
    idx = op.full_indices(rs)
    g = _geom(m)
    err_f = rel_l2(m.forward_project(vol), op.forward(vol.reshape(-1, R)[idx], idx, angles, g))
    err_b = rel_l2(m.sparse_back_project(sino, idx).cpu().numpy(), op.back(sino, idx, angles, g))
    print(f'R={R}: fwd {err_f:.2e} back {err_b:.2e}')
    assert err_f < TOL_PROJ and err_b < TOL_PROJ


def test_random_subsets_and_view_splits():
    """Small random pixel subsets take the split-view back path."""
    from oracle import parallel as op
    import mbirtorch_amd as mb
    print('[gpu] random subsets / view splits', flush=True)
    m, angles = _model((360, 16, 96))
    rs = m.get_params('recon_shape')
    rng = np.random.default_rng(5)
    np.random.seed(3)
    part = mb.gen_pixel_partition(rs, 64)
    idx = part[7].astype(np.int64)
    P = len(idx)
    ns = m.projector().back_splits(P)
    assert ns > 1, ns
    sino = rng.standard_normal((360, 16, 96)).astype(np.float32)
    vals = rng.standard_normal((P, 16)).astype(np.float32)
    g = _geom(m)
    got_b = m.sparse_back_project(sino, idx).cpu().numpy()
    got_f = m.sparse_forward_project(vals, idx).cpu().numpy()
    eb = rel_l2(got_b, op.back(sino, idx, angles, g))
    ef = rel_l2(got_f, op.forward(vals, idx, angles, g))
    print(f'subset P={P} splits={ns}: back {eb:.2e} fwd {ef:.2e}')
    assert eb < TOL_PROJ and ef < TOL_PROJ


def test_back_deterministic():
    m, _ = _model((90, 16, 64))
    rs = m.get_params('recon_shape')
    sino = np.random.default_rng(0).random((90, 16, 64)).astype(np.float32)
    a = m.back_project(sino)
    b = m.back_project(sino)
    assert np.array_equal(a, b)


def test_adjoint_identity_fp64():
    import mbirtorch_amd as mb
    print('[gpu] adjoint identity', flush=True)
    torch.manual_seed(0)
    sino_shape = (48, 40, 32)
    m, _ = _model(sino_shape)
    rs = m.get_params('recon_shape')
    idx = torch.as_tensor(mb.gen_full_indices(rs), dtype=torch.int64, device='cuda')
    x = torch.rand((idx.shape[0], rs[2]), device='cuda')
    y = torch.rand(sino_shape, device='cuda')
    ax = m.sparse_forward_project(x, idx)
    aty = m.sparse_back_project(y, idx)
    lhs, rhs = mb.dot_f64(ax, y), mb.dot_f64(x, aty)
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f'adjoint: <Ax,y>={lhs:.10e} <x,ATy>={rhs:.10e} rel={rel:.2e}')
    assert rel < 1e-6


def test_weighted_back_and_hessian():
    from oracle import parallel as op
    print('[gpu] weighted back / hessian', flush=True)
    m, angles = _model((60, 8, 48))
    V, R, C = m.get_params('sinogram_shape')
    rs = m.get_params('recon_shape')
    rng = np.random.default_rng(9)
    w = rng.random((V, R, C)).astype(np.float32) + 0.5
    e = rng.standard_normal((V, R, C)).astype(np.float32)
    idx = op.full_indices(rs)
    ops = m.projector()
    idx_t = torch.as_tensor(idx, device='cuda')
    got = ops.back(m._vrc_to_vcr(torch.as_tensor(e, device='cuda')), idx_t,
                   weights=m._vrc_to_vcr(torch.as_tensor(w, device='cuda'))).cpu().numpy()
    ref = op.back((w * e).astype(np.float64), idx, angles, _geom(m))
    assert rel_l2(got, ref) < TOL_PROJ
    hess = m.compute_hessian_diagonal(weights=w)
    href = np.zeros((rs[0] * rs[1], rs[2]))
    allidx = np.arange(rs[0] * rs[1])
    href = op.back(w, allidx, angles, _geom(m), coeff_power=2).reshape(rs)
    err = rel_l2(hess, href)
    print(f'weighted back {rel_l2(got, ref):.2e}, hessian {err:.2e}')
    assert err < TOL_PROJ


def test_fbp_filter_matches_oracle():
    from oracle import filters as of
    print('[gpu] fbp filter', flush=True)
    m, _ = _model((30, 8, 96))
    V, R, C = m.get_params('sinogram_shape')
    sino = np.random.default_rng(4).standard_normal((V, R, C)).astype(np.float32)
    got = m.fbp_filter(sino)
    ref = of.fbp_filter_parallel(sino, 1.0, 1.0)
    err = rel_l2(got, ref)
    print(f'fbp filter rel L2 = {err:.2e}')
    assert err < TOL_PROJ


def test_autograd_gradient_is_back_projection():
    import mbirtorch_amd as mb
    print('[gpu] autograd', flush=True)
    torch.manual_seed(0)
    sino_shape = (24, 16, 16)
    m, _ = _model(sino_shape)
    rs = m.get_params('recon_shape')
    vol = torch.rand(tuple(rs), device='cuda', requires_grad=True)
    y = torch.rand(sino_shape, device='cuda')
    sino = mb.forward_project_differentiable(m, vol)
    torch.sum(sino * y).backward()
    expected = m.back_project(y, output_sharded=True)
    rel = float((vol.grad - expected).abs().max() / expected.abs().max())
    assert rel < 1e-5, rel


def test_autograd_cpu_f64_leaf():
    import mbirtorch_amd as mb
    m, _ = _model((12, 4, 16))
    rs = m.get_params('recon_shape')
    vol = torch.rand(tuple(rs), dtype=torch.float64, requires_grad=True)
    sino = mb.forward_project_differentiable(m, vol)
    sino.sum().backward()
    assert vol.grad is not None and vol.grad.dtype == torch.float64
    assert vol.grad.device.type == 'cpu'


def test_autograd_finite_differences():
    """Loss ||Ax - y||^2 accumulated in FP64; central differences at h=1e-3
    on random voxels against the autograd gradient (hand-written A^T)."""
    import mbirtorch_amd as mb
    print('[gpu] finite differences', flush=True)
    m, _ = _model((36, 8, 32))
    rs = m.get_params('recon_shape')
    rng = np.random.default_rng(11)
    x0 = torch.as_tensor(rng.random(rs).astype(np.float32), device='cuda')
    y = torch.as_tensor(rng.random((36, 8, 32)).astype(np.float32), device='cuda')

    def loss(x):
        r = (mb.forward_project_differentiable(m, x) - y).double()
        return torch.sum(r * r)

    x = x0.clone().requires_grad_(True)
    loss(x).backward()
    grad = x.grad
    mask = mb.get_2d_ror_mask(rs)
    rows, cols = np.nonzero(mask)
    pick = rng.choice(len(rows), size=16, replace=False)
    h = 1e-3
    worst = 0.0
    for k in pick:
        i, j, s = int(rows[k]), int(cols[k]), int(rng.integers(rs[2]))
        xp = x0.clone()
        xm = x0.clone()
        xp[i, j, s] += h
        xm[i, j, s] -= h
        with torch.no_grad():
            fd = float((loss(xp) - loss(xm)) / (2 * h))
        g = float(grad[i, j, s])
        worst = max(worst, abs(fd - g) / max(1.0, abs(g)))
    print(f'finite differences: worst relative error {worst:.2e}')
    assert worst < 5e-3


@pytest.mark.parametrize('sino_shape,params', [
    ((90, 16, 64), {}), ((37, 5, 50), {}), ((40, 24, 64), {}), ((36, 32, 96), {}),
    ((30, 48, 64), {}), ((20, 64, 128), {}), ((24, 28, 64), {}),
    ((30, 32, 64), dict(delta_voxel=2.0)),       # 2*psf+1 taps: n_p stored per entry
    ((30, 32, 64), dict(det_channel_offset=2.37)),
])
def test_forward_ray_and_binned_gathers_agree(sino_shape, params):
    """The full-list (ray) and arbitrary-list (binned) forward gathers sum the
    same weights; they differ only in summation order.  Covers every lane
    configuration of both forms (rows per chunk 8 .. 64, two channels or four
    per unit) and the long-bin placement (a dense list at 90 degrees puts an
    image row on one channel)."""
    import mbirtorch_amd as mb
    print(f'[gpu] ray vs binned {sino_shape} {params}', flush=True)
    m, _ = _model(sino_shape, **params)
    rs = m.get_params('recon_shape')
    idx = m.full_indices_device()
    vals = torch.rand((idx.shape[0], rs[2]), device='cuda')
    ops = m.projector()
    a = ops.forward(vals, idx)
    b = ops.forward(vals, idx, runs=False)
    err = rel_l2(a.cpu().numpy(), b.cpu().numpy())
    print(f'ray vs binned rel L2 {err:.2e}')
    assert err < 1e-6
