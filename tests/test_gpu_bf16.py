"""bfloat16 storage of the error sinogram and weights: the bf16 kernels
compute exactly what the float32 kernels compute on the widened values
(arithmetic is float32 in both), updates round to nearest even, and a VCD
reconstruction with bf16 storage stays within NRMSE 5e-3 of the float32
path and of the FP64 oracle (bfloat16 keeps 8 mantissa bits, a relative
rounding of 2^-9 per error-sinogram update; measured ~1.2e-3 parallel and
~1.8e-3 cone on these cases)."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _bf16_round(a):
    """float32 -> nearest-even bfloat16 -> float32, on the host."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def _model(V=40, R=8, C=48):
    import mbirtorch_amd as mb
    angles = np.linspace(0, np.pi, V, endpoint=False)
    return mb.ParallelBeamModel((V, R, C), angles)


def test_conversion_rounds_to_nearest_even():
    from mbirtorch_amd.tomography_model import _to_bf16, _to_f32
    print('[gpu] bf16 conversion', flush=True)
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 1e3,
                        np.float32([0.0, -0.0, 1.0, 1.00390625, 1.005859375, 3e-39])])
    t = torch.as_tensor(x, device='cuda')
    back = _to_f32(_to_bf16(t)).cpu().numpy()
    assert np.array_equal(back.view(np.uint32), _bf16_round(x).view(np.uint32))
    # torch's own conversion is round-to-nearest-even too
    assert np.array_equal(back, t.to(torch.bfloat16).float().cpu().numpy())


@pytest.mark.parametrize('weighted,R', [(False, 8), (True, 8), (True, 32), (True, 64)])
def test_parallel_bf16_kernels_equal_f32_on_widened_values(weighted, R):
    import mbirtorch_amd as mb
    from mbirtorch_amd import _lib
    from mbirtorch_amd.tomography_model import _to_bf16, _to_f32
    print(f'[gpu] bf16 parallel kernels weighted={weighted} R={R}', flush=True)
    V, C = 40, 48
    m = _model(V=V, R=R, C=C)
    ops = m.projector()
    rs = tuple(m.get_params('recon_shape'))
    g = torch.Generator(device='cuda').manual_seed(1)
    e16 = _to_bf16(torch.randn((V, C, R), device='cuda', generator=g))
    e32 = _to_f32(e16)
    w16 = _to_bf16(torch.rand((V, C, R), device='cuda', generator=g)) if weighted else None
    w32 = _to_f32(w16) if weighted else None
    np.random.seed(3)
    idx = torch.as_tensor(mb.gen_pixel_partition(rs, 8)[2].astype(np.int64), device='cuda')
    P = int(idx.shape[0])
    lib, st = _lib.lib(), _lib.stream_ptr()
    ns = ops.back_splits(P)
    b32 = torch.empty(ns * P * R, device='cuda')
    b16 = torch.empty_like(b32)
    ops.back_into(e32, idx, b32, ns, weights=w32)
    _lib.check(lib.mbir_par_back_bf16(_lib.ptr(e16), _lib.ptr(w16), _lib.ptr(idx), P,
                                      _lib.ptr(ops.table), ops.geom, ns, _lib.ptr(b16), st), 'b')
    assert torch.equal(b16, b32)
    delta = torch.randn((P, R), device='cuda', generator=g)
    n = ops.fwd_partials_count(ray=False)
    p32 = torch.empty(2 * n, dtype=torch.float64, device='cuda')
    p16 = torch.empty_like(p32)
    ops.fwd_linesearch(delta, idx, e32, w32, p32)
    _lib.check(lib.mbir_par_fwd_linesearch_bf16(_lib.ptr(delta), _lib.ptr(idx), P,
                                                _lib.ptr(ops.table), ops.geom, _lib.ptr(e16),
                                                _lib.ptr(w16), None, _lib.ptr(p16), st), 'ls')
    assert torch.equal(p16, p32)
    alpha = torch.full((1,), 0.37, device='cuda')
    ops.fwd_update(delta, idx, alpha, e32)
    _lib.check(lib.mbir_par_fwd_update_bf16(_lib.ptr(delta), _lib.ptr(idx), P,
                                            _lib.ptr(ops.table), ops.geom, _lib.ptr(alpha),
                                            _lib.ptr(e16), st), 'up')
    assert np.array_equal(_to_f32(e16).cpu().numpy(), _bf16_round(e32.cpu().numpy()))


def _vcd_case():
    from mbirtorch_amd.phantom import random_ellipsoids, add_noise
    m = _model(V=48, R=8, C=40)
    rs = tuple(m.get_params('recon_shape'))
    vol = random_ellipsoids(rs, 5, seed=2, axes_range=(2.0, 9.0))
    sino = add_noise(m.forward_project(vol), 0.01, seed=3)
    return m, sino


def test_parallel_vcd_bf16_storage_close_to_f32_and_oracle():
    from oracle import parallel as op
    from oracle import vcd
    print('[gpu] bf16 parallel vcd', flush=True)
    out = {}
    for storage in ('float32', 'bfloat16'):
        m, sino = _vcd_case()
        m.set_params(p=1.2, q=2.0, T=1.0, sinogram_storage=storage)
        np.random.seed(0)
        out[storage], _ = m.recon(sino, max_iterations=6, stop_threshold_change_pct=0.0,
                                  print_logs=False, logfile_path=None)
    m, sino = _vcd_case()
    rs = tuple(m.get_params('recon_shape'))
    geom = dict(num_rows=rs[0], num_cols=rs[1], num_channels=40, delta_det_channel=1.0,
                det_channel_offset=0.0, delta_voxel=1.0, delta_voxel_row=1.0)
    np.random.seed(0)
    ref, _ = vcd.vcd_recon(vcd.ParallelSystem(m.get_params('angles'), geom, rs), sino,
                           dict(p=1.2, q=2.0, T=1.0), 6)

    def nrmse(a, b):
        return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))
    e_f32 = nrmse(out['float32'], ref)
    e_bf16 = nrmse(out['bfloat16'], ref)
    e_mix = nrmse(out['bfloat16'], out['float32'])
    print(f'NRMSE vs FP64 oracle: float32 {e_f32:.2e}, bfloat16 {e_bf16:.2e}; '
          f'bf16 vs f32 {e_mix:.2e}')
    assert e_f32 < 1e-4
    assert e_bf16 < 5e-3 and e_mix < 5e-3


def test_cone_vcd_bf16_storage_close_to_f32():
    import mbirtorch_amd as mb
    from mbirtorch_amd.phantom import random_ellipsoids, add_noise
    print('[gpu] bf16 cone vcd', flush=True)
    out = {}
    V, R, C = 36, 8, 24
    angles = np.linspace(0, 2 * np.pi, V, endpoint=False)
    zs = np.linspace(-2, 2, V)
    for storage in ('float32', 'bfloat16'):
        m = mb.ConeBeamModel((V, R, C), angles, 80.0, 40.0, helical_z_shifts=zs)
        rs = tuple(m.get_params('recon_shape'))
        vol = random_ellipsoids(rs, 5, seed=3, axes_range=(2.0, 6.0))
        sino = add_noise(m.forward_project(vol), 0.01, seed=4)
        m.set_params(sinogram_storage=storage)
        np.random.seed(0)
        out[storage], _ = m.recon(sino, max_iterations=4, stop_threshold_change_pct=0.0,
                                  print_logs=False, logfile_path=None)
    a, b = out['bfloat16'], out['float32']
    e = float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))
    print(f'cone bf16 vs f32 NRMSE {e:.2e}')
    assert e < 5e-3
