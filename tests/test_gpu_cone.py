"""GPU parity of the cone-beam HIP projectors against the FP64 cone oracle
(oracle/cone.py, pinned by brute force in tests/test_oracle_cone.py):
projections within 1e-5 relative L2, the adjoint identity with FP64 inner
products, FDK, and a seeded cone VCD (DC damping) against the oracle."""

import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _cone_case(kind, V=24, R=10, C=24):
    import mbirtorch_amd as mb
    from oracle import cone as oc
    angles = np.linspace(0, 2 * np.pi, V, endpoint=False) + 0.05
    zs = np.linspace(-3, 3, V) if kind in ('helical', 'thick') else None
    curved = kind == 'curved'
    sdd, sid = 80.0, 40.0
    m = mb.ConeBeamModel((V, R, C), angles, sdd, sid, helical_z_shifts=zs,
                         use_curved_detector=curved)
    # 'thick': slices 2.5 voxels high, as in the helical mixed-precision
    # geometry -- tap radius 2, the 5-tap windows.
    vsa = 2.5 if kind == 'thick' else 1.0
    if kind == 'thick':
        m.set_params(voxel_slice_aspect=vsa)
    rs = tuple(m.get_params('recon_shape'))
    g = oc.geometry((V, R, C), angles, sdd, sid, recon_shape=rs,
                    delta_voxel=m.get_params('delta_voxel'), z_shifts=zs, curved=curved,
                    voxel_slice_aspect=vsa,
                    recon_slice_offset=m.get_params('recon_slice_offset'))
    return m, g, rs


@pytest.mark.parametrize('kind', ['flat', 'curved', 'helical', 'thick'])
def test_cone_forward_back_match_oracle(kind):
    from oracle import cone as oc
    from oracle import parallel as op
    print(f'[gpu] cone parity {kind}', flush=True)
    m, g, rs = _cone_case(kind)
    assert m.get_psf_radii() == oc.psf_radii(g)
    assert m.get_psf_radii()[0] == (2 if kind == 'thick' else 1)
    idx = op.full_indices(rs)
    rng = np.random.default_rng(0)
    vol = rng.random(rs).astype(np.float32)
    sino = m.forward_project(vol)
    ref = oc.forward(vol.reshape(-1, rs[2])[idx], idx, g)
    ef = rel_l2(sino, ref)
    y = rng.random(sino.shape).astype(np.float32)
    back = m.sparse_back_project(y, idx).cpu().numpy()
    eb = rel_l2(back, oc.back(y, idx, g))
    hess = m.sparse_back_project(y, idx, coeff_power=2).cpu().numpy()
    eh = rel_l2(hess, oc.back(y, idx, g, coeff_power=2))
    print(f'cone {kind}: fwd {ef:.2e} back {eb:.2e} hess {eh:.2e}')
    assert ef < TOL and eb < TOL and eh < TOL


def test_cone_adjoint_fp64():
    import mbirtorch_amd as mb
    print('[gpu] cone adjoint', flush=True)
    m, g, rs = _cone_case('helical', V=32, R=12, C=32)
    idx = torch.as_tensor(mb.gen_full_indices(rs), device='cuda')
    x = torch.rand((idx.shape[0], rs[2]), device='cuda')
    y = torch.rand((32, 12, 32), device='cuda')
    lhs = mb.dot_f64(m.sparse_forward_project(x, idx), y)
    rhs = mb.dot_f64(x, m.sparse_back_project(y, idx))
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f'cone adjoint rel {rel:.2e}')
    assert rel < 1e-6


def test_cone_random_subset():
    from oracle import cone as oc
    import mbirtorch_amd as mb
    print('[gpu] cone subset', flush=True)
    m, g, rs = _cone_case('flat', V=64)
    np.random.seed(4)
    idx = mb.gen_pixel_partition(rs, 16)[3].astype(np.int64)
    rng = np.random.default_rng(1)
    vals = rng.random((len(idx), rs[2])).astype(np.float32)
    got = m.sparse_forward_project(vals, idx).cpu().numpy()
    assert rel_l2(got, oc.forward(vals, idx, g)) < TOL
    y = rng.random((64, 10, 24)).astype(np.float32)
    got = m.sparse_back_project(y, idx).cpu().numpy()
    assert rel_l2(got, oc.back(y, idx, g)) < TOL


def test_fdk_matches_oracle():
    from oracle import cone as oc
    from oracle import filters as of
    from oracle import parallel as op
    print('[gpu] fdk', flush=True)
    for kind in ('flat', 'helical'):
        m, g, rs = _cone_case(kind)
        rng = np.random.default_rng(2)
        sino = rng.random((24, 10, 24)).astype(np.float32)
        got = m.recon_fdk(sino)
        taps = of.ramp_filter(24) * oc.fdk_scale(g) * (np.pi / 24)
        filt = of.apply_row_filter(sino, taps, row_weight=oc.fdk_row_weight(g))
        idx = op.full_indices(rs)
        ref = np.zeros((rs[0] * rs[1], rs[2]))
        ref[idx] = oc.back(filt, idx, g)
        ref = ref.reshape(rs)
        if kind == 'helical':
            ref = ref * oc.helical_z_weight(g)[None, None, :]
        err = rel_l2(got, ref)
        print(f'fdk {kind}: {err:.2e}')
        assert err < TOL


def test_cone_vcd_dc_damping_matches_oracle():
    """Seeded cone recon (FDK init, DC damping on) against the FP64 oracle VCD
    with the same operator and damping profile."""
    from oracle import cone as oc
    from oracle import parallel as op
    from oracle import vcd
    from mbirtorch_amd.phantom import random_ellipsoids, add_noise
    print('[gpu] cone vcd', flush=True)
    m, g, rs = _cone_case('flat', V=36, R=8, C=24)
    vol = random_ellipsoids(rs, 5, seed=3, axes_range=(2.0, 6.0))
    idx = op.full_indices(rs)
    sino = add_noise(oc.forward(vol.reshape(-1, rs[2])[idx], idx, g), 0.01, seed=4)
    np.random.seed(0)
    rec, d = m.recon(sino, max_iterations=3, stop_threshold_change_pct=0.0,
                     print_logs=False, logfile_path=None)
    np.random.seed(0)
    ref, st = vcd.vcd_recon(vcd.ConeSystem(g), sino, {}, 3)
    fm = d['recon_params']['fm_rmse']
    e = float(np.sqrt(np.mean((rec - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
    print(f'cone vcd: NRMSE {e:.2e}, fm_rmse {fm} vs {st["fm_rmse"]}')
    assert e < 1e-4
    np.testing.assert_allclose(fm, st['fm_rmse'], rtol=1e-4)


def test_cone_gather_many_rows_and_channel_blocks_adjoint():
    """A shape with several 64-channel blocks and two row chunks (the
    gather's tiling) and a dense full-region list (long bins): the forward is
    the exact adjoint of the oracle-checked back projector, and the two
    output layouts agree."""
    import mbirtorch_amd as mb
    from mbirtorch_amd import _lib
    print('[gpu] cone gather tiling', flush=True)
    V, R, C = 12, 80, 150
    angles = np.linspace(0, 2 * np.pi, V, endpoint=False) + 0.1
    m = mb.ConeBeamModel((V, R, C), angles, 400.0, 200.0)
    rs = tuple(m.get_params('recon_shape'))
    idx = m.full_indices_device()
    x = torch.rand((idx.shape[0], rs[2]), device='cuda')
    y = torch.rand((V, R, C), device='cuda')
    ops = m.projector()
    fv = ops.forward(x, idx, layout=_lib.MBIR_LAYOUT_VRC)
    fc = ops.forward(x, idx, layout=_lib.MBIR_LAYOUT_VCR)
    assert torch.equal(fv, fc.permute(0, 2, 1))
    lhs = mb.dot_f64(fv, y)
    rhs = mb.dot_f64(x, m.sparse_back_project(y, idx))
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f'cone gather adjoint rel {rel:.2e} (recon {rs})')
    assert rel < 1e-6
