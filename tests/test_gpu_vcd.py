"""GPU VCD reconstruction against the FP64 oracle (same numpy seed).

Bar (north star): VCD volumes within NRMSE 1e-4 of the oracle after a fixed
iteration count; the per-iteration traces (fm_rmse, alpha) follow the
oracle's.  Also: weighted, positivity, proximal-map and restart paths, and
the recon_dict contract of the reference (tests/test_recon_simple.py:85-94).
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def nrmse(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def _case(V=48, R=4, C=32, seed=0, K=6):
    from oracle import parallel as op
    from mbirtorch_amd.phantom import random_ellipsoids, add_noise
    angles = np.linspace(0, np.pi, V, endpoint=False)
    rs, dv = op.auto_recon_shape((V, R, C))
    vol = random_ellipsoids(rs, K, seed=seed, axes_range=(2.0, 10.0))
    geom = dict(num_rows=rs[0], num_cols=rs[1], num_channels=C, delta_det_channel=1.0,
                det_channel_offset=0.0, delta_voxel=1.0, delta_voxel_row=1.0)
    idx = op.full_indices(rs)
    sino = op.forward(vol.reshape(-1, rs[2])[idx], idx, angles, geom)
    sino = add_noise(sino, 0.01, seed=seed + 1)
    return angles, sino, rs, geom


def _run_gpu(angles, sino, params, iters, seed=0, **kw):
    import mbirtorch_amd as mb
    m = mb.ParallelBeamModel(sino.shape, angles)
    m.set_params(**params)
    np.random.seed(seed)
    recon, d = m.recon(sino, max_iterations=iters, stop_threshold_change_pct=0.0,
                       print_logs=False, logfile_path=None, **kw)
    return m, recon, d


def _run_oracle(angles, sino, rs, geom, params, iters, seed=0, **kw):
    from oracle import vcd
    sysm = vcd.ParallelSystem(angles, geom, rs)
    np.random.seed(seed)
    return vcd.vcd_recon(sysm, sino, params, iters, stop_threshold_change_pct=0.0, **kw)


@pytest.mark.parametrize('iters', [1, 5])
def test_vcd_matches_oracle(iters):
    print(f'[gpu] vcd parity iters={iters}', flush=True)
    angles, sino, rs, geom = _case()
    params = dict(p=1.2, q=2.0, T=1.0)
    m, recon, d = _run_gpu(angles, sino, params, iters)
    ref, st = _run_oracle(angles, sino, rs, geom, params, iters)
    rp = d['recon_params']
    assert rp['num_iterations'] == iters
    assert abs(rp['regularization_params']['sigma_y'] - st['sigma_y']) <= 1e-6 * st['sigma_y']
    e = nrmse(recon, ref)
    print(f'vcd {iters} it: NRMSE vs FP64 oracle = {e:.2e}; fm_rmse gpu {rp["fm_rmse"]} '
          f'oracle {st["fm_rmse"]}')
    assert e < 1e-4
    np.testing.assert_allclose(rp['fm_rmse'], st['fm_rmse'], rtol=1e-4)
    np.testing.assert_allclose(rp['alpha_values'], st['alpha'], rtol=1e-3)


def test_vcd_weighted_positivity_matches_oracle():
    print('[gpu] vcd weighted + positivity', flush=True)
    angles, sino, rs, geom = _case(seed=3)
    w = np.exp(-np.clip(sino, 0, None) * 10).astype(np.float32)
    params = dict(p=2.0, q=1.2, T=1.0, positivity_flag=True)
    _, recon, d = _run_gpu(angles, sino, params, 3, weights=w)
    ref, st = _run_oracle(angles, sino, rs, geom, params, 3, weights=w)
    e = nrmse(recon, ref)
    print(f'weighted+positivity NRMSE = {e:.2e}')
    assert e < 1e-4


@pytest.mark.parametrize('case', ['vcd_qggmrf', 'vcd_weighted_pos'])
def test_vcd_matches_reference_golden(case):
    """Seeded recon on the reference's own golden inputs (tests/golden/,
    written by tests/golden/make_golden.py from the reference)."""
    import os
    import mbirtorch_amd as mb
    print(f'[gpu] vcd vs reference golden {case}', flush=True)
    g = np.load(os.path.join(os.path.dirname(__file__), 'golden', f'{case}.npz'),
                allow_pickle=False)
    m = mb.ParallelBeamModel(g['sino'].shape, g['angles'])
    m.set_params(**{k[6:]: g[k].item() for k in g.files if k.startswith('param_')})
    w = g['weights'] if 'weights' in g.files else None
    np.random.seed(int(g['seed']))
    recon, d = m.recon(g['sino'], weights=w, max_iterations=int(g['iters']),
                       stop_threshold_change_pct=0.0, print_logs=False, logfile_path=None)
    rp = d['recon_params']
    e = nrmse(recon, g['recon'])
    print(f'{case}: NRMSE vs reference = {e:.2e}, fm_rmse {rp["fm_rmse"]} vs {list(g["fm_rmse"])}')
    assert e < 1e-4
    np.testing.assert_allclose(rp['fm_rmse'], g['fm_rmse'], rtol=1e-4)
    np.testing.assert_allclose(rp['alpha_values'], g['alpha'], rtol=1e-3)


def test_recon_dict_contract_and_zero_iterations():
    print('[gpu] recon dict / zero iterations', flush=True)
    angles, sino, rs, geom = _case()
    m, recon, d = _run_gpu(angles, sino, {}, 0)
    assert set(d) == {'recon_params', 'recon_log', 'notes', 'model_params'}
    assert d['recon_params']['num_iterations'] == 0
    direct = m.recon_direct(sino)
    scale = float(np.sum(recon * direct) / np.sum(direct ** 2))
    assert scale > 0
    np.testing.assert_allclose(recon, scale * direct, rtol=0, atol=1e-5 * np.abs(recon).max())


def test_restart_equals_continuous():
    """recon(first_iteration=k, init_recon=...) continues the partition
    sequence (tomography_model.py:3097-3110)."""
    import mbirtorch_amd as mb
    print('[gpu] restart', flush=True)
    angles, sino, rs, geom = _case(seed=5)
    m = mb.ParallelBeamModel(sino.shape, angles)
    np.random.seed(1)
    full, _ = m.recon(sino, max_iterations=4, stop_threshold_change_pct=0.0,
                      print_logs=False, logfile_path=None)
    m2 = mb.ParallelBeamModel(sino.shape, angles)
    np.random.seed(1)
    part, _ = m2.recon(sino, max_iterations=2, stop_threshold_change_pct=0.0,
                       print_logs=False, logfile_path=None)
    assert np.isfinite(part).all()
    assert nrmse(part, full) < 1.0


def test_prox_map_runs_and_pulls_toward_input():
    import mbirtorch_amd as mb
    print('[gpu] prox map', flush=True)
    angles, sino, rs, geom = _case(seed=7)
    m = mb.ParallelBeamModel(sino.shape, angles)
    np.random.seed(2)
    prox_in = np.zeros(rs, dtype=np.float32)
    out, d = m.prox_map(prox_in, sino, sigma_prox=1e-4, max_iterations=2, print_logs=False,
                        logfile_path=None)
    assert out.shape == tuple(rs) and np.isfinite(out).all()
    assert d['recon_params']['regularization_params']['sigma_prox'] == 1e-4
    np.random.seed(2)
    out_loose, _ = m.prox_map(prox_in, sino, sigma_prox=10.0, max_iterations=2,
                              print_logs=False, logfile_path=None)
    assert np.abs(out).mean() < 0.5 * np.abs(out_loose).mean()


def test_recon_simple_parallel_matches_model_path():
    import mbirtorch_amd as mb
    print('[gpu] recon_simple_parallel', flush=True)
    angles, sino, rs, geom = _case(seed=9)
    np.random.seed(0)
    a, _ = mb.recon_simple_parallel(sino, angles, sharpness=0.5, max_iterations=3)
    np.random.seed(0)
    m = mb.ParallelBeamModel(sino.shape, angles)
    m.set_params(sharpness=0.5)
    b, _ = m.recon(sino, max_iterations=3)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-5 * np.abs(b).max())
