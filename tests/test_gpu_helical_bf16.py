"""The helical mixed-precision configuration, 64x64x32 sub-case (4 turns at
pitch 1.0, magnification 2, 128 views, 16x96 detector; written by
tests/golden/make_helical_subcase.py with the FP64 oracle): the float32 and
the bfloat16-storage reconstructions after 6 VCD iterations (same seed) and
their NRMSE against the FP64 oracle's."""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'helical_subcase.npz')


@pytest.mark.skipif(not os.path.exists(GOLD), reason='helical sub-case fixture not generated')
def test_helical_subcase_f32_and_bf16_against_fp64_oracle():
    import mbirtorch_amd as mb
    print('[gpu] helical sub-case', flush=True)
    g = np.load(GOLD, allow_pickle=False)
    sino = g['sino']
    V, R, C = sino.shape
    ref = g['recon']
    out = {}
    for storage in ('float32', 'bfloat16'):
        m = mb.ConeBeamModel((V, R, C), g['angles'], float(g['sdd']), float(g['sid']),
                             helical_z_shifts=g['z_shifts'])
        m.set_params(recon_shape=ref.shape, delta_voxel=float(g['delta_voxel']),
                     voxel_slice_aspect=float(g['voxel_slice_aspect']),
                     sinogram_storage=storage)
        assert tuple(m.get_params('recon_shape')) == ref.shape
        np.random.seed(int(g['seed']))
        out[storage], d = m.recon(sino, max_iterations=int(g['iters']),
                                  stop_threshold_change_pct=0.0, print_logs=False,
                                  logfile_path=None)

    def nrmse(a, b):
        return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))
    e32, e16 = nrmse(out['float32'], ref), nrmse(out['bfloat16'], ref)
    print(f'helical 64x64x32 NRMSE vs FP64 oracle: float32 {e32:.2e}, bfloat16 {e16:.2e}; '
          f'bf16 vs f32 {nrmse(out["bfloat16"], out["float32"]):.2e}')
    assert e32 < 1e-4
    assert e16 < 5e-3
