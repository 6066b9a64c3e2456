"""CPU-only checks of the host side: the C-ABI library loads and exports every
symbol include/mbir_hip.h declares; the model API's parameter semantics;
the product's partition generator against the reference's golden partitions;
the host view table; and loud failure without a GPU."""

import os
import re

import numpy as np
import pytest
import torch

import mbirtorch_amd as mb
from mbirtorch_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, 'include', 'mbir_hip.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(mbir_[a-z0-9_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = _declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)
    assert lib.mbir_version() >= 1


def test_geometry_struct_layout():
    assert torch is not None
    import ctypes
    assert ctypes.sizeof(_lib.ParGeom) == 16 * 4
    assert ctypes.sizeof(_lib.VcdParams) == (12 + 14) * 4 + 4 * 4


def test_size_queries_need_no_gpu():
    lib = _lib.load()
    g = _lib.ParGeom(num_rows=512, num_cols=512, num_slices=64, num_channels=512,
                     num_views=720, psf_radius=1, two_tap=1, unit_channel=1,
                     delta_det_channel=1.0, delta_voxel=1.0, delta_voxel_row=1.0)
    assert lib.mbir_par_back_splits(g, 205012) == 1
    assert lib.mbir_par_back_splits(g, 1602) > 1
    # Both forward forms: one workgroup per (view, 128 channels, 64 rows).
    assert lib.mbir_par_fwd_partials_count(g, 0) == 720 * 4
    assert lib.mbir_par_fwd_partials_count(g, 1) == 720 * 4
    assert lib.mbir_reduce_blocks(10 ** 6) >= 1


def test_model_geometry_matches_reference_rules():
    angles = np.linspace(0, np.pi, 720, endpoint=False)
    m = mb.ParallelBeamModel((720, 64, 512), angles)
    assert tuple(m.get_params('recon_shape')) == (512, 512, 64)
    assert m.get_psf_radius() == 1
    assert m.get_params('delta_voxel') == 1.0
    m = mb.ParallelBeamModel((180, 8, 128), np.linspace(0, np.pi, 180, endpoint=False))
    assert tuple(m.get_params('recon_shape')) == (128, 128, 8)


def test_param_semantics():
    m = mb.ParallelBeamModel((10, 4, 16), np.linspace(0, np.pi, 10, endpoint=False))
    assert m.get_params(['p', 'q', 'T']) == [2.0, 1.2, 1.0]
    with pytest.raises(NameError):
        m.get_params('nope')
    with pytest.raises(ValueError):
        m.set_params(nope=1)
    with pytest.warns(UserWarning):
        m.set_params(sigma_x=0.3)
    assert m.get_params('auto_regularize_flag') is False
    with pytest.warns(UserWarning):
        m.set_params(sharpness=0.5)
    assert m.get_params('auto_regularize_flag') is True
    m.set_params(p=1.2, q=2.0)
    assert m.get_params('p') == 1.2
    with pytest.raises(ValueError):
        mb.ParallelBeamModel((10, 4, 16), np.linspace(0, np.pi, 9))
    m.set_params(voxel_slice_aspect=2.0)
    with pytest.raises(ValueError):
        m.verify_valid_params()


def test_partitions_match_reference_golden():
    g = np.load(os.path.join(ROOT, 'tests', 'golden', 'partitions.npz'), allow_pickle=False)
    for seed, shape in ((0, (32, 32, 4)), (7, (40, 28, 2))):
        np.random.seed(seed)
        parts = mb.gen_set_of_pixel_partitions(shape, [1, 2, 4, 8, 16, 32, 64, 128, 128, 128,
                                                       128])
        for k, p in enumerate(parts):
            assert np.array_equal(p.numpy(), g[f's{seed}_p{k}']), (seed, k)
        assert np.array_equal(np.random.permutation(16), g[f's{seed}_next'])


def test_partition_sequence():
    seq = mb.gen_partition_sequence([2, 4, 6] + [7, 8, 9, 10] * 25, 10)
    assert list(seq) == [2, 4, 6, 7, 8, 9, 10, 7, 8, 9]
    assert list(mb.gen_partition_sequence([0, 1], 4)) == [0, 1, 1, 1]


def test_view_table_uses_float32_chain():
    from mbirtorch_amd.parallel_beam import parallel_view_table
    angles = np.linspace(0, np.pi, 37, endpoint=False).astype(np.float32)
    t = parallel_view_table(angles, 1.0, 0.8, 0.9).numpy()
    c = np.cos(angles.astype(np.float64))
    assert np.allclose(t[:, 0], c, atol=1e-6)
    fp = np.maximum(np.abs(t[:, 0]) * np.float32(1.0), np.abs(t[:, 1]) * np.float32(0.8))
    assert np.array_equal(t[:, 2], (fp / np.float32(0.9)).astype(np.float32))
    assert np.array_equal(t[:, 4], ((t[:, 2] + np.float32(1)) / np.float32(2)))
    assert np.array_equal(t[:, 5], np.minimum(t[:, 2], np.float32(1)))


def test_autoregularization_matches_reference_golden():
    g = np.load(os.path.join(ROOT, 'tests', 'golden', 'autoreg.npz'), allow_pickle=False)
    for name in ('a', 'b'):
        sino, angles = g[f'{name}_sino'], g[f'{name}_angles']
        w = np.exp(-np.clip(sino, 0, None) * 10).astype(np.float32)
        for tag, weights, sharp in (('u', None, 1.0), ('w', w, 0.7)):
            m = mb.ParallelBeamModel(sino.shape, angles)
            m.set_params(sharpness=sharp)
            with np.errstate(all='ignore'):
                import warnings
                with warnings.catch_warnings():
                    warnings.simplefilter('ignore')
                    v = m.auto_set_regularization_params(sino, weights=weights)
            got = [v['sigma_y'], v['sigma_x'], v['sigma_prox']]
            np.testing.assert_allclose(got, g[f'{name}_{tag}'], rtol=1e-7)
            # A tensor sinogram is view-sampled where it lives: same values.
            m2 = mb.ParallelBeamModel(sino.shape, angles)
            m2.set_params(sharpness=sharp)
            with np.errstate(all='ignore'), warnings.catch_warnings():
                warnings.simplefilter('ignore')
                v2 = m2.auto_set_regularization_params(
                    torch.as_tensor(sino), weights=None if weights is None
                    else torch.as_tensor(weights))
            assert [v2['sigma_y'], v2['sigma_x'], v2['sigma_prox']] == got


def test_input_checks_on_tensors():
    from mbirtorch_amd.tomography_model import _all_finite
    a = torch.zeros(3, 4)
    assert _all_finite(a) and _all_finite(a.numpy())
    a[1, 2] = float('nan')
    assert not _all_finite(a) and not _all_finite(a.numpy())
    a[1, 2] = float('inf')
    assert not _all_finite(a)


@pytest.mark.skipif(torch.cuda.is_available(), reason='checks the no-GPU error path')
def test_projection_without_gpu_fails_loudly():
    m = mb.ParallelBeamModel((10, 4, 16), np.linspace(0, np.pi, 10, endpoint=False))
    with pytest.raises(RuntimeError, match='HIP'):
        m.forward_project(np.zeros(m.get_params('recon_shape'), np.float32))


def test_phantom_is_seeded_and_leaves_global_rng_alone():
    from mbirtorch_amd.phantom import random_ellipsoids
    np.random.seed(5)
    a = random_ellipsoids((32, 32, 8), 4, seed=1, axes_range=(2, 8))
    after = np.random.random()
    np.random.seed(5)
    b = random_ellipsoids((32, 32, 8), 4, seed=1, axes_range=(2, 8))
    assert np.array_equal(a, b) and np.random.random() == after
    assert a.max() > 0 and a.min() >= 0


def test_row_runs_table():
    from mbirtorch_amd.parallel_beam import row_runs
    import mbirtorch_amd as mb
    rs = (12, 10, 1)
    idx = torch.as_tensor(mb.gen_full_indices(rs))
    runs = row_runs(idx, 12, 10)
    assert runs is not None and int(runs[:, 2].sum()) == idx.shape[0]
    np.random.seed(0)
    sub = torch.as_tensor(mb.gen_pixel_partition(rs, 4)[0].astype(np.int64))
    assert row_runs(sub, 12, 10) is None


def test_struct_layouts_match_the_c_header(tmp_path):
    """sizeof/offsetof of every ABI struct, from the C compiler, equal the
    ctypes mirrors (catches field drift between include/mbir_hip.h and
    _lib.py)."""
    import ctypes
    import shutil
    import subprocess
    cc = shutil.which('gcc') or shutil.which('cc')
    if cc is None:
        pytest.skip('no C compiler')
    structs = {'mbir_par_geom': _lib.ParGeom, 'mbir_cone_geom': _lib.ConeGeom,
               'mbir_vcd_params': _lib.VcdParams, 'mbir_vcd_job': _lib.VcdJob}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mbir_hip.h"',
             'int main(void) {']
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'printf("{cname}.{field} %zu\\n", offsetof({cname}, {field}));')
    lines += ['return 0;', '}']
    src = tmp_path / 'layout.c'
    src.write_text('\n'.join(lines))
    exe = tmp_path / 'layout'
    subprocess.run([cc, '-I', os.path.join(ROOT, 'include'), str(src), '-o', str(exe)],
                   check=True)
    got = dict(line.split() for line in subprocess.run(
        [str(exe)], check=True, capture_output=True, text=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for field, _ in py._fields_:
            assert int(got[f'{cname}.{field}']) == getattr(py, field).offset, (cname, field)


def test_memory_ledger_closed_form_and_admission():
    from mbirtorch_amd.serving import MemoryLedger, request_device_bytes
    g = [1, 2, 4, 8, 16, 32, 64, 128, 128, 128, 128]
    small = request_device_bytes((90, 8, 128), (128, 128, 8), g)
    big = request_device_bytes((720, 32, 512), (512, 512, 32), g)
    assert big['total'] > 50 * small['total']
    # The resident part holds at least the error sinogram, recon and Hessian.
    assert big['resident'] >= 4 * 720 * 32 * 512 + 2 * 4 * 512 * 512 * 32
    assert request_device_bytes((90, 8, 128), (128, 128, 8), g, weighted=True)['resident'] \
        == small['resident'] + 4 * 90 * 8 * 128
    led = MemoryLedger(1000, margin=0.2)
    assert led.capacity == 800
    assert led.try_reserve(500) and not led.try_reserve(400) and led.try_reserve(300)
    led.release(500)
    assert led.used == 300 and led.fits_alone(800) and not led.fits_alone(801)


def test_runner_and_engine_reject_bad_arguments_before_any_launch():
    """mbir_vcd_run validates the job, the step count and every subset id on
    the host (it returns MBIR_ERR_ARG before a launch, so fake device pointers
    are never touched); without a device the request engine is not created."""
    import ctypes
    lib = _lib.lib()
    g = _lib.ParGeom()
    job = _lib.VcdJob()
    fake = ctypes.c_void_p(0x1000)
    job.geometry = _lib.MBIR_GEOMETRY_PARALLEL
    job.num_subsets = 2
    job.sino_dtype = _lib.MBIR_DTYPE_F32
    job.par_geom = ctypes.pointer(g)
    job.prm.num_slices = 4
    for name in ('view_table', 'flat_recon', 'error_vcr', 'hessian', 'back_parts', 'delta',
                 'dir_parts', 'fwd_parts', 'alpha', 'apply_scratch', 'red_scratch'):
        setattr(job, name, fake)
    sizes = (ctypes.c_int32 * 2)(8, 8)
    pix = (ctypes.c_void_p * 2)(0x2000, 0x3000)
    job.subset_sizes = sizes
    job.subset_pixels = ctypes.cast(pix, ctypes.POINTER(ctypes.c_void_p))
    stats = ctypes.c_void_p(0x4000)
    bad = (ctypes.c_int32 * 1)(2)  # subset id out of range
    assert lib.mbir_vcd_run(ctypes.byref(job), ctypes.cast(bad, ctypes.c_void_p), 1, stats,
                            0, None) == _lib.MBIR_ERR_ARG
    neg = (ctypes.c_int32 * 1)(-1)
    assert lib.mbir_vcd_run(ctypes.byref(job), ctypes.cast(neg, ctypes.c_void_p), 1, stats,
                            0, None) == _lib.MBIR_ERR_ARG
    assert lib.mbir_vcd_run(ctypes.byref(job), None, -1, stats, 0, None) == _lib.MBIR_ERR_ARG
    assert lib.mbir_vcd_run(ctypes.byref(job), None, 0, None, 0, None) == _lib.MBIR_ERR_ARG
    job.sino_dtype = 7
    assert lib.mbir_vcd_run(ctypes.byref(job), None, 0, stats, 0, None) == _lib.MBIR_ERR_ARG
    if not torch.cuda.is_available():
        assert not lib.mbir_server_create(4)
    assert not lib.mbir_server_create(0) and not lib.mbir_server_create(65)


def test_direction_partials_capacity_covers_every_subset_size():
    """The engine sizes the direction partials for the largest count over a
    run's subset sizes, which at 128x128x8 is not the largest subset's."""
    from mbirtorch_amd import _lib
    from mbirtorch_amd.tomography_model import _dir_partials_capacity
    lib = _lib.lib()
    rs, S = (128, 128, 8), 8
    sizes = [int(mb.gen_pixel_partition(rs, g).shape[1]) for g in (1, 2, 4, 8, 16, 32, 64, 128)]
    counts = [int(lib.mbir_vcd_direction_partials_count(P, S)) for P in sizes]
    cap = _dir_partials_capacity(sizes, S)
    assert cap == max(counts) and all(c <= cap for c in counts)
    assert max(counts) > counts[0]  # the largest subset alone would under-size it
