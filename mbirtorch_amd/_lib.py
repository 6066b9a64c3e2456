"""ctypes binding of libmbir_hip.so (declared in include/mbir_hip.h).

The library is built in-tree by ``mbirtorch_amd.build_lib``.  There is no
fallback: every projection, filter and VCD step runs through these entry
points, and a missing library or a missing GPU raises immediately.
"""

import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libmbir_hip.so')

MBIR_LAYOUT_VRC = 0
MBIR_LAYOUT_VCR = 1

# Status codes of include/mbir_hip.h.
MBIR_OK = 0
MBIR_ERR_SHAPE = -1
MBIR_ERR_GEOMETRY = -2
MBIR_ERR_LAUNCH = -3
MBIR_ERR_WORKSPACE = -4
MBIR_ERR_ARG = -5

_ERRORS = {-1: (ValueError, 'shape mismatch'), -2: (ValueError, 'invalid geometry'),
           -3: (RuntimeError, 'kernel launch failed'),
           -4: (RuntimeError, 'workspace too small'), -5: (ValueError, 'invalid argument')}


class ParGeom(ctypes.Structure):
    _fields_ = [('num_rows', ctypes.c_int32), ('num_cols', ctypes.c_int32),
                ('num_slices', ctypes.c_int32), ('num_channels', ctypes.c_int32),
                ('num_views', ctypes.c_int32), ('psf_radius', ctypes.c_int32),
                ('two_tap', ctypes.c_int32), ('unit_channel', ctypes.c_int32),
                ('delta_det_channel', ctypes.c_float),
                ('det_channel_offset', ctypes.c_float),
                ('delta_voxel', ctypes.c_float), ('delta_voxel_row', ctypes.c_float),
                ('det_center', ctypes.c_float), ('row_center', ctypes.c_float),
                ('col_center', ctypes.c_float), ('pad', ctypes.c_float)]


class ConeGeom(ctypes.Structure):
    _fields_ = [('num_rows', ctypes.c_int32), ('num_cols', ctypes.c_int32),
                ('num_slices', ctypes.c_int32), ('num_det_rows', ctypes.c_int32),
                ('num_channels', ctypes.c_int32), ('num_views', ctypes.c_int32),
                ('psf_radius', ctypes.c_int32), ('bp_psf_radius', ctypes.c_int32),
                ('curved', ctypes.c_int32), ('unit_channel', ctypes.c_int32),
                ('pad0', ctypes.c_int32), ('pad1', ctypes.c_int32),
                ('delta_voxel', ctypes.c_float), ('delta_voxel_row', ctypes.c_float),
                ('delta_voxel_slice', ctypes.c_float), ('delta_det_channel', ctypes.c_float),
                ('delta_det_row', ctypes.c_float), ('det_channel_offset', ctypes.c_float),
                ('det_row_offset', ctypes.c_float), ('source_detector_dist', ctypes.c_float),
                ('source_iso_dist', ctypes.c_float), ('inv_magnification', ctypes.c_float),
                ('det_center_channel', ctypes.c_float), ('det_center_row', ctypes.c_float),
                ('row_center', ctypes.c_float), ('col_center', ctypes.c_float),
                ('slice_center', ctypes.c_float), ('half_slice_span', ctypes.c_float),
                ('area', ctypes.c_float), ('pad2', ctypes.c_float)]


class VcdParams(ctypes.Structure):
    _fields_ = [('b', ctypes.c_float * 6), ('two_b', ctypes.c_float * 6),
                ('sigma_x', ctypes.c_float), ('p', ctypes.c_float), ('q', ctypes.c_float),
                ('T', ctypes.c_float), ('a_min', ctypes.c_float), ('t_sigma', ctypes.c_float),
                ('q_minus_p', ctypes.c_float), ('p_minus_2', ctypes.c_float),
                ('q_over_p', ctypes.c_float), ('two_sigma_p', ctypes.c_float),
                ('fm_constant', ctypes.c_float), ('neg_fm_constant', ctypes.c_float),
                ('inv_sigma_prox2', ctypes.c_float), ('max_alpha', ctypes.c_float),
                ('num_rows', ctypes.c_int32), ('num_cols', ctypes.c_int32),
                ('num_slices', ctypes.c_int32), ('use_prox', ctypes.c_int32)]


MBIR_GEOMETRY_PARALLEL = 0
MBIR_GEOMETRY_CONE = 1
MBIR_DTYPE_F32 = 0
MBIR_DTYPE_BF16 = 1


class VcdJob(ctypes.Structure):
    """mbir_vcd_job: one reconstruction's device state and subset tables."""
    _fields_ = [('geometry', ctypes.c_int32), ('positivity', ctypes.c_int32),
                ('num_subsets', ctypes.c_int32), ('sino_dtype', ctypes.c_int32),
                ('par_geom', ctypes.POINTER(ParGeom)), ('cone_geom', ctypes.POINTER(ConeGeom)),
                ('view_table', ctypes.c_void_p), ('prm', VcdParams),
                ('flat_recon', ctypes.c_void_p), ('error_vcr', ctypes.c_void_p),
                ('weights_vcr', ctypes.c_void_p), ('hessian', ctypes.c_void_p),
                ('prox', ctypes.c_void_p), ('dc_profile', ctypes.c_void_p),
                ('subset_pixels', ctypes.POINTER(ctypes.c_void_p)),
                ('subset_sizes', ctypes.POINTER(ctypes.c_int32)),
                ('subset_runs', ctypes.POINTER(ctypes.c_void_p)),
                ('back_parts', ctypes.c_void_p), ('back_parts_capacity', ctypes.c_int64),
                ('delta', ctypes.c_void_p), ('dir_parts', ctypes.c_void_p),
                ('fwd_parts', ctypes.c_void_p), ('alpha', ctypes.c_void_p),
                ('apply_scratch', ctypes.c_void_p), ('dc_scratch', ctypes.c_void_p),
                ('red_scratch', ctypes.c_void_p), ('sino_size', ctypes.c_int64),
                ('recon_size', ctypes.c_int64), ('adelta', ctypes.c_void_p)]


_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_GP = ctypes.POINTER(ParGeom)
_VP = ctypes.POINTER(VcdParams)
_CP = ctypes.POINTER(ConeGeom)

# name -> (restype, argtypes)
SIGNATURES = {
    'mbir_version': (ctypes.c_int, []),
    'mbir_par_forward': (ctypes.c_int, [_P, _P, _I32, _P, _P, _GP, _I32, _P, _P]),
    'mbir_par_back_splits': (_I32, [_GP, _I32]),
    'mbir_par_back': (ctypes.c_int, [_P, _P, _P, _I32, _P, _GP, _I32, _I32, _P, _P]),
    'mbir_reduce_splits': (ctypes.c_int, [_P, _I32, _I64, _P, _P]),
    'mbir_par_fwd_partials_count': (_I32, [_GP, _I32]),
    'mbir_par_fwd_linesearch': (ctypes.c_int, [_P, _P, _I32, _P, _P, _GP, _P, _P, _P, _P, _P]),
    'mbir_par_fwd_update': (ctypes.c_int, [_P, _P, _I32, _P, _P, _GP, _P, _P, _P]),
    'mbir_cone_forward': (ctypes.c_int, [_P, _P, _I32, _P, _CP, _I32, _P, _P]),
    'mbir_cone_back_splits': (_I32, [_CP, _I32]),
    'mbir_cone_back': (ctypes.c_int, [_P, _P, _P, _I32, _P, _CP, _I32, _I32, _P, _P]),
    'mbir_cone_fwd_partials_count': (_I32, [_CP]),
    'mbir_cone_fwd_linesearch': (ctypes.c_int, [_P, _P, _I32, _P, _CP, _P, _P, _P, _P, _P]),
    'mbir_cone_fwd_update': (ctypes.c_int, [_P, _P, _I32, _P, _CP, _P, _P, _P]),
    'mbir_vcd_direction_partials_count': (_I32, [_I32, _I32]),
    'mbir_vcd_direction_dc_scratch': (_I64, [_I32, _I32]),
    'mbir_vcd_direction_dc': (ctypes.c_int, [_P, _P, _I32, _P, _I32, _P, _P, _VP, _P, _P, _P,
                                             _P, _P]),
    'mbir_vcd_direction': (ctypes.c_int, [_P, _P, _I32, _P, _I32, _P, _P, _VP, _P, _P, _P]),
    'mbir_vcd_alpha': (ctypes.c_int, [_P, _I32, _P, _I32, _VP, _P, _P, _P]),
    'mbir_vcd_positivity': (ctypes.c_int, [_P, _P, _I32, _I32, _P, _P, _P]),
    'mbir_vcd_apply_scratch': (ctypes.c_int64, [_I32, _I32]),
    'mbir_vcd_error_update': (ctypes.c_int, [_P, _P, _P, _I64, _P]),
    'mbir_vcd_error_update_bf16': (ctypes.c_int, [_P, _P, _P, _I64, _P]),
    'mbir_par_back_bf16': (ctypes.c_int, [_P, _P, _P, _I32, _P, _GP, _I32, _P, _P]),
    'mbir_par_fwd_linesearch_bf16': (ctypes.c_int, [_P, _P, _I32, _P, _GP, _P, _P, _P, _P, _P]),
    'mbir_par_fwd_update_bf16': (ctypes.c_int, [_P, _P, _I32, _P, _GP, _P, _P, _P]),
    'mbir_cone_back_bf16': (ctypes.c_int, [_P, _P, _P, _I32, _P, _CP, _I32, _P, _P]),
    'mbir_cone_fwd_linesearch_bf16': (ctypes.c_int, [_P, _P, _I32, _P, _CP, _P, _P, _P, _P, _P]),
    'mbir_cone_fwd_update_bf16': (ctypes.c_int, [_P, _P, _I32, _P, _CP, _P, _P, _P]),
    'mbir_sumsq_bf16': (ctypes.c_int, [_P, _P, _I64, _P, _P, _P]),
    'mbir_f32_to_bf16': (ctypes.c_int, [_P, _P, _I64, _P]),
    'mbir_bf16_to_f32': (ctypes.c_int, [_P, _P, _I64, _P]),
    'mbir_vcd_run': (ctypes.c_int, [ctypes.POINTER(VcdJob), _P, _I32, _P, _I32, _P]),
    'mbir_server_create': (ctypes.c_void_p, [_I32]),
    'mbir_server_submit': (ctypes.c_int, [_P, ctypes.POINTER(VcdJob), _P, _P, _I32, _P, _I64]),
    'mbir_server_poll': (_I32, [_P, _P, _P, _I32, _I32]),
    'mbir_server_in_flight': (_I32, [_P]),
    'mbir_server_destroy': (None, [_P]),
    'mbir_vcd_apply': (ctypes.c_int, [_P, _P, _I32, _I32, _P, _P, _P, _P, _P]),
    'mbir_vcd_finish': (ctypes.c_int, [_P, _P, _I64, _P, _P, _I32, _I32, _P, _P, _I32, _P,
                                       _I32, _VP, _P, _P, _P, _P]),
    'mbir_vcd_finish_bf16': (ctypes.c_int, [_P, _P, _I64, _P, _P, _I32, _I32, _P, _P, _I32,
                                            _P, _I32, _VP, _P, _P, _P, _P]),
    'mbir_vcd_block_stats_rows': (_I32, [_I32]),
    'mbir_vcd_block_stats_reduce': (ctypes.c_int, [_P, _I32, _I32, _P, _P]),
    'mbir_reduce_blocks': (_I32, [_I64]),
    'mbir_sumsq_f64': (ctypes.c_int, [_P, _P, _I64, _P, _P, _P]),
    'mbir_sumabs_f64': (ctypes.c_int, [_P, _I64, _P, _P, _P]),
    'mbir_dot_f64': (ctypes.c_int, [_P, _P, _I64, _P, _P, _P]),
    'mbir_row_filter': (ctypes.c_int, [_P, _P, _P, _I32, _I32, _I32, _I32, _P, _P]),
    'mbir_transpose_sino': (ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _P, _P]),
}

_lib = None
_lock = threading.Lock()


def load(path=LIB_PATH):
    """Load the library (no GPU needed) and bind every declared symbol."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise RuntimeError(
                f'libmbir_hip.so not found at {path}; build it with '
                '`python -m mbirtorch_amd.build_lib` (there is no CPU fallback)')
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def lib():
    return _lib if _lib is not None else load()


def check(status, what):
    if status != 0:
        exc, msg = _ERRORS.get(status, (RuntimeError, 'unknown error'))
        raise exc(f'{what}: {msg} (status {status})')


def require_gpu(t=None):
    """Raise unless a HIP device is present (and ``t`` lives on it)."""
    if not torch.cuda.is_available():
        raise RuntimeError('mbirtorch_amd needs a HIP (MI355X) device; none is visible '
                           '(there is no CPU fallback)')
    if t is not None and t.device.type != 'cuda':
        raise ValueError(f'expected a device tensor, got one on {t.device}')


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
