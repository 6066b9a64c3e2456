"""ParallelBeamModel over the gfx950 parallel-beam projectors.

Constructor, parameters, recon geometry, psf radius, validity checks, FBP and
recon_simple_parallel follow parallel_beam.py:97-379,672-722 of the reference.
The projector itself (``ParallelProjector``) replaces the reference's
per-view-batch bodies (parallel_beam.py:56-94, triton_parallel.py) with the
HIP entry points of libmbir_hip.so, always over the whole view range.
"""

import numpy as np
import torch

from . import _lib
from .tomography_model import TomographyModel


def parallel_view_table(angles, delta_voxel, delta_voxel_row, delta_det_channel):
    """Per-view float32 constants (cos, sin, W, ws, (W+1)/2, min(W,1), 0, 0),
    each formed with the reference's float32 operations and order
    (parallel_beam.py:38-51, horizontal_fan.py:68-69) on the host CPU."""
    a = torch.as_tensor(np.asarray(angles, dtype=np.float32))
    c, s = torch.cos(a), torch.sin(a)
    fp = torch.maximum(c.abs() * delta_voxel, s.abs() * delta_voxel_row)
    W = fp / delta_det_channel
    ws = (delta_voxel_row * delta_voxel) / fp
    halfw = (W + 1.0) / 2.0
    lmax = torch.clamp(W, max=1.0)
    z = torch.zeros_like(c)
    return torch.stack([c, s, W, ws, halfw, lmax, z, z], dim=1).contiguous()


def row_runs(idx, num_rows, num_cols):
    """Image-row run table (int32 (num_rows, 3): list position of the first
    listed pixel of the row, its column, run length) when every image row's
    listed pixels are consecutive columns in increasing list order, else
    None.  One device-to-host synchronisation."""
    P = int(idx.shape[0])
    if P == 0:
        return None
    rows = torch.div(idx, num_cols, rounding_mode='floor')
    cols = idx - rows * num_cols
    if P > 1:
        d = idx[1:] - idx[:-1]
        same = rows[1:] == rows[:-1]
        ok = torch.all(d > 0) & torch.all(~same | (d == 1))
    else:
        ok = torch.ones((), dtype=torch.bool, device=idx.device)
    ok = ok & (rows[0] >= 0) & (rows[-1] < num_rows)
    if not bool(ok):
        return None
    counts = torch.bincount(rows, minlength=num_rows)
    starts = torch.cumsum(counts, 0) - counts
    jmin = torch.where(counts > 0, cols[starts.clamp(max=P - 1)], torch.zeros_like(counts))
    return torch.stack([starts, jmin, counts], dim=1).to(torch.int32).contiguous()


class ParallelProjector:
    """Device projector for fixed parallel-beam parameters."""

    uses_row_runs = True
    geometry_id = _lib.MBIR_GEOMETRY_PARALLEL

    def __init__(self, model, device):
        (ddc, dco, dv, vra) = model.get_params(['delta_det_channel', 'det_channel_offset',
                                                'delta_voxel', 'voxel_row_aspect'])
        V, R, C = model.get_params('sinogram_shape')
        rs = model.get_params('recon_shape')
        dvr = vra * dv
        angles = np.asarray(model.get_params('angles'), dtype=np.float32)
        table = parallel_view_table(angles, dv, dvr, ddc)
        g = _lib.ParGeom()
        g.num_rows, g.num_cols, g.num_slices = int(rs[0]), int(rs[1]), int(rs[2])
        g.num_channels, g.num_views = int(C), int(V)
        g.psf_radius = int(model.get_psf_radius())
        g.two_tap = int(bool(torch.all(table[:, 2] <= 1.0)))
        g.unit_channel = int(np.float32(ddc) == np.float32(1.0))
        g.delta_det_channel = np.float32(ddc)
        g.det_channel_offset = np.float32(dco)
        g.delta_voxel = np.float32(dv)
        g.delta_voxel_row = np.float32(dvr)
        g.det_center = np.float32((C - 1) / 2.0)
        g.row_center = np.float32((rs[0] - 1) / 2.0)
        g.col_center = np.float32((rs[1] - 1) / 2.0)
        self.geom = g
        self.table_host = table
        self.table = table.to(device)
        self.device = device
        self.shape = (int(V), int(R), int(C))
        self.num_slices = int(rs[2])

    # sizes
    def back_splits(self, P):
        return int(_lib.lib().mbir_par_back_splits(self.geom, int(P)))

    def fwd_partials_count(self, ray=False):
        return int(_lib.lib().mbir_par_fwd_partials_count(self.geom, int(bool(ray))))

    def _check_values(self, values, idx):
        if values.dim() != 2 or values.shape[1] != self.num_slices:
            raise ValueError(f'voxel values must be (num_pixels, {self.num_slices}); '
                             f'got {tuple(values.shape)}')
        if idx.dim() != 1 or idx.shape[0] != values.shape[0]:
            raise ValueError('pixel_indices must be 1-D with one entry per voxel cylinder')

    def forward(self, values, idx, layout=_lib.MBIR_LAYOUT_VRC, runs=None):
        """Forward projection of cylinders ``values`` at ``idx``.  ``runs`` is
        the image-row run table of ``idx`` (see row_runs); None derives it,
        False forces the binned gather."""
        _lib.require_gpu(values)
        values = values.contiguous()
        self._check_values(values, idx)
        if runs is None:
            runs = row_runs(idx, self.geom.num_rows, self.geom.num_cols)
        V, R, C = self.shape
        shape = (V, R, C) if layout == _lib.MBIR_LAYOUT_VRC else (V, C, R)
        out = torch.empty(shape, dtype=torch.float32, device=values.device)
        _lib.check(_lib.lib().mbir_par_forward(
            _lib.ptr(values), _lib.ptr(idx), int(idx.shape[0]),
            _lib.ptr(runs) if runs is not False and runs is not None else None,
            _lib.ptr(self.table), self.geom, layout, _lib.ptr(out), _lib.stream_ptr()),
            'parallel forward')
        return out

    def back_into(self, sino_vcr, idx, parts, ns, weights=None):
        _lib.check(_lib.lib().mbir_par_back(
            _lib.ptr(sino_vcr), _lib.ptr(weights), _lib.ptr(idx), int(idx.shape[0]),
            _lib.ptr(self.table), self.geom, 1, ns, _lib.ptr(parts), _lib.stream_ptr()),
            'parallel back')

    def back(self, sino_vcr, idx, coeff_power=1, weights=None):
        _lib.require_gpu(sino_vcr)
        V, R, C = self.shape
        if tuple(sino_vcr.shape) != (V, C, R):
            raise ValueError(f'channel-major sinogram must be {(V, C, R)}; '
                             f'got {tuple(sino_vcr.shape)}')
        P = int(idx.shape[0])
        ns = self.back_splits(P)
        parts = torch.empty((ns, P, R), dtype=torch.float32, device=sino_vcr.device)
        _lib.check(_lib.lib().mbir_par_back(
            _lib.ptr(sino_vcr), _lib.ptr(weights), _lib.ptr(idx), P, _lib.ptr(self.table),
            self.geom, int(coeff_power), ns, _lib.ptr(parts), _lib.stream_ptr()),
            'parallel back')
        if ns == 1:
            return parts[0]
        out = torch.empty((P, R), dtype=torch.float32, device=sino_vcr.device)
        _lib.check(_lib.lib().mbir_reduce_splits(_lib.ptr(parts), ns, P * R, _lib.ptr(out),
                                                 _lib.stream_ptr()), 'reduce splits')
        return out

    def fwd_linesearch(self, delta, idx, e_vcr, w_vcr, partials, runs=None, adelta=None):
        """FP64 line-search partials of A(delta) against the error sinogram;
        ``runs`` (row_runs of ``idx``, or None) selects the ray-driven form;
        ``adelta`` (channel-major float32 sinogram, or None) receives A(delta)."""
        _lib.check(_lib.lib().mbir_par_fwd_linesearch(
            _lib.ptr(delta), _lib.ptr(idx), int(idx.shape[0]), _lib.ptr(runs),
            _lib.ptr(self.table), self.geom, _lib.ptr(e_vcr), _lib.ptr(w_vcr),
            _lib.ptr(adelta), _lib.ptr(partials), _lib.stream_ptr()),
            'parallel forward line search')

    def fwd_update(self, delta, idx, alpha, e_vcr, runs=None):
        _lib.check(_lib.lib().mbir_par_fwd_update(
            _lib.ptr(delta), _lib.ptr(idx), int(idx.shape[0]), _lib.ptr(runs),
            _lib.ptr(self.table), self.geom, _lib.ptr(alpha), _lib.ptr(e_vcr),
            _lib.stream_ptr()), 'parallel forward update')


class ParallelBeamModel(TomographyModel):
    """Parallel-beam geometry (parallel_beam.py:97-379).

    Args:
        sinogram_shape (tuple): (views, rows, channels).
        angles (ndarray): view angles in radians, one per view.
        view_batch_size, compile_mode: accepted for API compatibility; the
            HIP projectors always process the whole view range per launch.
    """

    rows_track_slices = True

    def __init__(self, sinogram_shape, angles, view_batch_size=None, compile_mode='auto'):
        angles = np.asarray(angles, dtype=np.float32)
        super().__init__(sinogram_shape, view_batch_size=view_batch_size,
                         compile_mode=compile_mode, geometry_type='parallel',
                         view_params_name='angles', angles=angles)

    def get_magnification(self):
        return 1.0

    def get_psf_radius(self):
        ddc, dv, vra = self.get_params(['delta_det_channel', 'delta_voxel', 'voxel_row_aspect'])
        return int(np.ceil(np.ceil(max(dv, vra * dv) / ddc) / 2))

    def auto_set_recon_geometry(self, no_compile=False, no_warning=False):
        ddr, ddc, vra = self.get_params(['delta_det_row', 'delta_det_channel',
                                         'voxel_row_aspect'])
        mag = self.get_magnification()
        dv = ddc / mag
        dvr = vra * dv
        _, rows, channels = self.get_params('sinogram_shape')
        nr = int(np.ceil(channels * ddc / (dvr * mag)))
        nc = int(np.ceil(channels * ddc / (dv * mag)))
        ns = int(np.round(rows * ((ddr / dv) / mag)))
        self.set_params(no_compile=no_compile, no_warning=no_warning,
                        recon_shape=(nr, nc, ns), delta_voxel=dv)

    def verify_valid_params(self):
        super().verify_valid_params()
        sino_shape, angles, vra, vsa = self.get_params(['sinogram_shape', 'angles',
                                                        'voxel_row_aspect',
                                                        'voxel_slice_aspect'])
        if vra <= 0:
            raise ValueError('Voxel row aspect ratio must be positive. \n'
                             f'Got {vra} for voxel_row_aspect.')
        if vsa != 1.0:
            raise ValueError('Setting voxel slice aspect ratio is not supported for parallel '
                             f'beam model. \nGot {vsa}.')
        if np.asarray(angles).shape[0] != sino_shape[0]:
            raise ValueError('Number of view dependent parameter vectors must equal the '
                             'number of views.')
        rs = self.get_params('recon_shape')
        if rs[2] != sino_shape[1]:
            raise ValueError('Number of recon slices must match number of sinogram rows. \n'
                             f'Got {rs} for recon_shape and {sino_shape} for sinogram_shape')

    def _projector_key(self):
        names = ['sinogram_shape', 'recon_shape', 'delta_det_channel', 'det_channel_offset',
                 'delta_voxel', 'voxel_row_aspect']
        vals = tuple(tuple(v) if isinstance(v, (list, tuple)) else v
                     for v in self.get_params(names))
        angles = np.asarray(self.get_params('angles'), dtype=np.float32)
        return vals + (angles.tobytes(), str(self.torch_device))

    def _make_projector(self):
        return ParallelProjector(self, self.torch_device)

    # -- FBP ------------------------------------------------------------------
    def fbp_filter(self, sinogram, filter_name='ramp', output_sharded=False):
        dv, vra = self.get_params(['delta_voxel', 'voxel_row_aspect'])
        scale = 1.0 / (dv * (vra * dv))
        return self._apply_direct_recon_filter(sinogram, filter_name, filter_scale=scale,
                                               output_sharded=output_sharded)

    def _direct_recon_flat(self, sinogram, filter_name='ramp'):
        dv, vra = self.get_params(['delta_voxel', 'voxel_row_aspect'])
        scale = 1.0 / (dv * (vra * dv))
        filt = self._apply_direct_recon_filter(sinogram, filter_name, filter_scale=scale,
                                               layout=_lib.MBIR_LAYOUT_VCR)
        idx = self.full_indices_device()
        cyl = self.projector().back(filt, idx)
        return self._scatter_flat(cyl, idx)

    def recon_fbp(self, sinogram, filter_name='ramp', output_sharded=False):
        rs = self.get_params('recon_shape')
        recon = self._direct_recon_flat(sinogram, filter_name).reshape(rs[0], rs[1], -1)
        return recon if output_sharded else self._to_host(recon)

    def recon_direct(self, sinogram, filter_name='ramp', output_sharded=False):
        return self.recon_fbp(sinogram, filter_name=filter_name, output_sharded=output_sharded)

    def direct_filter(self, sinogram, filter_name='ramp', output_sharded=False):
        return self.fbp_filter(sinogram, filter_name=filter_name, output_sharded=output_sharded)


def recon_simple_parallel(sinogram, angles, weights=None, sharpness=1.0, max_iterations=15):
    """One-call parallel-beam reconstruction (parallel_beam.py:672-722)."""
    if torch.is_tensor(angles):
        angles = angles.detach().cpu().numpy()
    model = ParallelBeamModel(tuple(sinogram.shape), angles)
    model.set_params(sharpness=sharpness)
    return model.recon(sinogram, weights=weights, max_iterations=max_iterations)
