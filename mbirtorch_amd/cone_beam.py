"""ConeBeamModel over the gfx950 cone-beam projectors.

Constructor, parameters, magnification, psf radii, auto recon geometry
(with helical travel and axial padding), validity checks, FDK, the helical
FDK z weight, DC damping and recon_simple_cone follow cone_beam.py:302-1190
of the reference.  ``ConeProjector`` replaces the per-view-batch bodies
(cone_beam.py:161-299, triton_cone.py) with the HIP entry points of
libmbir_hip.so, always over the whole view range.
"""

import math
import warnings

import numpy as np
import torch

from . import _lib
from .tomography_model import TomographyModel
from .vcd_utils import get_2d_ror_mask

# (a, b, p, c) of the per-slice damping s_k = (c t^p + a b^p) / (t^p + b^p)
# (cone_beam.py:39-45); None disables it.
DC_DAMPING_DEFAULT = (0.25, 100.0, 0.7, 0.5)


def cone_view_table(view_params, recon_slice_offset):
    """Per-view float32 (cos, sin, angle, recon_slice_offset - z_shift),
    formed with torch float32 ops on the host (cone_beam.py:71-72, 152)."""
    vp = torch.as_tensor(np.asarray(view_params, dtype=np.float32))
    a, zs = vp[:, 0], vp[:, 1]
    z_off = recon_slice_offset - zs
    return torch.stack([torch.cos(a), torch.sin(a), a, z_off], dim=1).contiguous()


def _support_radius(recon_shape, dvr, dv, use_ror_mask=True):
    """vcd_utils.py:82-94."""
    hr = 0.5 * recon_shape[0] * dvr
    hc = 0.5 * recon_shape[1] * dv
    if use_ror_mask is True:
        return float(max(hr, hc))
    return float(np.hypot(hr, hc))


class ConeProjector:
    """Device projector for fixed cone-beam parameters."""

    geometry_id = _lib.MBIR_GEOMETRY_CONE

    def __init__(self, model, device):
        V, R, C = model.get_params('sinogram_shape')
        rs = model.get_params('recon_shape')
        (ddr, ddc, dro, dco, sdd, dv, vra, vsa, rso, curved) = model.get_params(
            ['delta_det_row', 'delta_det_channel', 'det_row_offset', 'det_channel_offset',
             'source_detector_dist', 'delta_voxel', 'voxel_row_aspect', 'voxel_slice_aspect',
             'recon_slice_offset', 'use_curved_detector'])
        mag = model.get_magnification()
        psf, bp = model.get_psf_radii()
        f32 = np.float32
        g = _lib.ConeGeom()
        g.num_rows, g.num_cols, g.num_slices = int(rs[0]), int(rs[1]), int(rs[2])
        g.num_det_rows, g.num_channels, g.num_views = int(R), int(C), int(V)
        g.psf_radius, g.bp_psf_radius = int(psf), int(bp)
        g.curved = int(bool(curved))
        g.unit_channel = int(f32(ddc) == f32(1.0))
        dvr, dvs = vra * dv, vsa * dv
        g.delta_voxel, g.delta_voxel_row, g.delta_voxel_slice = f32(dv), f32(dvr), f32(dvs)
        g.delta_det_channel, g.delta_det_row = f32(ddc), f32(ddr)
        g.det_channel_offset, g.det_row_offset = f32(dco), f32(dro)
        g.source_detector_dist = f32(sdd)
        g.source_iso_dist = f32(sdd / mag) if not math.isinf(sdd) else f32(np.inf)
        g.inv_magnification = f32(1.0 / mag)
        g.det_center_channel = f32((C - 1) / 2.0)
        g.det_center_row = f32((R - 1) / 2.0)
        g.row_center = f32((rs[0] - 1) / 2.0)
        g.col_center = f32((rs[1] - 1) / 2.0)
        g.slice_center = f32((rs[2] - 1) / 2.0)
        g.half_slice_span = f32(dvs * (rs[2] - 1) / 2.0)
        g.area = f32(dvr * dv)
        self.geom = g
        self.table = cone_view_table(model.get_params('view_params_array'), rso).to(device)
        self.device = device
        self.shape = (int(V), int(R), int(C))
        self.num_slices = int(rs[2])

    def back_splits(self, P):
        return int(_lib.lib().mbir_cone_back_splits(self.geom, int(P)))

    def fwd_partials_count(self, ray=False):
        del ray  # one cone forward form
        return int(_lib.lib().mbir_cone_fwd_partials_count(self.geom))

    def forward(self, values, idx, layout=_lib.MBIR_LAYOUT_VRC, runs=None):
        """Forward projection of cylinders ``values`` at ``idx``.  ``runs`` (the
        image-row run table the parallel projector uses to pick its gather) is
        accepted for the shared projector interface; the cone forward has one
        form."""
        del runs
        _lib.require_gpu(values)
        values = values.contiguous()
        if values.dim() != 2 or values.shape[1] != self.num_slices:
            raise ValueError(f'voxel values must be (num_pixels, {self.num_slices}); '
                             f'got {tuple(values.shape)}')
        if idx.dim() != 1 or idx.shape[0] != values.shape[0]:
            raise ValueError('pixel_indices must be 1-D with one entry per voxel cylinder')
        V, R, C = self.shape
        shape = (V, R, C) if layout == _lib.MBIR_LAYOUT_VRC else (V, C, R)
        out = torch.empty(shape, dtype=torch.float32, device=values.device)
        _lib.check(_lib.lib().mbir_cone_forward(
            _lib.ptr(values), _lib.ptr(idx), int(idx.shape[0]), _lib.ptr(self.table),
            self.geom, layout, _lib.ptr(out), _lib.stream_ptr()), 'cone forward')
        return out

    def back_into(self, sino_vcr, idx, parts, ns, weights=None):
        _lib.check(_lib.lib().mbir_cone_back(
            _lib.ptr(sino_vcr), _lib.ptr(weights), _lib.ptr(idx), int(idx.shape[0]),
            _lib.ptr(self.table), self.geom, 1, ns, _lib.ptr(parts), _lib.stream_ptr()),
            'cone back')

    def back(self, sino_vcr, idx, coeff_power=1, weights=None):
        _lib.require_gpu(sino_vcr)
        V, R, C = self.shape
        if tuple(sino_vcr.shape) != (V, C, R):
            raise ValueError(f'channel-major sinogram must be {(V, C, R)}; '
                             f'got {tuple(sino_vcr.shape)}')
        P = int(idx.shape[0])
        S = self.num_slices
        ns = self.back_splits(P)
        parts = torch.empty((ns, P, S), dtype=torch.float32, device=sino_vcr.device)
        _lib.check(_lib.lib().mbir_cone_back(
            _lib.ptr(sino_vcr), _lib.ptr(weights), _lib.ptr(idx), P, _lib.ptr(self.table),
            self.geom, int(coeff_power), ns, _lib.ptr(parts), _lib.stream_ptr()), 'cone back')
        if ns == 1:
            return parts[0]
        out = torch.empty((P, S), dtype=torch.float32, device=sino_vcr.device)
        _lib.check(_lib.lib().mbir_reduce_splits(_lib.ptr(parts), ns, P * S, _lib.ptr(out),
                                                 _lib.stream_ptr()), 'reduce splits')
        return out

    def fwd_linesearch(self, delta, idx, e_vcr, w_vcr, partials, runs=None, adelta=None):
        del runs  # one cone forward form
        _lib.check(_lib.lib().mbir_cone_fwd_linesearch(
            _lib.ptr(delta), _lib.ptr(idx), int(idx.shape[0]), _lib.ptr(self.table),
            self.geom, _lib.ptr(e_vcr), _lib.ptr(w_vcr), _lib.ptr(adelta), _lib.ptr(partials),
            _lib.stream_ptr()), 'cone forward line search')

    def fwd_update(self, delta, idx, alpha, e_vcr, runs=None):
        del runs
        _lib.check(_lib.lib().mbir_cone_fwd_update(
            _lib.ptr(delta), _lib.ptr(idx), int(idx.shape[0]), _lib.ptr(self.table),
            self.geom, _lib.ptr(alpha), _lib.ptr(e_vcr), _lib.stream_ptr()),
            'cone forward update')


class ConeBeamModel(TomographyModel):
    """Cone-beam geometry: flat or curved detector, circular or helical scan
    (cone_beam.py:302-1131).

    Args:
        sinogram_shape (tuple): (views, rows, channels).
        angles (ndarray): view angles in radians.
        source_detector_dist (float): source-to-detector distance (ALU).
        source_iso_dist (float): source-to-iso distance (ALU).
        helical_z_shifts (ndarray, optional): per-view z shift (helical scan).
        use_curved_detector (bool): cylindrical detector of radius SDD.
    """

    _dc_damping = DC_DAMPING_DEFAULT

    def __init__(self, sinogram_shape, angles, source_detector_dist, source_iso_dist,
                 helical_z_shifts=None, use_curved_detector=False, view_batch_size=None,
                 compile_mode='auto'):
        angles = np.asarray(angles, dtype=np.float32).flatten()
        if helical_z_shifts is None:
            zs = np.zeros_like(angles)
        else:
            zs = np.asarray(helical_z_shifts, dtype=np.float32).flatten()
        if zs.shape != angles.shape:
            raise ValueError('Incompatible view dependent vector lengths: all view-dependent '
                             'vectors must have the same length.')
        vpa = np.stack([angles, zs], axis=1)
        super().__init__(sinogram_shape, view_batch_size=view_batch_size,
                         compile_mode=compile_mode, geometry_type='cone',
                         view_params_name='view_params_array', view_params_array=vpa,
                         source_detector_dist=source_detector_dist,
                         source_iso_dist=source_iso_dist, recon_slice_offset=0.0,
                         axial_pad_fraction=0.0, use_curved_detector=use_curved_detector)

    def get_magnification(self):
        sdd, sid = self.get_params(['source_detector_dist', 'source_iso_dist'])
        if np.isinf(sdd):
            return 1
        return sdd / sid

    def verify_valid_params(self):
        super().verify_valid_params()
        sino_shape, vpa = self.get_params(['sinogram_shape', 'view_params_array'])
        vra, vsa = self.get_params(['voxel_row_aspect', 'voxel_slice_aspect'])
        V, R = sino_shape[:2]
        if vpa is None:
            raise ValueError('view_params_array was not set.')
        if tuple(np.shape(vpa)) != (V, 2):
            raise ValueError('Number of view dependent parameter vectors must equal the '
                             'number of views.')
        if vra <= 0 or vsa <= 0:
            raise ValueError('Voxel aspect ratios must be positive.')
        sdd, ddr, dro = self.get_params(['source_detector_dist', 'delta_det_row',
                                         'det_row_offset'])
        if ddr * R / 2 + abs(dro) > sdd:
            warnings.warn('Cone angle is more than 45 degrees.  This will likely produce '
                          'recon artifacts.')

    def get_psf_radii(self):
        """(psf_radius, bp_psf_radius), cone_beam.py:519-553."""
        (ddr, ddc, sdd, rs, dv, vra, vsa) = self.get_params(
            ['delta_det_row', 'delta_det_channel', 'source_detector_dist', 'recon_shape',
             'delta_voxel', 'voxel_row_aspect', 'voxel_slice_aspect'])
        mag = self.get_magnification()
        dvr, dvs = vra * dv, vsa * dv
        ddet = min(ddr, ddc)
        if np.isinf(sdd):
            mmax = mmin = 1
        else:
            sid = sdd / mag
            half = max(0.5 * rs[1] * dv, 0.5 * rs[0] * dvr)
            mmax = sdd / (sid - half)
            mmin = sdd / (sid + half)
        psf = int(np.ceil(np.ceil(max(dv, dvr, dvs) * mmax / ddet) / 2))
        bp = int(np.ceil(np.ceil(ddet / (mmin * min(dv, dvr, dvs))) / 2))
        return psf, bp

    def get_psf_radius(self):
        return self.get_psf_radii()[0]

    @staticmethod
    def detector_mn_to_uv(m, n, delta_det_channel, delta_det_row, det_channel_offset,
                          det_row_offset, num_det_rows, num_det_channels):
        v = (m - (num_det_rows - 1) / 2.0) * delta_det_row - det_row_offset
        u = (n - (num_det_channels - 1) / 2.0) * delta_det_channel - det_channel_offset
        return u, v

    def auto_set_recon_geometry(self, no_compile=False, no_warning=False):
        """cone_beam.py:573-658."""
        ddr, ddc = self.get_params(['delta_det_row', 'delta_det_channel'])
        vra, vsa = self.get_params(['voxel_row_aspect', 'voxel_slice_aspect'])
        mag = self.get_magnification()
        dv = ddc / mag
        dvr, dvs = vra * dv, vsa * dv
        _, R, C = self.get_params('sinogram_shape')
        nr = int(np.round(C * ((ddc / dvr) / mag)))
        nc = int(np.round(C * ((ddc / dv) / mag)))
        zs = np.asarray(self.get_params('view_params_array'))[:, 1]
        zmin, zmax = float(np.min(zs)), float(np.max(zs))
        H_iso = R * (ddr / mag)
        ns = max(1, int(np.ceil((H_iso + (zmax - zmin)) / dvs)))
        rso = 0.5 * (zmin + zmax)
        sdd, dro, dco, use_ror = self.get_params(['source_detector_dist', 'det_row_offset',
                                                  'det_channel_offset', 'use_ror_mask'])
        support = _support_radius((nr, nc), dvr, dv, use_ror_mask=use_ror)
        _, v_lo = self.detector_mn_to_uv(-0.5, 0.0, ddc, ddr, dco, dro, R, C)
        _, v_hi = self.detector_mn_to_uv(R - 0.5, 0.0, ddc, ddr, dco, dro, R, C)
        v_bot = max(float(v_lo), float(v_hi))
        v_top = min(float(v_lo), float(v_hi))
        zpv = 1.0 / float(mag)
        if not np.isinf(sdd):
            zpv += support / float(sdd)
        ex_bot = max(0.0, v_bot * zpv - float(H_iso) / 2)
        ex_top = max(0.0, -v_top * zpv - float(H_iso) / 2)
        apf = self.get_params('axial_pad_fraction')
        if isinstance(apf, (tuple, list)):
            if len(apf) != 2:
                raise ValueError('axial_pad_fraction must be a float or a (top_fraction, '
                                 f'bottom_fraction) pair; got {apf!r}.')
            f_top, f_bot = (float(f) for f in apf)
        else:
            f_top = f_bot = float(apf)
        if f_top < 0 or f_bot < 0:
            raise ValueError(f'axial_pad_fraction must be >= 0; got {apf!r}.')
        n_top = int(np.ceil(f_top * ex_top / dvs))
        n_bot = int(np.ceil(f_bot * ex_bot / dvs))
        ns += n_top + n_bot
        rso += 0.5 * (n_bot - n_top) * float(dvs)
        self.set_params(no_compile=no_compile, no_warning=no_warning, recon_shape=(nr, nc, ns),
                        delta_voxel=dv, recon_slice_offset=rso)

    def _projector_key(self):
        names = ['sinogram_shape', 'recon_shape', 'delta_det_channel', 'delta_det_row',
                 'det_channel_offset', 'det_row_offset', 'delta_voxel', 'voxel_row_aspect',
                 'voxel_slice_aspect', 'source_detector_dist', 'source_iso_dist',
                 'recon_slice_offset', 'use_curved_detector']
        vals = tuple(tuple(v) if isinstance(v, (list, tuple)) else v
                     for v in self.get_params(names))
        vpa = np.asarray(self.get_params('view_params_array'), dtype=np.float32)
        return vals + (vpa.tobytes(), str(self.torch_device))

    def _make_projector(self):
        return ConeProjector(self, self.torch_device)

    # -- DC damping -----------------------------------------------------------
    def _dc_damping_slice_profile(self):
        """Per-slice s_k as a float32 device tensor, or None when disabled
        (cone_beam.py:424-475)."""
        cfg = self._dc_damping
        if cfg is None:
            return None
        rs = self.get_params('recon_shape')
        dv, vsa, oz, R = self.get_params(['delta_voxel', 'voxel_slice_aspect',
                                          'recon_slice_offset', 'source_iso_dist'])
        zs = np.asarray(self.get_params('view_params_array'))[:, 1]
        a, b, p, c = cfg
        nz = rs[2]
        L = rs[0] * dv
        dz = vsa * dv
        z = (np.arange(nz) - (nz - 1) / 2.0) * dz + oz

        def profile(t):
            return (c * t ** p + a * b ** p) / (t ** p + b ** p)

        if zs.max() - zs.min() == 0:
            s = profile(L * np.abs(z - zs[0]) / (R * dz))
        else:
            t = L * np.abs(z[:, None] - zs[None, :]) / (R * dz)
            s = profile(t).mean(axis=1)
        return torch.as_tensor(s.astype(np.float32), device=self.torch_device)

    def _direction_profile(self):
        return self._dc_damping_slice_profile()

    # -- FDK ------------------------------------------------------------------
    def _fdk_weight_and_scale(self):
        _, R, C = (int(x) for x in self.get_params('sinogram_shape'))
        (sdd, dv, ddr, ddc, vra, vsa, dro, dco) = self.get_params(
            ['source_detector_dist', 'delta_voxel', 'delta_det_row', 'delta_det_channel',
             'voxel_row_aspect', 'voxel_slice_aspect', 'det_row_offset', 'det_channel_offset'])
        vol = dv * (vra * dv) * (vsa * dv)
        M0 = self.get_magnification()
        mg, ng = np.meshgrid(np.arange(R), np.arange(C), indexing='ij')
        u, v = self.detector_mn_to_uv(mg, ng, ddc, ddr, dco, dro, R, C)
        wmap = sdd / np.sqrt(sdd ** 2 + u ** 2 + v ** 2)
        return wmap.astype(np.float32), ddr / (vol * M0)

    def fdk_filter(self, sinogram, filter_name='ramp', output_sharded=False):
        """FDK filtering: cosine pre-weight, ramp filter, scale (cone_beam.py:661-695)."""
        wmap, scale = self._fdk_weight_and_scale()
        return self._apply_direct_recon_filter(sinogram, filter_name, filter_scale=scale,
                                               output_sharded=output_sharded, row_weight=wmap)

    def _helical_z_weight_row(self):
        V, R, _ = self.get_params('sinogram_shape')
        zs = np.asarray(self.get_params('view_params_array'))[:, 1]
        dv, vsa, rs, rso, ddr = self.get_params(['delta_voxel', 'voxel_slice_aspect',
                                                 'recon_shape', 'recon_slice_offset',
                                                 'delta_det_row'])
        M0 = self.get_magnification()
        dvs = vsa * dv
        ns = rs[2]
        zk = dvs * (np.arange(ns) - (ns - 1) / 2.0) + rso
        half = 0.5 * R * ddr / M0
        cov = np.sum(np.abs(zk[:, None] - zs[None, :]) <= half, axis=1)
        return np.where(cov > 0, V / np.maximum(cov, 1), 0.0)

    def helical_fdk_z_weight(self, recon, sinogram=None):
        w = torch.as_tensor(self._helical_z_weight_row().astype(np.float32), device=recon.device)
        return recon * w[None, None, :]

    def _direct_recon_flat(self, sinogram, filter_name='ramp'):
        wmap, scale = self._fdk_weight_and_scale()
        filt = self._apply_direct_recon_filter(sinogram, filter_name, filter_scale=scale,
                                               row_weight=wmap, layout=_lib.MBIR_LAYOUT_VCR)
        idx = self.full_indices_device()
        flat = self._scatter_flat(self.projector().back(filt, idx), idx)
        zs = np.asarray(self.get_params('view_params_array'))[:, 1]
        if float(np.max(zs) - np.min(zs)) > 0:
            w = torch.as_tensor(self._helical_z_weight_row().astype(np.float32),
                                device=flat.device)
            flat = flat * w[None, :]
        return flat

    def recon_fdk(self, sinogram, filter_name='ramp', output_sharded=False):
        rs = self.get_params('recon_shape')
        recon = self._direct_recon_flat(sinogram, filter_name).reshape(rs[0], rs[1], -1)
        return recon if output_sharded else self._to_host(recon)

    def recon_direct(self, sinogram, filter_name='ramp', output_sharded=False):
        return self.recon_fdk(sinogram, filter_name=filter_name, output_sharded=output_sharded)

    def direct_filter(self, sinogram, filter_name='ramp', output_sharded=False):
        return self.fdk_filter(sinogram, filter_name=filter_name, output_sharded=output_sharded)


def recon_simple_cone(sinogram, angles, source_detector_dist, source_iso_dist, weights=None,
                      sharpness=1.0, max_iterations=15):
    """One-call cone-beam reconstruction (cone_beam.py:1134-1190)."""
    if torch.is_tensor(angles):
        angles = angles.detach().cpu().numpy()
    model = ConeBeamModel(tuple(sinogram.shape), angles, source_detector_dist=source_detector_dist,
                          source_iso_dist=source_iso_dist)
    model.set_params(sharpness=sharpness)
    return model.recon(sinogram, weights=weights, max_iterations=max_iterations)
