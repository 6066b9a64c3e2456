"""TomographyModel: the reference's model API over the MI355X HIP library.

Public surface (tomography_model.py:1877-2054, 3015-3204 of the reference):
``forward_project``, ``back_project``, ``sparse_forward_project``,
``sparse_back_project``, ``compute_hessian_diagonal``, ``recon``,
``recon_direct``, ``prox_map``, ``auto_set_regularization_params``,
``get_forward_model_loss``, ``get_recon_dict``, ``set_params``/``get_params``.
Arrays cross the boundary as numpy (default) or as device tensors
(``output_sharded=True``); inside, everything stays on the GPU:

* sinogram-like state is kept channel-major (views, channels, rows);
* the volume is flat (rows*cols, slices);
* one VCD subset step is six launches (back projection, prior + direction,
  forward + line search, step size, forward + error update, recon update)
  with the step size kept on the device, so the subset loop never waits on
  the host; the host synchronises once per iteration for the statistics,
  as the reference does (tomography_model.py:2936-2946).

The VCD host loop consumes numpy's global RNG exactly like the reference
(partitions up front, one permutation per iteration), so a seeded run follows
the reference's subset order.
"""

import datetime
import ctypes
import math
import os
import warnings

import numpy as np
import torch

from . import _lib
from . import vcd_utils
from ._params import (AUTO_REGULARIZATION_PARAM_NAMES, RECON_PARAM_NAMES, Param,
                      ParameterHandler)

F32_EPS = float(np.finfo(np.float32).eps)


def _dev():
    _lib.require_gpu()
    return torch.device('cuda', torch.cuda.current_device())


def _b_from_nbr_wts(nbr_wts):
    b = np.array([nbr_wts, nbr_wts]).T.flatten()
    return tuple(b / np.sum(b))


class _Reducer:
    """FP64 reductions on the device into a small double buffer."""

    def __init__(self, device):
        self.device = device
        self._scratch = None

    def scratch(self, n):
        nb = 3 * int(_lib.lib().mbir_reduce_blocks(int(n)))
        if self._scratch is None or self._scratch.numel() < nb:
            self._scratch = torch.empty(max(nb, 3 * 2048), dtype=torch.float64,
                                        device=self.device)
        return self._scratch

    def sumsq(self, a, w, out):
        n = a.numel()
        _lib.check(_lib.lib().mbir_sumsq_f64(_lib.ptr(a), _lib.ptr(w), n,
                                             _lib.ptr(self.scratch(n)), _lib.ptr(out),
                                             _lib.stream_ptr()), 'sumsq')

    def sumabs(self, a, out):
        n = a.numel()
        _lib.check(_lib.lib().mbir_sumabs_f64(_lib.ptr(a), n, _lib.ptr(self.scratch(n)),
                                              _lib.ptr(out), _lib.stream_ptr()), 'sumabs')

    def dot(self, a, b, out):
        n = a.numel()
        _lib.check(_lib.lib().mbir_dot_f64(_lib.ptr(a), _lib.ptr(b), n,
                                           _lib.ptr(self.scratch(n)), _lib.ptr(out),
                                           _lib.stream_ptr()), 'dot')


def dot_f64(a, b):
    """<a, b> accumulated in FP64 on the device (adjoint tests, line search)."""
    a = a.contiguous()
    b = b.contiguous()
    out = torch.zeros(1, dtype=torch.float64, device=a.device)
    _Reducer(a.device).dot(a.reshape(-1), b.reshape(-1), out)
    return float(out.item())


class TomographyModel(ParameterHandler):
    """Base class for the geometries (see ParallelBeamModel, ConeBeamModel)."""

    def __init__(self, sinogram_shape, view_batch_size=None, compile_mode='auto', **kwargs):
        super().__init__()
        self.view_batch_size = view_batch_size
        self.compile_mode = compile_mode
        self._ops = None
        self._ops_key = None
        self._idx_cache = None
        self.prox_data = None
        for key, val in kwargs.items():
            self.params[key] = Param(val, True)
        self.set_params(no_compile=True, no_warning=True,
                        sinogram_shape=tuple(int(s) for s in sinogram_shape))
        self.auto_set_recon_geometry(no_compile=True, no_warning=True)
        self.verify_valid_params()

    # -- geometry hooks ---------------------------------------------------------
    def get_magnification(self):
        raise NotImplementedError

    def get_psf_radius(self):
        raise NotImplementedError

    def auto_set_recon_geometry(self, no_compile=False, no_warning=False):
        raise NotImplementedError

    def _projector_key(self):
        """Every parameter the device projector depends on (read per call)."""
        raise NotImplementedError

    def _make_projector(self):
        raise NotImplementedError

    def refresh_device_bindings(self):
        self._ops = None
        self._ops_key = None
        self._idx_cache = None

    @property
    def torch_device(self):
        return _dev()

    def projector(self):
        """The device projector for the CURRENT parameters (rebuilt on change)."""
        key = self._projector_key()
        if self._ops is None or self._ops_key != key:
            self._ops = self._make_projector()
            self._ops_key = key
        return self._ops

    # -- placement --------------------------------------------------------------
    def _sino_to_device(self, sinogram):
        V = self.get_params('sinogram_shape')[0]
        t = torch.as_tensor(sinogram, dtype=torch.float32, device=self.torch_device)
        if t.shape[0] != V:
            raise ValueError('Cannot place the sinogram: its view axis has size '
                             f'{t.shape[0]}, but the model expects {V} views.')
        return t.contiguous()

    def _recon_to_device(self, recon):
        S = self.get_params('recon_shape')[2]
        t = torch.as_tensor(recon, dtype=torch.float32, device=self.torch_device)
        if t.shape[-1] != S:
            raise ValueError('Cannot place the reconstruction: its slice axis has size '
                             f'{t.shape[-1]}, but the model expects {S} slices.')
        return t.contiguous()

    @staticmethod
    def _to_host(t):
        return t.detach().cpu().numpy()

    def full_indices_device(self):
        recon_shape, use_ror_mask = self.get_params(['recon_shape', 'use_ror_mask'])
        if not isinstance(use_ror_mask, bool):
            return torch.as_tensor(vcd_utils.gen_full_indices(recon_shape, use_ror_mask),
                                   dtype=torch.int64, device=self.torch_device)
        key = (tuple(recon_shape), use_ror_mask, str(self.torch_device))
        if self._idx_cache is None or self._idx_cache[0] != key:
            idx = torch.as_tensor(vcd_utils.gen_full_indices(recon_shape, use_ror_mask),
                                  dtype=torch.int64, device=self.torch_device)
            self._idx_cache = (key, idx)
        return self._idx_cache[1]

    def _vrc_to_vcr(self, sino):
        V, R, C = sino.shape
        out = torch.empty((V, C, R), dtype=torch.float32, device=sino.device)
        _lib.check(_lib.lib().mbir_transpose_sino(_lib.ptr(sino), V, R, C, 1, _lib.ptr(out),
                                                  _lib.stream_ptr()), 'transpose')
        return out

    def _vcr_to_vrc(self, sino):
        V, C, R = sino.shape
        out = torch.empty((V, R, C), dtype=torch.float32, device=sino.device)
        _lib.check(_lib.lib().mbir_transpose_sino(_lib.ptr(sino), V, R, C, 0, _lib.ptr(out),
                                                  _lib.stream_ptr()), 'transpose')
        return out

    # -- projections ------------------------------------------------------------
    def sparse_forward_project(self, voxel_values, pixel_indices):
        """Cylinders (P, slices) at ``pixel_indices`` -> sinogram (V, R, C) tensor."""
        values = self._recon_to_device(voxel_values)
        idx = torch.as_tensor(pixel_indices, dtype=torch.int64, device=self.torch_device)
        return self.projector().forward(values, idx.contiguous(), layout=_lib.MBIR_LAYOUT_VRC)

    def sparse_back_project(self, sinogram, pixel_indices, coeff_power=1):
        """Sinogram (V, R, C) -> cylinders (P, slices) tensor at ``pixel_indices``."""
        sino = self._sino_to_device(sinogram)
        idx = torch.as_tensor(pixel_indices, dtype=torch.int64, device=self.torch_device)
        return self.projector().back(self._vrc_to_vcr(sino), idx.contiguous(),
                                     coeff_power=coeff_power)

    def full_index_runs_device(self):
        """Image-row run table of the full index list (cached with it)."""
        idx = self.full_indices_device()
        cache = getattr(self, '_runs_cache', None)
        if cache is None or cache[0] is not idx:
            from .parallel_beam import row_runs
            rs = self.get_params('recon_shape')
            self._runs_cache = (idx, row_runs(idx, rs[0], rs[1]))
        return self._runs_cache[1]

    def forward_project(self, recon, output_sharded=False):
        recon = self._recon_to_device(recon)
        idx = self.full_indices_device()
        values = recon.reshape(-1, recon.shape[-1])[idx].contiguous()
        sino = self.projector().forward(values, idx, layout=_lib.MBIR_LAYOUT_VRC,
                                        runs=self._full_runs_or_false())
        return sino if output_sharded else self._to_host(sino)

    def _full_runs_or_false(self):
        runs = self.full_index_runs_device()
        return False if runs is None else runs

    def _scatter_flat(self, cyl, idx):
        recon_shape = self.get_params('recon_shape')
        flat = torch.zeros((recon_shape[0] * recon_shape[1], cyl.shape[-1]),
                           dtype=torch.float32, device=cyl.device)
        flat[idx] = cyl
        return flat

    def back_project(self, sinogram, output_sharded=False):
        recon_shape = self.get_params('recon_shape')
        sino = self._sino_to_device(sinogram)
        idx = self.full_indices_device()
        cyl = self.projector().back(self._vrc_to_vcr(sino), idx)
        recon = self._scatter_flat(cyl, idx).reshape(tuple(recon_shape[:2]) + (cyl.shape[-1],))
        return recon if output_sharded else self._to_host(recon)

    def compute_hessian_diagonal(self, weights=None, output_sharded=False, indices=None):
        """Back projection of the weights (ones by default) with squared
        coefficients (tomography_model.py:1951-2032)."""
        sino_shape, recon_shape = self.get_params(['sinogram_shape', 'recon_shape'])
        if weights is None:
            w_vcr = torch.ones((sino_shape[0], sino_shape[2], sino_shape[1]),
                               dtype=torch.float32, device=self.torch_device)
        else:
            if tuple(np.shape(weights)) != tuple(sino_shape):
                raise ValueError('Weights must be constant or an array compatible with '
                                 f'sinogram\nGot weights.shape = {tuple(np.shape(weights))}, '
                                 f'but sinogram.shape = {tuple(sino_shape)}')
            w_vcr = self._vrc_to_vcr(self._sino_to_device(weights))
        n_grid = recon_shape[0] * recon_shape[1]
        dense = indices is None
        idx = (torch.arange(n_grid, dtype=torch.int64, device=self.torch_device) if dense
               else torch.as_tensor(indices, dtype=torch.int64,
                                    device=self.torch_device).contiguous())
        cyl = self.projector().back(w_vcr, idx, coeff_power=2)
        if dense:
            hess = cyl.reshape(recon_shape[0], recon_shape[1], -1)
        else:
            hess = self._scatter_flat(cyl, idx).reshape(recon_shape[0], recon_shape[1], -1)
        return hess if output_sharded else self._to_host(hess)

    def get_voxels_at_indices(self, recon, indices):
        return recon.reshape((-1, recon.shape[-1]))[indices]

    # -- direct recon -------------------------------------------------------------
    def _apply_direct_recon_filter(self, sinogram, filter_name, filter_scale,
                                   output_sharded=False, row_weight=None, layout=None):
        """Ramp row filter with scale*pi/V folded into the taps
        (tomography_model.py:2279-2321, tomography_utils.py:18-78)."""
        sino = self._sino_to_device(sinogram)
        V, R, C = sino.shape
        taps = generate_direct_recon_filter(C, filter_name)
        taps = taps * np.float32(filter_scale * (np.pi / V))
        taps_t = torch.as_tensor(taps, dtype=torch.float32, device=sino.device)
        rw = None if row_weight is None else torch.as_tensor(
            row_weight, dtype=torch.float32, device=sino.device).contiguous()
        out_layout = _lib.MBIR_LAYOUT_VRC if layout is None else layout
        shape = (V, R, C) if out_layout == _lib.MBIR_LAYOUT_VRC else (V, C, R)
        out = torch.empty(shape, dtype=torch.float32, device=sino.device)
        _lib.check(_lib.lib().mbir_row_filter(_lib.ptr(sino), _lib.ptr(taps_t), _lib.ptr(rw),
                                              V, R, C, out_layout, _lib.ptr(out),
                                              _lib.stream_ptr()), 'row filter')
        if layout is not None:
            return out
        return out if output_sharded else self._to_host(out)

    def _direct_recon_flat(self, sinogram):
        """Filtered back projection onto the flat (rows*cols, slices) volume."""
        raise NotImplementedError

    def recon_direct(self, sinogram, filter_name='ramp', output_sharded=False):
        """Direct (filtered back projection) reconstruction; geometry specific."""
        raise NotImplementedError

    def _direction_profile(self):
        """Per-slice damping profile for the VCD direction (None: the plain
        preconditioned gradient, tomography_model.py:2420-2432)."""
        return None

    def direct_recon(self, sinogram, filter_name='ramp', output_sharded=False):
        """Alias of :meth:`recon_direct`."""
        return self.recon_direct(sinogram, filter_name=filter_name,
                                 output_sharded=output_sharded)

    # -- auto-regularization ------------------------------------------------------
    def auto_set_regularization_params(self, sinogram, weights=None):
        """sigma_y, sigma_x, sigma_prox from a <=20-view subsample (host numpy,
        tomography_model.py:2057-2227)."""
        if self.get_params('auto_regularize_flag'):
            small = self.subsample_views(sinogram)
            small_w = 1 if weights is None else self.subsample_views(weights)
            ind = self._get_sino_indicator(small, verbose=self.get_params('verbose'))
            self._check_lateral_truncation(ind)
            self.auto_set_sigma_y(small, ind, small_w)
            std = self._get_estimate_of_recon_std(small, ind)
            self.auto_set_sigma_x(std)
            self.auto_set_sigma_prox(std)
        vals = [float(v) for v in self.get_params(list(AUTO_REGULARIZATION_PARAM_NAMES))]
        return dict(zip(AUTO_REGULARIZATION_PARAM_NAMES, vals))

    def _check_lateral_truncation(self, ind):
        if np.all(ind):
            return
        frac = float(np.mean(np.logical_or(ind[:, :, 0], ind[:, :, -1])))
        if frac > 0.02 and self.get_params('verbose') > 0:
            warnings.warn(f"Lateral FoV truncation detected: the object support reaches the "
                          f"detector's edge channels in {frac:.0%} of the sampled view-rows.  "
                          f"Consider using scale_recon_shape(s, s) where s >= 1.1 to improve "
                          f"image quality.")

    def auto_set_sigma_y(self, sinogram, sino_indicator, weights=1):
        snr_db = self.get_params('snr_db')
        mag = self.get_magnification()
        dv, ddc = self.get_params(['delta_voxel', 'delta_det_channel'])
        rms = float(np.average(weights * np.asarray(sinogram) ** 2, None, sino_indicator) ** 0.5)
        rel = 10 ** (-snr_db / 20)
        pitch = dv / (ddc / mag)
        sigma_y = np.float32(rel * rms * (pitch ** 0.5))
        self.set_params(no_warning=True, sigma_y=float(sigma_y), auto_regularize_flag=True)

    def auto_set_sigma_x(self, recon_std):
        s = np.float32(0.2 * (2 ** self.get_params('sharpness')) * recon_std)
        self.set_params(no_warning=True, sigma_x=float(s), auto_regularize_flag=True)

    def auto_set_sigma_prox(self, recon_std):
        s = np.float32(0.2 * (2 ** self.get_params('sharpness')) * recon_std)
        self.set_params(no_warning=True, sigma_prox=float(s), auto_regularize_flag=True)

    @staticmethod
    def subsample_views(array, max_views_to_use=20):
        """Every (n // m)-th view as a host array; a tensor is sliced where it
        lives and only the sample is copied."""
        n = array.shape[0]
        m = min(max_views_to_use, n)
        step = max(n // m, 1)
        if torch.is_tensor(array):
            return array.detach()[::step].cpu().numpy()
        return np.array(array[::step])

    @staticmethod
    def _get_sino_indicator(sinogram, verbose=1):
        sinogram = np.asarray(sinogram)
        if np.iscomplexobj(sinogram):
            raise TypeError('sinogram must be real-valued; got complex dtype.')
        if not np.isfinite(sinogram).all():
            raise ValueError('sinogram contains NaN and/or Inf values.')
        left, right = vcd_utils.estimate_background_cluster_boundaries(sinogram)
        threshold = right + (right - left)
        mx = np.max(sinogram)
        if mx <= 0:
            if verbose > 0:
                warnings.warn('Sinogram contains no positive values. This may lead to a '
                              'contrast reversed reconstruction.')
            return np.ones_like(sinogram, dtype=np.int8)
        if mx < threshold:
            if verbose > 0:
                warnings.warn('\nUnable to determine sinogram background. This may affect '
                              'regularization.\n')
            return np.ones_like(sinogram, dtype=np.int8)
        med = np.median(sinogram[sinogram >= threshold])
        return np.int8(sinogram >= 0.25 * med)

    def _get_estimate_of_recon_std(self, sinogram, sino_indicator):
        ddc, dv, recon_shape = self.get_params(['delta_det_channel', 'delta_voxel',
                                                'recon_shape'])
        mag = self.get_magnification()
        typical = np.average(np.abs(sinogram), weights=sino_indicator)
        tpl_space = (2 * recon_shape[0] * recon_shape[1]) / (recon_shape[0] + recon_shape[1]) * dv
        tpl_sino = sinogram.shape[-1] * ddc / mag
        return typical / np.minimum(tpl_space, tpl_sino)

    # -- losses -----------------------------------------------------------------
    @staticmethod
    def get_forward_model_loss(error_sinogram, sigma_y, weights=None, normalize=True):
        e = torch.as_tensor(error_sinogram, dtype=torch.float32)
        if weights is None:
            w, avg = 1, 1
        elif np.ndim(weights) == 0:
            w, avg = weights, weights
        else:
            w = torch.as_tensor(weights, dtype=torch.float32, device=e.device)
            avg = torch.mean(w)
        if normalize:
            return torch.sqrt(torch.sum(e * e * w) / (avg * float(e.numel()))) / sigma_y
        return (1.0 / (2 * sigma_y ** 2)) * torch.sum((e * e) * w)

    # -- VCD ----------------------------------------------------------------------
    def _vcd_params(self, prox=False):
        b = _b_from_nbr_wts(self.get_params('qggmrf_nbr_wts'))
        sigma_x, p, q, T, sigma_y, sigma_prox, max_alpha = self.get_params(
            ['sigma_x', 'p', 'q', 'T', 'sigma_y', 'sigma_prox', 'max_alpha'])
        rs = self.get_params('recon_shape')
        prm = _lib.VcdParams()
        for i in range(6):
            prm.b[i] = np.float32(b[i])
            prm.two_b[i] = np.float32(b[i] * 2.0)
        f32 = np.float32
        prm.sigma_x, prm.p, prm.q, prm.T = f32(sigma_x), f32(p), f32(q), f32(T)
        prm.a_min = f32(T * sigma_x * F32_EPS)
        prm.t_sigma = f32(T * sigma_x)
        prm.q_minus_p = f32(q - p)
        prm.p_minus_2 = f32(p - 2)
        prm.q_over_p = f32(q / p)
        prm.two_sigma_p = f32(2 * sigma_x ** p)
        fm = 1.0 / (sigma_y ** 2.0)
        prm.fm_constant = f32(fm)
        prm.neg_fm_constant = f32(-fm)
        prm.inv_sigma_prox2 = f32(1.0 / (sigma_prox ** 2.0))
        prm.max_alpha = f32(max_alpha)
        prm.num_rows, prm.num_cols, prm.num_slices = int(rs[0]), int(rs[1]), int(rs[2])
        prm.use_prox = 1 if prox else 0
        return prm

    def _initial_error_state(self, y_vcr, x_flat, w_vcr, scale):
        """e = y - alpha*A x0, x0 *= alpha with the weighted least-squares scale
        (tomography_model.py:1666-1712); inner products in FP64."""
        idx = self.full_indices_device()
        fwd = self.projector().forward(x_flat[idx].contiguous(), idx,
                                       layout=_lib.MBIR_LAYOUT_VCR,
                                       runs=self._full_runs_or_false())
        wf = fwd if w_vcr is None else w_vcr * fwd
        nrm = dot_f64(wf, fwd)
        alpha = dot_f64(wf, y_vcr) / nrm if (nrm > 0 and scale) else 1
        del wf
        e = y_vcr - alpha * fwd
        del fwd
        return e, alpha * x_flat if alpha != 1 else x_flat

    def _vcd_recon(self, sinogram, partitions, partition_sequence, stop_threshold_change_pct,
                   weights=None, init_recon=None, prox_input=None, first_iteration=0,
                   init_error_sinogram=None, fm_hessian=None, return_checkpoint=False):
        """The VCD engine on the device (tomography_model.py:2699-3013)."""
        engine, hess = self._prepare_vcd(sinogram, partitions, weights=weights,
                                         init_recon=init_recon, prox_input=prox_input,
                                         init_error_sinogram=init_error_sinogram,
                                         fm_hessian=fm_hessian)
        recon_shape, sino_shape = self.get_params(['recon_shape', 'sinogram_shape'])
        S = recon_shape[2]
        verbose, sigma_y = self.get_params(['verbose', 'sigma_y'])
        total_sino = math.prod(sino_shape)
        max_iters = partition_sequence.size
        fm_rmse = np.zeros(max_iters)
        nmae = np.zeros(max_iters)
        alpha_values = np.zeros(max_iters)
        dnorm = np.zeros((max_iters, S))
        num_iters = 0
        self.logger.info('Starting VCD iterations')
        for i in range(max_iters):
            pidx = int(partition_sequence[i])
            num_subsets = partitions[pidx].shape[0]
            order = np.random.permutation(num_subsets)
            st = engine.iteration(pidx, order, total_sino)
            fm_rmse[i] = st['fm_loss'] / sigma_y
            nmae[i] = st['ell1'] / st['recon_l1'] if st['recon_l1'] else float('nan')
            alpha_values[i] = st['alpha_sum'] / num_subsets
            dnorm[i] = np.sqrt(st['sumsq'])
            if verbose >= 1:
                self.logger.info('\nAfter iteration {} of a max of {}: Pct change={:.4f}, '
                                 'Forward loss={:.4f}'.format(i + first_iteration,
                                                              max_iters + first_iteration,
                                                              100 * nmae[i], fm_rmse[i]))
                self.logger.info(f'Relative step size (alpha)={alpha_values[i]:.2f}, '
                                 f"Error sino RMSE={st['es_rmse']:.4f}")
                self.logger.info('Number subsets = {}'.format(num_subsets))
            num_iters += 1
            if nmae[i] < stop_threshold_change_pct / 100:
                self.logger.warning('Change threshold stopping condition reached')
                break
        recon3d = engine.x.reshape(recon_shape[0], recon_shape[1], S)
        losses = (fm_rmse[:num_iters], np.zeros(num_iters), nmae[:num_iters],
                  alpha_values[:num_iters], dnorm[:num_iters])
        if return_checkpoint:
            ckpt = {'error_sinogram': self._vcr_to_vrc(_to_f32(engine.e)), 'fm_hessian':
                    hess.reshape(recon_shape[0], recon_shape[1], S)}
            return recon3d, losses, ckpt
        return recon3d, losses

    def _prepare_vcd(self, sinogram, partitions, weights=None, init_recon=None, prox_input=None,
                     init_error_sinogram=None, fm_hessian=None):
        """Device state of a VCD run -- initial recon, error sinogram, Hessian
        diagonal, parameters -- as a _VcdEngine (tomography_model.py:2699-2850).
        Returns (engine, hessian)."""
        self.verify_valid_params()
        recon_shape, sino_shape = self.get_params(['recon_shape', 'sinogram_shape'])
        if tuple(np.shape(sinogram)) != tuple(sino_shape):
            raise ValueError('sinogram does not have the shape in sinogram_shape. \n'
                             f'Expected {tuple(sino_shape)}, got {tuple(np.shape(sinogram))}.')
        dev = self.torch_device
        S = recon_shape[2]
        n_grid = recon_shape[0] * recon_shape[1]
        ops = self.projector()
        constant_weights = weights is None
        w_vcr = None if constant_weights else self._vrc_to_vcr(self._sino_to_device(weights))
        if init_error_sinogram is not None and init_recon is None:
            raise ValueError('init_error_sinogram requires init_recon (the pair must be a '
                             'consistent resume state; see the docstring).')
        scale = init_recon is None
        if init_recon is None:
            self.logger.info('Starting direct recon for initial reconstruction')
            x_flat = self._direct_recon_flat(sinogram)
        elif isinstance(init_recon, int):
            x_flat = torch.full((n_grid, S), float(init_recon), dtype=torch.float32, device=dev)
        else:
            if tuple(np.shape(init_recon)) != tuple(recon_shape):
                raise ValueError('init_recon does not have the correct shape. Expected '
                                 f'{tuple(recon_shape)}, got {tuple(np.shape(init_recon))}.')
            x_flat = self._recon_to_device(init_recon).reshape(n_grid, S).clone()
        if init_error_sinogram is not None:
            self.logger.info('Resuming from init_error_sinogram')
            e_vcr = self._vrc_to_vcr(self._sino_to_device(init_error_sinogram))
        else:
            self.logger.info('Initializing error sinogram')
            y_vcr = self._vrc_to_vcr(self._sino_to_device(sinogram))
            e_vcr, x_flat = self._initial_error_state(y_vcr, x_flat, w_vcr, scale)
            del y_vcr
        x_flat = x_flat.contiguous()
        prox_flat = None
        if prox_input is not None:
            if tuple(np.shape(prox_input)) != tuple(recon_shape):
                raise ValueError('prox_input does not have the correct size. \n'
                                 f'Expected {tuple(recon_shape)}, got shape '
                                 f'{tuple(np.shape(prox_input))} for prox_input shape.')
            prox_flat = self._recon_to_device(prox_input).reshape(n_grid, S).contiguous()
        if fm_hessian is None:
            self.logger.info('Computing Hessian diagonal')
            idx = (None if self.get_params('use_ror_mask') is False
                   else self.full_indices_device())
            hw = w_vcr if w_vcr is not None else torch.ones_like(e_vcr)
            hidx = idx if idx is not None else torch.arange(n_grid, device=dev)
            cyl = ops.back(hw, hidx, coeff_power=2)
            hess = self._scatter_flat(cyl, hidx)
            del hw, cyl
        else:
            self.logger.info('Using precomputed Hessian diagonal')
            hess = self._recon_to_device(fm_hessian).reshape(n_grid, S).contiguous()

        prm = self._vcd_params(prox=prox_input is not None)
        storage = self.get_params('sinogram_storage')
        if storage not in ('float32', 'bfloat16'):
            raise ValueError(f"sinogram_storage must be 'float32' or 'bfloat16'; got {storage!r}")
        if storage == 'bfloat16':
            e_vcr = _to_bf16(e_vcr)
            if w_vcr is not None:
                w_vcr = _to_bf16(w_vcr)
        engine = _VcdEngine(self, ops, prm, x_flat, e_vcr, w_vcr, hess, prox_flat, partitions)
        return engine, hess

    def initialize_recon(self, sinogram, weights=None, init_recon=None, max_iterations=15,
                         first_iteration=0, logfile_path='~/.mbirtorch/logs/recon.log',
                         print_logs=True):
        """Partitions, partition sequence, input checks and auto-regularization
        (tomography_model.py:3015-3095)."""
        if first_iteration == 0 or self.log_buffer is None:
            self.setup_logger(logfile_path=logfile_path, print_logs=print_logs)
        from . import __version__
        self.logger.info('MBIRTorch-AMD Version = {}'.format(__version__))
        recon_shape, granularity, use_ror_mask = self.get_params(
            ['recon_shape', 'granularity', 'use_ror_mask'])
        partitions = vcd_utils.gen_set_of_pixel_partitions(recon_shape, granularity,
                                                           device=self.torch_device,
                                                           use_ror_mask=use_ror_mask)
        seq = vcd_utils.gen_partition_sequence(self.get_params('partition_sequence'),
                                               max_iterations=max_iterations)
        seq = seq[first_iteration:]
        # Input checks where the data live: a device tensor is checked on the
        # device and only the auto-regularization's view sample is copied out.
        sino_chk = sinogram.detach() if torch.is_tensor(sinogram) else np.asarray(sinogram)
        if (sino_chk.is_complex() if torch.is_tensor(sino_chk) else np.iscomplexobj(sino_chk)):
            raise TypeError('sinogram must be real-valued; got complex dtype.')
        if not _all_finite(sino_chk):
            raise ValueError('sinogram contains NaN and/or Inf values.')
        if weights is not None:
            w_chk = weights.detach() if torch.is_tensor(weights) else np.asarray(weights)
            if not _all_finite(w_chk):
                raise ValueError('weights contains NaN and/or Inf values.')
            if bool((w_chk < 0).any()):
                raise ValueError('weights contain negative values.')
            if bool((w_chk == 0).all()):
                raise ValueError('all weights are zero.')
        reg = self.auto_set_regularization_params(sino_chk, weights=weights)
        return sinogram, weights, init_recon, partitions, seq, granularity, reg

    def recon(self, sinogram, weights=None, init_recon=None, max_iterations=15,
              stop_threshold_change_pct=0.2, first_iteration=0,
              logfile_path='~/.mbirtorch/logs/recon.log', print_logs=True,
              output_sharded=False):
        """MBIR reconstruction by multi-granular VCD (tomography_model.py:3097-3204).
        Returns (recon, recon_dict)."""
        (sinogram, weights, init_recon, partitions, seq, granularity,
         reg) = self.initialize_recon(sinogram, weights, init_recon, max_iterations,
                                      first_iteration, logfile_path=logfile_path,
                                      print_logs=print_logs)
        with torch.no_grad():
            recon, losses = self._vcd_recon(sinogram, partitions, seq,
                                            stop_threshold_change_pct, weights=weights,
                                            init_recon=init_recon,
                                            first_iteration=first_iteration)
        recon_params = dict(zip(RECON_PARAM_NAMES, [
            len(losses[0]), granularity, [int(v) for v in seq],
            [float(v) for v in losses[0]], [0], reg,
            [100 * float(v) for v in losses[2]], [float(v) for v in losses[3]],
            [[float(v) for v in row] for row in losses[4]]]))
        if logfile_path:
            self.logger.info('Logs written to {}'.format(
                os.path.abspath(os.path.expanduser(logfile_path))))
        for h in list(self.logger.handlers):
            h.flush()
        self.close_log_file()
        notes = 'Reconstruction completed: {}\n\n'.format(datetime.datetime.now())
        recon_dict = self.get_recon_dict(recon_params, notes=notes)
        return (recon if output_sharded else self._to_host(recon)), recon_dict

    def prox_map(self, prox_input, sinogram, sigma_prox=None, weights=None, init_recon=None,
                 do_initialization=True, stop_threshold_change_pct=0.2, max_iterations=3,
                 first_iteration=0, logfile_path='~/.mbirtorch/logs/prox.log', print_logs=True,
                 output_sharded=False):
        """Proximal map for Plug-and-Play (tomography_model.py:3251-3372): VCD
        with the prior (x - prox_input)^2 / (2 sigma_prox^2).  Initialization
        (partitions, auto-regularization) is cached across calls unless
        do_initialization; an explicit sigma_prox overrides the automatic one
        for this call only.  Returns (recon, recon_dict)."""
        if do_initialization or self.prox_data is None:
            (sinogram, weights, init_recon, partitions, seq, granularity,
             reg) = self.initialize_recon(sinogram, weights, init_recon, max_iterations,
                                          first_iteration, logfile_path=logfile_path,
                                          print_logs=print_logs)
            self.prox_data = (partitions, seq, granularity, reg)
        else:
            partitions, seq, granularity, reg = self.prox_data
        saved_sigma_prox = self.get_params('sigma_prox')
        if sigma_prox is not None:
            reg = dict(reg, sigma_prox=sigma_prox)
            self.set_params(no_warning=True, sigma_prox=sigma_prox,
                            auto_regularize_flag=self.get_params('auto_regularize_flag'))
        try:
            with torch.no_grad():
                recon, losses = self._vcd_recon(sinogram, partitions, seq,
                                                stop_threshold_change_pct, weights=weights,
                                                init_recon=init_recon, prox_input=prox_input,
                                                first_iteration=first_iteration)
        finally:
            self.set_params(no_warning=True, sigma_prox=saved_sigma_prox)
        recon_params = dict(zip(RECON_PARAM_NAMES, [
            len(losses[0]), granularity, [int(v) for v in seq],
            [float(v) for v in losses[0]], [0], reg,
            [100 * float(v) for v in losses[2]], [float(v) for v in losses[3]],
            [[float(v) for v in row] for row in losses[4]]]))
        for h in list(self.logger.handlers):
            h.flush()
        self.close_log_file()
        notes = 'Proximal map completed: {}\n\n'.format(datetime.datetime.now())
        recon_dict = self.get_recon_dict(recon_params, notes=notes)
        return (recon if output_sharded else self._to_host(recon)), recon_dict

    def get_recon_dict(self, recon_params=None, notes=None, save_log=True, save_model=True):
        d = {'recon_params': recon_params if recon_params is not None
             else '# Recon params not saved.'}
        d['recon_log'] = (self.log_buffer.getvalue() if (self.log_buffer is not None and save_log)
                          else '# Log info not saved.')
        d['notes'] = notes if notes is not None else '# No notes saved'
        d['model_params'] = ({k: v.val for k, v in self.params.items()} if save_model
                             else '# Model not saved')
        return d


def generate_direct_recon_filter(num_channels, filter_name='ramp'):
    """Space-domain ramp taps of length 2C-1 (tomography_utils.py:18-26)."""
    if filter_name not in ['ramp']:
        raise ValueError('Unsupported filter. Supported filters are: ramp.')
    n = np.arange(-num_channels + 1, num_channels)
    return ((1 / 2) * np.sinc(n) - (1 / 4) * (np.sinc(n / 2)) ** 2).astype(np.float32)


def _to_bf16(t):
    """float32 device tensor -> bfloat16 (round to nearest even) via the library."""
    out = torch.empty(t.shape, dtype=torch.bfloat16, device=t.device)
    _lib.check(_lib.lib().mbir_f32_to_bf16(_lib.ptr(t.contiguous()), _lib.ptr(out), t.numel(),
                                           _lib.stream_ptr()), 'to bf16')
    return out


def _to_f32(t):
    if t.dtype == torch.float32:
        return t
    out = torch.empty(t.shape, dtype=torch.float32, device=t.device)
    _lib.check(_lib.lib().mbir_bf16_to_f32(_lib.ptr(t.contiguous()), _lib.ptr(out), t.numel(),
                                           _lib.stream_ptr()), 'to f32')
    return out


def _dir_partials_capacity(sizes, num_slices):
    """Partial triples the direction step needs over subsets of these sizes:
    the largest count, which need not be the largest subset's (smaller
    subsets spread each element over more lanes and workgroups)."""
    lib = _lib.lib()
    return max(int(lib.mbir_vcd_direction_partials_count(int(P), int(num_slices)))
               for P in set(sizes))


def _all_finite(a):
    """True when every entry is finite (a tensor is reduced on its device)."""
    if torch.is_tensor(a):
        return bool(torch.isfinite(a).all())
    return bool(np.isfinite(a).all())


class _VcdEngine:
    """Device-resident VCD state handed to the native schedule runner
    (mbir_vcd_run): one library call per iteration issues every subset step
    on the current stream; the host synchronises once per iteration to read
    the statistics."""

    def __init__(self, model, ops, prm, x, e, w, hess, prox, partitions):
        self.model, self.ops, self.prm = model, ops, prm
        self.x, self.e, self.w, self.hess, self.prox = x, e, w, hess, prox
        self.partitions = partitions
        S = self.S = prm.num_slices
        dev = self.dev = x.device
        lib = _lib.lib()
        sizes = [int(p.shape[1]) for p in partitions]
        max_p = max(sizes)
        # Subset tables: subset id = offset[partition] + k.
        self.offsets = np.cumsum([0] + [int(p.shape[0]) for p in partitions])[:-1]
        n_sub = int(sum(int(p.shape[0]) for p in partitions))
        self._pix = (ctypes.c_void_p * n_sub)()
        self._sizes = (ctypes.c_int32 * n_sub)()
        self._runs = (ctypes.c_void_p * n_sub)()
        self._keep = []
        row_runs = None
        if getattr(ops, 'uses_row_runs', False):
            from .parallel_beam import row_runs
        rs = model.get_params('recon_shape')
        for pi, part in enumerate(partitions):
            part = part.contiguous()
            self._keep.append(part)
            n_k, P = int(part.shape[0]), int(part.shape[1])
            runs = None
            if row_runs is not None and n_k == 1:
                # A single-subset partition is the whole region of
                # reconstruction in row order: the ray-driven forward applies.
                runs = row_runs(part[0], rs[0], rs[1])
                if runs is not None:
                    self._keep.append(runs)
            for k in range(n_k):
                sid = int(self.offsets[pi]) + k
                self._pix[sid] = part.data_ptr() + k * P * part.element_size()
                self._sizes[sid] = P
                self._runs[sid] = None if runs is None else runs.data_ptr()
        back_cap = max(ops.back_splits(P) * P * S for P in sizes)
        self.back_parts = torch.empty(back_cap, dtype=torch.float32, device=dev)
        self.delta = torch.empty(max_p * S, dtype=torch.float32, device=dev)
        n_dir = _dir_partials_capacity(sizes, S)
        self.dir_parts = torch.empty(3 * n_dir, dtype=torch.float64, device=dev)
        n_fwd = max(ops.fwd_partials_count(ray=False), ops.fwd_partials_count(ray=True))
        self.fwd_parts = torch.empty(2 * n_fwd, dtype=torch.float64, device=dev)
        self.alpha = torch.zeros(1, dtype=torch.float32, device=dev)
        # Per-workgroup statistics rows, carried to each iteration's end
        # (which clears them): zero before the first step.
        n_apply = int(lib.mbir_vcd_apply_scratch(max_p, S))
        self.apply_scratch = torch.zeros(max(n_apply, 1), dtype=torch.float64, device=dev)
        # A(delta) of the current subset, kept by the line search so the error
        # update is one streaming pass (sinogram-sized float32).
        self.adelta = torch.empty(e.numel(), dtype=torch.float32, device=dev)
        n_red = 3 * int(lib.mbir_reduce_blocks(max(e.numel(), x.numel())))
        self.red_scratch = torch.empty(n_red, dtype=torch.float64, device=dev)
        self.profile = model._direction_profile()
        self.dc_scratch = None
        if self.profile is not None:
            n = int(lib.mbir_vcd_direction_dc_scratch(max_p, S))
            self.dc_scratch = torch.empty(n, dtype=torch.float32, device=dev)
        # [sumsq per slice (S), ell1, alpha_sum, wsq, sq, wsum, recon_l1]
        self.stats = torch.zeros(S + 6, dtype=torch.float64, device=dev)

        job = _lib.VcdJob()
        job.geometry = ops.geometry_id
        job.positivity = int(bool(model.get_params('positivity_flag')))
        job.num_subsets = n_sub
        job.sino_dtype = (_lib.MBIR_DTYPE_BF16 if e.dtype == torch.bfloat16
                          else _lib.MBIR_DTYPE_F32)
        if ops.geometry_id == _lib.MBIR_GEOMETRY_PARALLEL:
            job.par_geom = ctypes.pointer(ops.geom)
        else:
            job.cone_geom = ctypes.pointer(ops.geom)
        job.view_table = ops.table.data_ptr()
        job.prm = prm
        job.flat_recon = x.data_ptr()
        job.error_vcr = e.data_ptr()
        job.weights_vcr = None if w is None else w.data_ptr()
        job.hessian = hess.data_ptr()
        job.prox = None if prox is None else prox.data_ptr()
        job.dc_profile = None if self.profile is None else self.profile.data_ptr()
        job.subset_pixels = self._pix
        job.subset_sizes = self._sizes
        job.subset_runs = self._runs
        job.back_parts = self.back_parts.data_ptr()
        job.back_parts_capacity = back_cap
        job.delta = self.delta.data_ptr()
        job.dir_parts = self.dir_parts.data_ptr()
        job.fwd_parts = self.fwd_parts.data_ptr()
        job.alpha = self.alpha.data_ptr()
        job.apply_scratch = self.apply_scratch.data_ptr()
        job.dc_scratch = None if self.dc_scratch is None else self.dc_scratch.data_ptr()
        job.red_scratch = self.red_scratch.data_ptr()
        job.sino_size = e.numel()
        job.recon_size = x.numel()
        job.adelta = self.adelta.data_ptr()
        self.job = job

    def iteration(self, pidx, order, total_sino):
        """All subset steps of one iteration in `order`, then the statistics."""
        S = self.S
        ids = (np.asarray(order, dtype=np.int64) + int(self.offsets[pidx])).astype(np.int32)
        self.stats.zero_()
        _lib.check(_lib.lib().mbir_vcd_run(
            ctypes.byref(self.job), ids.ctypes.data_as(ctypes.c_void_p), int(ids.size),
            _lib.ptr(self.stats), 1, _lib.stream_ptr()), 'vcd iteration')
        h = self.stats.cpu().numpy()
        wsq, sq, wsum = h[S + 2], h[S + 3], h[S + 4]
        avg_w = 1.0 if self.w is None else wsum / total_sino
        return dict(sumsq=h[:S], ell1=h[S], alpha_sum=h[S + 1],
                    fm_loss=math.sqrt(wsq / (avg_w * total_sino)),
                    es_rmse=math.sqrt(sq / total_sino), recon_l1=h[S + 5])
