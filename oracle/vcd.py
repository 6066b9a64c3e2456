"""VCD reconstruction oracle (FP64 arithmetic, numpy legacy RNG) -- test infrastructure.

Restates the reference's multi-granular vectorized coordinate descent:

* pixel partitions from numpy's GLOBAL legacy RNG (permutation, then choice
  without replacement; one subset skips the RNG)   -- vcd_utils.py:97-193
* auto-regularization on a <=20-view subsample, 400-bin background
  histogram, 0.25 * median indicator, sigma_y / sigma_x / sigma_prox
                                                    -- tomography_model.py:2057-2227,
                                                       vcd_utils.py:263-306
* initial error state with the optimal scale alpha -- tomography_model.py:1666-1712
* subset update: prior, back projection, direction, prior line terms,
  forward of delta, lin/quad, clamped alpha, positivity, in-place update
                                                    -- tomography_model.py:54-87,2508-2657
* partition iterator (np.random.permutation per iteration) and the per
  iteration statistics                              -- tomography_model.py:2662-2697,
                                                       2920-2987, 2347-2357

The operator comes in as a ``system`` object (see ``ParallelSystem``) so the
same loop serves parallel and cone geometries.

Pinning: the partition RNG sequence and the auto-regularization values are
pinned by the reference's own outputs written to tests/golden/ by
tests/golden/make_golden.py (see that script); the converged fixed point for
a quadratic prior (p = q = 2) is pinned by an FP64 conjugate-gradient solve
of the normal equations on a tiny cell (tests/test_oracle_vcd.py).  The
exact per-iteration trajectory is pinned only by the golden trace.
"""

import warnings

import numpy as np

from . import filters as _filters
from . import parallel as _par
from . import qggmrf as _qg

F32_EPS = float(np.finfo(np.float32).eps)

DEFAULTS = dict(
    sigma_y=1.0, sigma_x=1.0, sigma_prox=1.0, p=2.0, q=1.2, T=1.0,
    qggmrf_nbr_wts=[1.0, 1.0, 1.0], auto_regularize_flag=True,
    positivity_flag=False, snr_db=30.0, sharpness=1.0,
    granularity=[1, 2, 4, 8, 16, 32, 64, 128, 128, 128, 128],
    partition_sequence=[2, 4, 6] + [7, 8, 9, 10] * 25, max_alpha=1.5,
    use_ror_mask=True)


# -- partitions (vcd_utils.py:97-193) ----------------------------------------
def gen_pixel_partition(recon_shape, num_subsets, use_ror_mask=True):
    nr, nc = recon_shape[:2]
    indices = np.arange(nr * nc, dtype=np.int32)
    if use_ror_mask:
        indices = indices[_par.ror_mask(recon_shape).flatten() == 1]
    if num_subsets > len(indices):
        num_subsets = len(indices)
    if num_subsets == 1:
        return np.sort(indices).reshape(1, -1)
    per = int(np.ceil(len(indices) / num_subsets))
    extra = num_subsets * per - len(indices)
    indices = np.random.permutation(indices)
    extra_idx = np.random.choice(indices[:(num_subsets - 1) * per], size=extra,
                                 replace=False)
    indices = np.concatenate((indices, extra_idx))
    return np.sort(indices.reshape(num_subsets, -1), axis=1)


def gen_set_of_pixel_partitions(recon_shape, granularity, use_ror_mask=True):
    return [gen_pixel_partition(recon_shape, g, use_ror_mask).astype(np.int64)
            for g in granularity]


def gen_partition_sequence(seq, max_iterations):
    seq = np.array(seq)
    if max_iterations > seq.size:
        return np.concatenate((seq, np.full(max_iterations - seq.size, seq[-1])))
    return seq[:max_iterations]


# -- auto-regularization -------------------------------------------------------
def subsample_views(a, max_views=20):
    """tomography_model.py:2161-2214 (single-array branch)."""
    v = a.shape[0]
    step = max(v // min(max_views, v), 1)
    return np.array(a[::step])


def background_cluster_boundaries(sino):
    """vcd_utils.py:263-306."""
    hist, edges = np.histogram(np.asarray(sino).ravel(), bins=400)
    centers = 0.5 * (edges[:-1] + edges[1:])
    peaks = []
    if len(hist) > 1 and hist[0] > hist[1]:
        peaks.append(0)
    for i in range(1, len(hist) - 1):
        if hist[i] >= hist[i - 1] and hist[i] > hist[i + 1]:
            peaks.append(i)
    if not peaks:
        k = int(np.argmin(np.abs(centers)))
    else:
        k = min(peaks, key=lambda i: abs(centers[i]))
    cutoff = 0.1 * hist[k]
    lo = hi = k
    while lo > 0 and hist[lo] > cutoff:
        lo -= 1
    while hi < len(hist) - 1 and hist[hi] > cutoff:
        hi += 1
    return centers[lo], centers[hi]


def sino_indicator(sino):
    """tomography_model.py:2216-2250."""
    sino = np.asarray(sino)
    left, right = background_cluster_boundaries(sino)
    threshold = right + (right - left)
    mx = np.max(sino)
    if mx <= 0 or mx < threshold:
        return np.ones_like(sino, dtype=np.int8)
    med = np.median(sino[sino >= threshold])
    return np.int8(sino >= 0.25 * med)


def auto_regularization(sino, weights, recon_shape, delta_voxel,
                        delta_det_channel, magnification, snr_db, sharpness):
    """Returns (sigma_y, sigma_x, sigma_prox) as python floats of float32
    values -- tomography_model.py:2057-2276."""
    small = subsample_views(np.asarray(sino))
    sw = 1 if weights is None else subsample_views(np.asarray(weights))
    ind = sino_indicator(small)
    signal_rms = float(np.average(sw * small ** 2, None, ind) ** 0.5)
    rel_noise = 10 ** (-snr_db / 20)
    rel_pitch = delta_voxel / (delta_det_channel / magnification)
    sigma_y = np.float32(rel_noise * signal_rms * rel_pitch ** 0.5)
    typical = np.average(np.abs(small), weights=ind)
    tpl_space = (2 * recon_shape[0] * recon_shape[1]) / (recon_shape[0] + recon_shape[1]) \
        * delta_voxel
    tpl_sino = small.shape[-1] * delta_det_channel / magnification
    recon_std = typical / np.minimum(tpl_space, tpl_sino)
    sigma_x = np.float32(0.2 * (2 ** sharpness) * recon_std)
    return float(sigma_y), float(sigma_x), float(sigma_x)


# -- the operators ---------------------------------------------------------------
class ParallelSystem:
    """Parallel-beam operator for the VCD oracle: the full ROR system matrix
    A (V*C, P) in FP64 acting on each slice; sinogram vectors are the
    (V*C, R) matrix of a (V, R, C) sinogram, flattened."""

    def __init__(self, angles, geom, recon_shape, use_ror_mask=True):
        self.angles = np.asarray(angles, dtype=np.float32)
        self.geom = geom
        self.recon_shape = tuple(recon_shape)
        self.pix = _par.full_indices(recon_shape, use_ror_mask)
        self.pos = -np.ones(recon_shape[0] * recon_shape[1], dtype=np.int64)
        self.pos[self.pix] = np.arange(len(self.pix))
        psf = _par.psf_radius(geom['delta_voxel'], geom['delta_voxel_row'],
                              geom['delta_det_channel'])
        self.A = _par.system_matrix(self.pix, self.angles, geom, psf).tocsc()
        self.A2 = _par.system_matrix(self.pix, self.angles, geom, psf,
                                     coeff_power=2).tocsc()
        self.V, self.C = len(angles), geom['num_channels']
        self.S = recon_shape[2]
        self.magnification = 1.0

    def cols(self, idx):
        p = self.pos[np.asarray(idx)]
        if np.any(p < 0):
            raise ValueError('index outside the ROR')
        return p

    def to_vec(self, s):
        V, R, C = s.shape
        return np.asarray(s, dtype=np.float64).transpose(0, 2, 1).reshape(-1)

    def from_vec(self, v):
        return v.reshape(self.V, self.C, self.S).transpose(0, 2, 1)

    def fwd(self, idx, d):
        return np.asarray(self.A[:, self.cols(idx)] @ d).reshape(-1)

    def back(self, idx, ev):
        return np.asarray(self.A[:, self.cols(idx)].T @ ev.reshape(-1, self.S))

    def hessian(self, wv):
        h = np.zeros((self.recon_shape[0] * self.recon_shape[1], self.S))
        h[self.pix] = self.A2.T @ wv.reshape(-1, self.S)
        return h

    def direct_recon(self, sino):
        """FBP: filter, then the adjoint at the ROR pixels -- parallel_beam.py:321-368."""
        f = _filters.fbp_filter_parallel(sino, self.geom['delta_voxel'],
                                         self.geom['delta_voxel_row'])
        flat = np.zeros((self.recon_shape[0] * self.recon_shape[1], self.S))
        flat[self.pix] = self.back(self.pix, self.to_vec(f))
        return flat

    def direction(self, fg, pg, fh, ph):
        return -((fg + pg) / (fh + ph))


class ConeSystem:
    """Cone-beam operator for the VCD oracle: A (V*R*C, P*S) from
    oracle/cone.py, FDK initialisation and the DC-damped direction
    (cone_beam.py:48-55, 476-496)."""

    def __init__(self, g, use_ror_mask=True, dc_damping=True):
        from . import cone as _cone
        self._cone = _cone
        self.g = g
        self.recon_shape = tuple(g['recon_shape'])
        self.S = self.recon_shape[2]
        self.pix = _par.full_indices(self.recon_shape, use_ror_mask)
        self.pos = -np.ones(self.recon_shape[0] * self.recon_shape[1], dtype=np.int64)
        self.pos[self.pix] = np.arange(len(self.pix))
        self.A = _cone.system_matrix(self.pix, g).tocsc()
        self.A2 = _cone.system_matrix(self.pix, g, coeff_power=2).tocsc()
        self.geom = dict(delta_voxel=g['dv'], delta_det_channel=g['ddc'])
        self.magnification = g['M']
        self.s_prof = _cone.dc_damping_profile(g) if dc_damping else None

    def _cols(self, idx):
        p = self.pos[np.asarray(idx)]
        if np.any(p < 0):
            raise ValueError('index outside the ROR')
        return (p[:, None] * self.S + np.arange(self.S)[None, :]).reshape(-1)

    def to_vec(self, s):
        return np.asarray(s, dtype=np.float64).reshape(-1)

    def from_vec(self, v):
        return v.reshape(self.g['V'], self.g['R'], self.g['C'])

    def fwd(self, idx, d):
        return np.asarray(self.A[:, self._cols(idx)] @ np.asarray(d).reshape(-1))

    def back(self, idx, ev):
        return np.asarray(self.A[:, self._cols(idx)].T @ ev).reshape(len(idx), self.S)

    def hessian(self, wv):
        h = np.zeros((self.recon_shape[0] * self.recon_shape[1], self.S))
        h[self.pix] = (self.A2.T @ wv).reshape(-1, self.S)
        return h

    def direct_recon(self, sino):
        c = self._cone
        g = self.g
        taps = _filters.ramp_filter(g['C']) * c.fdk_scale(g) * (np.pi / g['V'])
        f = _filters.apply_row_filter(sino, taps, row_weight=c.fdk_row_weight(g))
        flat = np.zeros((self.recon_shape[0] * self.recon_shape[1], self.S))
        flat[self.pix] = self.back(self.pix, self.to_vec(f))
        if float(np.max(g['z_shifts']) - np.min(g['z_shifts'])) > 0:
            flat = flat * c.helical_z_weight(g)[None, :]
        return flat

    def direction(self, fg, pg, fh, ph):
        if self.s_prof is None:
            return -((fg + pg) / (fh + ph))
        return self._cone.dc_damped_direction(fg, pg, fh, ph, self.s_prof)


def vcd_recon(system, sino, params, max_iterations, weights=None,
              init_recon=None, stop_threshold_change_pct=0.0,
              first_iteration=0, return_state=False):
    """Seeded callers must set np.random.seed before calling.  Returns
    (recon (rows, cols, slices), stats dict); FP64 throughout."""
    prm = dict(DEFAULTS)
    prm.update(params)
    rs = system.recon_shape
    S = rs[2]
    partitions = gen_set_of_pixel_partitions(rs, prm['granularity'],
                                             prm['use_ror_mask'])
    seq = gen_partition_sequence(prm['partition_sequence'], max_iterations)
    seq = seq[first_iteration:]
    if prm['auto_regularize_flag']:
        sy, sx, sp = auto_regularization(
            sino, weights, rs, system.geom['delta_voxel'],
            system.geom['delta_det_channel'], system.magnification,
            prm['snr_db'], prm['sharpness'])
        prm.update(sigma_y=sy, sigma_x=sx, sigma_prox=sp)
    y = system.to_vec(sino)
    w = None if weights is None else system.to_vec(weights)
    # Initial state (tomography_model.py:2803-2833, 1666-1712).
    if init_recon is None:
        x = system.direct_recon(sino)
        fwd = system.fwd(system.pix, x[system.pix])
        wf = fwd if w is None else w * fwd
        nrm = np.sum(wf * fwd)
        alpha = np.sum(wf * y) / nrm if nrm > 0 else 1.0
        e = y - alpha * fwd
        x = alpha * x
    else:
        x = np.asarray(init_recon, dtype=np.float64).reshape(-1, S).copy()
        e = y - system.fwd(system.pix, x[system.pix])
    # Hessian diagonal at the ROR pixels (tomography_model.py:2872-2889).
    hess = system.hessian(np.ones_like(y) if w is None else w)
    fm = 1.0 / prm['sigma_y'] ** 2
    b = _qg.b_from_nbr_wts(prm['qggmrf_nbr_wts'])
    n_sino = y.size
    stats = dict(fm_rmse=[], nmae=[], alpha=[], delta_norm_per_slice=[],
                 sigma_y=prm['sigma_y'], sigma_x=prm['sigma_x'])
    for it in range(len(seq)):
        part = partitions[seq[it]]
        order = np.random.permutation(part.shape[0])
        ell1 = 0.0
        asum = 0.0
        sumsq = np.zeros(S)
        for k in order:
            idx = part[k]
            gp, hp = _qg.grad_hess_at_indices(x, rs, idx, b, prm['sigma_x'],
                                              prm['p'], prm['q'], prm['T'])
            we = e if w is None else w * e
            fg = -fm * system.back(idx, we)
            fh = fm * hess[idx]
            d = system.direction(fg, gp, fh, hp)
            pl = np.sum(gp * d)
            pq = np.sum(hp * d * d)
            ad = system.fwd(idx, d)
            if w is None:
                lin, quad = fm * np.sum(e * ad), fm * np.sum(ad * ad)
            else:
                lin, quad = fm * np.sum(w * e * ad), fm * np.sum(ad * ad * w)
            alpha = (lin - pl) / (quad + pq + F32_EPS)
            alpha = min(max(alpha, F32_EPS), prm['max_alpha'])
            if prm['positivity_flag']:
                d = np.maximum(-x[idx] / (alpha + F32_EPS), d)
                ad = system.fwd(idx, d)
            ds = alpha * d
            # index_add_ semantics: a pixel listed twice in one subset (the
            # padding can repeat one) receives both updates.
            np.add.at(x, idx, ds)
            sumsq += np.sum(ds * ds, axis=0)
            e -= alpha * ad
            ell1 += np.sum(np.abs(ds))
            asum += alpha
        wavg = 1.0 if w is None else np.mean(w)
        ww = 1.0 if w is None else w
        fm_rmse = np.sqrt(np.sum(e * e * ww) / (wavg * n_sino)) / prm['sigma_y']
        l1 = np.sum(np.abs(x))
        nmae = ell1 / l1 if l1 else float('nan')
        stats['fm_rmse'].append(float(fm_rmse))
        stats['nmae'].append(float(nmae))
        stats['alpha'].append(float(asum / part.shape[0]))
        stats['delta_norm_per_slice'].append(np.sqrt(sumsq))
        if nmae < stop_threshold_change_pct / 100:
            break
    recon = x.reshape(rs[0], rs[1], S)
    if return_state:
        return recon, stats, system.from_vec(e), hess
    return recon, stats
