"""Parallel-beam projector oracle (FP64 by default) -- test infrastructure.

Restates the reference's parallel-beam geometry chain and the shared
horizontal-fan trapezoid rule:

* geometry chain ``_parallel_hfan_math``      -- parallel_beam.py:22-53
* tap rule ``tap_weights``                    -- horizontal_fan.py:21-24,51-71
* forward body (scatter over taps)            -- horizontal_fan.py:74-107,
                                                 parallel_beam.py:56-73
* back body (gather over taps, coeff_power)   -- horizontal_fan.py:110-154,
                                                 parallel_beam.py:76-94
* psf radius / auto recon geometry            -- parallel_beam.py:172-206
* ROR mask and full index list                -- vcd_utils.py:18-67,196-202

Here the operator is assembled as one explicit sparse matrix A of shape
(V*C, P*1) per call, so forward is ``A @ values`` and back is ``A.T @ sino``
(``A**coeff_power`` entrywise for the Hessian diagonal).

Pinning (independent of this restatement), in tests/test_oracle_parallel.py:
* brute force: every matrix entry equals an element-by-element FP64 interval
  intersection (projected box footprint of width W against detector cell
  [n-1/2, n+1/2]) times the in-plane area / footprint length, on tiny cells;
* mass conservation: without detector-edge truncation, sum over channels of
  A(v, p, .) equals delta_voxel * delta_voxel_row / delta_det_channel;
* known answers: a voxel at iso at angle 0 lands on the centre channel of an
  odd detector with weight 1; at 45 degrees the centre tap weight is 1;
* the psf radius bounds every nonzero tap (no tap is dropped).
"""

import numpy as np
import scipy.sparse as sp


def psf_radius(delta_voxel, delta_voxel_row, delta_det_channel):
    """parallel_beam.py:172-180."""
    max_footprint = max(delta_voxel, delta_voxel_row)
    return int(np.ceil(np.ceil(max_footprint / delta_det_channel) / 2))


def auto_recon_shape(sinogram_shape, delta_det_channel=1.0, delta_det_row=1.0,
                     voxel_row_aspect=1.0):
    """parallel_beam.py:182-206 (magnification 1).  Returns (recon_shape,
    delta_voxel)."""
    delta_voxel = delta_det_channel / 1.0
    delta_voxel_row = voxel_row_aspect * delta_voxel
    _, rows, channels = sinogram_shape
    nr = int(np.ceil(channels * delta_det_channel / delta_voxel_row))
    nc = int(np.ceil(channels * delta_det_channel / delta_voxel))
    ns = int(np.round(rows * (delta_det_row / delta_voxel)))
    return (nr, nc, ns), delta_voxel


def ror_mask(recon_shape):
    """Inscribed-ellipse region of reconstruction, vcd_utils.py:50-67."""
    nr, nc = recon_shape[:2]
    rc, cc = (nr - 1) / 2, (nc - 1) / 2
    col = (np.arange(nc) - cc)[None, :]
    row = (np.arange(nr) - rc)[:, None]
    return (col / cc) ** 2 + (row / rc) ** 2 <= 1.0


def full_indices(recon_shape, use_ror_mask=True):
    """Sorted flat pixel indices inside the ROR, vcd_utils.py:196-202."""
    n = recon_shape[0] * recon_shape[1]
    idx = np.arange(n, dtype=np.int64)
    if use_ror_mask:
        idx = idx[ror_mask(recon_shape).ravel()]
    return idx


def hfan(pix, angles, geom, dtype=np.float64, trig=None):
    """Per (view, pixel) hfan data (n_p, centers, W, ws) -- parallel_beam.py:22-53.

    ``geom`` is a dict with num_rows, num_cols, num_channels,
    delta_det_channel, det_channel_offset, delta_voxel, delta_voxel_row.
    ``dtype`` float32 evaluates each operation in single precision in the
    reference's order (no fused multiply-add); ``trig`` optionally supplies
    (cos, sin) per view, e.g. from torch's float32 kernels.
    """
    t = dtype
    pix = np.asarray(pix, dtype=np.int64)
    nr, nc, C = geom['num_rows'], geom['num_cols'], geom['num_channels']
    row = (pix // nc).astype(t)
    col = (pix % nc).astype(t)
    y_t = t(geom['delta_voxel_row']) * (row - t((nr - 1) / 2.0))
    x_t = t(geom['delta_voxel']) * (col - t((nc - 1) / 2.0))
    if trig is None:
        a = np.asarray(angles, dtype=np.float32).astype(t)
        c, s = np.cos(a), np.sin(a)
    else:
        c, s = (np.asarray(v, dtype=t) for v in trig)
    x = c[:, None] * x_t[None, :] - s[:, None] * y_t[None, :]
    n_p = (x + t(geom['det_channel_offset'])) / t(geom['delta_det_channel']) \
        + t((C - 1) / 2.0)
    fp = np.maximum(np.abs(c) * t(geom['delta_voxel']),
                    np.abs(s) * t(geom['delta_voxel_row']))
    W = fp / t(geom['delta_det_channel'])
    ws = t(geom['delta_voxel_row'] * geom['delta_voxel']) / fp
    centers = np.rint(n_p).astype(np.int64)
    return n_p, centers, W[:, None], ws[:, None]


def tap_weight(n_p, n, W, ws, num_channels):
    """Trapezoid weight of tap n, zeroed off the detector; horizontal_fan.py:51-71.
    Returns (A, n clamped)."""
    t = n_p.dtype.type
    nf = n.astype(n_p.dtype)
    A = np.maximum((W + t(1.0)) / t(2.0) - np.abs(n_p - nf), t(0.0))
    A = np.minimum(A, np.minimum(W, t(1.0))) * ws
    A = A * ((n >= 0) & (n < num_channels)).astype(n_p.dtype)
    return A, np.clip(n, 0, num_channels - 1)


def system_matrix(pix, angles, geom, psf, coeff_power=1, dtype=np.float64,
                  trig=None):
    """Sparse A (V*C, P): row v*C + n, column = position in ``pix``."""
    n_p, centers, W, ws = hfan(pix, angles, geom, dtype=dtype, trig=trig)
    V, P = n_p.shape
    C = geom['num_channels']
    rows, cols, vals = [], [], []
    col_idx = np.broadcast_to(np.arange(P)[None, :], (V, P))
    vbase = (np.arange(V) * C)[:, None]
    for off in range(-psf, psf + 1):
        A, n = tap_weight(n_p, centers + off, W, ws, C)
        if coeff_power != 1:
            A = A ** coeff_power
        rows.append((vbase + n).ravel())
        cols.append(col_idx.ravel())
        vals.append(A.astype(np.float64).ravel())
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    vals = np.concatenate(vals)
    return sp.csr_matrix((vals, (rows, cols)), shape=(V * C, P))


def forward(values, pix, angles, geom, psf=None, dtype=np.float64, trig=None):
    """Sparse forward projection: values (P, R) -> sinogram (V, R, C), FP64
    accumulation.  horizontal_fan.py:74-107 + parallel_beam.py:56-73."""
    if psf is None:
        psf = psf_radius(geom['delta_voxel'], geom['delta_voxel_row'],
                         geom['delta_det_channel'])
    values = np.asarray(values, dtype=np.float64)
    A = system_matrix(pix, angles, geom, psf, dtype=dtype, trig=trig)
    V, C = len(angles), geom['num_channels']
    out = (A @ values).reshape(V, C, values.shape[1])
    return np.ascontiguousarray(out.transpose(0, 2, 1))


def back(sino, pix, angles, geom, psf=None, coeff_power=1, dtype=np.float64,
         trig=None):
    """Sparse back projection: sinogram (V, R, C) -> (P, R), FP64 accumulation.
    horizontal_fan.py:110-154 + parallel_beam.py:76-94."""
    if psf is None:
        psf = psf_radius(geom['delta_voxel'], geom['delta_voxel_row'],
                         geom['delta_det_channel'])
    sino = np.asarray(sino, dtype=np.float64)
    V, R, C = sino.shape
    A = system_matrix(pix, angles, geom, psf, coeff_power=coeff_power,
                      dtype=dtype, trig=trig)
    s = sino.transpose(0, 2, 1).reshape(V * C, R)
    return np.asarray(A.T @ s)


def brute_force_matrix(pix, angles, geom):
    """Dense A (V*C, P) built entry by entry from the geometry itself: the
    overlap length of the projected box footprint [n_p - W/2, n_p + W/2] with
    detector cell [n - 1/2, n + 1/2], times in-plane area / footprint length.
    FP64, pure python loops -- tiny cells only."""
    import math
    nr, nc, C = geom['num_rows'], geom['num_cols'], geom['num_channels']
    dv, dvr = geom['delta_voxel'], geom['delta_voxel_row']
    ddc, dco = geom['delta_det_channel'], geom['det_channel_offset']
    V, P = len(angles), len(pix)
    A = np.zeros((V * C, P))
    for v, th in enumerate(np.asarray(angles, dtype=np.float32)):
        th = float(th)
        c, s = math.cos(th), math.sin(th)
        fp = max(abs(c) * dv, abs(s) * dvr)
        for j, p in enumerate(pix):
            i_row, i_col = divmod(int(p), nc)
            yt = dvr * (i_row - (nr - 1) / 2)
            xt = dv * (i_col - (nc - 1) / 2)
            u = c * xt - s * yt                      # detector coordinate (ALU)
            lo = (u - fp / 2 + dco) / ddc + (C - 1) / 2
            hi = (u + fp / 2 + dco) / ddc + (C - 1) / 2
            for n in range(C):
                ov = min(hi, n + 0.5) - max(lo, n - 0.5)
                if ov > 0:
                    A[v * C + n, j] = ov * dv * dvr / fp
    return A
