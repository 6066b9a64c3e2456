"""Cone-beam projector oracle (FP64) -- test infrastructure.

Restates the reference's cone geometry as two separable fans:

* rotated pixel (x, y) and per-pixel magnification 1/(1/M - y/SDD)
                                               -- cone_beam.py:59-76
* horizontal fan data (flat: u = mag*x, theta = atan2(u, SDD), W scaled by
  1/cos(theta); curved: u = SDD*atan2(x, SID - y), theta = u/SDD)
                                               -- cone_beam.py:79-112
* vertical affine map m(k) = m0 + W_r * k with W_r = mag*dz/drow and the
  helical z offset                             -- cone_beam.py:115-158
* cone-angle factor 1/cos(phi), phi = atan2(mag*z_k, SDD), applied to the
  voxel values (forward) and to the vertical weight (back)
                                               -- cone_beam.py:198-207,279-293
* trapezoid weights in both directions, zero off the detector
                                               -- horizontal_fan.py:21-24
* psf radii, auto recon geometry, FDK pre-weight / scale, helical z weight,
  DC damping profile                           -- cone_beam.py:519-658,661-796,
                                                  424-496

The system matrix entry for sinogram (v, m, n) and voxel (p, k) is
    A_h(v,p,n) * T(m0 + W_r*k - m; W_r) / cos(phi(v,p,k)),
    T(d; W) = min(max((W+1)/2 - |d|, 0), min(1, W)).
The reference's forward gathers voxel taps per detector row and its back
gathers row taps per voxel; both evaluate exactly this weight, so with tap
sets that cover every nonzero weight they are one matrix and its transpose.
Here every nonzero entry is assembled, with no tap truncation.

Pinning (tests/test_oracle_cone.py): brute force -- every entry equals the
product of the two interval overlaps (horizontal box footprint of width W_c
against the channel cell, vertical width W_r against the row cell) times the
weight scale and 1/cos(phi) computed element by element in FP64; the
reference's tap-radius formulas cover every nonzero entry (no tap dropped);
SDD -> infinity with zero z shift reduces to the parallel-beam oracle;
mass conservation horizontally (sum over channels of A_h = area/ddc * mag/cos
theta for a flat detector).
"""

import math

import numpy as np
import scipy.sparse as sp

from . import parallel as _par


def geometry(sinogram_shape, angles, sdd, sid, recon_shape=None, delta_voxel=None,
             z_shifts=None, curved=False, delta_det_channel=1.0, delta_det_row=1.0,
             det_channel_offset=0.0, det_row_offset=0.0, voxel_row_aspect=1.0,
             voxel_slice_aspect=1.0, recon_slice_offset=None):
    """Geometry dict; defaults follow ConeBeamModel's auto geometry
    (cone_beam.py:573-658, axial_pad_fraction 0)."""
    V, R, C = sinogram_shape
    M = 1.0 if math.isinf(sdd) else sdd / sid
    zs = np.zeros(V) if z_shifts is None else np.asarray(z_shifts, dtype=np.float64)
    if delta_voxel is None:
        delta_voxel = delta_det_channel / M
    if recon_shape is None:
        dvr = voxel_row_aspect * delta_voxel
        dvs = voxel_slice_aspect * delta_voxel
        nr = int(np.round(C * ((delta_det_channel / dvr) / M)))
        nc = int(np.round(C * ((delta_det_channel / delta_voxel) / M)))
        H_iso = R * (delta_det_row / M)
        travel = float(np.max(zs) - np.min(zs))
        ns = max(1, int(np.ceil((H_iso + travel) / dvs)))
        recon_shape = (nr, nc, ns)
    if recon_slice_offset is None:
        recon_slice_offset = 0.5 * (float(np.min(zs)) + float(np.max(zs)))
    return dict(V=V, R=R, C=C, angles=np.asarray(angles, dtype=np.float32).astype(np.float64),
                z_shifts=np.asarray(zs, dtype=np.float32).astype(np.float64), sdd=float(sdd),
                sid=float(sid), M=M, curved=bool(curved), recon_shape=tuple(recon_shape),
                dv=float(delta_voxel), dvr=float(voxel_row_aspect * delta_voxel),
                dvs=float(voxel_slice_aspect * delta_voxel), ddc=float(delta_det_channel),
                ddr=float(delta_det_row), dco=float(det_channel_offset),
                dro=float(det_row_offset), rso=float(recon_slice_offset))


def psf_radii(g):
    """(psf_radius, bp_psf_radius) -- cone_beam.py:519-553."""
    nr, nc, _ = g['recon_shape']
    ddet = min(g['ddr'], g['ddc'])
    if math.isinf(g['sdd']):
        mmax = mmin = 1.0
    else:
        sid = g['sdd'] / g['M']
        half = max(0.5 * nc * g['dv'], 0.5 * nr * g['dvr'])
        mmax = g['sdd'] / (sid - half)
        mmin = g['sdd'] / (sid + half)
    maxp = max(g['dv'], g['dvr'], g['dvs'])
    minp = min(g['dv'], g['dvr'], g['dvs'])
    psf = int(np.ceil(np.ceil(maxp * mmax / ddet) / 2))
    bp = int(np.ceil(np.ceil(ddet / (mmin * minp)) / 2))
    return psf, bp


def pixel_data(pix, g):
    """Per (view, pixel): n_p, W_c, ws, mag, m0, W_r, plus z_offset per view."""
    nr, nc, ns = g['recon_shape']
    pix = np.asarray(pix, dtype=np.int64)
    row = (pix // nc).astype(np.float64)
    col = (pix % nc).astype(np.float64)
    yt = g['dvr'] * (row - (nr - 1) / 2.0)
    xt = g['dv'] * (col - (nc - 1) / 2.0)
    a = g['angles'][:, None]
    c, s = np.cos(a), np.sin(a)
    x = c * xt[None] - s * yt[None]
    y = s * xt[None] + c * yt[None]
    sdd = g['sdd']
    mag = 1.0 / (1.0 / g['M'] - (np.zeros_like(y) if math.isinf(sdd) else y / sdd))
    if not g['curved']:
        u = mag * x
        theta = np.arctan2(u, sdd) if not math.isinf(sdd) else np.zeros_like(u)
    else:
        sid = sdd / g['M']
        u = sdd * np.arctan2(x, sid - y)
        theta = u / sdd
    n_p = (u + g['dco']) / g['ddc'] + (g['C'] - 1) / 2.0
    fp = np.maximum(np.abs(np.cos(a - theta)) * g['dv'], np.abs(np.sin(a - theta)) * g['dvr'])
    Wc = mag * (fp / g['ddc'])
    if not g['curved']:
        Wc = Wc / np.cos(theta)
    ws = (g['dvr'] * g['dv']) / fp
    z_off = g['rso'] - g['z_shifts']
    Wr = mag * g['dvs'] / g['ddr']
    z0 = z_off[:, None] - g['dvs'] * (ns - 1) / 2.0
    m0 = (mag * z0 + g['dro']) / g['ddr'] + (g['R'] - 1) / 2.0
    return n_p, Wc, ws, mag, m0, Wr, z_off


def _trap(d, W):
    return np.minimum(np.maximum((W + 1.0) / 2.0 - np.abs(d), 0.0), np.minimum(W, 1.0))


def system_matrix(pix, g, coeff_power=1):
    """Sparse A (V*R*C, P*S): rows in sinogram VRC order, columns voxel
    (position in pix, slice)."""
    n_p, Wc, ws, mag, m0, Wr, z_off = pixel_data(pix, g)
    V, R, C = g['V'], g['R'], g['C']
    ns = g['recon_shape'][2]
    P = len(pix)
    kk = np.arange(ns, dtype=np.float64)
    rows, cols, vals = [], [], []
    hw_c = int(np.ceil(np.max((Wc + 1) / 2))) + 1
    for v in range(V):
        z = g['dvs'] * (kk - (ns - 1) / 2.0) + z_off[v]
        for j in range(P):
            inv_cos = (np.ones(ns) if math.isinf(g['sdd'])
                       else np.sqrt(1.0 + (mag[v, j] * z / g['sdd']) ** 2))
            nc0 = int(np.floor(n_p[v, j]))
            nn = np.arange(nc0 - hw_c, nc0 + hw_c + 1)
            ah = _trap(n_p[v, j] - nn, Wc[v, j]) * ws[v, j]
            keep = (ah > 0) & (nn >= 0) & (nn < C)
            nn, ah = nn[keep], ah[keep]
            if nn.size == 0:
                continue
            mk = m0[v, j] + Wr[v, j] * kk                    # (S,)
            hw_r = int(np.ceil((Wr[v, j] + 1) / 2)) + 1
            # all slices at once: rows floor(m(k)) - hw_r .. + hw_r
            mm = np.floor(mk).astype(np.int64)[:, None] + np.arange(-hw_r, hw_r + 1)[None, :]
            av = _trap(mk[:, None] - mm, Wr[v, j]) * inv_cos[:, None]
            ks, ts = np.nonzero((av > 0) & (mm >= 0) & (mm < R))
            if ks.size == 0:
                continue
            w = av[ks, ts][:, None] * ah[None, :]
            if coeff_power != 1:
                w = w ** coeff_power
            rr = (v * R + mm[ks, ts])[:, None] * C + nn[None, :]
            rows.append(rr.ravel())
            cols.append(np.broadcast_to((j * ns + ks)[:, None], rr.shape).ravel())
            vals.append(w.ravel())
    rows = np.concatenate(rows) if rows else np.zeros(0, np.int64)
    cols = np.concatenate(cols) if cols else np.zeros(0, np.int64)
    vals = np.concatenate(vals) if vals else np.zeros(0)
    return sp.csr_matrix((vals, (rows, cols)), shape=(V * R * C, P * ns))


def forward(values, pix, g):
    """values (P, S) -> sinogram (V, R, C), FP64."""
    A = system_matrix(pix, g)
    out = A @ np.asarray(values, dtype=np.float64).reshape(-1)
    return out.reshape(g['V'], g['R'], g['C'])


def back(sino, pix, g, coeff_power=1):
    """sinogram (V, R, C) -> (P, S), FP64."""
    A = system_matrix(pix, g, coeff_power=coeff_power)
    return (A.T @ np.asarray(sino, dtype=np.float64).reshape(-1)).reshape(len(pix), -1)


def brute_force_matrix(pix, g):
    """Dense A entry by entry in FP64 from interval overlaps: horizontal
    footprint of width W_c against the channel cell and vertical footprint
    of width W_r against the row cell, times ws and 1/cos(phi)."""
    n_p, Wc, ws, mag, m0, Wr, z_off = pixel_data(pix, g)
    V, R, C = g['V'], g['R'], g['C']
    ns = g['recon_shape'][2]
    A = np.zeros((V * R * C, len(pix) * ns))

    def overlap(center, width, cell):
        return max(0.0, min(center + width / 2, cell + 0.5) - max(center - width / 2, cell - 0.5))

    for v in range(V):
        for j in range(len(pix)):
            for k in range(ns):
                z = g['dvs'] * (k - (ns - 1) / 2.0) + z_off[v]
                vv = mag[v, j] * z
                inv_cos = 1.0 if math.isinf(g['sdd']) else math.sqrt(1 + (vv / g['sdd']) ** 2)
                mk = m0[v, j] + Wr[v, j] * k
                for m in range(R):
                    ov_r = overlap(mk, Wr[v, j], m)
                    if ov_r <= 0:
                        continue
                    for n in range(C):
                        ov_c = overlap(n_p[v, j], Wc[v, j], n)
                        if ov_c > 0:
                            A[(v * R + m) * C + n, j * ns + k] = ov_c * ws[v, j] * ov_r * inv_cos
    return A


def reference_tap_sets_complete(pix, g):
    """True when the reference's tap windows (channel: round(n_p) +- psf;
    back vertical: round(m_p) +- psf; forward vertical: round(k_m) +- bp_psf)
    contain every nonzero weight."""
    psf, bp = psf_radii(g)
    n_p, Wc, ws, mag, m0, Wr, _ = pixel_data(pix, g)
    ok_c = np.all((Wc + 1) / 2 <= psf + 0.5 + 1e-12)
    ok_r = np.all((Wr + 1) / 2 <= psf + 0.5 + 1e-12)
    # forward: rows gather voxel taps; nonzero |k - k_m| < (W_r+1)/(2 W_r)
    ok_k = np.all((Wr + 1) / (2 * Wr) <= bp + 0.5 + 1e-12)
    return bool(ok_c and ok_r and ok_k)


def fdk_row_weight(g):
    """FDK cosine pre-weight (R, C): SDD / sqrt(SDD^2 + u^2 + v^2) -- cone_beam.py:661-695."""
    m = np.arange(g['R'])[:, None]
    n = np.arange(g['C'])[None, :]
    v = (m - (g['R'] - 1) / 2.0) * g['ddr'] - g['dro']
    u = (n - (g['C'] - 1) / 2.0) * g['ddc'] - g['dco']
    return g['sdd'] / np.sqrt(g['sdd'] ** 2 + u ** 2 + v ** 2)


def fdk_scale(g):
    return g['ddr'] / (g['dv'] * g['dvr'] * g['dvs'] * g['M'])


def helical_z_weight(g):
    """Per-slice V / coverage weight (cone_beam.py:753-796)."""
    ns = g['recon_shape'][2]
    zk = g['dvs'] * (np.arange(ns) - (ns - 1) / 2.0) + g['rso']
    half = 0.5 * g['R'] * g['ddr'] / g['M']
    vis = np.abs(zk[:, None] - g['z_shifts'][None, :]) <= half
    cov = np.sum(vis, axis=1)
    return np.where(cov > 0, g['V'] / np.maximum(cov, 1), 0.0)


def dc_damping_profile(g, cfg=(0.25, 100.0, 0.7, 0.5)):
    """Per-slice s_k (cone_beam.py:424-475)."""
    a, b, p, c = cfg
    nr, _, ns = g['recon_shape']
    L = nr * g['dv']
    dz = g['dvs']
    z = (np.arange(ns) - (ns - 1) / 2.0) * dz + g['rso']
    R = g['sid']
    zs = g['z_shifts']

    def prof(t):
        return (c * t ** p + a * b ** p) / (t ** p + b ** p)

    if zs.max() - zs.min() == 0:
        return prof(L * np.abs(z - zs[0]) / (R * dz))
    t = L * np.abs(z[:, None] - zs[None, :]) / (R * dz)
    return prof(t).mean(axis=1)


def dc_damped_direction(fg, pg, fh, ph, s_row):
    """cone_beam.py:48-55."""
    gsum = fg + pg
    hinv = 1.0 / (fh + ph)
    gbar = np.sum(hinv * gsum, axis=0) / np.sum(hinv, axis=0)
    return -(gsum - ((1.0 - s_row) * gbar)[None, :]) * hinv
