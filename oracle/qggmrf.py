"""qGGMRF prior oracle (FP64) -- test infrastructure.

Restates the reference's qGGMRF surrogate gradient / Hessian diagonal at a
set of pixel cylinders and its loss:

* neighbour weights b (row+, row-, col+, col-, slice+, slice-) -- qggmrf.py:22-32
* 2*b_tilde = rho'(delta)/delta with the |delta| floor T*sigma_x*eps32
                                                               -- qggmrf.py:35-74
* gradient / Hessian at indices: reflected slice ends, clamped in-plane
  neighbours                                                   -- qggmrf.py:77-183
* proximal-map gradient                                        -- qggmrf.py:186-206
* loss rho summed over np.diff pairs                           -- qggmrf.py:209-252

Pinning (tests/test_oracle_qggmrf.py): the gradient equals FP64 central
differences of ``loss`` at interior voxels (for p, q in the configs' ranges);
each neighbour's surrogate coefficient times delta equals rho'(delta) from
differentiating the closed-form rho; p = q = 2 gives the constant Hessian
sum(b) * 2 / (2 sigma_x^2) ... (rho''(0) for the quadratic case); boundary
voxels see zero deltas across the clamped / reflected edge.
"""

import numpy as np

F32_EPS = float(np.finfo(np.float32).eps)


def b_from_nbr_wts(nbr_wts):
    """qggmrf.py:22-32."""
    b = np.array([nbr_wts, nbr_wts], dtype=np.float64).T.flatten()
    return tuple(b / np.sum(b))


def two_b_tilde(delta, sigma_x, p, q, T):
    """rho'(delta)/delta (Table 8.1 surrogate), |delta| floored at
    T*sigma_x*eps32 -- qggmrf.py:35-57 (times 2)."""
    a_min = T * sigma_x * F32_EPS
    ad = np.maximum(np.abs(delta), a_min)
    ds = ad / (T * sigma_x)
    ds_qp = ds ** (q - p)
    num = ad ** (p - 2) / (2 * sigma_x ** p) * ds_qp * ((q / p) + ds_qp)
    return 2.0 * num / (1 + ds_qp) ** 2


def rho(delta, sigma_x, p, q, T):
    """qGGMRF potential, qggmrf.py:236-244."""
    a_min = T * sigma_x * F32_EPS
    ad = np.maximum(np.abs(delta), a_min)
    ds_qp = (ad / (T * sigma_x)) ** (q - p)
    return (ad ** p) / (p * sigma_x ** p) * ds_qp / (1 + ds_qp)


def grad_hess_at_indices(flat_recon, recon_shape, pix, b, sigma_x, p, q, T):
    """Gradient and Hessian surrogate at cylinders ``pix`` of the flat
    (rows*cols, slices) recon -- qggmrf.py:77-183.  Returns (P, S) arrays."""
    x = np.asarray(flat_recon, dtype=np.float64)
    pix = np.asarray(pix, dtype=np.int64)
    nr, nc = recon_shape[0], recon_shape[1]
    cyl = x[pix]                                       # (P, S)
    # Slice direction: deltas v[j]-v[j-1], zero at both reflected ends.
    d = np.concatenate([np.zeros((len(pix), 1)), np.diff(cyl, axis=1),
                        np.zeros((len(pix), 1))], axis=1)   # (P, S+1)
    c2 = two_b_tilde(d, sigma_x, p, q, T)
    cd = c2 * d
    grad = -b[4] * cd[:, 1:] + b[5] * cd[:, :-1]
    hess = b[4] * c2[:, 1:] + b[5] * c2[:, :-1]
    row, col = pix // nc, pix % nc
    for (dr, dc), bv in (((1, 0), b[0]), ((-1, 0), b[1]),
                         ((0, 1), b[2]), ((0, -1), b[3])):
        r = np.clip(row + dr, 0, nr - 1)
        c = np.clip(col + dc, 0, nc - 1)
        delta = cyl - x[r * nc + c]
        c2 = bv * two_b_tilde(delta, sigma_x, p, q, T)
        grad = grad + c2 * delta
        hess = hess + c2
    return grad, hess


def prox_gradient_at_indices(flat_recon, prox_input, pix, sigma_prox):
    """qggmrf.py:186-206."""
    return (np.asarray(flat_recon)[pix] - np.asarray(prox_input)[pix]) / sigma_prox ** 2


def loss(full_recon, b, sigma_x, p, q, T):
    """Prior loss over all neighbour pairs -- qggmrf.py:209-252."""
    x = np.asarray(full_recon, dtype=np.float64)
    b_axis = [(b[j] + b[j + 1]) / (2 * sum(b)) for j in (0, 2, 4)]
    total = 0.0
    for axis in range(3):
        total += float(np.sum(b_axis[axis] * rho(np.diff(x, axis=axis),
                                                 sigma_x, p, q, T)))
    return total
