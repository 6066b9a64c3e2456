"""Host-side VCD helpers: ROR mask, pixel partitions, weights.

The partitions are drawn from numpy's GLOBAL legacy RNG with exactly the
reference's call sequence (vcd_utils.py:97-193): np.random.permutation of the
ROR indices, then np.random.choice(..., replace=False) for the padding (which
consumes the RNG even when the padding is empty), one subset skipping the
RNG altogether, each subset sorted.  A seeded run therefore draws the same
subsets in the same order as the reference.
"""

import warnings

import numpy as np
import torch


def get_2d_ror_mask(recon_shape, *, use_ror_mask=True, crop_radius_pixels=0,
                    crop_radius_fraction=0.0):
    """Inscribed-ellipse region of reconstruction (vcd_utils.py:18-79)."""
    if use_ror_mask is False:
        if crop_radius_pixels != 0 and crop_radius_fraction != 0.0:
            raise ValueError('crop_radius_pixels and crop_radius_fraction must be zero '
                             'if use_ror_mask is set to False.')
        return np.ones(recon_shape[:2], dtype=bool)
    if use_ror_mask is True:
        if crop_radius_pixels != 0 and crop_radius_fraction != 0.0:
            raise ValueError('Only one of crop_radius_pixels and crop_radius_fraction '
                             'can be nonzero.')
        nr, nc = recon_shape[:2]
        rc, cc = (nr - 1) / 2, (nc - 1) / 2
        rr = rc - max(int(rc * crop_radius_fraction), crop_radius_pixels, 0)
        cr = cc - max(int(cc * crop_radius_fraction), crop_radius_pixels, 0)
        cols, rows = np.meshgrid(np.arange(nc) - cc, np.arange(nr) - rc)
        return (cols / cr) ** 2 + (rows / rr) ** 2 <= 1.0
    mask = np.asarray(use_ror_mask)
    if mask.shape != tuple(recon_shape[:2]):
        raise ValueError(f'Custom use_ror_mask must have shape recon_shape[:2]. '
                         f'Got mask.shape={mask.shape}, expected {tuple(recon_shape[:2])}.')
    if not np.all((mask == 0) | (mask == 1)):
        raise ValueError('Custom use_ror_mask must contain only 0s and 1s.')
    return mask


def _ror_indices(recon_shape, use_ror_mask):
    """Flat (row-major) indices of the region of reconstruction, int32."""
    nr, nc = recon_shape[:2]
    indices = np.arange(nr * nc, dtype=np.int32)
    if use_ror_mask is not False:
        indices = indices[get_2d_ror_mask(recon_shape, use_ror_mask=use_ror_mask).flatten() == 1]
    return indices


def gen_pixel_partition(recon_shape, num_subsets, use_ror_mask=True):
    return _partition_of(_ror_indices(recon_shape, use_ror_mask), num_subsets)


def _partition_of(indices, num_subsets):
    """The subsets of one partition of `indices` (vcd_utils.py:97-157)."""
    if num_subsets > len(indices):
        num_subsets = len(indices)
        warnings.warn('\nThe number of partition subsets is greater than the number of '
                      'pixels in the region of reconstruction.  \nReducing the number '
                      'of subsets to equal the number of indices.')
    if num_subsets == 1:
        return np.sort(indices).reshape(1, -1)
    per_subset = int(np.ceil(len(indices) / num_subsets))
    num_extra = num_subsets * per_subset - len(indices)
    indices = np.random.permutation(indices)
    extra = np.random.choice(indices[:(num_subsets - 1) * per_subset], size=num_extra,
                             replace=False)
    indices = np.concatenate((indices, extra))
    return np.sort(indices.reshape(num_subsets, indices.size // num_subsets), axis=1)


def gen_set_of_pixel_partitions(recon_shape, granularity, device=None, use_ror_mask=True):
    # The region's index list is the same for every granularity: formed once.
    indices = _ror_indices(recon_shape, use_ror_mask)
    return [torch.as_tensor(np.ascontiguousarray(_partition_of(indices, g)),
                            dtype=torch.int64, device=device) for g in granularity]


def gen_partition_sequence(partition_sequence, max_iterations):
    seq = np.array(partition_sequence)
    if max_iterations > seq.size:
        return np.concatenate((seq, np.full(max_iterations - seq.size, seq[-1])))
    return seq[:max_iterations]


def gen_full_indices(recon_shape, use_ror_mask=True):
    return gen_pixel_partition(recon_shape, num_subsets=1, use_ror_mask=use_ror_mask)[0]


def gen_weights(sinogram, weight_type):
    """Noise-model weights (vcd_utils.py:205-260)."""
    xp = torch if isinstance(sinogram, torch.Tensor) else np
    if weight_type == 'unweighted':
        return xp.ones_like(sinogram)
    if weight_type == 'transmission':
        return xp.exp(-sinogram)
    if weight_type == 'transmission_root':
        return xp.exp(-sinogram / 2)
    if weight_type == 'emission':
        return 1.0 / (xp.abs(sinogram) + 0.1)
    raise Exception('gen_weights: undefined weight_type {}'.format(weight_type))


def estimate_background_cluster_boundaries(sinogram):
    """Background cluster bounds from a 400-bin histogram (vcd_utils.py:263-306)."""
    hist, edges = np.histogram(np.asarray(sinogram).ravel(), bins=400)
    centers = 0.5 * (edges[:-1] + edges[1:])
    peaks = []
    if len(hist) > 1 and hist[0] > hist[1]:
        peaks.append(0)
    for i in range(1, len(hist) - 1):
        if hist[i] >= hist[i - 1] and hist[i] > hist[i + 1]:
            peaks.append(i)
    if not peaks:
        k = int(np.argmin(np.abs(centers - 0.0)))
    else:
        k = min(peaks, key=lambda i: abs(centers[i] - 0.0))
    cutoff = 0.1 * hist[k]
    lo = hi = k
    while lo > 0 and hist[lo] > cutoff:
        lo -= 1
    while hi < len(hist) - 1 and hist[hi] > cutoff:
        hi += 1
    return centers[lo], centers[hi]
