"""Direct-reconstruction (FBP / FDK) filter oracle (FP64) -- test infrastructure.

* ramp filter h[n] = 1/2 sinc(n) - 1/4 sinc(n/2)^2, n in [-C+1, C-1]
                                                    -- tomography_utils.py:18-26
* scale by filter_scale * pi / V folded into h     -- tomography_model.py:2279-2321
* per-row 'valid' convolution (output length C), optional per-detector
  pre-weight                                        -- tomography_utils.py:29-78
* parallel FBP scale 1/(dv * dvr)                   -- parallel_beam.py:294-319
* cone FDK cosine pre-weight and scale             -- cone_beam.py:661-695

Pinning (tests/test_oracle_filters.py): the taps equal the closed-form
Ram-Lak kernel h[0] = 1/4, h[odd n] = -1/(pi^2 n^2), h[even n != 0] = 0; the
row convolution equals an explicit O(C^2) sum and numpy.convolve 'valid'.
"""

import numpy as np


def ramp_filter(num_channels):
    """Length 2C-1 ramp taps, FP64."""
    n = np.arange(-num_channels + 1, num_channels)
    return 0.5 * np.sinc(n) - 0.25 * np.sinc(n / 2) ** 2


def apply_row_filter(sino, taps, row_weight=None):
    """'valid' convolution of every (view, row) with ``taps`` (length 2C-1):
    out[v, r, c] = sum_k in[v, r, k] * taps[c - k + C - 1]."""
    x = np.asarray(sino, dtype=np.float64)
    if row_weight is not None:
        x = x * np.asarray(row_weight, dtype=np.float64)[None]
    V, R, C = x.shape
    rows = x.reshape(-1, C)
    out = np.empty_like(rows)
    for i in range(rows.shape[0]):
        out[i] = np.convolve(rows[i], taps)[C - 1:2 * C - 1]
    return out.reshape(V, R, C)


def fbp_filter_parallel(sino, delta_voxel, delta_voxel_row):
    """Parallel FBP filtering: taps * (1/(dv*dvr)) * pi/V."""
    V, _, C = np.shape(sino)
    taps = ramp_filter(C) * (1.0 / (delta_voxel * delta_voxel_row)) * (np.pi / V)
    return apply_row_filter(sino, taps)
