"""Seeded synthetic phantoms for tests and benchmarks.

``random_ellipsoids``: K ellipsoids from ``np.random.default_rng(seed)`` (the
global numpy RNG is never touched, so the VCD partition sequence of a later
seeded recon is unaffected): centres uniform in the inner 80% of each axis,
semi-axes uniform in [a_min, a_max] voxels, a uniformly random 3-D rotation,
attenuation uniform in [0, mu_max] per ALU, summed, then multiplied by the
inscribed-ellipse ROR mask so the phantom lies where the model reconstructs.
The reference builds its phantoms from the same pieces (utilities.py:64-98
add_ellipsoid, utilities.py:1307-1341 _gen_ellipsoid).
"""

import numpy as np

from .vcd_utils import get_2d_ror_mask


def _random_rotation(rng):
    """Uniform random rotation matrix (QR of a Gaussian matrix, sign-fixed)."""
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))[None, :]
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def random_ellipsoids(recon_shape, num_ellipsoids=10, seed=0, axes_range=(5.0, 30.0),
                      mu_max=0.02, ror_mask=True, dtype=np.float32):
    nr, nc, ns = recon_shape
    rng = np.random.default_rng(seed)
    vol = np.zeros(recon_shape, dtype=np.float64)
    for _ in range(num_ellipsoids):
        center = np.array([rng.uniform(0.1 * n, 0.9 * n) for n in (nr, nc, ns)])
        axes = rng.uniform(axes_range[0], axes_range[1], size=3)
        rot = _random_rotation(rng)
        mu = rng.uniform(0.0, mu_max)
        # Evaluate only inside the bounding box of the largest semi-axis.
        rad = float(np.max(axes))
        lo = [max(0, int(np.floor(center[k] - rad))) for k in range(3)]
        hi = [min(n, int(np.ceil(center[k] + rad)) + 1) for k, n in enumerate((nr, nc, ns))]
        d0 = np.arange(lo[0], hi[0], dtype=np.float64)[:, None, None] - center[0]
        d1 = np.arange(lo[1], hi[1], dtype=np.float64)[None, :, None] - center[1]
        d2 = np.arange(lo[2], hi[2], dtype=np.float64)[None, None, :] - center[2]
        q = 0.0
        for k in range(3):
            u = rot[0, k] * d0 + rot[1, k] * d1 + rot[2, k] * d2
            q = q + (u / axes[k]) ** 2
        vol[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] += mu * (q <= 1.0)
    if ror_mask:
        vol *= get_2d_ror_mask(recon_shape)[:, :, None]
    return vol.astype(dtype)


def add_noise(sinogram, rel_std=0.01, seed=0):
    """Gaussian noise of std rel_std * max|sinogram| from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    s = np.asarray(sinogram, dtype=np.float64)
    noisy = s + rng.normal(0.0, rel_std * np.max(np.abs(s)), size=s.shape)
    return noisy.astype(np.float32)
