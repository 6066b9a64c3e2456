"""Pin the parallel-beam oracle (oracle/parallel.py) independently of its own
formula: brute-force interval intersection, mass conservation, known answers,
tap-range completeness, and the FP32 geometry chain vs torch's float32 ops."""

import numpy as np
import pytest

from oracle import parallel as op


def _geom(nr, nc, C, ddc=1.0, dco=0.0, dv=1.0, dvr=1.0):
    return dict(num_rows=nr, num_cols=nc, num_channels=C, delta_det_channel=ddc,
                det_channel_offset=dco, delta_voxel=dv, delta_voxel_row=dvr)


GEOMS = {
    'square': _geom(8, 8, 11),
    'offset': _geom(8, 10, 13, dco=1.3),
    'wide_voxel': _geom(6, 6, 17, dv=2.0),
    'row_aspect': _geom(9, 7, 12, dvr=0.6),
    'fine_channel': _geom(7, 7, 15, ddc=0.75),
    'coarse_channel': _geom(8, 8, 6, ddc=1.7),
}


@pytest.mark.parametrize('name', list(GEOMS))
def test_system_matrix_equals_brute_force(name):
    g = GEOMS[name]
    angles = np.linspace(0, np.pi, 13, endpoint=False) + 0.01
    pix = np.arange(g['num_rows'] * g['num_cols'])
    psf = op.psf_radius(g['delta_voxel'], g['delta_voxel_row'], g['delta_det_channel'])
    A = op.system_matrix(pix, angles, g, psf).toarray()
    B = op.brute_force_matrix(pix, angles, g)
    assert np.abs(A - B).max() < 1e-12 * max(1.0, np.abs(B).max())


def test_mass_conservation():
    """Without edge truncation every pixel deposits dv*dvr/ddc per view."""
    g = _geom(16, 16, 64, dv=1.0, dvr=0.8, ddc=0.9)
    angles = np.linspace(0, np.pi, 37, endpoint=False)
    pix = op.full_indices((16, 16, 1))
    psf = op.psf_radius(1.0, 0.8, 0.9)
    A = op.system_matrix(pix, angles, g, psf).toarray().reshape(37, 64, -1)
    col_sums = A.sum(axis=1)
    np.testing.assert_allclose(col_sums, 0.8 / 0.9, rtol=1e-12)


def test_known_answers():
    # Odd detector, voxel at iso, angle 0: one channel, weight 1.
    g = _geom(3, 3, 5)
    A = op.system_matrix(np.array([4]), np.array([0.0]), g, 1).toarray()
    assert A[2, 0] == pytest.approx(1.0, abs=1e-7)
    assert np.count_nonzero(np.abs(A) > 1e-7) == 1
    # 45 degrees: W = 0.7071, ws = 1.4142, centre tap weight min(1,W)*ws = 1.
    A = op.system_matrix(np.array([4]), np.array([np.pi / 4]), g, 1).toarray()
    assert A[2, 0] == pytest.approx(1.0, abs=1e-6)
    assert A.sum() == pytest.approx(1.0, abs=1e-6)


def test_forward_back_are_transposes():
    g = _geom(10, 10, 14, dco=0.4)
    angles = np.linspace(0, np.pi, 9, endpoint=False)
    pix = op.full_indices((10, 10, 3))
    rng = np.random.default_rng(0)
    x = rng.standard_normal((len(pix), 3))
    y = rng.standard_normal((9, 3, 14))
    ax = op.forward(x, pix, angles, g)
    aty = op.back(y, pix, angles, g)
    assert np.sum(ax * y) == pytest.approx(np.sum(x * aty), rel=1e-12)


def test_no_nonzero_tap_outside_psf():
    """Every nonzero trapezoid overlap lies within the reference's tap range,
    so the truncated tap loop drops nothing (brute force has no tap limit)."""
    for name, g in GEOMS.items():
        angles = np.linspace(0, np.pi, 31)
        pix = np.arange(g['num_rows'] * g['num_cols'])
        psf = op.psf_radius(g['delta_voxel'], g['delta_voxel_row'], g['delta_det_channel'])
        A = op.system_matrix(pix, angles, g, psf).toarray()
        B = op.brute_force_matrix(pix, angles, g)
        assert np.allclose(A, B, atol=1e-12), name


def test_two_tap_rule():
    """When W <= 1 the nonzero taps are floor(n_p) and floor(n_p)+1 (the
    kernel's two-tap specialisation)."""
    g = _geom(12, 12, 20, dco=0.37)
    angles = np.linspace(0, np.pi, 50, endpoint=False)
    pix = np.arange(144)
    n_p, centers, W, ws = op.hfan(pix, angles, g)
    assert np.all(W <= 1.0)
    A = op.system_matrix(pix, angles, g, 1).toarray().reshape(50, 20, 144)
    base = np.floor(n_p).astype(int)
    for v in range(50):
        for j in range(144):
            nz = np.nonzero(np.abs(A[v, :, j]) > 0)[0]
            assert set(nz) <= {base[v, j], base[v, j] + 1}


def test_fp32_chain_matches_torch_float32():
    """The oracle's float32 mode reproduces torch's float32 ops bit for bit
    when given torch's cos/sin (the host table uses the same ops)."""
    import torch
    g = _geom(64, 64, 64, dco=0.25)
    angles = np.linspace(0, np.pi, 30, endpoint=False).astype(np.float32)
    a = torch.as_tensor(angles)
    c, s = torch.cos(a).numpy(), torch.sin(a).numpy()
    pix = np.arange(64 * 64)
    n_p, centers, W, ws = op.hfan(pix, angles, g, dtype=np.float32, trig=(c, s))
    # torch restatement of the same chain (float32 everywhere)
    ri = torch.as_tensor(pix // 64).float()
    ci = torch.as_tensor(pix % 64).float()
    yt = 1.0 * (ri - 63 / 2.0)
    xt = 1.0 * (ci - 63 / 2.0)
    x = torch.as_tensor(c)[:, None] * xt[None] - torch.as_tensor(s)[:, None] * yt[None]
    npt = (x + 0.25) / 1.0 + 63 / 2.0
    assert np.array_equal(n_p, npt.numpy())


def test_auto_recon_shape_and_psf():
    assert op.auto_recon_shape((720, 64, 512)) == ((512, 512, 64), 1.0)
    assert op.auto_recon_shape((180, 8, 128)) == ((128, 128, 8), 1.0)
    assert op.psf_radius(1.0, 1.0, 1.0) == 1
    assert op.psf_radius(2.0, 1.0, 1.0) == 1
    assert op.psf_radius(3.0, 1.0, 1.0) == 2
    assert len(op.full_indices((512, 512, 64))) == 205012
    assert len(op.full_indices((128, 128, 8))) == 12644
