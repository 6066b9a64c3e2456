"""Pin the cone-beam oracle (oracle/cone.py): element-by-element brute force,
tap-set completeness of the reference's radii, the parallel-beam limit, and
FDK / helical / DC-damping closed forms."""

import math

import numpy as np
import pytest

from oracle import cone as oc
from oracle import parallel as op


def _g(curved=False, helical=False, **kw):
    V, R, C = 6, 10, 12
    angles = np.linspace(0, 2 * np.pi, V, endpoint=False) + 0.1
    zs = np.linspace(-2, 2, V) if helical else None
    return oc.geometry((V, R, C), angles, sdd=48.0, sid=24.0, z_shifts=zs, curved=curved,
                       recon_shape=(6, 6, 5), **kw)


@pytest.mark.parametrize('curved,helical', [(False, False), (True, False), (False, True)])
def test_cone_matrix_equals_brute_force(curved, helical):
    g = _g(curved=curved, helical=helical)
    pix = op.full_indices(g['recon_shape'])
    A = oc.system_matrix(pix, g).toarray()
    B = oc.brute_force_matrix(pix, g)
    assert np.abs(A).max() > 0
    assert np.abs(A - B).max() < 1e-12 * np.abs(B).max()


def test_reference_tap_radii_cover_every_weight():
    for g in (_g(), _g(curved=True), _g(helical=True),
              oc.geometry((8, 256, 512), np.linspace(0, 2 * np.pi, 8, endpoint=False),
                          sdd=2048.0, sid=1024.0)):
        pix = op.full_indices(g['recon_shape'])[:50]
        assert oc.reference_tap_sets_complete(pix, g)
    g = oc.geometry((8, 256, 512), np.linspace(0, 2 * np.pi, 8, endpoint=False),
                    sdd=2048.0, sid=1024.0)
    assert g['recon_shape'] == (512, 512, 256)
    assert oc.psf_radii(g) == (1, 1)


def test_parallel_limit():
    """SDD = inf, no z shift: cone reduces to parallel beam slice by slice."""
    V, R, C = 7, 4, 10
    angles = np.linspace(0, np.pi, V, endpoint=False) + 0.05
    g = oc.geometry((V, R, C), angles, sdd=math.inf, sid=math.inf, recon_shape=(8, 8, 4))
    pix = op.full_indices(g['recon_shape'])
    rng = np.random.default_rng(0)
    x = rng.standard_normal((len(pix), 4))
    got = oc.forward(x, pix, g)
    pg = dict(num_rows=8, num_cols=8, num_channels=C, delta_det_channel=1.0,
              det_channel_offset=0.0, delta_voxel=1.0, delta_voxel_row=1.0)
    ref = op.forward(x, pix, angles, pg)
    np.testing.assert_allclose(got, ref, atol=1e-12)


def test_forward_back_transpose_and_hessian():
    g = _g(helical=True)
    pix = op.full_indices(g['recon_shape'])
    rng = np.random.default_rng(1)
    x = rng.standard_normal((len(pix), 5))
    y = rng.standard_normal((6, 10, 12))
    assert np.sum(oc.forward(x, pix, g) * y) == pytest.approx(np.sum(x * oc.back(y, pix, g)),
                                                              rel=1e-12)
    A = oc.system_matrix(pix, g).toarray()
    H = oc.back(np.ones((6, 10, 12)), pix, g, coeff_power=2).reshape(-1)
    np.testing.assert_allclose(H, (A ** 2).sum(axis=0), rtol=1e-12)


def test_fdk_weight_and_dc_profile_closed_forms():
    g = _g()
    w = oc.fdk_row_weight(g)
    assert w.shape == (10, 12) and w.max() <= 1.0
    assert w[4, 5] == pytest.approx(48 / math.sqrt(48 ** 2 + 0.25 + 0.25))
    s = oc.dc_damping_profile(g)
    # s(t) = (c t^p + a b^p) / (t^p + b^p): s(0) = a, increasing toward c
    assert np.all((s >= 0.25 - 1e-12) & (s <= 0.5 + 1e-12))
    zw = oc.helical_z_weight(_g(helical=True))
    assert np.all(zw >= 1.0)
