"""Pin the qGGMRF oracle (oracle/qggmrf.py): gradient = central differences of
the closed-form loss; each surrogate coefficient times delta = rho'(delta)
from differentiating rho; quadratic case has the constant Hessian
1/(2 sigma_x^2); clamped / reflected boundaries."""

import numpy as np
import pytest

from oracle import qggmrf as qg


def _rho_prime_fd(d, sx, p, q, T, h=1e-7):
    return (qg.rho(d + h, sx, p, q, T) - qg.rho(d - h, sx, p, q, T)) / (2 * h)


@pytest.mark.parametrize('p,q', [(2.0, 1.2), (1.2, 2.0), (1.1, 2.0), (2.0, 2.0)])
def test_surrogate_coefficient_is_rho_prime_over_delta(p, q):
    sx, T = 0.7, 1.3
    d = np.array([-2.1, -0.5, -0.03, 0.02, 0.4, 1.7, 3.0])
    lhs = qg.two_b_tilde(d, sx, p, q, T) * d
    rhs = _rho_prime_fd(d, sx, p, q, T)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-6)


@pytest.mark.parametrize('p,q', [(2.0, 1.2), (1.2, 2.0)])
def test_gradient_is_loss_derivative(p, q):
    rng = np.random.default_rng(0)
    shape = (7, 6, 5)
    x = rng.standard_normal(shape)
    b = qg.b_from_nbr_wts([1.0, 0.5, 2.0])
    sx, T = 0.8, 1.0
    flat = x.reshape(-1, shape[2])
    pix = np.arange(shape[0] * shape[1])
    g, _ = qg.grad_hess_at_indices(flat, shape, pix, b, sx, p, q, T)
    h = 1e-6
    for (i, j, s) in [(3, 2, 2), (1, 1, 1), (5, 4, 3), (0, 0, 0), (6, 5, 4), (2, 0, 2)]:
        xp, xm = x.copy(), x.copy()
        xp[i, j, s] += h
        xm[i, j, s] -= h
        fd = (qg.loss(xp, b, sx, p, q, T) - qg.loss(xm, b, sx, p, q, T)) / (2 * h)
        assert g[i * shape[1] + j, s] == pytest.approx(fd, rel=1e-5, abs=1e-8)


def test_quadratic_prior_constant_hessian():
    rng = np.random.default_rng(1)
    shape = (5, 5, 4)
    x = rng.standard_normal(shape).reshape(-1, 4)
    b = qg.b_from_nbr_wts([1.0, 1.0, 1.0])
    sx = 0.6
    _, hss = qg.grad_hess_at_indices(x, shape, np.arange(25), b, sx, 2.0, 2.0, 1.0)
    np.testing.assert_allclose(hss, 1.0 / (2 * sx ** 2), rtol=1e-12)


def test_constant_volume_has_zero_gradient():
    x = np.full((6 * 6, 3), 2.5)
    b = qg.b_from_nbr_wts([1.0, 1.0, 1.0])
    g, _ = qg.grad_hess_at_indices(x, (6, 6, 3), np.arange(36), b, 1.0, 2.0, 1.2, 1.0)
    assert np.abs(g).max() == 0.0


def test_b_normalisation():
    b = qg.b_from_nbr_wts([1.0, 2.0, 3.0])
    assert sum(b) == pytest.approx(1.0)
    assert b[0] == b[1] and b[2] == b[3] and b[4] == b[5]
    assert b[4] == pytest.approx(3.0 / 12.0)
