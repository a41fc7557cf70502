"""Pins the oracle's B-splines and PME moduli (CPU)."""

import math

import numpy as np
import pytest

from oracle import bspline


def _m6_pieces(u):
    """Order-6 cardinal B-spline as six explicit quintic pieces (torchff/pme_ref.py:30-59).

    An independent formulation of the same function: the oracle evaluates it by
    the Cox-de Boor recursion, this evaluates the closed piecewise polynomial.
    """
    out = np.zeros_like(u)
    a = u
    p1, p2, p3 = (u - 1) ** 5, (u - 2) ** 5, (u - 3) ** 5
    pieces = [
        a ** 5 / 120,
        a ** 5 / 120 - p1 / 20,
        a ** 5 / 120 + p2 / 8 - p1 / 20,
        a ** 5 / 120 - p3 / 6 + p2 / 8 - p1 / 20,
        a ** 5 / 24 - a ** 4 + 19 * a ** 3 / 2 - 89 * a ** 2 / 2 + 409 * a / 4 - 1829 / 20,
        -a ** 5 / 120 + a ** 4 / 4 - 3 * a ** 3 + 18 * a ** 2 - 54 * a + 324 / 5,
    ]
    for k, piece in enumerate(pieces):
        sel = (u >= k) & (u < k + 1)
        out[sel] = piece[sel]
    return out


def _dm6_pieces(u):
    """Derivative pieces of the order-6 spline (torchff/pme_ref.py:61-86)."""
    out = np.zeros_like(u)
    a = u
    q1, q2 = (u - 1) ** 4, (u - 2) ** 4
    pieces = [
        a ** 4 / 24,
        a ** 4 / 24 - q1 / 4,
        a ** 4 / 24 + 5 * q2 / 8 - q1 / 4,
        -5 * a ** 4 / 12 + 6 * a ** 3 - 63 * a ** 2 / 2 + 71 * a - 231 / 4,
        5 * a ** 4 / 24 - 4 * a ** 3 + 57 * a ** 2 / 2 - 89 * a + 409 / 4,
        -a ** 4 / 24 + a ** 3 - 9 * a ** 2 + 36 * a - 54,
    ]
    for k, piece in enumerate(pieces):
        sel = (u >= k) & (u < k + 1)
        out[sel] = piece[sel]
    return out


def test_order6_recursion_matches_reference_pieces():
    u = np.linspace(-0.5, 6.5, 7001)
    np.testing.assert_allclose(bspline.bspline_M(u, 6), _m6_pieces(u), rtol=0, atol=1e-12)
    np.testing.assert_allclose(bspline.bspline_dM(u, 6), _dm6_pieces(u), rtol=0, atol=1e-11)


@pytest.mark.parametrize("p", [3, 4, 5, 6, 7, 8])
def test_partition_of_unity(p):
    u = np.random.default_rng(p).uniform(0, 1, 200)
    tot = sum(bspline.bspline_M(u + k, p) for k in range(p))
    dtot = sum(bspline.bspline_dM(u + k, p) for k in range(p))
    np.testing.assert_allclose(tot, 1.0, atol=1e-14)
    np.testing.assert_allclose(dtot, 0.0, atol=1e-13)


def test_integer_values():
    np.testing.assert_allclose(bspline.bspline_M(np.arange(1.0, 5.0), 5) * 24, [1, 11, 11, 1], atol=1e-13)
    np.testing.assert_allclose(bspline.bspline_M(np.arange(1.0, 6.0), 6) * 120, [1, 26, 66, 26, 1], atol=1e-12)


def test_derivative_by_finite_difference():
    u = np.random.default_rng(0).uniform(0.01, 4.99, 100)
    h = 1e-6
    fd = (bspline.bspline_M(u + h, 5) - bspline.bspline_M(u - h, 5)) / (2 * h)
    np.testing.assert_allclose(bspline.bspline_dM(u, 5), fd, atol=1e-8)
    fd2 = (bspline.bspline_dM(u + h, 5) - bspline.bspline_dM(u - h, 5)) / (2 * h)
    np.testing.assert_allclose(bspline.bspline_d2M(u, 5), fd2, atol=1e-6)


@pytest.mark.parametrize("K", [10, 24, 33])
def test_order6_moduli_match_reference_cosine_form(K):
    # torchff/bsplines.py:20-54: b(k) = sum_{m=-3}^{2} c(m+3) cos(2 pi m k / K), c = (0,1,26,66,26,1)/120
    c = {1: 1 / 120, 2: 26 / 120, 3: 66 / 120, 4: 26 / 120, 5: 1 / 120}
    k = np.arange(K)
    b = sum(c.get(m + 3, 0.0) * np.cos(2 * math.pi * m * k / K) for m in range(-3, 3))
    np.testing.assert_allclose(bspline.theta_sq(K, 6), b * b, rtol=1e-12)


def test_order5_moduli_nyquist_fixed():
    th = bspline.theta_sq(96, 5)
    assert np.all(np.isfinite(th)) and np.all(th > 0)
    # direct evaluation away from Nyquist: |sum_j M5(j+1) e^{2 pi i k j/K}|^2
    k = 7
    w = np.array([1, 11, 11, 1]) / 24
    z = sum(w[j] * np.exp(2j * math.pi * k * j / 96) for j in range(4))
    assert abs(th[k] - abs(z) ** 2) < 1e-14


@pytest.mark.parametrize("p", [4, 5, 6])
def test_anchor_covers_nearest_points(p):
    s = np.random.default_rng(1).uniform(-5, 50, 1000)
    m0, u0 = bspline.anchor(s, p)
    st = bspline.stencil(p)
    u = u0[:, None] + st[None, :]
    assert np.all(u >= 0) and np.all(u <= p)
    # weights sum to one: the stencil holds the full support
    np.testing.assert_allclose(bspline.bspline_M(u, p).sum(1), 1.0, atol=1e-13)
    # weight of grid point m0 + j equals M_p(g - s + p/2)
    g = m0[:, None] + st[None, :]
    np.testing.assert_allclose(u, g - s[:, None] + p / 2, atol=1e-12)
