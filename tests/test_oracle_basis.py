"""Pins the oracle's barycentric basis (bli_impl.hpp:8-77) with closed forms.

- Lagrange cardinality: L_j(x_k) = delta_jk at the Chebyshev-Lobatto nodes.
- Identity with the product form of the Lagrange polynomial
  L_i(x) = prod_{j != i} (x - c_j) / (c_i - c_j): fixes the barycentric
  weights (1/2, (-1)^i, (-1)^n/2; Berrut & Trefethen, SIAM Rev. 46 (2004)),
  so a wrong weight, sign or node fails.
- Partition of unity and reproduction of polynomials of degree <= n.
"""
import numpy as np
import pytest

from oracle import pyoracle as O


def nodes(lo, hi, n):
    return np.cos(np.pi * np.arange(n + 1) / n) * 0.5 * (hi - lo) + 0.5 * (hi + lo)


@pytest.mark.parametrize("n", range(1, 11))
def test_cardinal_at_nodes(n):
    lo, hi = -0.3, 1.7
    c = nodes(lo, hi, n)
    for k in range(n + 1):
        b = O.interp_vals_bli(c[k], lo, hi, n)
        e = np.zeros(n + 1)
        e[k] = 1.0
        assert np.array_equal(b, e)


@pytest.mark.parametrize("n", range(1, 11))
def test_equals_product_form(n):
    rng = np.random.default_rng(n)
    lo, hi = -2.0, 0.5
    c = nodes(lo, hi, n)
    for x in rng.uniform(lo, hi, 25):
        b = O.interp_vals_bli(x, lo, hi, n)
        ref = np.array([np.prod([(x - c[j]) / (c[i] - c[j]) for j in range(n + 1) if j != i])
                        for i in range(n + 1)])
        assert np.allclose(b, ref, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("n", range(1, 11))
def test_partition_of_unity_and_polynomial_reproduction(n):
    rng = np.random.default_rng(100 + n)
    lo, hi = 0.25, 0.75
    c = nodes(lo, hi, n)
    coef = rng.standard_normal(n + 1)
    p = np.polynomial.polynomial.Polynomial(coef)
    for x in rng.uniform(lo - 0.1, hi + 0.1, 20):
        b = O.interp_vals_bli(x, lo, hi, n)
        assert abs(b.sum() - 1.0) < 1e-13
        assert abs(b @ p(c) - p(x)) < 1e-10 * max(1.0, abs(p(x)))


@pytest.mark.parametrize("n", [1, 4, 8, 10])
def test_shifted_points_are_the_nodes(n):
    lo, hi = -1.25, 3.0
    s = O.bli_points_shift(lo, hi, n)
    # same nodes as the basis up to the rounding of cos(pi/n*i) vs cos(pi*i/n) (bli_impl.hpp:25 vs :74)
    assert np.allclose(s, nodes(lo, hi, n), rtol=0, atol=4e-16 * 4)
    assert s[0] == pytest.approx(hi, abs=1e-15) and s[-1] == pytest.approx(lo, abs=1e-15)


def test_coincidence_tolerance_is_absolute_epsilon():
    # |x - c_i| < DBL_EPSILON selects the unit vector (bli_impl.hpp:36-45)
    n, lo, hi = 6, -1.0, 1.0
    b = O.interp_vals_bli(1.0 - 1e-17, lo, hi, n)
    assert b[0] == 1.0 and b[1:].sum() == 0.0
    b = O.interp_vals_bli(1.0 - 1e-12, lo, hi, n)
    assert b[0] != 1.0 and abs(b[0] - 1.0) < 1e-10
