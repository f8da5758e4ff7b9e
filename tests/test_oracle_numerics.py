"""Pins of the oracle's building blocks (CPU only).

gauleg / GaussQuad   (gauss.F90:81-122, 307-384): numpy's Gauss-Legendre,
                     exactness on polynomials, the log-map identity.
LIP basis            (lip.F90:118-415): cardinality, polynomial reproduction,
                     closed-form integrals.
Grid1D               (grid.F90:446-567, 1398-1424): point placement, weights
                     integrate polynomials exactly.
expo_0..6            (exponentials.F90:695-1297): the reference's in-source
                     debug values (t=2, p=1, a=0, b=inf) and brute-force
                     quadrature over both branches (t<0.1 Gauss, t>=0.1 erf).
"""
import math

import numpy as np
import pytest
from scipy.integrate import quad

import oracle as O


def test_gauleg_matches_numpy():
    for n in (1, 2, 5, 10, 20, 25, 40):
        x, w = O.gauleg(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.max(np.abs(x - xr)) < 1e-15
        # the reference takes the weight from the derivative evaluated before
        # its last Newton step, which costs a few 1e-15 (absolute) against numpy
        assert np.max(np.abs(w - wr)) < 1e-14


def test_gauss_quad_linear_and_log_parts():
    tp, tw = O.gauss_quad(20, 16, 0.0, 2.0, 500.0)
    # linear part integrates t^k on [0,2] exactly up to degree 2n-1
    for k in (0, 3, 11, 39):
        assert abs(np.dot(tw[:20], tp[:20] ** k) - 2.0 ** (k + 1) / (k + 1)) < 1e-12 * 2.0 ** (k + 1)
    # log part: int_2^500 dt/t = log(250) (the map makes 1/t a constant integrand)
    assert abs(np.dot(tw[20:], 1.0 / tp[20:]) - math.log(250.0)) < 1e-13
    # 2/sqrt(pi) int_0^inf exp(-t^2 r^2) dt = 1/r, up to the coda pi/tlog^2-sized tail
    w = O.coulomb_weights(tw)
    for r in (0.5, 1.0, 3.0):
        approx = np.dot(w, np.exp(-(tp * r) ** 2))
        tail = 2 / math.sqrt(math.pi) * quad(lambda t: math.exp(-(t * r) ** 2), 500.0, np.inf)[0]
        assert abs(approx + tail - 1.0 / r) < 2e-4 / r


@pytest.mark.parametrize("nlip,gt", [(3, 1), (5, 1), (7, 1), (6, 2), (7, 2)])
def test_lip_cardinality_and_reproduction(nlip, gt):
    c = O.lip_coeffs(nlip, gt)
    first, last = O.lip_first_last(nlip, gt)
    if gt == 1:
        nodes = np.arange(nlip) - (nlip + 1) // 2 + 1
    else:
        nodes = O.gauss_lobatto_grid(nlip, first, last)
    assert nodes[0] == first and nodes[-1] == last
    V = np.vander(nodes, nlip)  # decreasing powers, as the reference stores them
    assert np.max(np.abs(c @ V.T - np.eye(nlip))) < 1e-12
    # integrals of each basis polynomial over [first, last]
    ints = O.lip_integrals(nlip, gt)
    pw = np.arange(nlip - 1, -1, -1)
    exact = c @ ((last ** (pw + 1) - first ** (pw + 1)) / (pw + 1))
    assert np.max(np.abs(ints - exact)) < 1e-12
    assert abs(ints.sum() - (last - first)) < 1e-12


def test_lip_coefficients_small_case():
    # 3-point basis on (-1, 0, 1): x(x-1)/2, 1-x^2, x(x+1)/2 (decreasing powers)
    c = O.lip_coeffs(3, 1)
    assert np.allclose(c, [[0.5, -0.5, 0.0], [-1.0, 0.0, 1.0], [0.5, 0.5, 0.0]], atol=0, rtol=0)


def test_grid_points_and_weights():
    ax = O.Axis.ncell_fixed(-8.0, 8.0, 7, 16)
    assert ax.n == 97
    assert np.max(np.abs(ax.points - np.linspace(-8, 8, 97))) < 1e-13
    w = ax.ints()
    # weights integrate polynomials up to degree nlip-1 exactly over the axis
    for k in (0, 2, 6):
        exact = (8.0 ** (k + 1) - (-8.0) ** (k + 1)) / (k + 1)
        assert abs(np.dot(w, ax.points ** k) - exact) < 1e-10 * max(1.0, abs(exact))
    nc, step = O.Axis.equidistant(-12.0, 12.0, 7, 0.1).ncell, None
    assert nc == 40


EXPO_DEBUG = {  # exponentials.F90 debug comments: t=2, p=1, a=0, b=inf
    0: lambda sp, e4, E: 0.25 * sp + 0.25 * E * sp,
    1: lambda sp, e4, E: 0.25 * sp + 0.125 * e4 + 0.25 * E * sp,
    2: lambda sp, e4, E: 9 / 32 * sp + 0.125 * e4 + 9 / 32 * E * sp,
    3: lambda sp, e4, E: 11 / 32 * sp + 5 / 32 * e4 + 11 / 32 * E * sp,
    4: lambda sp, e4, E: 115 / 256 * sp + 13 / 64 * e4 + 115 / 256 * E * sp,
    5: lambda sp, e4, E: 159 / 256 * sp + 9 / 32 * e4 + 159 / 256 * E * sp,
    6: lambda sp, e4, E: 1847 / 2048 * sp + 209 / 512 * e4 + 1847 / 2048 * E * sp,
}


@pytest.mark.parametrize("n", range(7))
def test_expo_in_source_debug_values(n):
    # b = 30 stands for infinity: exp(-4*29^2) underflows to 0 and erf saturates
    v = O.expo(n, 0.0, 30.0, 2.0, 1.0)
    ref = EXPO_DEBUG[n](math.sqrt(math.pi), math.exp(-4.0), math.erf(2.0))
    assert abs(v - ref) <= 4e-16 * ref


@pytest.mark.parametrize("t,p", [(0.1, 0.3), (0.35, -1.7), (1.0, 2.5), (2.0, 0.0), (5.0, 3.4),
                                 (0.02, 1.0), (0.07, -5.0), (0.0999, 2.0)])
def test_integrals_vs_quadrature(t, p):
    a, b = -3.0, 3.0
    vals = O.integrals(a, b, t, p, 7)
    for i in range(7):
        ref = quad(lambda x: x ** i * math.exp(-(t * (x - p)) ** 2), a, b, epsabs=1e-15,
                   epsrel=1e-13, limit=400)[0]
        # values(n-i) holds the i-th moment (exponentials.F90:647-666); the
        # closed forms cancel to ~4e-12 relative at the t = 0.1 branch edge
        assert abs(vals[6 - i] - ref) <= 2e-11 * max(1.0, abs(ref)), (i, vals[6 - i], ref)


def test_gauss_branch_is_used_below_t_01():
    # t < 0.1 uses the 48-point rule; its value differs from the closed form by
    # rounding only, and the switch happens exactly at t = 0.1
    t, p = 0.0999999, 0.7
    g = O.gauss_expo(3, -3.0, 3.0, t, p)
    c = O.expo(3, -3.0, 3.0, t, p)
    assert O.integrals(-3.0, 3.0, t, p, 7)[3] == g
    assert abs(g - c) < 1e-9 * abs(c)
    assert O.integrals(-3.0, 3.0, 0.1, p, 7)[3] == O.expo(3, -3.0, 3.0, 0.1, p)
