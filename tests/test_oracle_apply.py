"""Pins of the oracle's operator matrices and tensor apply (CPU only).

End to end, the oracle reproduces the reference's own golden self-interaction
energies (tests/gauss_test.F90 TEST 1-3, committed under tests/golden/) --
grid, LIP basis, quadrature, both integral branches, the apply loop
(function3d.F90:4301-4314), the coda (:4327-4332) and the LIP-weighted dot
product all enter those numbers.  The remaining tests pin single pieces by
mathematics: partition of unity of the matrices, brute-force tensor sums,
analytic Coulomb and Yukawa potentials of Gaussians.
"""
import math

import numpy as np
import pytest
from scipy.special import erf, erfcx

import oracle as O
from _helpers import gauss_generator_axes, load_golden, random_gaussians, rel_maxdiff


@pytest.mark.parametrize("case", load_golden("gauss_test_selfint.json")["cases"],
                         ids=lambda c: c["name"])
def test_reference_selfint_goldens(case):
    axes = gauss_generator_axes(case["q"], case["alpha"], case["centers"], case["step"],
                                case["nlip"], case["thres"])
    tp, tw = O.gauss_quad(10, 8, 0.0, 2.0, 500.0)
    fx, fy, fz = O.coulomb3d_matrices(axes, axes, tp)
    x, y, z = (a.points for a in axes)
    rho = O.gauss_density(case["q"], case["alpha"], case["centers"], x, y, z)
    V = O.transform_cube(fx, fy, fz, O.coulomb_weights(tw), O.coda(500.0), rho)
    e = O.cube_integral(axes[0].ints(), axes[1].ints(), axes[2].ints(), rho, V)
    ref = case["selfint_numerical"]
    assert abs(e - ref) <= 5e-15 * abs(ref), (e, ref)


def test_matrix_rows_integrate_the_gaussian_exactly():
    # sum_i P_i = 1 on every cell, so each row of F sums to the integral of
    # exp(-t^2 (x_o - x')^2) over the input axis; x-weighted rows give the
    # first moment.  Covers shared cell end points, h scaling, both branches.
    ax_in = O.Axis.ncell_fixed(-4.0, 4.0, 7, 8)
    ax_out = O.Axis.ncell_fixed(-5.0, 3.0, 5, 10)
    tp = np.array([0.05, 0.3, 0.7, 2.0, 9.0, 40.0])
    F = O.exponentials(ax_in, ax_out, tp, transpose=False)  # (out, in, T)
    a, b = ax_in.points[0], ax_in.points[-1]
    for it, t in enumerate(tp):
        for io, xo in enumerate(ax_out.points):
            m0 = math.sqrt(math.pi) / (2 * t) * (erf(t * (b - xo)) - erf(t * (a - xo)))
            m1 = xo * m0 + (math.exp(-(t * (a - xo)) ** 2) - math.exp(-(t * (b - xo)) ** 2)) / (2 * t * t)
            assert abs(F[io, :, it].sum() - m0) < 1e-11 * max(1.0, m0)
            assert abs(F[io, :, it] @ ax_in.points - m1) < 1e-10 * max(1.0, abs(m1))


def test_transposed_storage_of_y_and_z_matrices():
    ax = O.Axis.ncell_fixed(-3.0, 3.0, 5, 4)
    tp = np.array([0.2, 1.5, 30.0])
    fx = O.exponentials(ax, ax, tp, transpose=False)
    fy = O.exponentials(ax, ax, tp, transpose=True)
    assert np.array_equal(fy, fx.transpose(1, 0, 2))


def test_transform_cube_matches_brute_force_sum():
    rng = np.random.default_rng(7)
    ax = [O.Axis.ncell_fixed(-2.0, 2.0, 5, 2), O.Axis.ncell_fixed(-2.0, 1.0, 5, 3),
          O.Axis.ncell_fixed(-1.0, 2.0, 5, 1)]
    tp, tw = O.gauss_quad(3, 2, 0.0, 2.0, 50.0)
    fx, fy, fz = O.coulomb3d_matrices(ax, ax, tp)
    w = O.coulomb_weights(tw)
    cube = np.asfortranarray(rng.standard_normal([a.n for a in ax]))
    out = O.transform_cube(fx, fy, fz, w, 0.0, cube)
    ref = np.einsum("t,aet,fbt,gct,efg->abc", w, fx, fy, fz, cube)
    assert rel_maxdiff(out, ref) < 1e-14
    coda = O.coda(50.0)
    out2 = O.transform_cube(fx, fy, fz, w, coda, cube)
    assert np.max(np.abs(out2 - out - coda * cube)) < 1e-15 * np.max(np.abs(out))


def test_coulomb_potential_of_gaussians_vs_erf():
    # 49^3 cube over [-8,8]^3, T=40 (20 linear + 20 log), smooth densities
    rng = np.random.default_rng(11)
    ax = O.Axis.ncell_fixed(-8.0, 8.0, 7, 8)
    q, alpha, c = random_gaussians(rng, 3, (0.3, 0.8))
    tp, tw = O.gauss_quad(20, 20, 0.0, 2.0, 500.0)
    fx, fy, fz = O.coulomb3d_matrices([ax] * 3, [ax] * 3, tp)
    rho = O.gauss_density(q, alpha, c, ax.points, ax.points, ax.points)
    V = O.transform_cube(fx, fy, fz, O.coulomb_weights(tw), O.coda(500.0), rho)
    Va = O.gauss_potential(q, alpha, c, ax.points, ax.points, ax.points)
    assert rel_maxdiff(V, Va) < 2e-4


def yukawa_gaussian(alpha, k, r):
    """Potential of a normalised Gaussian under exp(-k r)/r (closed form with
    erfcx for stability); reduces to erf(sqrt(a) r)/r as k -> 0."""
    sa = math.sqrt(alpha)
    r = np.maximum(r, 1e-12)
    u = k / (2 * sa)
    # e^{k^2/4a} [e^{-kr} erfc(u - sa r) - e^{kr} erfc(u + sa r)] / (2r)
    x1, x2 = u - sa * r, u + sa * r
    t1 = np.exp(-(sa * r) ** 2) * erfcx(x1)
    t1 = np.where(x1 < 0, np.exp(u * u - k * r) * (2 - erfcx(-x1) * np.exp(-x1 * x1)), t1)
    t2 = np.exp(-(sa * r) ** 2) * erfcx(x2)
    return (t1 - t2) / (2 * r)


def test_helmholtz_weights_and_potential_vs_yukawa():
    tp, tw = O.gauss_quad(20, 20, 0.0, 2.0, 500.0)
    E = -0.8  # k^2 = 1.6
    w = O.helmholtz_weights(tp, tw, E)
    assert np.allclose(w, 2 / math.sqrt(math.pi) * tw * np.exp(2 * E / (4 * tp ** 2)), rtol=1e-15)
    assert np.array_equal(O.helmholtz_weights(tp, tw, -0.001), O.coulomb_weights(tw))
    ax = O.Axis.ncell_fixed(-8.0, 8.0, 7, 8)
    fx, fy, fz = O.coulomb3d_matrices([ax] * 3, [ax] * 3, tp)
    alpha, cen = 0.6, np.array([0.3, -0.2, 0.1])
    rho = O.gauss_density([1.0], [alpha], cen, ax.points, ax.points, ax.points)
    V = O.transform_cube(fx, fy, fz, w, 0.0, rho)
    X, Y, Z = np.meshgrid(ax.points, ax.points, ax.points, indexing="ij")
    r = np.sqrt((X - cen[0]) ** 2 + (Y - cen[1]) ** 2 + (Z - cen[2]) ** 2)
    Va = yukawa_gaussian(alpha, math.sqrt(-2 * E), r)
    assert rel_maxdiff(V, Va) < 1e-3
    # k -> 0 limit of the closed form is the Coulomb potential
    assert np.max(np.abs(yukawa_gaussian(alpha, 1e-9, r) - erf(math.sqrt(alpha) * r) / r)) < 1e-7
