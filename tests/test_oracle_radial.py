"""Oracle of the bubbles' radial transforms (row f2), pinned on the CPU:

* the LIP cumulative integrals (lip.F90:364-485, evaluators.F90:1263-1320)
  integrate piecewise polynomials of degree <= nlip-1 exactly on the
  non-uniform radial grids (grid.F90:569-603);
* the modified spherical Bessel functions (helmholtz3d.F90:879-1210, REAL128
  series and recursions) equal scipy.special.spherical_in/_kn to 1e-15;
* the radial Poisson solution (coulomb.F90:222-336) of r^l exp(-Z r) equals
  its closed form (incomplete gamma functions) to the LIP integration error;
* the radial Helmholtz convolution (helmholtz3d.F90:519-657): for l = 0 and a
  normalised Gaussian it is the Yukawa potential of the Gaussian (closed
  form, itself pinned by its k -> 0 limit), and as k -> 0 it tends to the
  Coulomb transform for every l (8k i_l(kr<) k_l(kr>) -> 4pi/(2l+1)
  r<^l / r>^(l+1)), which pins its normalisation.
"""
import math

import numpy as np
import pytest
from scipy.special import gamma, gammainc, spherical_in, spherical_kn

import oracle as O


def radial(Z=3.0, ncell=33, nlip=7, r_max=20.0):
    st = O.radial_steps(Z, ncell, nlip, r_max)
    return st, O.Axis(0.0, st, nlip).points


@pytest.mark.parametrize("nlip", [5, 7])
def test_inout_integrals_exact_for_polynomials(nlip):
    st, r = radial(2.0, 20, nlip)
    for deg in range(nlip):
        f = r ** deg
        inw, outw = O.integrate_inout(st, nlip, f)
        ref_out = r ** (deg + 1) / (deg + 1)
        ref_in = (r[-1] ** (deg + 1) - r ** (deg + 1)) / (deg + 1)
        scale = r[-1] ** (deg + 1)
        assert np.max(np.abs(outw - ref_out)) < 1e-13 * scale, deg
        assert np.max(np.abs(inw - ref_in)) < 1e-13 * scale, deg
    # degree nlip is not exact (the pin would catch a wrong-degree basis)
    inw, outw = O.integrate_inout(st, nlip, r ** nlip)
    assert np.max(np.abs(outw - r ** (nlip + 1) / (nlip + 1))) > 1e-10 * r[-1] ** (nlip + 1)


def test_lip_inout_integral_tables_add_up():
    for nlip in (3, 5, 7):
        ii, io = O.lip_inout_integrals(nlip)
        full = O.lip_integrals(nlip)
        # int_first^x_j + int_x_j^last = int over the cell, node by node
        for j in range(nlip - 2):
            assert np.allclose(io[:, j] + ii[:, j + 1], full, atol=1e-14)
        assert np.allclose(io[:, -1], full, atol=1e-14)
        assert np.allclose(ii[:, 0], full, atol=1e-14)


@pytest.mark.parametrize("x", [1e-6, 1e-3, 0.3, 2.5, 17.0, 60.0])
def test_bessels_match_scipy(x):
    lmax = 10
    bi, bk = O.bessel_i(lmax, x), O.bessel_k(lmax, x)
    ri = np.array([spherical_in(l, x) for l in range(lmax + 1)])
    rk = np.array([spherical_kn(l, x) for l in range(lmax + 1)])
    assert np.max(np.abs(bi / ri - 1.0)) < 2e-15 * (1 if x > 1e-3 else 10)
    assert np.max(np.abs(bk / rk - 1.0)) < 2e-15 * (1 if x > 1e-3 else 10)
    assert O.bessel_i(3, 0.0).tolist() == [1.0, 0.0, 0.0, 0.0]
    assert O.bessel_k(3, 0.0).tolist() == [0.0, 0.0, 0.0, 0.0]


def slater_bubble(lmax, Z, r, c=None):
    nlm = (lmax + 1) ** 2
    bf = np.zeros((len(r), nlm), order="F")
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            bf[:, l * l + l + m] = (1.0 if c is None else c[l * l + l + m]) * r ** l * np.exp(-Z * r)
    return bf


def test_coulomb_radial_matches_closed_form():
    Z, lmax = 2.0, 4
    st, r = radial(Z)
    bf = slater_bubble(lmax, Z, r)
    V = O.bubble_coulomb(lmax, 0, st, 7, bf)
    rr = np.where(r > 0, r, 1.0)
    for l in range(lmax + 1):
        n = 2 * l + 2
        inner = gammainc(n + 1, Z * r) * gamma(n + 1) / Z ** (n + 1)
        outer = (1 + Z * r) * np.exp(-Z * r) / Z ** 2
        ref = 4 * math.pi / (2 * l + 1) * (np.where(r > 0, rr ** (-(l + 1)), 0.0) * inner
                                           + r ** l * outer)
        for m in range(-l, l + 1):
            err = np.max(np.abs(V[:, l * l + l + m] - ref)) / np.max(np.abs(ref))
            assert err < 3e-6, (l, m, err)  # LIP integration error on 33 cells (r^-(l+1) near 0)
    # the error is the radial integration's: it falls ~64x per halving of the cells
    errs = []
    for ncell in (33, 66):
        st2, r2 = radial(Z, ncell)
        V2 = O.bubble_coulomb(1, 0, st2, 7, slater_bubble(1, Z, r2))[:, 1]
        ref = 4 * math.pi / 3 * (np.where(r2 > 0, 1 / np.where(r2 > 0, r2, 1) ** 2, 0.0)
                                 * gammainc(5, Z * r2) * gamma(5) / Z ** 5
                                 + r2 * (1 + Z * r2) * np.exp(-Z * r2) / Z ** 2)
        errs.append(np.max(np.abs(V2 - ref)) / np.max(np.abs(ref)))
    assert errs[1] < errs[0] / 20
    # input stored with k = -1 (r^-1 f, as nuclear potentials are): the
    # transform multiplies by r^(l+2+k), so r^-1 (r f) gives the k = 0 result
    Vk = O.bubble_coulomb(lmax, -1, st, 7, bf * r[:, None])
    assert np.max(np.abs(Vk[1:] - V[1:])) < 1e-13 * np.max(np.abs(V))


def test_helmholtz_radial_l0_gaussian_is_yukawa():
    alpha, E = 1.3, -0.8
    k = math.sqrt(-2 * E)
    st, r = radial(1.0, 60)
    f = (alpha / math.pi) ** 1.5 * np.exp(-alpha * r ** 2)
    V = O.bubble_helmholtz(0, 0, st, 7, E, f[:, None])[:, 0]
    ref = np.array([O.yukawa_gauss_potential([1.0], [alpha], [[0.0, 0.0, 0.0]], k, [x], [0.0],
                                             [0.0])[0, 0, 0] for x in r])
    err = np.max(np.abs(V[1:] - ref[1:])) / np.max(np.abs(ref))
    err0 = abs(V[0] - ref[0]) / np.max(np.abs(ref))
    print(f"Helmholtz l=0 Gaussian vs Yukawa closed form: {err:.2e} (r = 0: {err0:.2e})")
    assert err < 1e-7
    # at r = 0 the integrand's first value is extrapolated (third order)
    assert err0 < 1e-6


def test_helmholtz_radial_tends_to_coulomb():
    Z, lmax = 2.5, 4
    st, r = radial(Z)
    rng = np.random.default_rng(4)
    bf = slater_bubble(lmax, Z, r, rng.uniform(-1, 1, (lmax + 1) ** 2))
    C = O.bubble_coulomb(lmax, 0, st, 7, bf)
    H = O.bubble_helmholtz(lmax, 0, st, 7, -5e-15, bf)
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            lm = l * l + l + m
            err = np.max(np.abs(H[:, lm] - C[:, lm])) / np.max(np.abs(C[:, lm]))
            assert err < 1e-6, (l, m, err)
    # and a finite energy really screens: |V| drops below the Coulomb tail
    H1 = O.bubble_helmholtz(lmax, 0, st, 7, -1.0, bf)
    assert abs(H1[-10, 0]) < 0.1 * abs(C[-10, 0])


def test_exp_weighted_inout_integrals():
    """The exponentially fitted scheme's cell weights: at mu = 0 the plain
    LIP in/out integrals, and for mu > 0 exact for polynomials of degree <= 6
    against exp(-mu |t - t_j|) (scipy quad)."""
    from scipy.integrate import quad
    ii0, io0 = O.exp_inout(7, 0.0)
    ii, io = O.lip_inout_integrals(7)
    assert np.max(np.abs(ii0 - ii)) < 1e-13 and np.max(np.abs(io0 - io)) < 1e-13
    nodes = np.arange(7.0)
    for mu in (0.3, 4.0, 17.0):
        wi, wo = O.exp_inout(7, mu)
        for q in range(7):
            for j in range(6):
                ref_in = quad(lambda t: math.exp(-mu * (t - j)) * t ** q, j, 6, epsabs=0,
                              epsrel=1e-13, limit=200)[0]
                ref_out = quad(lambda t: math.exp(-mu * (j + 1 - t)) * t ** q, 0, j + 1,
                               epsabs=0, epsrel=1e-13, limit=200)[0]
                got_in, got_out = np.dot(wi[:, j], nodes ** q), np.dot(wo[:, j], nodes ** q)
                # relative to the sum's terms (nodes^6 reaches 6^6)
                s_in = np.dot(np.abs(wi[:, j]), nodes ** q)
                s_out = np.dot(np.abs(wo[:, j]), nodes ** q)
                assert abs(got_in - ref_in) < 1e-13 * s_in, (mu, q, j, got_in, ref_in)
                assert abs(got_out - ref_out) < 1e-13 * s_out, (mu, q, j, got_out, ref_out)


@pytest.mark.parametrize("E", [-2.0, -20.0])
def test_helmholtz_radial_deep_energies_fitted(E):
    """kappa L >> 1 on the outer cells of a Z = 8 radial grid (33 cells to
    20 bohr, the last 12.9 bohr long): the reference's scheme returns errors
    of order exp(kappa L) (7.6 of max|V| at E = -2, 1e21 at E = -20 for a
    diffuse Gaussian); the exponentially fitted form stays at the
    interpolation error of the smooth parts.  l = 0: the Yukawa potential of
    a normalised Gaussian (closed form); l = 2: the defining integrals with
    scipy's Bessel functions and quad."""
    from scipy.integrate import quad
    k = math.sqrt(-2 * E)
    st, r = radial(8.0, 33)
    assert 6 * st.max() * k > 4.0
    for alpha in (0.3, 3.0):
        f = (alpha / math.pi) ** 1.5 * np.exp(-alpha * r ** 2)
        V = O.bubble_helmholtz(0, 0, st, 7, E, f[:, None])[:, 0]
        ref = np.array([O.yukawa_gauss_potential([1.0], [alpha], [[0.0, 0.0, 0.0]], k, [x], [0.0],
                                                 [0.0])[0, 0, 0] for x in r])
        err = np.max(np.abs(V - ref)) / np.max(np.abs(ref))
        print(f"E {E} alpha {alpha}: l = 0 fitted vs closed form {err:.2e}")
        assert err < 1e-6
    # l = 2, f = r^2 exp(-r^2)
    lmax, alpha = 2, 1.0
    bf = np.zeros((len(r), 9), order="F")
    bf[:, 6] = r ** 2 * np.exp(-alpha * r ** 2)
    V = O.bubble_helmholtz(lmax, 0, st, 7, E, bf)[:, 6]
    fl = lambda s: s ** 2 * math.exp(-alpha * s * s)
    for i in range(5, len(r) - 1, 23):
        x = r[i]
        a = quad(lambda s: s * s * spherical_in(2, k * s) * fl(s), 0, x, epsrel=1e-12)[0]
        b = quad(lambda s: s * s * spherical_kn(2, k * s) * fl(s), x, max(x, 12.0) + 8,
                 epsrel=1e-12, limit=200)[0]
        ref = 8 * k * (spherical_kn(2, k * x) * a + spherical_in(2, k * x) * b)
        assert abs(V[i] - ref) < 1e-6 * np.max(np.abs(V)), (E, x, V[i], ref)
