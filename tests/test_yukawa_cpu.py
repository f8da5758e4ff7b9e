"""Host math of the GBFMM-Helmholtz (Yukawa) far field, no GPU needed.

* dage_bessel_modified_spherical against the oracle's REAL128 series
  (helmholtz3d.F90's Bessel collections, restated with __float128) and
  scipy's spherical_in / spherical_kn;
* the addition theorem the expansions rest on, exp(-k|x-y|)/|x-y| =
  (2k/pi) sum (2l+1) i_l(k|y|) k_l(k|x|) P_l(cos g), with scipy alone;
* dage_fmm_yukawa_interaction_matrix: local expansion of the field of point
  charges in a source box, built from multipoles formed here (scipy Bessel
  functions times the oracle's solid harmonics), against the direct sum of
  exp(-k r)/r at points of the target box: the error falls with lmax
  (truncation), and for a distant box it is at rounding level;
* dage_fmm_yukawa_translation_matrix: R_lm(z + d) = sum A R_l'm'(z), and the
  multipoles of shifted charges, to the re-expansion's truncation.
"""
import ctypes
import math

import numpy as np
import pytest
from scipy.special import eval_legendre, spherical_in, spherical_kn

import oracle as O
from dage_amd import _lib

_dp = ctypes.POINTER(ctypes.c_double)


def lib_bessel(lmax, x):
    i = np.zeros(lmax + 1)
    k = np.zeros(lmax + 1)
    _lib.call("dage_bessel_modified_spherical", lmax, float(x), i.ctypes.data_as(_dp),
              k.ctypes.data_as(_dp))
    return i, k


@pytest.mark.parametrize("x", [1e-3, 0.37, 2.5, 11.0, 19.99, 20.01, 47.0, 250.0])
def test_bessel_against_oracle_and_scipy(x):
    lmax = 16
    i, k = lib_bessel(lmax, x)
    oi, ok = O.bessel_i(lmax, x), O.bessel_k(lmax, x)
    ls = np.arange(lmax + 1)
    si, sk = spherical_in(ls, x), spherical_kn(ls, x)  # k_0 = (pi/2) e^-x / x, as here
    nz = oi != 0.0
    assert np.max(np.abs(i[nz] - oi[nz]) / np.abs(oi[nz])) < 1e-13
    assert np.max(np.abs(k - ok) / np.abs(ok)) < 1e-13
    big = si > 1e-250
    assert np.max(np.abs(i[big] - si[big]) / si[big]) < 1e-12
    assert np.max(np.abs(k - sk) / sk) < 1e-12


def regular(lmax, kappa, v):
    """R_lm(v) = i_l(kappa|v|) S_lm(v)/|v|^l (scipy and the oracle's harmonics)."""
    r = float(np.linalg.norm(v))
    S = O.solid_harmonics(lmax, *v)
    out = np.zeros((lmax + 1) ** 2)
    for l in range(lmax + 1):
        f = spherical_in(l, kappa * r) / r ** l
        out[l * l:(l + 1) ** 2] = S[l * l:(l + 1) ** 2] * f
    return out


def test_addition_theorem_normalisation():
    kappa = 0.9
    x = np.array([2.0, -1.5, 2.5])
    y = np.array([0.3, 0.4, -0.2])
    g = float(np.dot(x, y) / (np.linalg.norm(x) * np.linalg.norm(y)))
    rx, ry = np.linalg.norm(x), np.linalg.norm(y)
    series = sum((2 * l + 1) * spherical_in(l, kappa * ry) * spherical_kn(l, kappa * rx) *
                 eval_legendre(l, g) for l in range(40))
    exact = math.exp(-kappa * np.linalg.norm(x - y)) / np.linalg.norm(x - y)
    assert abs(2 * kappa / math.pi * series - exact) < 1e-14 * exact


def m2l_error(lmax, kappa, D, box, rng):
    N = (lmax + 1) ** 2
    rho = math.sqrt(3) / 2 * box
    T = np.zeros(N * N)
    _lib.call("dage_fmm_yukawa_interaction_matrix", lmax, kappa, np.asarray(D, float).ctypes.data_as(_dp),
              rho, T.ctypes.data_as(_dp))
    T = T.reshape((N, N), order="F")
    src = rng.uniform(-box / 2, box / 2, (12, 3))
    q = rng.standard_normal(12)
    M = sum(qq * regular(lmax, kappa, s) for qq, s in zip(q, src))
    L = T @ M
    tg = rng.uniform(-box / 2, box / 2, (8, 3))
    exact = np.array([sum(qq * math.exp(-kappa * np.linalg.norm(t - s - D)) /
                          np.linalg.norm(t - s - D) for qq, s in zip(q, src)) for t in tg])
    # the unscreened scale of the same charges: the near-field magnitude the
    # far field is added to in a GBFMM potential
    scale = max(sum(abs(qq) / np.linalg.norm(t - s - D) for qq, s in zip(q, src)) for t in tg)
    approx = np.array([L @ regular(lmax, kappa, t) for t in tg])
    err = float(np.max(np.abs(approx - exact)))
    return err / float(np.max(np.abs(exact))), err / scale


@pytest.mark.parametrize("kappa", [0.15, 1.2, 3.0])
def test_yukawa_m2l_against_direct_sum(kappa):
    """The expansion converges with lmax; relative to the screened far field
    itself it converges more slowly for larger kappa * box (i_l(kappa a) does
    not fall with l below l ~ kappa a), while that far field shrinks as
    exp(-kappa R): against the unscreened scale the lmax-12 error stays below
    1e-7 for every kappa."""
    rng = np.random.default_rng(int(kappa * 100))
    box = 3.0
    near = np.array([2 * box, box, -box])      # closest well-separated offset
    errs = [m2l_error(l, kappa, near, box, np.random.default_rng(7)) for l in (4, 8, 12)]
    print(f"kappa {kappa}: M2L error at lmax 4/8/12 (of the far field, of the unscreened "
          f"scale): {[(f'{a:.1e}', f'{b:.1e}') for a, b in errs]}")
    assert errs[2][0] < errs[1][0] < errs[0][0]
    assert errs[2][1] < 1e-7
    far = np.array([5 * box, -4 * box, 3 * box])
    e_far = m2l_error(12, kappa, far, box, rng)
    print(f"kappa {kappa}: distant offset, lmax 12: {e_far[0]:.2e} / {e_far[1]:.2e}")
    assert e_far[1] < 1e-13


@pytest.mark.parametrize("kappa,tol", [(0.15, 1e-12), (1.2, 1e-12), (3.0, 1e-6)])
def test_yukawa_shift(kappa, tol):
    """Unlike a solid harmonic, a shifted i_l(kappa r) C_lm has components of
    every degree: the lmax-10 re-expansion over a shift of 1.3 is exact to
    rounding at kappa <= 1.2 and to 2e-8 at kappa 3."""
    lmax = 10
    N = (lmax + 1) ** 2
    box = 3.0
    d = np.array([box / 4, -box / 4, box / 4])
    rho = math.sqrt(3) / 2 * box / 2
    A = np.zeros(N * N)
    _lib.call("dage_fmm_yukawa_translation_matrix", lmax, kappa, d.ctypes.data_as(_dp), rho,
              A.ctypes.data_as(_dp))
    A = A.reshape((N, N), order="F")
    rng = np.random.default_rng(3)
    for z in rng.uniform(-box / 4, box / 4, (5, 3)):
        ref = regular(lmax, kappa, z + d)
        assert np.max(np.abs(A @ regular(lmax, kappa, z) - ref)) < tol * np.max(np.abs(ref))
    # multipoles of child charges moved to the parent centre
    src = rng.uniform(-box / 4, box / 4, (10, 3))
    q = rng.standard_normal(10)
    Mc = sum(qq * regular(lmax, kappa, s) for qq, s in zip(q, src))
    Mp = sum(qq * regular(lmax, kappa, s + d) for qq, s in zip(q, src))
    assert np.max(np.abs(A @ Mc - Mp)) < tol * np.max(np.abs(Mp))


def test_yukawa_matrices_reject_bad_geometry():
    T = np.zeros(4)
    with pytest.raises(_lib.DageError):
        _lib.call("dage_fmm_yukawa_interaction_matrix", 1, 1.0,
                  np.array([1.0, 0.0, 0.0]).ctypes.data_as(_dp), 0.6, T.ctypes.data_as(_dp))
    with pytest.raises(_lib.DageError):
        lib_bessel(4, 0.0)


def test_fmm_matrices_at_lmax22():
    """lmax 22, the reference's GBFMM-Helmholtz far-field order: the 1/r
    interaction matrix equals the oracle's, and the Yukawa M2L matrix
    reproduces the direct sum to rounding of the unscreened scale."""
    L, N = 22, 23 * 23
    dR = np.array([3.0, -2.0, 1.0])
    T = np.zeros(N * N)
    _lib.call("dage_fmm_interaction_matrix", L, dR.ctypes.data_as(_dp), T.ctypes.data_as(_dp))
    To = O.fmm_interaction(L, dR)
    assert np.max(np.abs(T.reshape((N, N), order="F") - To)) < 1e-14 * np.max(np.abs(To))
    box = 3.0
    e = m2l_error(L, 1.2, np.array([2 * box, box, -box]), box, np.random.default_rng(7))
    print(f"Yukawa M2L at lmax 22 (kappa 1.2): {e[0]:.1e} of the far field, {e[1]:.1e} of "
          f"the unscreened scale")
    assert e[1] < 1e-13
