"""The Yukawa-Gaussian closed form that checks Helmholtz applies, pinned by
mathematics: the k -> 0 limit is the Coulomb erf potential, and radial
quadrature of the Yukawa kernel reproduces it, r = 0 included."""
import numpy as np
from scipy.integrate import quad

import oracle as O


def test_k_to_zero_is_coulomb():
    x = np.linspace(-3, 3, 7)
    a = O.yukawa_gauss_potential([1.0, 0.5], [1.3, 4.0], [[0.1, 0.2, -0.3], [1, 0, 0]], 1e-10,
                                 x, x, x)
    b = O.gauss_potential([1.0, 0.5], [1.3, 4.0], [[0.1, 0.2, -0.3], [1, 0, 0]], x, x, x)
    assert np.max(np.abs(a - b)) < 1e-8


def test_radial_quadrature():
    al, k = 1.3, 0.9
    rho = lambda s: (al / np.pi) ** 1.5 * np.exp(-al * s * s)
    phi0 = 4 * np.pi * quad(lambda s: rho(s) * s * np.exp(-k * s), 0, 20, limit=200)[0]
    c0 = O.yukawa_gauss_potential([1.0], [al], [[0, 0, 0]], k, [0.0], [0.0], [0.0])[0, 0, 0]
    assert abs(phi0 - c0) < 1e-12
    for r in (0.3, 1.0, 2.5, 6.0):
        v = quad(lambda s: rho(s) * s * (np.exp(-k * abs(r - s)) - np.exp(-k * (r + s))), 0, 20,
                 limit=200, points=[r])[0] * 2 * np.pi / (k * r)
        c = O.yukawa_gauss_potential([1.0], [al], [[0, 0, 0]], k, [r], [0.0], [0.0])[0, 0, 0]
        assert abs(v - c) < 1e-11 * max(1.0, abs(v)), r
