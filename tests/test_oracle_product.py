"""Oracle of the bubble product (row f1), pinned on the CPU.

* YProduct coefficients (harmonic.F90:884-1000): for random directions u,
  sum_LM c(lm1, lm2 -> LM) Y_LM(u) = Y_lm1(u) Y_lm2(u) with the pinned Racah
  solid harmonics, for every pair up to l 4 x 4 (1e-13);
* selection rules: c vanishes unless |l1-l2| <= L <= l1+l2 with L + l1 + l2
  even; Y_00 is the identity;
* the product of two bubble sets evaluated at points whose distance from the
  centre is a radial grid node (no radial interpolation) equals the product
  of the two sets' values there (1e-13), through the oracle's
  Bubbles_eval_3dgrid.
"""
import numpy as np

import oracle as O


def test_yproduct_reproduces_pointwise_products():
    rng = np.random.default_rng(0)
    l1, l2 = 4, 3
    c = O.yproduct(l1, l2)
    for _ in range(20):
        u = rng.standard_normal(3)
        u /= np.linalg.norm(u)
        Y = O.solid_harmonics(l1 + l2, *u)
        lhs = np.einsum("ijo,o->ij", c, Y)
        rhs = np.outer(Y[:(l1 + 1) ** 2], Y[:(l2 + 1) ** 2])
        assert np.max(np.abs(lhs - rhs)) < 1e-13


def test_yproduct_selection_rules():
    c = O.yproduct(3, 3)
    for i in range(16):
        li = int(np.sqrt(i))
        for j in range(16):
            lj = int(np.sqrt(j))
            for o in range(49):
                lo = int(np.sqrt(o))
                if abs(li - lj) > lo or lo > li + lj or (li + lj + lo) % 2:
                    assert c[i, j, o] == 0.0
    # Y_00 = 1 times Y_j is Y_j
    assert np.allclose(c[0, :, :16], np.eye(16), atol=1e-14)


def test_bubble_product_on_radial_nodes():
    rng = np.random.default_rng(3)
    l1, l2 = 2, 3
    st = O.radial_steps(2.0, 20, 7, 15.0)
    r = O.Axis(0.0, st, 7).points
    bf1 = np.asfortranarray(rng.standard_normal((len(r), (l1 + 1) ** 2)))
    bf2 = np.asfortranarray(rng.standard_normal((len(r), (l2 + 1) ** 2)))
    prod = O.bubble_product(l1, bf1, l2, bf2)
    centre = np.array([0.3, -0.2, 0.1])
    for node in (5, 37, 88):
        u = rng.standard_normal(3)
        u /= np.linalg.norm(u)
        p = centre + r[node] * u
        vals = []
        for lmax, bf in ((l1, bf1), (l2, bf2), (l1 + l2, prod)):
            bs = O.BubbleSet(centre[None, :], [st], [bf], lmax, 0, 7)
            vals.append(O.bubbles_eval_grid(bs, [p[0]], [p[1]], [p[2]])[0, 0, 0])
        assert abs(vals[2] - vals[0] * vals[1]) < 1e-12 * max(1.0, abs(vals[2]))
