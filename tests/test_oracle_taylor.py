"""The oracle's Taylor-series bubbles (Function3D_get_taylor_series_bubbles,
function3d.F90:1338-1404; Bubbles_get_contaminants, bubbles.F90:2844-2903;
Bubbles_make_taylor, bubbles.F90:1435-1533) and the product that uses them,
pinned without the library:

* cube contaminants: on a cube sampled from a polynomial of degree <= 6 per
  axis the LIP interpolant is the polynomial, so its derivatives at any
  point (a cell boundary included) are the analytic ones;
* bubble contaminants: with radial functions that are polynomials of degree
  <= 6 the interpolant is smooth everywhere, and the derivatives at the
  other centres must match 4th-order finite differences of the bubbles
  evaluated on a stencil (k = 0 and k = -1);
* make_taylor: at radial nodes along random directions the Taylor bubble
  equals w(r) sum d_kappa / |kappa|! (x - R)^kappa evaluated directly;
* the product: cube + bubbles is the product of the two functions on the
  grid, and around a nucleus whose -Z/r bubble multiplies a smooth function
  the cube is close to a quadratic (fit residual below 1/5 of the plain
  product's, whose cube keeps the -Z f/r singularity; 1/70 and 1/10 here).
"""
import math

import numpy as np
import pytest

import oracle as O

# a polynomial in (x, y, z): {(i, j, k): coefficient}
POLY = {(0, 0, 0): 0.7, (1, 0, 0): 0.3, (0, 1, 0): -0.2, (0, 0, 1): 0.5, (2, 0, 0): 0.1,
        (1, 1, 0): -0.4, (0, 1, 1): 0.7, (0, 0, 2): -0.2, (3, 1, 0): 0.05, (1, 2, 2): -0.03,
        (4, 0, 1): 0.01, (2, 3, 1): 0.02}


def poly_eval(x, y, z, d=(0, 0, 0)):
    tot = 0.0
    for (i, j, k), c in POLY.items():
        if i < d[0] or j < d[1] or k < d[2]:
            continue
        f = c * math.perm(i, d[0]) * math.perm(j, d[1]) * math.perm(k, d[2])
        tot = tot + f * x ** (i - d[0]) * y ** (j - d[1]) * z ** (k - d[2])
    return tot


def test_cart_indices_order():
    assert O.cart_indices(2) == [(0, 0, 0), (0, 0, 1), (0, 0, 2), (0, 1, 0), (0, 1, 1),
                                 (0, 2, 0), (1, 0, 0), (1, 0, 1), (1, 1, 0), (2, 0, 0)]


def test_cube_contaminants_of_a_polynomial():
    ax = O.Axis.ncell_fixed(-3.0, 3.0, 7, 3)  # cells of length 2: boundaries at -1, 1
    P = ax.points
    X, Y, Z = np.meshgrid(P, P, P, indexing="ij")
    cube = np.asfortranarray(poly_eval(X, Y, Z))
    centres = np.array([[0.37, -1.21, 2.05], [1.0, -1.0, 0.0], [-2.9, 2.999, 0.5],
                        [3.5, 0.0, 0.0]])
    got = O.cube_contaminants([ax] * 3, cube, centres, 2)
    for c, R in enumerate(centres[:3]):
        for d, kp in enumerate(O.cart_indices(2)):
            ref = poly_eval(*R, d=kp)
            assert abs(got[d, c] - ref) < 1e-11 * max(1.0, abs(ref)), (c, kp, got[d, c], ref)
    assert np.all(got[:, 3] == 0.0)  # outside the grid


def poly_bubbles(rng, centres, lmax, k, ncell=20):
    steps = [O.radial_steps(z, ncell, 7, 12.0) for z in (1.0, 2.0, 3.0)[:len(centres)]]
    bfs = []
    for st in steps:
        r = O.Axis(0.0, st, 7).points
        cols = []
        for lm in range((lmax + 1) ** 2):
            a = rng.uniform(-1, 1, 7)
            cols.append(sum(a[j] * (r / 6.0) ** j for j in range(7)))
        bfs.append(np.asfortranarray(np.array(cols).T))
    return O.BubbleSet(np.asarray(centres, float), steps, bfs, lmax, k, 7)


@pytest.mark.parametrize("k", [0, -1])
def test_bubble_contaminants_against_finite_differences(k):
    rng = np.random.default_rng(3 - k)
    centres = [[0.0, 0.0, 0.0], [1.3, -0.4, 0.9], [-0.8, 1.1, -1.2]]
    bub = poly_bubbles(rng, centres, 3, k)
    got = O.bubble_contaminants(bub, 2)
    h = 2e-3
    off = h * np.arange(-2, 3)
    for ib, R in enumerate(centres):
        others = O.BubbleSet(bub.centers, bub.steps,
                             [np.zeros_like(f) if i == ib else f for i, f in enumerate(bub.bf)],
                             bub.lmax, bub.k, 7)
        F = O.bubbles_eval_grid(others, R[0] + off, R[1] + off, R[2] + off)
        d1 = np.array([1.0, -8.0, 0.0, 8.0, -1.0]) / (12 * h)
        d2 = np.array([-1.0, 16.0, -30.0, 16.0, -1.0]) / (12 * h * h)
        e0 = np.array([0, 0, 1.0, 0, 0])
        ops = {0: e0, 1: d1, 2: d2}

        def fd(kp):
            if sum(kp) == 2 and max(kp) == 1:   # mixed: first differences on two axes
                w = [d1 if kp[a] else e0 for a in range(3)]
            else:
                w = [ops[kp[a]] for a in range(3)]
            return np.einsum("ijk,i,j,k->", F, *w)

        scale = max(abs(fd(kp)) for kp in O.cart_indices(2))
        for d, kp in enumerate(O.cart_indices(2)):
            ref = fd(kp)
            assert abs(got[d, ib] - ref) < 1e-6 * scale, (k, ib, kp, got[d, ib], ref)


@pytest.mark.parametrize("k", [0, -1])
def test_make_taylor_reproduces_the_series(k):
    rng = np.random.default_rng(11)
    centres = [[0.2, -0.1, 0.4], [1.5, 0.7, -0.9]]
    bub = poly_bubbles(rng, centres, 1, k)
    contam = rng.uniform(-1, 1, (10, 2))
    T = O.make_taylor(bub, contam, 2)
    assert T.lmax == 2 and T.k == k
    for i, R in enumerate(centres):
        only = O.BubbleSet(T.centers, T.steps,
                           [f if j == i else np.zeros_like(f) for j, f in enumerate(T.bf)],
                           T.lmax, T.k, 7)
        r_nodes = O.Axis(0.0, T.steps[i], 7).points[1:80]
        for _ in range(4):
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            for r in r_nodes[::7]:
                p = np.asarray(R) + r * u
                got = O.bubbles_eval_grid(only, [p[0]], [p[1]], [p[2]])[0, 0, 0]
                ref = float(O.taylor_weight(r)) * sum(
                    contam[d, i] / math.factorial(sum(kp)) * (r * u[0]) ** kp[0]
                    * (r * u[1]) ** kp[1] * (r * u[2]) ** kp[2]
                    for d, kp in enumerate(O.cart_indices(2)))
                assert abs(got - ref) < 1e-11 * max(1.0, abs(ref)), (i, r, got, ref)


def test_product_with_taylor_terms():
    ax = O.Axis.ncell_fixed(-4.0, 4.0, 7, 8)   # 49 points, h = 1/6
    P = ax.points
    # two nuclei between grid points (a k < 0 bubble is 0 at r = 0 by the
    # reference's convention, which makes a grid point on a nucleus special)
    h = P[1] - P[0]
    centres = np.array([[P[18] + 0.31 * h, P[24] - 0.22 * h, P[24] + 0.4 * h],
                        [P[30] - 0.27 * h, P[24] + 0.35 * h, P[27] + 0.12 * h]])
    steps = [O.radial_steps(z, 25, 7, 12.0) for z in (2.0, 1.0)]
    vb = [np.full((len(O.Axis(0.0, st, 7).points), 1), -z, order="F")
          for st, z in zip(steps, (2.0, 1.0))]
    V = O.BubbleSet(centres, steps, vb, 0, -1, 7)
    X, Y, Z = np.meshgrid(P, P, P, indexing="ij")
    c2 = np.asfortranarray(np.exp(-0.5 * ((X - 0.3) ** 2 + (Y + 0.2) ** 2 + Z ** 2)))
    rng = np.random.default_rng(5)
    bf2 = []
    for st in steps:
        r = O.Axis(0.0, st, 7).points
        bf2.append(np.asfortranarray(np.outer(np.exp(-1.5 * r), rng.uniform(-1, 1, 9))
                                     * np.array([r ** int(math.isqrt(q)) for q in range(9)]).T))
    f2 = O.BubbleSet(centres, steps, bf2, 2, 0, 7)
    zero = np.zeros_like(c2)
    cube, bo = O.function3d_multiply_taylor([ax] * 3, zero, V, c2, f2, 2)
    assert bo.lmax == 4 and bo.k == -1
    BV, B2 = O.bubbles_eval_grid(V, P, P, P), O.bubbles_eval_grid(f2, P, P, P)
    whole = cube + O.bubbles_eval_grid(bo, P, P, P)
    assert np.max(np.abs(whole - BV * (c2 + B2))) < 1e-12 * np.max(np.abs(BV * (c2 + B2)))
    # the plain product's cube (no Taylor terms)
    plain_b = O.BubbleSet(centres, steps, [O.bubble_product(0, vb[i], 2, bf2[i]) for i in range(2)],
                          2, -1, 7)
    plain = BV * (c2 + B2) - O.bubbles_eval_grid(plain_b, P, P, P)
    # least-squares quadratic fit of the cube around each nucleus (points
    # beyond the first radial cell, where the k = -1 values are interpolated,
    # and within 3h): the plain product's cube carries -Z f(R)/r, which no
    # polynomial fits; with the Taylor terms what is left is -Z (f - T)/r =
    # O(r^2)
    X, Y, Z = np.meshgrid(P, P, P, indexing="ij")
    for b, R in enumerate(centres):
        d = np.sqrt((X - R[0]) ** 2 + (Y - R[1]) ** 2 + (Z - R[2]) ** 2)
        sel = (d > 6 * steps[b][0]) & (d < 3 * h)
        x, y, z = X[sel] - R[0], Y[sel] - R[1], Z[sel] - R[2]
        A = np.stack([np.ones_like(x), x, y, z, x * x, y * y, z * z, x * y, x * z, y * z], 1)
        res = {}
        for name, c in (("taylor", cube), ("plain", plain)):
            coef = np.linalg.lstsq(A, c[sel], rcond=None)[0]
            res[name] = np.max(np.abs(A @ coef - c[sel]))
        print(f"nucleus {b}: quadratic-fit residual of the cube {res['taylor']:.3e} with Taylor "
              f"terms, {res['plain']:.3e} without")
        assert res["taylor"] < 0.2 * res["plain"]
