"""Taylor-series bubbles and the product of cusp-type functions on the GPU
(row f1: Function3D_get_taylor_series_bubbles, function3d.F90:1338-1404;
Bubbles_make_taylor, bubbles.F90:1435-1533; BubblesMultiplier_multiply,
bubbles.F90:3567-3740) against the oracle (pinned in test_oracle_taylor.py):
contaminants (cube + bubbles, k = 0 and k = -1), Taylor bubbles, and the
product's bubbles and cube, all to 1e-11 relative (host libm erfc, pow and
summation order against numpy/scipy: ~1e-12 measured); Function3D.multiply of
two cusp-type functions takes this path by itself.
"""
import ctypes

import numpy as np
import pytest

import dage_amd
import dage_amd.function3d as F3
import oracle as O
from _helpers import rel_maxdiff

pytestmark = pytest.mark.gpu


def setup(ctx, k, seed):
    rng = np.random.default_rng(seed)
    ax = O.Axis.ncell_fixed(-4.0, 4.0, 7, 6)  # 37^3
    P = ax.points
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
    centres = np.array([[0.37, -0.41, 0.22], [-1.13, 0.9, -0.55], [1.4, 1.1, 0.8]])
    Z = [2.0, 1.0, 3.0]
    steps = [O.radial_steps(z, 25, 7, 12.0) for z in Z]
    lmax = 2
    bfs = []
    b = dage_amd.Bubbles(ctx, 3, lmax, k, 7)
    for i in range(3):
        r = O.Axis(0.0, steps[i], 7).points
        bf = np.asfortranarray(np.outer(np.exp(-Z[i] * r), rng.uniform(-1, 1, (lmax + 1) ** 2))
                               * np.array([r ** int(np.sqrt(q)) for q in range((lmax + 1) ** 2)]).T)
        b.set_bubble(i, centres[i], Z[i], steps[i], bf)
        bfs.append(bf)
    X, Y, Zc = np.meshgrid(P, P, P, indexing="ij")
    cube = np.asfortranarray(np.exp(-0.4 * ((X - 0.2) ** 2 + (Y + 0.3) ** 2 + (Zc - 0.1) ** 2))
                             * (1 + 0.3 * X - 0.2 * Y * Zc))
    f = F3.Function3D(g, dage_amd.DeviceCube(ctx, cube.shape, cube), b, type=F3.F3D_TYPE_CUSP)
    bset = O.BubbleSet(centres, steps, bfs, lmax, k, 7)
    return ax, g, f, cube, bset


@pytest.mark.parametrize("k", [0, -1])
def test_contaminants_and_taylor_bubbles(ctx, k):
    ax, g, f, cube, bset = setup(ctx, k, 3)
    got = f.contaminants(2)
    ref = O.bubble_contaminants(bset, 2) + O.cube_contaminants([ax] * 3, cube, bset.centers, 2)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"k {k}: contaminants vs oracle {err:.2e}")
    assert err < 1e-11
    T = f.taylor_series_bubbles()
    Tref = O.make_taylor(bset, ref, 2)
    assert T.lmax == 2 and T.k == k
    for i in range(3):
        e = rel_maxdiff(T.get(i)[3], Tref.bf[i])
        assert e < 1e-11, (i, e)
    T.close()
    f.close()
    g.close()


def test_product_of_cusp_functions(ctx):
    ax, g, f1, c1, b1 = setup(ctx, -1, 5)
    _, g2, f2, c2, b2 = setup(ctx, 0, 6)
    f2.grid = g
    prod = f1.multiply(f2)  # both cusp-type: the Taylor path
    assert prod.bubbles.lmax == 4 and prod.bubbles.k == -1
    cref, bref = O.function3d_multiply_taylor([ax] * 3, c1, b1, c2, b2, 2)
    for i in range(3):
        e = rel_maxdiff(prod.bubbles.get(i)[3], bref.bf[i])
        assert e < 1e-11, (i, e)
    ec = rel_maxdiff(prod.cube.download(), cref)
    print(f"product with Taylor terms: cube vs oracle {ec:.2e}")
    assert ec < 1e-11
    # the plain product when a factor is not cusp-type
    f2.type = F3.F3D_TYPE_NONE
    plain = f1.multiply(f2)
    assert plain.bubbles.lmax == 4
    P = ax.points
    whole_t = prod.cube.download() + O.bubbles_eval_grid(bref, P, P, P)
    pb = O.BubbleSet(b1.centers, b1.steps, [plain.bubbles.get(i)[3] for i in range(3)], 4, -1, 7)
    whole_p = plain.cube.download() + O.bubbles_eval_grid(pb, P, P, P)
    assert rel_maxdiff(whole_t, whole_p) < 1e-12  # same function, different split
    with pytest.raises(dage_amd.DageError):
        f1.taylor_series_bubbles(3)
    for obj in (prod, plain, f1, f2, g, g2):
        obj.close()
