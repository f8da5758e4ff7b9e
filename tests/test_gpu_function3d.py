"""Function3D-level operations on the GPU (rows A8, A24, f3).

* operate_on (Operator3D_operate_on_Function3D, function3d.F90:3727-3792;
  Helmholtz3D_operate, helmholtz3d.F90:446-481): the result's cube equals the
  operator's cube transform bit for bit, its bubbles equal the oracle's
  radial transform (1e-13), a Helmholtz operator is left at energy 0;
* inject_extra_bubbles / dispose_extra_bubbles (function3d.F90:2198-2240):
  the l > lmax part lands in the cube exactly as the oracle's
  Bubbles_eval_3dgrid from lmin = lmax + 1 (1e-12), the l <= lmax part stays;
* dump / load (.fun) round trip of a device-resident function.
"""
import numpy as np
import pytest

import dage_amd
import dage_amd.function3d as F3
import oracle as O
from _helpers import random_gaussians, rel_maxdiff, slater_bubbles

pytestmark = pytest.mark.gpu


def make_function(ctx, seed=1, lmax=4):
    rng = np.random.default_rng(seed)
    ax = O.Axis.ncell_fixed(-5.0, 5.0, 7, 8)  # 49^3
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
    q, a, c = random_gaussians(rng, 3)
    rho = O.gauss_density(q, a, c, ax.points, ax.points, ax.points)
    bset, Z = slater_bubbles(rng, 2, lmax=lmax, box=2.0)
    bub = dage_amd.Bubbles(ctx, 2, lmax, 0, 7)
    for i in range(2):
        bub.set_bubble(i, bset.centers[i], Z[i], bset.steps[i], bset.bf[i])
    f = F3.Function3D(g, dage_amd.DeviceCube(ctx, rho.shape, rho), bub,
                      type=F3.F3D_TYPE_CUSP, label="rho")
    return f, ax, g, rho, bset


def test_operate_on_coulomb_and_helmholtz(ctx):
    f, ax, g, rho, bset = make_function(ctx)
    quad = dage_amd.GaussQuad(10, 10, tlin=2.0, tlog=500.0)
    op = dage_amd.Coulomb3D(ctx, g, g, quad)
    res = F3.operate_on(op, f)
    assert np.array_equal(res.cube.download(), op.transform_cube(rho))
    for i in range(2):
        ref = O.bubble_coulomb(4, 0, bset.steps[i], 7, bset.bf[i])
        assert rel_maxdiff(res.bubbles.get(i)[3], ref) < 1e-13
    res.close()
    E = -0.6
    hop = dage_amd.Helmholtz3D(ctx, g, g, quad, energy=E)
    cube_ref = hop.transform_cube(rho)
    hop.set_energy(E)
    res = F3.operate_on(hop, f)
    assert np.array_equal(res.cube.download(), cube_ref)
    for i in range(2):
        ref = O.bubble_helmholtz(4, 0, bset.steps[i], 7, E, bset.bf[i])
        assert rel_maxdiff(res.bubbles.get(i)[3], ref) < 1e-13
    # Helmholtz3D_operate leaves the operator at energy 0 (Coulomb weights)
    assert np.array_equal(hop.get_weights(), op.get_weights())
    res.close()
    only = F3.operate_on(op, f, only_cube=True)
    assert only.bubbles is None
    only.close()
    for obj in (op, hop):
        obj.close()
    f.close()
    g.close()


def test_inject_and_dispose_extra_bubbles(ctx):
    f, ax, g, rho, bset = make_function(ctx, seed=2, lmax=4)
    f.inject_extra_bubbles(2)
    ref = rho + O.bubbles_eval_grid(bset, ax.points, ax.points, ax.points, lmin=3)
    err = rel_maxdiff(f.cube.download(), ref)
    print(f"inject_extra_bubbles(2): {err:.2e}")
    assert err < 1e-12
    assert f.bubbles.lmax == 2
    for i in range(2):
        assert np.array_equal(f.bubbles.get(i)[3], bset.bf[i][:, :9])
    f.dispose_extra_bubbles(0)
    assert f.bubbles.lmax == 0 and np.array_equal(f.bubbles.get(1)[3], bset.bf[1][:, :1])
    f.close()
    g.close()


def test_dump_and_load_roundtrip(ctx, tmp_path):
    f, ax, g, rho, bset = make_function(ctx, seed=3)
    path = f.dump(str(tmp_path), "orbital_001")
    assert path.endswith("orbital_001.fun")
    F3.write_grid(g, str(tmp_path / "grid.g3d"))
    g2 = F3.read_grid(ctx, str(tmp_path / "grid.g3d"))
    assert g2.get_shape() == g.get_shape()
    assert np.array_equal(g2.get_coord(0), g.get_coord(0))
    h = F3.Function3D.load(g2, str(tmp_path), "orbital_001.fun", mould=f.bubbles)
    assert h.label == "rho" and h.type == F3.F3D_TYPE_CUSP
    assert np.array_equal(h.cube.download(), rho)
    for i in range(2):
        c, q, st, bf = h.bubbles.get(i)
        assert np.array_equal(bf, bset.bf[i]) and np.array_equal(c, bset.centers[i])
    h.close()
    f.close()
    g2.close()
    g.close()


def test_bubble_and_function_products(ctx):
    """The product (Function3D_multiply_sub without Taylor terms): bubbles
    against the oracle's product (1e-13), the cube against
    (c2 + B2)(c1 + B1) - B12 computed by the oracle on the grid (1e-12), and
    the whole product evaluated on the grid equals the product of the two
    functions evaluated on the grid exactly (bubbles + cube)."""
    f, ax, g, rho, bset = make_function(ctx, seed=4, lmax=2)
    rng = np.random.default_rng(44)
    q, a, c = random_gaussians(rng, 2)
    cube2 = O.gauss_density(q, a, c, ax.points, ax.points, ax.points)
    bf2 = [np.asfortranarray(rng.uniform(-1, 1, (len(O.Axis(0.0, st, 7).points), 16))
                             * np.exp(-O.Axis(0.0, st, 7).points)[:, None]) for st in bset.steps]
    b2 = dage_amd.Bubbles(ctx, 2, 3, 0, 7)
    for i in range(2):
        b2.set_bubble(i, bset.centers[i], 1.0, bset.steps[i], bf2[i])
    f2 = F3.Function3D(g, dage_amd.DeviceCube(ctx, cube2.shape, cube2), b2)
    prod = f.multiply(f2)
    assert prod.bubbles.lmax == 5
    bprod = []
    for i in range(2):
        ref = O.bubble_product(2, bset.bf[i], 3, bf2[i])
        bprod.append(ref)
        assert rel_maxdiff(prod.bubbles.get(i)[3], ref) < 1e-13
    bs2 = O.BubbleSet(bset.centers, bset.steps, bf2, 3, 0, 7)
    bs12 = O.BubbleSet(bset.centers, bset.steps, bprod, 5, 0, 7)
    P = ax.points
    B1 = O.bubbles_eval_grid(bset, P, P, P)
    B2 = O.bubbles_eval_grid(bs2, P, P, P)
    B12 = O.bubbles_eval_grid(bs12, P, P, P)
    ref_cube = (cube2 + B2) * (rho + B1) - B12
    got = prod.cube.download()
    err = rel_maxdiff(got, ref_cube)
    print(f"Function3D product cube vs oracle {err:.2e}")
    assert err < 1e-12
    whole = rel_maxdiff(got + B12, (cube2 + B2) * (rho + B1))
    assert whole < 1e-12
    for obj in (prod, f2, f):
        obj.close()
    g.close()
