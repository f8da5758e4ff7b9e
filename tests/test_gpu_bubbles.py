"""GPU parity of the fused bubbles inject (HP-C) against the oracle's
Bubbles_eval_3dgrid restatement.  Same algorithm in FP64 on both sides
(bisection cell lookup, Horner LIP polynomials, Racah recursion); the
tolerance 1e-12 of max|cube| covers the different summation orders."""
import numpy as np
import pytest

import dage_amd
import oracle as O
from _helpers import rel_maxdiff, slater_bubbles

pytestmark = pytest.mark.gpu


def lib_bubbles(ctx, bub):
    b = dage_amd.Bubbles(ctx, len(bub.steps), bub.lmax, bub.k, bub.nlip)
    for i in range(len(bub.steps)):
        b.set_bubble(i, bub.centers[i], 1.0, bub.steps[i], bub.bf[i])
    return b


@pytest.mark.parametrize("lmax,k,lmin,n", [(4, 0, 0, 49), (2, -1, 0, 37), (4, 0, 2, 41),
                                           (6, 1, 0, 33), (0, 0, 0, 25)])
def test_inject_vs_oracle(ctx, lmax, k, lmin, n):
    rng = np.random.default_rng(10 + lmax + n)
    bub, _ = slater_bubbles(rng, 4, lmax=lmax, k=k)
    ax = O.Axis.ncell_fixed(-12.0, 12.0, 7, (n - 1) // 6) if (n - 1) % 6 == 0 else \
        O.Axis.ncell_fixed(-12.0, 12.0, 5, (n - 1) // 4)
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, ax.nlip)
    base = np.asfortranarray(rng.standard_normal((ax.n,) * 3))
    cube = dage_amd.DeviceCube(ctx, base.shape, base)
    b = lib_bubbles(ctx, bub)
    b.inject(g, cube, lmin=lmin, multiplier=1.0)
    got = cube.download()
    ref = O.bubbles_eval_grid(bub, ax.points, ax.points, ax.points, lmin=lmin, cube=base)
    err = rel_maxdiff(got - base, ref - base)
    print(f"lmax={lmax} k={k} lmin={lmin} n={n}: {err:.3e}")
    assert err < 1e-12


def test_inject_multiplier_and_uneven_grid(ctx):
    rng = np.random.default_rng(77)
    bub, _ = slater_bubbles(rng, 3, lmax=3)
    axes = [O.Axis.ncell_fixed(-6.0, 6.0, 7, 6), O.Axis.ncell_fixed(-4.0, 5.0, 7, 5),
            O.Axis.ncell_fixed(-7.0, 2.0, 7, 4)]
    g = dage_amd.Grid3D(ctx, [a.qmin for a in axes], [a.steps for a in axes], 7)
    shape = tuple(a.n for a in axes)
    cube = dage_amd.DeviceCube(ctx, shape)
    cube.set_to_zero()
    b = lib_bubbles(ctx, bub)
    b.inject(g, cube, multiplier=-0.5)
    ref = O.bubbles_eval_grid(bub, *(a.points for a in axes))
    assert rel_maxdiff(cube.download(), -0.5 * ref) < 1e-12


def test_inject_host_buffer(ctx):
    """dage_bubbles_inject (Bubbles_eval_3dgrid on a host cube) equals the
    device inject bit for bit and the oracle to 1e-12."""
    rng = np.random.default_rng(78)
    bub, _ = slater_bubbles(rng, 3, lmax=4)
    ax = O.Axis.ncell_fixed(-6.0, 6.0, 7, 7)
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
    base = np.asfortranarray(rng.standard_normal((ax.n,) * 3))
    b = lib_bubbles(ctx, bub)
    host = b.eval_3dgrid(g, lmin=1, cube=base, multiplier=2.0)
    dev = dage_amd.DeviceCube(ctx, base.shape, base)
    b.inject(g, dev, lmin=1, multiplier=2.0)
    assert np.array_equal(host, dev.download())
    ref = O.bubbles_eval_grid(bub, ax.points, ax.points, ax.points, lmin=1)
    assert rel_maxdiff(host - base, 2.0 * ref) < 1e-12


def test_inject_errors(ctx):
    with pytest.raises(dage_amd.DageError):
        dage_amd.Bubbles(ctx, 2, 13)  # above the largest lmax (12)
    ax = O.Axis.ncell_fixed(-2.0, 2.0, 7, 2)
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
    cube = dage_amd.DeviceCube(ctx, (ax.n,) * 3)
    b = dage_amd.Bubbles(ctx, 2, 2)
    with pytest.raises(dage_amd.DageError) as e:
        b.inject(g, cube)
    assert "not been set" in str(e.value)
    with pytest.raises(ValueError):
        b.set_bubble(0, [0, 0, 0], 1.0, np.full(3, 0.1), np.zeros((5, 9)))


def test_inject_points_on_radial_cell_starts(ctx):
    """Cube points at exactly the radial cell starts: a bubble at a grid point
    whose radial cells are as wide as the cube spacing (0.25) puts the points
    along each axis through the centre on cell boundaries, where the cell
    lookup must pick the cell that starts there (grid.F90:1212-1246), and the
    far corners beyond the radial grid get nothing."""
    rng = np.random.default_rng(91)
    lmax = 3
    ax = O.Axis.ncell_fixed(-3.0, 3.0, 7, 4)  # 25 points, h = 0.25
    steps = np.full(10, 0.25 / 6)             # 10 cells of 0.25: r up to 2.5
    r = O.Axis(0.0, steps, 7).points
    c = rng.uniform(-1.0, 1.0, (lmax + 1) ** 2)
    bf = np.zeros((len(r), (lmax + 1) ** 2), order="F")
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            bf[:, l * l + l + m] = c[l * l + l + m] * (1.0 + r) ** l * np.exp(-r)
    center = np.array([[0.0, 0.5, -0.25]])
    bub = O.BubbleSet(center, [steps], [bf], lmax, 0, 7)
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
    cube = dage_amd.DeviceCube(ctx, (ax.n,) * 3)
    cube.set_to_zero()
    b = lib_bubbles(ctx, bub)
    b.inject(g, cube)
    ref = O.bubbles_eval_grid(bub, ax.points, ax.points, ax.points)
    got = cube.download()
    err = rel_maxdiff(got, ref)
    print(f"points on cell starts: {err:.3e}")
    assert err < 1e-12
    assert np.any(ref == 0.0) and np.array_equal(got == 0.0, ref == 0.0)


def test_inject_headline_size_planes(ctx):
    """The headline workload's inject (4 atoms, lmax 4, 199-point radial
    grids, 241^3 over [-12, 12]^3) checked against the oracle on 4 z-planes
    spread through the cube (the oracle on all 241 planes takes minutes)."""
    rng = np.random.default_rng(2410)
    bub, _ = slater_bubbles(rng, 4, lmax=4)
    ax = O.Axis.ncell_fixed(-12.0, 12.0, 7, 40)
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
    cube = dage_amd.DeviceCube(ctx, (ax.n,) * 3)
    cube.set_to_zero()
    b = lib_bubbles(ctx, bub)
    b.inject(g, cube)
    got = cube.download()
    planes = [0, 97, 120, 240]
    ref = O.bubbles_eval_grid(bub, ax.points, ax.points, ax.points[planes])
    err = rel_maxdiff(got[:, :, planes], ref)
    print(f"241^3 inject on planes {planes}: {err:.3e}")
    assert err < 1e-12
