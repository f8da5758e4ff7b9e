"""Bubbles' radial transforms and high-lmax inject on the GPU (rows f2, A24).

* dage_bubbles_coulomb_transform / _helmholtz_transform against the oracle's
  restatements of Coulomb3D_evaluate_bubbles_potential (coulomb.F90:222-336)
  and Helmholtz3D_transform_bubbles_sub (helmholtz3d.F90:519-657), bubble by
  bubble (1e-13 of max|V|; the GPU evaluates the Bessel functions in FP64
  where the reference uses REAL128);
* the transforms agree with the cube operators: the Coulomb3D / Helmholtz3D
  apply of an injected smooth bubble density equals the injected transformed
  bubbles on the cube (two independent discretisations; 1e-5);
* bubbles with lmax 8, injected from lmin 5 (the l > lmax part of a bubble
  product, inject_extra_bubbles, function3d.F90:2198-2240), against the
  oracle's Bubbles_eval_3dgrid (1e-12);
* bubble arithmetic and readback.
"""
import numpy as np
import pytest

import dage_amd
import oracle as O
from _helpers import rel_maxdiff, slater_bubbles

pytestmark = pytest.mark.gpu


def lib_bubbles(ctx, bset, Z=None, k=None):
    b = dage_amd.Bubbles(ctx, len(bset.steps), bset.lmax, bset.k if k is None else k, bset.nlip)
    for i, (st, bf) in enumerate(zip(bset.steps, bset.bf)):
        b.set_bubble(i, bset.centers[i], 1.0 if Z is None else Z[i], st, bf)
    return b


@pytest.mark.parametrize("k", [0, -1])
def test_coulomb_and_helmholtz_transforms_match_oracle(ctx, k):
    rng = np.random.default_rng(11 + k)
    bset, Z = slater_bubbles(rng, 3, lmax=4, k=k)
    b = lib_bubbles(ctx, bset, Z)
    vc = b.coulomb_transform()
    vh = b.helmholtz_transform(-0.7)
    for i in range(3):
        st, bf = bset.steps[i], bset.bf[i]
        ref_c = O.bubble_coulomb(4, k, st, 7, bf)
        ref_h = O.bubble_helmholtz(4, k, st, 7, -0.7, bf)
        got_c, got_h = vc.get(i)[3], vh.get(i)[3]
        ec, eh = rel_maxdiff(got_c, ref_c), rel_maxdiff(got_h, ref_h)
        print(f"k={k} bubble {i}: Coulomb {ec:.1e}, Helmholtz {eh:.1e}")
        assert ec < 1e-13 and eh < 1e-13
        c, q, st2, bf2 = b.get(i)
        assert np.array_equal(bf2, bf) and np.array_equal(st2, st) and q == Z[i]
    for obj in (vc, vh, b):
        obj.close()


@pytest.mark.parametrize("energy", [-31.0, -180.0])
def test_helmholtz_transform_deep_core_energies(ctx, energy):
    """Core-orbital energies (oxygen 1s ~ -20 Ha, and deeper): kappa r reaches
    158 and 379 on a 20-bohr radial grid, where the i_l series must not
    overflow (the reference sums it in REAL128) and the outer cells have
    kappa L >> 1, which takes the exponentially fitted form; against the
    oracle (REAL128 Bessel functions, quadrature weights), 1e-12 of max|V|,
    and against the Yukawa closed form of a diffuse Gaussian, 1e-6.  Beyond
    kappa r_max = 400 the series stops converging within its 300 terms:
    refused."""
    rng = np.random.default_rng(7)
    bset, Z = slater_bubbles(rng, 2, lmax=12, k=-1)
    b = lib_bubbles(ctx, bset, Z)
    vh = b.helmholtz_transform(energy)
    for i in range(2):
        ref = O.bubble_helmholtz(12, -1, bset.steps[i], 7, energy, bset.bf[i])
        got = vh.get(i)[3]
        assert np.all(np.isfinite(got))
        e = rel_maxdiff(got, ref)
        print(f"E {energy}: bubble {i} Helmholtz transform vs oracle {e:.1e}")
        assert e < 1e-12
    vh.close()
    # the fitted form against the closed form: the Yukawa potential of a
    # diffuse normalised Gaussian on a Z = 8 grid (last cell 12.9 bohr)
    st = O.radial_steps(8.0, 33, 7, 20.0)
    r = O.Axis(0.0, st, 7).points
    g = dage_amd.Bubbles(ctx, 1, 0, 0, 7)
    f = (0.3 / np.pi) ** 1.5 * np.exp(-0.3 * r ** 2)
    g.set_bubble(0, [0.0, 0.0, 0.0], 8.0, st, f[:, None])
    V = g.helmholtz_transform(energy)
    k = np.sqrt(-2 * energy)
    ref = np.array([O.yukawa_gauss_potential([1.0], [0.3], [[0.0, 0.0, 0.0]], k, [x], [0.0],
                                             [0.0])[0, 0, 0] for x in r])
    e = np.max(np.abs(V.get(0)[3][:, 0] - ref)) / np.max(np.abs(ref))
    print(f"E {energy}: diffuse Gaussian on the Z = 8 grid vs Yukawa closed form {e:.1e}")
    assert e < 1e-6
    V.close()
    g.close()
    qmax = max(O.Axis(0.0, st, 7).points[-1] for st in bset.steps)
    with pytest.raises(dage_amd.DageError):
        b.helmholtz_transform(-0.5 * (401.0 / qmax) ** 2)
    b.close()


def gaussian_bubbles(ctx, lmax, centers, alpha, coef, Z=1.0):
    bub = dage_amd.Bubbles(ctx, len(centers), lmax, 0, 7)
    sets = []
    for i, c in enumerate(centers):
        st = dage_amd.radial_grid_steps(Z, 60, 7, 12.0)
        r = O.Axis(0.0, st, 7).points
        bf = np.zeros((len(r), (lmax + 1) ** 2), order="F")
        for l in range(lmax + 1):
            for m in range(-l, l + 1):
                bf[:, l * l + l + m] = coef[i][l * l + l + m] * r ** l * np.exp(-alpha * r ** 2)
        bub.set_bubble(i, c, Z, st, bf)
        sets.append((st, bf))
    return bub


@pytest.mark.parametrize("energy", [None, -0.5])
def test_bubble_transforms_agree_with_cube_operators(ctx, energy):
    """Two discretisations of the same potential: the cube operator on the
    injected density and the radial transform injected afterwards."""
    ax = O.Axis.ncell_fixed(-6.0, 6.0, 7, 20)  # 121^3, h = 0.1
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
    rng = np.random.default_rng(5)
    lmax = 3
    centers = np.array([[0.4, -0.3, 0.2], [-0.5, 0.6, -0.1]])
    coef = rng.uniform(-1, 1, (2, (lmax + 1) ** 2))
    bub = gaussian_bubbles(ctx, lmax, centers, 1.5, coef)
    rho = dage_amd.DeviceCube(ctx, (ax.n,) * 3)
    bub.inject(g, rho)
    quad = dage_amd.GaussQuad(20, 20, tlin=2.0, tlog=500.0)
    if energy is None:
        op = dage_amd.Coulomb3D(ctx, g, g, quad)
        vb = bub.coulomb_transform()
    else:
        op = dage_amd.Helmholtz3D(ctx, g, g, quad, energy=energy)
        vb = bub.helmholtz_transform(energy)
    vcube = dage_amd.DeviceCube(ctx, (ax.n,) * 3)
    op.transform_cuda_cube(rho, vcube)
    vinj = dage_amd.DeviceCube(ctx, (ax.n,) * 3)
    vb.inject(g, vinj)
    a, b = vcube.download(), vinj.download()
    sl = slice(20, 101)  # away from the cube faces
    err = rel_maxdiff(a[sl, sl, sl], b[sl, sl, sl])
    print(f"energy {energy}: cube operator vs injected bubble transform {err:.2e}")
    assert err < 1e-5
    for obj in (vb, op, bub, rho, vcube, vinj, g):
        obj.close()


def test_high_lmax_inject_matches_oracle(ctx):
    rng = np.random.default_rng(8)
    bset, Z = slater_bubbles(rng, 3, lmax=8, box=2.0)
    b = lib_bubbles(ctx, bset, Z)
    ax = O.Axis.ncell_fixed(-4.0, 4.0, 7, 8)
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
    for lmin in (0, 5):
        d = dage_amd.DeviceCube(ctx, (ax.n,) * 3)
        b.inject(g, d, lmin=lmin)
        ref = O.bubbles_eval_grid(bset, ax.points, ax.points, ax.points, lmin=lmin)
        err = rel_maxdiff(d.download(), ref)
        print(f"lmax 8 inject from lmin {lmin}: {err:.2e}")
        assert err < 1e-12
        d.close()
    b.close()
    g.close()


def test_bubbles_axpby_and_like(ctx):
    rng = np.random.default_rng(9)
    bset, Z = slater_bubbles(rng, 2, lmax=2)
    x = lib_bubbles(ctx, bset, Z)
    y = dage_amd.Bubbles.like(x, lmax=4)
    assert not y.get(1)[3].any()
    y.axpby(2.0, x, 0.5)
    got = y.get(1)[3]
    assert np.array_equal(got[:, :9], 2.0 * bset.bf[1]) and not got[:, 9:].any()
    y.axpby(1.0, x, -1.0)
    assert np.allclose(y.get(0)[3][:, :9], -bset.bf[0], rtol=0, atol=1e-15)
    x.close()
    y.close()
