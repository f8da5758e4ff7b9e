"""GBFMM Coulomb potential on the GPU against the oracle's restatement of the
reference's GPU algorithm (gbfmm_coulomb3d.F90/.cu, gbfmm_parallel_info.F90).

Bar: potential within 1e-10 of max|V| of the oracle (FP64, summation order
differs); multipoles and local expansions within 1e-11 of their column
maxima.  The oracle's FMM matrices are pinned by the reference's
multipole_test goldens (test_oracle_fmm.py); its near field is the pinned
operator apply on sub-grids.  At a realistic size the result is also checked
against the analytic potential of the Gaussian density.
"""
import numpy as np
import pytest

import dage_amd
import oracle as O
from _helpers import random_gaussians, rel_maxdiff

pytestmark = pytest.mark.gpu


def grid_of(ctx, axes):
    return dage_amd.Grid3D(ctx, [a.qmin for a in axes], [a.steps for a in axes], axes[0].nlip)


def colrel(a, b):
    scale = np.maximum(np.max(np.abs(b), axis=0), 1e-300)
    return float(np.max(np.max(np.abs(a - b), axis=0) / scale))


@pytest.mark.parametrize("maxlevel,nlip,ncell,lmax", [(2, 5, 8, 6), (3, 3, 16, 4), (2, 7, 4, 12)])
def test_gbfmm_matches_oracle(ctx, maxlevel, nlip, ncell, lmax):
    print(f"gbfmm maxlevel={maxlevel} nlip={nlip} ncell={ncell} lmax={lmax}", flush=True)
    ax = O.Axis.ncell_fixed(-4.0, 4.0, nlip, ncell)
    rng = np.random.default_rng(100 + maxlevel)
    q_, a, c = random_gaussians(rng, 3, box=2.0)
    rho = O.gauss_density(q_, a, c, ax.points, ax.points, ax.points)
    gq = dage_amd.GaussQuad(6, 6)
    g = grid_of(ctx, [ax] * 3)
    fmm = dage_amd.GBFMMCoulomb3D(ctx, g, maxlevel=maxlevel, lmax=lmax, gauss=gq)
    pot = fmm.apply(rho)
    ref, qref, Lref = O.gbfmm_potential([ax] * 3, rho, maxlevel, lmax, gq.tp, gq.tw)
    assert rel_maxdiff(pot, ref) < 1e-10
    assert colrel(fmm.multipole_moments(), qref) < 1e-11
    assert colrel(fmm.local_expansions(), Lref) < 1e-11


def test_gbfmm_unequal_cells_vs_oracle(ctx):
    """Cells of unequal size: boxes differ in their points' positions, so the
    multipoles and local expansions run box by box (the equal-box route puts
    them on one shared harmonics matrix); both routes against the oracle."""
    rng = np.random.default_rng(41)
    nlip, ncell = 5, 8
    ax = O.Axis(-4.0, rng.uniform(0.8, 1.2, ncell) * 8.0 / (ncell * (nlip - 1)), nlip)
    q_, a, c = random_gaussians(rng, 3, box=2.0)
    rho = O.gauss_density(q_, a, c, ax.points, ax.points, ax.points)
    gq = dage_amd.GaussQuad(6, 6)
    g = grid_of(ctx, [ax] * 3)
    fmm = dage_amd.GBFMMCoulomb3D(ctx, g, maxlevel=2, lmax=6, gauss=gq)
    pot = fmm.apply(rho)
    ref, qref, Lref = O.gbfmm_potential([ax] * 3, rho, 2, 6, gq.tp, gq.tw)
    assert rel_maxdiff(pot, ref) < 1e-10
    assert colrel(fmm.multipole_moments(), qref) < 1e-11
    assert colrel(fmm.local_expansions(), Lref) < 1e-11
    fmm.close()


def test_gbfmm_uneven_axes(ctx):
    axes = [O.Axis.ncell_fixed(-4.0, 4.0, 5, 8), O.Axis.ncell_fixed(-3.0, 6.0, 5, 12),
            O.Axis(-2.5, np.array([0.3, 0.25, 0.35, 0.3]), 5)]
    rng = np.random.default_rng(7)
    q_, a, c = random_gaussians(rng, 2, box=1.5)
    rho = O.gauss_density(q_, a, c, *(x.points for x in axes))
    gq = dage_amd.GaussQuad(5, 5)
    g = grid_of(ctx, axes)
    fmm = dage_amd.GBFMMCoulomb3D(ctx, g, maxlevel=2, lmax=5, gauss=gq)
    pot = fmm.apply(rho)
    ref, _, _ = O.gbfmm_potential(axes, rho, 2, 5, gq.tp, gq.tw)
    assert rel_maxdiff(pot, ref) < 1e-10


def test_gbfmm_device_path_and_repeat(ctx):
    ax = O.Axis.ncell_fixed(-4.0, 4.0, 5, 8)
    rng = np.random.default_rng(9)
    q_, a, c = random_gaussians(rng, 2, box=2.0)
    rho = O.gauss_density(q_, a, c, ax.points, ax.points, ax.points)
    g = grid_of(ctx, [ax] * 3)
    fmm = dage_amd.GBFMMCoulomb3D(ctx, g, maxlevel=2, lmax=4, gauss=dage_amd.GaussQuad(5, 5))
    d_in = dage_amd.DeviceCube(ctx, rho.shape, rho)
    d_out = dage_amd.DeviceCube(ctx, rho.shape)
    fmm.apply_device(d_in, d_out)
    first = d_out.download()
    fmm.apply_device(d_in, d_out)  # outputs are overwritten, not accumulated
    assert np.array_equal(d_out.download(), first)
    assert np.array_equal(fmm.apply(rho), first)


def test_gbfmm_on_a_caller_stream(ctx):
    """The far field runs on the GBFMM object's own stream, forked from and
    joined back into the context stream: on a stream the caller set (created
    here through the HIP runtime), results are bit-identical to the default
    stream's, and the context stream alone orders them for the download."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    ax = O.Axis.ncell_fixed(-4.0, 4.0, 5, 8)
    rng = np.random.default_rng(19)
    q_, a, c = random_gaussians(rng, 2, box=2.0)
    rho = O.gauss_density(q_, a, c, ax.points, ax.points, ax.points)
    g = grid_of(ctx, [ax] * 3)
    fmm = dage_amd.GBFMMCoulomb3D(ctx, g, maxlevel=2, lmax=6, gauss=dage_amd.GaussQuad(8, 8))
    d_in = dage_amd.DeviceCube(ctx, rho.shape, rho)
    d_out = dage_amd.DeviceCube(ctx, rho.shape)
    fmm.apply_device(d_in, d_out)
    ref = d_out.download()
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        ctx.set_stream(stream.value)
        d_out.set_to_zero()
        for _ in range(3):
            fmm.apply_device(d_in, d_out)
        ctx.synchronize()
        got = d_out.download()
    finally:
        ctx.set_stream(None)
        assert hip.hipStreamSynchronize(stream) == 0
        assert hip.hipStreamDestroy(stream) == 0
    assert np.array_equal(got, ref)
    fmm.close()


def test_gbfmm_against_analytic(ctx):
    """121^3 points over [-6,6]^3 (nlip 7, 20 cells, 4x4x4 boxes), lmax 12,
    near field with the reference default GaussQuad(25, 25)."""
    ax = O.Axis.ncell_fixed(-6.0, 6.0, 7, 20)
    rng = np.random.default_rng(0xDA6E0002)
    q_, a, c = random_gaussians(rng, 4, alpha_range=(0.5, 5.0), box=2.0)
    rho = O.gauss_density(q_, a, c, ax.points, ax.points, ax.points)
    exact = O.gauss_potential(q_, a, c, ax.points, ax.points, ax.points)
    g = grid_of(ctx, [ax] * 3)
    fmm = dage_amd.GBFMMCoulomb3D(ctx, g, maxlevel=2, lmax=12)
    pot = fmm.apply(rho)
    err = rel_maxdiff(pot, exact)
    print(f"GBFMM vs analytic: {err:.3e}")
    assert err < 1e-4
    q = fmm.multipole_moments()
    total = O.cube_integral(ax.ints(), ax.ints(), ax.ints(), rho)
    assert abs(q[0, :64].sum() - total) < 1e-12 * abs(total)


def test_gbfmm_rejects_bad_arguments(ctx):
    ax = O.Axis.ncell_fixed(-4.0, 4.0, 5, 6)  # 6 cells do not split into 4 boxes
    g = grid_of(ctx, [ax] * 3)
    with pytest.raises(dage_amd.DageError):
        dage_amd.GBFMMCoulomb3D(ctx, g, maxlevel=2, lmax=4)
    ax = O.Axis.ncell_fixed(-4.0, 4.0, 5, 8)
    g = grid_of(ctx, [ax] * 3)
    with pytest.raises(dage_amd.DageError):
        dage_amd.GBFMMCoulomb3D(ctx, g, maxlevel=1, lmax=4)
    with pytest.raises(dage_amd.DageError):
        dage_amd.GBFMMCoulomb3D(ctx, g, maxlevel=2, lmax=23)
    # rho and pot may not overlap: refused before any kernel is launched
    fmm = dage_amd.GBFMMCoulomb3D(ctx, g, maxlevel=2, lmax=4)
    d = dage_amd.DeviceCube(ctx, (ax.n, ax.n, ax.n))
    with pytest.raises(dage_amd.DageError, match="overlap"):
        fmm.apply_device(d, d)
    d.close()
    fmm.close()


def test_gbfmm_mixed_near_field(ctx):
    """Split-BF16 near field (bf16x6) against the FP64 GBFMM: 1e-7 of max|V|."""
    ax = O.Axis.ncell_fixed(-6.0, 6.0, 7, 12)
    rng = np.random.default_rng(31)
    q_, a, c = random_gaussians(rng, 4, box=2.0)
    rho = O.gauss_density(q_, a, c, ax.points, ax.points, ax.points)
    g = grid_of(ctx, [ax] * 3)
    fmm = dage_amd.GBFMMCoulomb3D(ctx, g, maxlevel=2, lmax=8)
    v64 = fmm.apply(rho)
    v6 = fmm.apply(rho, precision=dage_amd.DAGE_PREC_BF16X6)
    v3 = fmm.apply(rho, precision=dage_amd.DAGE_PREC_BF16X3)
    print(f"GBFMM near field bf16x6 {rel_maxdiff(v6, v64):.2e}, bf16x3 {rel_maxdiff(v3, v64):.2e}")
    assert rel_maxdiff(v6, v64) < 1e-7
    assert rel_maxdiff(v3, v64) < 2e-5


def test_gbfmm_full_size_linearity_and_direct_apply(ctx):
    """At the headline size (241^3, 4x4x4 boxes, lmax 12, T = 50), where the
    oracle is too slow, properties that hold at any size: the GBFMM potential
    is linear in the density to rounding, and it agrees with the direct
    whole-grid apply of the same quadrature to the far-field truncation
    error (1-2e-5 measured)."""
    ax = O.Axis.ncell_fixed(-12.0, 12.0, 7, 40)
    rng = np.random.default_rng(241)
    q1, a1, c1 = random_gaussians(rng, 8, box=4.0)
    q2, a2, c2 = random_gaussians(rng, 8, box=4.0)
    rho1 = O.gauss_density(q1, a1, c1, ax.points, ax.points, ax.points)
    rho2 = O.gauss_density(q2, a2, c2, ax.points, ax.points, ax.points)
    g = grid_of(ctx, [ax] * 3)
    gq = dage_amd.GaussQuad(25, 25)
    fmm = dage_amd.GBFMMCoulomb3D(ctx, g, maxlevel=2, lmax=12, gauss=gq)
    p1, p2 = fmm.apply(rho1), fmm.apply(rho2)
    mix = fmm.apply(np.asfortranarray(2.0 * rho1 - 0.5 * rho2))
    lin = rel_maxdiff(mix, 2.0 * p1 - 0.5 * p2)
    op = dage_amd.Coulomb3D(ctx, g, g, gq, coda=False)  # the GBFMM near field has no coda
    direct = op.transform_cube(rho1)
    far = rel_maxdiff(p1, direct)
    print(f"241^3: linearity {lin:.2e}, GBFMM vs direct apply {far:.2e}")
    assert lin < 1e-12
    assert far < 1e-4
    op.close()
    fmm.close()
