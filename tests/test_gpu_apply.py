"""GPU parity of the FP64 Coulomb3D/Helmholtz3D apply against the oracle.

Tolerances:
  * apply with identical matrices (oracle matrices uploaded):   1e-12 of max|V|
  * matrices generated on the GPU vs the oracle's:               2e-9 of max|F_t|
    (the reference's closed forms cancel catastrophically for far entries;
    device erf/exp differ from glibc's by an ulp, which moves such entries
    by ~1e-9 absolute -- measured and reported by the test)
  * end to end (GPU matrices + GPU apply) on Gaussian densities: 1e-10 of
    max|V| (the FP64 parity bar)
  * reference goldens (tests/gauss_test.F90 TEST 1-3) through the GPU: 1e-13
"""
import math
import os

import numpy as np
import pytest

import dage_amd
import oracle as O
from _helpers import gauss_generator_axes, load_golden, random_gaussians, rel_maxdiff

pytestmark = pytest.mark.gpu


def lib_grid(ctx, axes, nlip):
    return dage_amd.Grid3D(ctx, [a.qmin for a in axes], [a.steps for a in axes], nlip)


def quad(nlin=20, nlog=20, tlog=500.0):
    return dage_amd.GaussQuad(nlin=nlin, nlog=nlog, tlin=2.0, tlog=tlog)


def test_grid_matches_oracle(ctx):
    axes = [O.Axis.ncell_fixed(-8.0, 8.0, 7, 16), O.Axis.ncell_fixed(-5.0, 7.0, 7, 9),
            O.Axis.ncell_fixed(-3.0, 2.0, 7, 5)]
    g = lib_grid(ctx, axes, 7)
    assert g.get_shape() == (97, 55, 31)
    for i, a in enumerate(axes):
        assert np.array_equal(g.get_coord(i), a.points)
        assert np.array_equal(g.get_ints(i), a.ints())


@pytest.mark.parametrize("nlip,ncell", [(7, 4), (5, 6)])
def test_matrices_vs_oracle(ctx, nlip, ncell):
    ax_in = O.Axis.ncell_fixed(-6.0, 6.0, nlip, ncell)
    ax_out = ax_in
    q = quad()
    g = lib_grid(ctx, [ax_in] * 3, nlip)
    op = dage_amd.Coulomb3D(ctx, g, g, q)
    fx, fy, fz = op.get_matrices()
    rx, ry, rz = O.coulomb3d_matrices([ax_in] * 3, [ax_out] * 3, q.tp)
    worst = 0.0
    for got, ref in ((fx, rx), (fy, ry), (fz, rz)):
        scale = np.max(np.abs(ref), axis=(0, 1))
        err = np.max(np.abs(got - ref), axis=(0, 1)) / scale
        worst = max(worst, float(err.max()))
    print(f"matrix max |dF|/max|F_t| = {worst:.3e}")
    assert worst < 2e-9
    # the closed forms' exact zeros (underflow) stay zero; the other zeros are
    # the flushed tails, entries at most 1e-20 of their row's largest (one
    # input point, one t); a 10x margin covers ulp-level differences of
    # entries next to the cutoff
    for got, ref, out_axis in ((fx, rx, 0), (fy, ry, 1), (fz, rz, 1)):
        rowmax = np.max(np.abs(ref), axis=out_axis, keepdims=True)
        assert np.all(got[ref == 0.0] == 0.0)
        flushed = (got == 0.0) & (ref != 0.0)
        assert np.all((np.abs(ref) <= 1e-19 * rowmax)[flushed])
        assert np.all(got[np.abs(ref) > 1e-19 * rowmax] != 0.0)
        print(f"flushed entries: {flushed.mean():.3f} of the matrix")


def run_both(ctx, axes, nlip, cube, q, coda=True, set_oracle_matrices=False, energy=None):
    g = lib_grid(ctx, axes, nlip)
    if energy is None:
        op = dage_amd.Coulomb3D(ctx, g, g, q, coda=coda)
    else:
        op = dage_amd.Helmholtz3D(ctx, g, g, q, energy=0.0, coda=coda)
    fx, fy, fz = O.coulomb3d_matrices(axes, axes, q.tp)
    if set_oracle_matrices:
        op.set_matrices(fx, fy, fz)
    w = O.coulomb_weights(q.tw)
    if energy is not None:
        op.set_energy(energy)
        w = O.helmholtz_weights(q.tp, q.tw, energy)
        assert np.array_equal(op.get_weights(), w)
    got = op.transform_cube(cube)
    ref = O.transform_cube(fx, fy, fz, w, O.coda(q.tlog) if coda else 0.0, cube)
    return got, ref, op


def test_apply_same_matrices_random_input(ctx):
    rng = np.random.default_rng(3)
    axes = [O.Axis.ncell_fixed(-4.0, 4.0, 7, 6), O.Axis.ncell_fixed(-3.0, 5.0, 7, 5),
            O.Axis.ncell_fixed(-2.0, 2.0, 7, 7)]
    cube = np.asfortranarray(rng.standard_normal([a.n for a in axes]))
    got, ref, _ = run_both(ctx, axes, 7, cube, quad(), set_oracle_matrices=True)
    err = rel_maxdiff(got, ref)
    print(f"random input, oracle matrices: {err:.3e}")
    assert err < 1e-12


@pytest.mark.parametrize("nlip,ncell,ngau", [(7, 8, 4), (5, 12, 4), (7, 16, 4)])
def test_apply_end_to_end_gaussians(ctx, nlip, ncell, ngau):
    rng = np.random.default_rng(100 + ncell)
    ax = O.Axis.ncell_fixed(-8.0, 8.0, nlip, ncell)
    q_, alpha, c = random_gaussians(rng, ngau)
    rho = O.gauss_density(q_, alpha, c, ax.points, ax.points, ax.points)
    got, ref, _ = run_both(ctx, [ax] * 3, nlip, rho, quad())
    err = rel_maxdiff(got, ref)
    print(f"N={ax.n} nlip={nlip}: GPU vs oracle {err:.3e}")
    assert err < 1e-10
    if ax.n == 97:  # config 1 shape: also against the analytic potential
        Va = O.gauss_potential(q_, alpha, c, ax.points, ax.points, ax.points)
        assert rel_maxdiff(got, Va) < 1e-3


def test_apply_no_coda_and_helmholtz(ctx):
    rng = np.random.default_rng(5)
    ax = O.Axis.ncell_fixed(-6.0, 6.0, 5, 10)
    q_, alpha, c = random_gaussians(rng, 3)
    rho = O.gauss_density(q_, alpha, c, ax.points, ax.points, ax.points)
    got, ref, _ = run_both(ctx, [ax] * 3, 5, rho, quad(), coda=False)
    assert rel_maxdiff(got, ref) < 1e-10
    for E in (-0.05, -0.8, -2.0):
        got, ref, _ = run_both(ctx, [ax] * 3, 5, rho, quad(), energy=E)
        assert rel_maxdiff(got, ref) < 1e-10, E


def test_apply_different_input_and_output_grids(ctx):
    # sub-grid output with no coda (the GBFMM near-field shape)
    rng = np.random.default_rng(9)
    ain = [O.Axis.ncell_fixed(-6.0, 6.0, 7, 10)] * 3
    aout = [O.Axis.ncell_fixed(-2.4, 1.2, 7, 3), O.Axis.ncell_fixed(-6.0, 0.0, 7, 5),
            O.Axis.ncell_fixed(0.0, 3.6, 7, 3)]
    q = quad()
    q_, alpha, c = random_gaussians(rng, 3)
    rho = O.gauss_density(q_, alpha, c, *(a.points for a in ain))
    gin, gout = lib_grid(ctx, ain, 7), lib_grid(ctx, aout, 7)
    op = dage_amd.Coulomb3D(ctx, gin, gout, q, coda=False)
    got = op.transform_cube(rho)
    fx = O.exponentials(ain[0], aout[0], q.tp, False)
    fy = O.exponentials(ain[1], aout[1], q.tp, True)
    fz = O.exponentials(ain[2], aout[2], q.tp, True)
    ref = O.transform_cube(fx, fy, fz, O.coulomb_weights(q.tw), 0.0, rho)
    assert got.shape == tuple(a.n for a in aout)
    assert rel_maxdiff(got, ref) < 1e-10
    with pytest.raises(dage_amd.DageError):
        dage_amd.Coulomb3D(ctx, gin, gout, q, coda=True)


@pytest.mark.parametrize("case", load_golden("gauss_test_selfint.json")["cases"],
                         ids=lambda c: c["name"])
def test_reference_goldens_through_gpu(ctx, case):
    axes = gauss_generator_axes(case["q"], case["alpha"], case["centers"], case["step"],
                                case["nlip"], case["thres"])
    g = lib_grid(ctx, axes, case["nlip"])
    op = dage_amd.Coulomb3D(ctx, g, g, dage_amd.GaussQuad(10, 8, 2.0, 500.0))
    rho = O.gauss_density(case["q"], case["alpha"], case["centers"], *(a.points for a in axes))
    V = op.transform_cube(rho)
    d_rho = dage_amd.DeviceCube(ctx, rho.shape, rho)
    d_v = dage_amd.DeviceCube(ctx, V.shape, V)
    e = g.integrate(d_rho, d_v)
    ref = case["selfint_numerical"]
    print(f"{case['name']}: {e!r} vs {ref!r}")
    assert abs(e - ref) < 1e-13 * abs(ref)


def test_device_apply_and_chunking(ctx, monkeypatch):
    rng = np.random.default_rng(21)
    ax = O.Axis.ncell_fixed(-5.0, 5.0, 7, 6)
    q = quad()
    g = lib_grid(ctx, [ax] * 3, 7)
    op = dage_amd.Coulomb3D(ctx, g, g, q)
    cube = np.asfortranarray(rng.standard_normal((ax.n,) * 3))
    host = op.transform_cube(cube)
    d_in = dage_amd.DeviceCube(ctx, cube.shape, cube)
    d_out = dage_amd.DeviceCube(ctx, cube.shape)
    op.transform_cuda_cube(d_in, d_out)
    assert np.array_equal(d_out.download(), host)
    d_out.set_to_zero()
    assert not d_out.download().any()
    # integral of the cube with the LIP weights matches the oracle
    w = [ax.ints()] * 3
    assert abs(g.integrate(d_in) - O.cube_integral(*w, cube)) < 1e-12 * np.abs(cube).sum()
    # a tiny workspace budget forces several t-chunks; result unchanged to rounding
    full = op.workspace_bytes()
    monkeypatch.setenv("DAGE_APPLY_WORKSPACE_MB", "1")
    op2 = dage_amd.Coulomb3D(ctx, g, g, q)
    chunked = op2.transform_cube(cube)
    assert rel_maxdiff(chunked, host) < 1e-14
    assert op2.workspace_bytes() < full


def test_apply_block_sparse_matrices(ctx, monkeypatch):
    """Banded matrices with zeroed 16-row panels and an all-zero t: the GEMMs
    skip K panels that are zero on a block's columns.  A skipped panel adds
    exactly zero, so every column-block count gives bit-identical results,
    and they match the oracle's dense apply."""
    rng = np.random.default_rng(77)
    axes = [O.Axis.ncell_fixed(-4.0, 4.0, 5, 15), O.Axis.ncell_fixed(-3.0, 3.0, 5, 11),
            O.Axis.ncell_fixed(-5.0, 5.0, 5, 18)]
    n = [a.n for a in axes]
    q = quad()
    T = q.get_npoints()
    mats = []
    for a in range(3):
        F = rng.standard_normal((n[a], n[a], T))
        i, j = np.meshgrid(np.arange(n[a]), np.arange(n[a]), indexing="ij")
        for t in range(T):
            bw = int(rng.integers(1, n[a]))
            F[np.abs(i - j) > bw, t] = 0.0
            if t % 3 == 0:  # zero a random 16-row panel of input points
                k0 = 16 * int(rng.integers(0, (n[a] + 15) // 16))
                F[:, k0:k0 + 16, t] = 0.0
        F[:, :, 7] = 0.0
        mats.append(np.asfortranarray(F))
    # library layouts: Fx (out, in, t); Fy, Fz (in, out, t)
    fx = mats[0]
    fy = np.asfortranarray(mats[1].transpose(1, 0, 2))
    fz = np.asfortranarray(mats[2].transpose(1, 0, 2))
    cube = np.asfortranarray(rng.standard_normal(n))
    g = lib_grid(ctx, axes, 5)
    ref = O.transform_cube(fx, fy, fz, O.coulomb_weights(q.tw), 0.0, cube, add_coda=False)
    outs = []
    for nb in ("1,1,1", "3,5,2", "8,8,8"):
        monkeypatch.setenv("DAGE_APPLY_NB", nb)
        op = dage_amd.Coulomb3D(ctx, g, g, q, coda=False)
        op.set_matrices(fx, fy, fz)
        outs.append(op.transform_cube(cube))
        op.close()
    err = rel_maxdiff(outs[0], ref)
    print(f"block-sparse matrices vs oracle: {err:.3e}")
    assert err < 1e-12
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_profiling_counts_kernel_launches(ctx):
    ax = O.Axis.ncell_fixed(-3.0, 3.0, 7, 3)
    g = lib_grid(ctx, [ax] * 3, 7)
    op = dage_amd.Coulomb3D(ctx, g, g, quad())
    ctx.profile(True)
    op.transform_cube(np.ones((ax.n,) * 3, order="F"))
    ms, n = ctx.profile_query("gemm_f64_zmode")
    ms_b, n_b = ctx.profile_query("band_zmode")  # a 19-point axis is banded at every t
    ctx.profile(False)
    assert n + n_b >= 1 and ms + ms_b > 0.0


def test_error_paths(ctx):
    ax = O.Axis.ncell_fixed(-3.0, 3.0, 7, 3)
    g = lib_grid(ctx, [ax] * 3, 7)
    op = dage_amd.Coulomb3D(ctx, g, g, quad())
    with pytest.raises(dage_amd.DageError) as e:
        op.transform_cube(np.ones((ax.n,) * 3, order="F"), precision=7)
    assert e.value.code == 5
    with pytest.raises(ValueError):
        op.transform_cube(np.ones((ax.n + 1, ax.n, ax.n), order="F"))
    ax5 = O.Axis.ncell_fixed(-3.0, 3.0, 5, 3)
    g5 = lib_grid(ctx, [ax5] * 3, 5)
    with pytest.raises(dage_amd.DageError):
        dage_amd.Coulomb3D(ctx, g, g5, quad(), coda=False)


def test_apply_box_banded_matrices_pitch_tiles(ctx, monkeypatch):
    """Near-field style matrices: entry (i, j) is nonzero only when the boxes
    of 30 points holding i and j are neighbours.  The column tiles then follow
    the boxes (a pitch of 30 or 60, not a multiple of 16); the result is
    bit-identical to 16-column tiles and matches the oracle."""
    rng = np.random.default_rng(78)
    ax = O.Axis.ncell_fixed(-4.0, 4.0, 7, 20)  # 121 points = 4 boxes of 30 (+1)
    n = ax.n
    q = quad()
    T = q.get_npoints()
    box = np.minimum(np.arange(n) // 30, 3)
    near = np.abs(box[:, None] - box[None, :]) <= 1
    mats = []
    for a in range(3):
        F = rng.standard_normal((n, n, T))
        F[~near, :] = 0.0
        mats.append(np.asfortranarray(F))
    fx = mats[0]
    fy = np.asfortranarray(mats[1].transpose(1, 0, 2))
    fz = np.asfortranarray(mats[2].transpose(1, 0, 2))
    cube = np.asfortranarray(rng.standard_normal((n,) * 3))
    g = lib_grid(ctx, [ax] * 3, 7)
    ref = O.transform_cube(fx, fy, fz, O.coulomb_weights(q.tw), 0.0, cube, add_coda=False)
    op = dage_amd.Coulomb3D(ctx, g, g, q, coda=False)
    op.set_matrices(fx, fy, fz)
    tiling = op.tiling()
    print("tiling", tiling)
    # x and y rows are one t each; the z pass's K runs over (zi, t), so its
    # panels straddle t and may not pay for a pitch
    assert all(p > 0 and p % 30 == 0 for p in tiling["pitch"][:2])
    got = op.transform_cube(cube)
    op.close()
    monkeypatch.setenv("DAGE_APPLY_NB", "2,2,2")
    op = dage_amd.Coulomb3D(ctx, g, g, q, coda=False)
    op.set_matrices(fx, fy, fz)
    assert op.tiling()["pitch"] == [0, 0, 0]
    grid16 = op.transform_cube(cube)
    op.close()
    err = rel_maxdiff(got, ref)
    print(f"box-banded matrices vs oracle: {err:.3e}")
    assert err < 1e-12
    assert np.array_equal(got, grid16)


def test_direct_z_pass_accumulates_over_t_chunks(ctx, monkeypatch):
    """One right-hand side, no coda and a z pass without K split: the z pass
    accumulates every t-chunk straight into the output (no partials, no
    reduction).  A 1 MB workspace forces one t-point per chunk at 37^3 (40
    accumulating chunks, K = 37, so no split); the result matches the oracle
    and the split-K path (full workspace), with dense and with box-banded
    (pitch-tiled) matrices."""
    rng = np.random.default_rng(91)
    ax = O.Axis.ncell_fixed(-6.0, 6.0, 7, 6)  # 37 points
    n = ax.n
    q = quad()
    T = q.get_npoints()
    g = lib_grid(ctx, [ax] * 3, 7)
    cube = np.asfortranarray(rng.standard_normal((n,) * 3))
    fx, fy, fz = O.coulomb3d_matrices([ax] * 3, [ax] * 3, q.tp)
    box = np.minimum(np.arange(n) // 12, 2)
    near = np.abs(box[:, None] - box[None, :]) <= 1
    banded = []
    for a in range(3):
        F = rng.standard_normal((n, n, T))
        F[~near, :] = 0.0
        banded.append(np.asfortranarray(F))
    bfx = banded[0]
    bfy = np.asfortranarray(banded[1].transpose(1, 0, 2))
    bfz = np.asfortranarray(banded[2].transpose(1, 0, 2))
    for label, mats in (("coulomb", (fx, fy, fz)), ("box-banded", (bfx, bfy, bfz))):
        ref = O.transform_cube(*mats, O.coulomb_weights(q.tw), 0.0, cube, add_coda=False)
        monkeypatch.delenv("DAGE_APPLY_WORKSPACE_MB", raising=False)
        op = dage_amd.Coulomb3D(ctx, g, g, q, coda=False)
        op.set_matrices(*mats)
        split = op.transform_cube(cube)
        full_ws = op.workspace_bytes()
        op.close()
        monkeypatch.setenv("DAGE_APPLY_WORKSPACE_MB", "1")
        op = dage_amd.Coulomb3D(ctx, g, g, q, coda=False)
        op.set_matrices(*mats)
        direct = op.transform_cube(cube)
        # V, U of one t-point and one unsplit z-pass partial (unused: the
        # z pass writes the output directly)
        assert op.workspace_bytes() == 3 * 8 * n ** 3 < full_ws
        op.close()
        e_ref, e_split = rel_maxdiff(direct, ref), rel_maxdiff(direct, split)
        print(f"{label}: direct z pass vs oracle {e_ref:.2e}, vs split-K path {e_split:.2e}")
        assert e_ref < 1e-12
        assert e_split < 1e-14


def test_flushed_tails_change_the_apply_at_rounding_level(ctx, monkeypatch):
    """The operator stores entries at most 1e-20 of their row's largest as
    zeros, and the passes skip the panels they leave empty.  On a 193-point
    axis (h = 0.1, the 385^3 headline's spacing) at T = 40 the flush empties
    a large share of the 16x16 blocks the exact zeros leave; the apply moves
    by less than 1e-15 of max|V| (the bound is 193 * 1e-20 per pass relative
    to max|F_t| max|g|), for a smooth density and for random input."""
    ax = O.Axis.ncell_fixed(-9.6, 9.6, 7, 32)
    assert ax.n == 193
    q = quad()
    g = lib_grid(ctx, [ax] * 3, 7)
    rng = np.random.default_rng(5)
    P = ax.points
    rho = O.gauss_density(np.array([1.0, -0.5]), np.array([1.5, 0.7]),
                          np.array([[0.3, -0.2, 0.1], [-3.0, 2.5, 4.0]]), P, P, P)
    rnd = np.asfortranarray(rng.uniform(-1.0, 1.0, (ax.n,) * 3))
    res, blocks = {}, {}
    # the same (MFMA) passes for both, so that only the flush differs (the
    # wider unflushed bands would leave fewer t-points to the banded passes,
    # and the low-rank factors of the small t would differ at 1e-14)
    monkeypatch.setenv("DAGE_APPLY_BAND", "0")
    monkeypatch.setenv("DAGE_APPLY_LOWRANK", "0")
    for cut in ("0", None):
        if cut is None:
            monkeypatch.delenv("DAGE_FLUSH_CUTOFF", raising=False)
        else:
            monkeypatch.setenv("DAGE_FLUSH_CUTOFF", cut)
        op = dage_amd.Helmholtz3D(ctx, g, g, q, energy=0.0)
        op.set_energy(-0.3)
        fx = op.get_matrices()[0]
        nz = np.abs(fx) > 0
        pad = np.zeros((208, 208, fx.shape[2]), bool)
        pad[:193, :193] = nz
        blocks[cut] = pad.reshape(13, 16, 13, 16, -1).any(axis=(1, 3)).mean()
        res[cut] = (op.transform_cube(rho), op.transform_cube(rnd))
        op.close()
    print(f"nonzero 16x16 blocks of Fx: {blocks['0']:.3f} exact zeros, {blocks[None]:.3f} flushed")
    assert blocks[None] < 0.85 * blocks["0"]
    for i, name in enumerate(("density", "random")):
        e = rel_maxdiff(res[None][i], res["0"][i])
        print(f"{name}: flushed vs unflushed apply {e:.2e}")
        assert e < 1e-15


@pytest.mark.parametrize("n,energy", [(97, None), (97, -0.4), (7, None), (61, -1.5)])
def test_banded_passes_match_the_mfma_passes(ctx, monkeypatch, n, energy):
    """t-points whose Fx_t, Fy_t and Fz_t keep every nonzero entry within 8
    points of the diagonal (for that t and all larger ones) take banded
    passes on the FP64 VALU.  At h = 0.1 and T = 40 the large-t tail is
    banded (every t on a 7-point axis); with DAGE_APPLY_BAND=0 every t takes
    the MFMA passes.  The same sums in another order: 1e-14 of max|V|; the
    oracle's dense apply with the operator's own matrices: 1e-12."""
    ax = O.Axis.ncell_fixed(-0.05 * (n - 1), 0.05 * (n - 1), 7, (n - 1) // 6)
    assert ax.n == n
    q = quad()
    T = q.get_npoints()
    g = lib_grid(ctx, [ax] * 3, 7)
    P = ax.points
    rho = O.gauss_density(np.array([1.0, 0.4]), np.array([2.0, 0.8]),
                          np.array([[0.2, -0.1, 0.3], [-0.9, 0.7, -0.4]]), P, P, P)
    res = {}
    for flag in ("2", "0"):  # 2: banded passes also where the grid is small
        monkeypatch.setenv("DAGE_APPLY_BAND", flag)
        if energy is None:
            op = dage_amd.Coulomb3D(ctx, g, g, q)
        else:
            op = dage_amd.Helmholtz3D(ctx, g, g, q, energy=0.0)
            op.set_energy(energy)
        res[flag] = (op.transform_cube(rho), op.tiling()["band_t0"], op.get_matrices())
        op.close()
    t0 = res["2"][1]
    print(f"n {n}: banded from t-point {t0} of {T}")
    assert res["0"][1] == T and 0 <= t0 < T
    if n == 7:
        assert t0 == 0
    e = rel_maxdiff(res["2"][0], res["0"][0])
    print(f"banded vs MFMA passes: {e:.2e}")
    assert e < 1e-14
    fx, fy, fz = res["2"][2]
    w = O.coulomb_weights(q.tw) if energy is None else O.helmholtz_weights(q.tp, q.tw, energy)
    ref = O.transform_cube(fx, fy, fz, w, O.coda(q.tlog), rho)
    e = rel_maxdiff(res["2"][0], ref)
    print(f"banded apply vs oracle: {e:.2e}")
    assert e < 1e-12


@pytest.mark.parametrize("n,energy", [(97, None), (97, -0.4), (193, -1.2)])
def test_lowrank_points_match_the_dense_passes(ctx, monkeypatch, n, energy):
    """The smallest t-points, whose 1D matrices are numerically of low rank
    (truncated at a residual column norm of 1e-14 of the largest), are
    applied through their factors (reduce each axis to the rank, expand
    back).  Against the dense passes (DAGE_APPLY_LOWRANK=0) the apply moves
    by the truncation, ~1e-14 of max|V| (bar 1e-13); against the oracle's
    dense apply with the operator's own matrices 1e-12."""
    ax = O.Axis.ncell_fixed(-0.05 * (n - 1), 0.05 * (n - 1), 7, (n - 1) // 6)
    assert ax.n == n
    q = quad()
    g = lib_grid(ctx, [ax] * 3, 7)
    P = ax.points
    rho = O.gauss_density(np.array([1.0, 0.4]), np.array([2.0, 0.8]),
                          np.array([[0.2, -0.1, 0.3], [-0.9, 0.7, -0.4]]), P, P, P)
    rng = np.random.default_rng(9)
    rnd = np.asfortranarray(rng.uniform(-1.0, 1.0, rho.shape))
    res = {}
    for flag in ("2", "0"):  # 2: also below the size where it pays
        monkeypatch.setenv("DAGE_APPLY_LOWRANK", flag)
        if energy is None:
            op = dage_amd.Coulomb3D(ctx, g, g, q)
        else:
            op = dage_amd.Helmholtz3D(ctx, g, g, q, energy=0.0)
            op.set_energy(energy)
        til = op.tiling()
        res[flag] = (op.transform_cube(rho), op.transform_cube(rnd), til, op.get_matrices())
        op.close()
    t1, kx = res["2"][2]["lowrank_t1"], res["2"][2]["lowrank_kx"]
    print(f"n {n}: {t1} low-rank t-points, x ranks summing to {kx}")
    assert res["0"][2]["lowrank_t1"] == 0 and t1 >= 3
    for i, name in enumerate(("density", "random")):
        e = rel_maxdiff(res["2"][i], res["0"][i])
        print(f"{name}: low-rank vs dense passes {e:.2e}")
        assert e < 1e-13
    if n <= 97:
        fx, fy, fz = res["2"][3]
        w = O.coulomb_weights(q.tw) if energy is None else O.helmholtz_weights(q.tp, q.tw, energy)
        ref = O.transform_cube(fx, fy, fz, w, O.coda(q.tlog), rho)
        e = rel_maxdiff(res["2"][0], ref)
        print(f"low-rank apply vs oracle: {e:.2e}")
        assert e < 1e-12


@pytest.mark.parametrize("n,energy", [(61, -0.7), (97, None)])
def test_fused_banded_xy_pass_is_bit_identical(ctx, monkeypatch, n, energy):
    """The banded x and y passes in one kernel (V kept in LDS) against the
    two separate banded kernels: the same sums in the same order, so the
    apply is bit-identical (axes of 61 and 97 points leave partial 56-point
    x tiles and 48-point y tiles)."""
    ax = O.Axis.ncell_fixed(-0.05 * (n - 1), 0.05 * (n - 1), 7, (n - 1) // 6)
    q = quad()
    g = lib_grid(ctx, [ax] * 3, 7)
    P = ax.points
    rng = np.random.default_rng(21)
    rho = O.gauss_density(np.array([1.0, -0.3]), np.array([1.7, 0.6]),
                          np.array([[0.1, 0.2, -0.3], [1.1, -0.8, 0.5]]), P, P, P)
    rho = np.asfortranarray(rho + 1e-3 * rng.uniform(-1.0, 1.0, rho.shape))
    monkeypatch.setenv("DAGE_APPLY_BAND", "2")
    if energy is None:
        op = dage_amd.Coulomb3D(ctx, g, g, q)
    else:
        op = dage_amd.Helmholtz3D(ctx, g, g, q, energy=0.0)
        op.set_energy(energy)
    assert op.tiling()["band_t0"] < q.get_npoints()
    res = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("DAGE_APPLY_BANDXY", flag)
        res[flag] = op.transform_cube(rho)
    op.close()
    ndiff = int(np.count_nonzero(res["1"] != res["0"]))
    print(f"n {n}: fused banded x/y vs separate: {ndiff} points differ")
    assert ndiff == 0


@pytest.mark.parametrize("n,energy", [(97, -0.5), (61, None)])
def test_t_grouped_x_pass_is_bit_identical(ctx, monkeypatch, n, energy):
    """The x pass with 4 t-points per block sharing each loaded input panel
    (32-column tiles of 4 t) against one t per block: the same K panels (the
    extra ones of the union over t add exact zeros) in the same order, so
    the apply is bit-identical; t-chunks of 3 and 4 (the last group of a
    chunk short)."""
    ax = O.Axis.ncell_fixed(-0.05 * (n - 1), 0.05 * (n - 1), 7, (n - 1) // 6)
    q = quad()
    g = lib_grid(ctx, [ax] * 3, 7)
    rng = np.random.default_rng(33)
    cube = np.asfortranarray(rng.uniform(-1.0, 1.0, (ax.n,) * 3))
    monkeypatch.setenv("DAGE_APPLY_NB", "2,2,2")
    for tchunk in ("3", "4"):
        monkeypatch.setenv("DAGE_APPLY_TCHUNK", tchunk)
        if energy is None:
            op = dage_amd.Coulomb3D(ctx, g, g, q)
        else:
            op = dage_amd.Helmholtz3D(ctx, g, g, q, energy=0.0)
            op.set_energy(energy)
        res = {}
        for flag in ("1", "0"):
            monkeypatch.setenv("DAGE_APPLY_TGROUP", flag)
            res[flag] = op.transform_cube(cube)
        op.close()
        ndiff = int(np.count_nonzero(res["1"] != res["0"]))
        print(f"n {n}, t-chunk {tchunk}: t-grouped vs one-t x pass: {ndiff} points differ")
        assert ndiff == 0


def test_apply_far_entries_dominate(ctx):
    """Two compact unit charges 20 bohr apart on a 161^3 grid: around the
    midpoint the density is zero to underflow and the whole potential comes
    from matrix entries 8-12 bohr off the diagonal, where the closed forms of
    the operator matrices cancel most.  GPU apply (GPU-generated matrices)
    against the oracle on that region, 1e-10 of the region's max|V|."""
    _far_entries(ctx, None)


@pytest.mark.parametrize("energy", [-0.05])
def test_apply_far_entries_helmholtz(ctx, energy):
    _far_entries(ctx, energy)


def _far_entries(ctx, energy):
    ax = O.Axis.ncell_fixed(-12.0, 12.0, 5, 40)  # 161 points, h = 0.15
    assert ax.n == 161
    q = quad()
    qs, alpha = np.ones(2), np.array([2.0, 3.0])
    centers = np.array([[-10.0, 0.3, -0.2], [10.0, -0.4, 0.1]])
    rho = O.gauss_density(qs, alpha, centers, ax.points, ax.points, ax.points)
    g = lib_grid(ctx, [ax] * 3, 5)
    if energy is None:
        op = dage_amd.Coulomb3D(ctx, g, g, q)
        w = O.coulomb_weights(q.tw)
    else:
        op = dage_amd.Helmholtz3D(ctx, g, g, q, energy=0.0)
        op.set_energy(energy)
        w = O.helmholtz_weights(q.tp, q.tw, energy)
    got = op.transform_cube(rho)
    # oracle on the block |x| <= 3, |y| <= 1.5, |z| <= 0.6 only (output
    # columns of the matrices restricted; the coda term is pointwise)
    xs = np.nonzero(np.abs(ax.points) <= 3.0)[0]
    ys = np.nonzero(np.abs(ax.points) <= 1.5)[0]
    zs = np.nonzero(np.abs(ax.points) <= 0.6)[0]
    fx, fy, fz = O.coulomb3d_matrices([ax] * 3, [ax] * 3, q.tp)
    ref = O.transform_cube(fx[xs], fy[:, ys], fz[:, zs], w, 0.0, rho, add_coda=False)
    if energy is None:
        ref += O.coda(q.tlog) * rho[np.ix_(xs, ys, zs)]
    blk = got[np.ix_(xs, ys, zs)]
    assert np.max(np.abs(rho[np.ix_(xs, ys, zs)])) < 1e-30 * np.max(np.abs(rho))
    err = rel_maxdiff(blk, ref)
    print(f"far-entry block (E={energy}): max|V| there {np.max(np.abs(ref)):.4e}, "
          f"GPU vs oracle {err:.3e}")
    assert err < 1e-10
    if energy is None:
        Va = O.gauss_potential(qs, alpha, centers, ax.points[xs], ax.points[ys], ax.points[zs])
        assert rel_maxdiff(blk, Va) < 1e-4
