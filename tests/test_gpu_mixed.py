"""Mixed-precision (split-BF16) apply against the oracle and the FP64 path.

bf16x6: 3-term bf16 splits, 6 MFMA products, FP32 sums of 32 products
promoted to FP64, FP32 intermediates.  Accuracy bar: <= 1e-7 of max|V|
on the configs' Gaussian densities (CPU emulation of the same arithmetic:
5.2e-8 at 97^3, T=40).  bf16x3 is the faster 2-term variant (~2e-6).
"""
import numpy as np
import pytest

import dage_amd
import oracle as O
from _helpers import random_gaussians, rel_maxdiff

pytestmark = pytest.mark.gpu

BF16X6, BF16X3 = dage_amd.DAGE_PREC_BF16X6, dage_amd.DAGE_PREC_BF16X3


def setup(ctx, axes, nlip, coda=True):
    g = dage_amd.Grid3D(ctx, [a.qmin for a in axes], [a.steps for a in axes], nlip)
    q = dage_amd.GaussQuad(nlin=20, nlog=20, tlin=2.0, tlog=500.0)
    return g, q, dage_amd.Coulomb3D(ctx, g, g, q, coda=coda)


@pytest.mark.parametrize("ncell", [8, 16])
def test_bf16x6_config1_shape(ctx, ncell):
    rng = np.random.default_rng(0xDA6E0001)
    ax = O.Axis.ncell_fixed(-8.0, 8.0, 7, ncell)
    q_, alpha, c = random_gaussians(rng, 4)
    rho = O.gauss_density(q_, alpha, c, ax.points, ax.points, ax.points)
    g, q, op = setup(ctx, [ax] * 3, 7)
    v64 = op.transform_cube(rho)
    v6 = op.transform_cube(rho, precision=BF16X6)
    v3 = op.transform_cube(rho, precision=BF16X3)
    e6, e3 = rel_maxdiff(v6, v64), rel_maxdiff(v3, v64)
    print(f"N={ax.n}: bf16x6 {e6:.3e}  bf16x3 {e3:.3e} (vs FP64 GPU)")
    assert e6 < 1e-7
    assert e3 < 2e-5
    fx, fy, fz = O.coulomb3d_matrices([ax] * 3, [ax] * 3, q.tp)
    ref = O.transform_cube(fx, fy, fz, O.coulomb_weights(q.tw), O.coda(500.0), rho)
    assert rel_maxdiff(v6, ref) < 1e-7


def test_bf16x6_uneven_helmholtz_nocoda(ctx):
    rng = np.random.default_rng(8)
    axes = [O.Axis.ncell_fixed(-6.0, 6.0, 5, 10), O.Axis.ncell_fixed(-5.0, 4.0, 5, 9),
            O.Axis.ncell_fixed(-4.0, 6.0, 5, 11)]
    q_, alpha, c = random_gaussians(rng, 3)
    rho = O.gauss_density(q_, alpha, c, *(a.points for a in axes))
    g = dage_amd.Grid3D(ctx, [a.qmin for a in axes], [a.steps for a in axes], 5)
    q = dage_amd.GaussQuad(nlin=20, nlog=20, tlin=2.0, tlog=500.0)
    op = dage_amd.Helmholtz3D(ctx, g, g, q, energy=-1.3, coda=False)
    v64 = op.transform_cube(rho)
    v6 = op.transform_cube(rho, precision=BF16X6)
    assert rel_maxdiff(v6, v64) < 1e-7


def test_mixed_device_path_and_chunking(ctx, monkeypatch):
    rng = np.random.default_rng(12)
    ax = O.Axis.ncell_fixed(-5.0, 5.0, 7, 6)
    q_, alpha, c = random_gaussians(rng, 2)
    rho = O.gauss_density(q_, alpha, c, ax.points, ax.points, ax.points)
    g, q, op = setup(ctx, [ax] * 3, 7)
    d_in = dage_amd.DeviceCube(ctx, rho.shape, rho)
    d_out = dage_amd.DeviceCube(ctx, rho.shape)
    op.transform_cuda_cube(d_in, d_out, precision=BF16X6)
    full = d_out.download()
    monkeypatch.setenv("DAGE_APPLY_WORKSPACE_MB", "1")
    op2 = dage_amd.Coulomb3D(ctx, g, g, q)
    chunked = op2.transform_cube(rho, precision=BF16X6)
    # t-chunks change the split-K partition of the z pass, so the FP32 chunk
    # sums round differently: agreement at the scheme's own error level
    assert rel_maxdiff(chunked, full) < 1e-7
    assert rel_maxdiff(full, op.transform_cube(rho)) < 1e-7


def test_mixed_follows_set_matrices(ctx):
    """The split-BF16 planes are rebuilt when the matrices are replaced."""
    rng = np.random.default_rng(21)
    ax = O.Axis.ncell_fixed(-4.0, 4.0, 5, 6)
    q_, alpha, c = random_gaussians(rng, 2)
    rho = O.gauss_density(q_, alpha, c, ax.points, ax.points, ax.points)
    g, q, op = setup(ctx, [ax] * 3, 5, coda=False)
    v1 = op.transform_cube(rho, precision=BF16X6)
    fx, fy, fz = op.get_matrices()
    op.set_matrices(2.0 * fx, fy, fz)
    v2 = op.transform_cube(rho, precision=BF16X6)
    assert rel_maxdiff(v2, 2.0 * v1) < 1e-7


def test_mixed_block_sparse_matrices(ctx, monkeypatch):
    """Split-BF16 passes skip 32-deep K steps that are zero on a block's
    columns: bit-identical across column-block counts, and within the bf16x6
    bar of the FP64 apply of the same banded, panel-zeroed matrices."""
    rng = np.random.default_rng(78)
    axes = [O.Axis.ncell_fixed(-4.0, 4.0, 5, 15), O.Axis.ncell_fixed(-3.0, 3.0, 5, 11),
            O.Axis.ncell_fixed(-5.0, 5.0, 5, 18)]
    n = [a.n for a in axes]
    q = dage_amd.GaussQuad(nlin=20, nlog=20, tlin=2.0, tlog=500.0)
    T = q.get_npoints()
    mats = []
    for a in range(3):
        i, j = np.meshgrid(np.arange(n[a]), np.arange(n[a]), indexing="ij")
        F = np.exp(-0.05 * (i - j) ** 2)[:, :, None] * rng.uniform(0.5, 1.5, (1, 1, T))
        for t in range(T):
            bw = int(rng.integers(2, n[a]))
            F[np.abs(i - j) > bw, t] = 0.0
            if t % 3 == 0:
                k0 = 32 * int(rng.integers(0, (n[a] + 31) // 32))
                F[:, k0:k0 + 32, t] = 0.0
        mats.append(np.asfortranarray(F))
    fx = mats[0]
    fy = np.asfortranarray(mats[1].transpose(1, 0, 2))
    fz = np.asfortranarray(mats[2].transpose(1, 0, 2))
    cube = np.asfortranarray(1.0 + rng.uniform(0.0, 1.0, n))
    g = dage_amd.Grid3D(ctx, [a.qmin for a in axes], [a.steps for a in axes], 5)
    outs = []
    for nb in ("1,1,1", "3,5,2", "8,8,8"):
        monkeypatch.setenv("DAGE_APPLY_NB", nb)
        op = dage_amd.Coulomb3D(ctx, g, g, q, coda=False)
        op.set_matrices(fx, fy, fz)
        if nb == "1,1,1":
            ref = op.transform_cube(cube)
        outs.append(op.transform_cube(cube, precision=BF16X6))
        op.close()
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    err = rel_maxdiff(outs[0], ref)
    print(f"block-sparse bf16x6 vs FP64: {err:.3e}")
    assert err < 2e-7  # toy shapes: the bar of test_gpu_batch.py (1.1e-7 measured here)


def test_mixed_pitch_tiles(ctx, monkeypatch):
    """Box-banded matrices (nonzero only between neighbouring 30-point boxes):
    the split-BF16 passes use the operator's box-following column tiles; the
    K-step order inside a tile is unchanged, so the result is bit-identical
    to 16-column tiles, and within the bf16x6 bar of the FP64 apply."""
    rng = np.random.default_rng(79)
    ax = O.Axis.ncell_fixed(-4.0, 4.0, 7, 20)  # 121 points = 4 boxes of 30 (+1)
    n = ax.n
    q = dage_amd.GaussQuad(nlin=20, nlog=20, tlin=2.0, tlog=500.0)
    T = q.get_npoints()
    box = np.minimum(np.arange(n) // 30, 3)
    near = np.abs(box[:, None] - box[None, :]) <= 1
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    mats = []
    for a in range(3):
        F = np.exp(-0.01 * (i - j) ** 2)[:, :, None] * rng.uniform(0.5, 1.5, (1, 1, T))
        F[~near, :] = 0.0
        mats.append(np.asfortranarray(F))
    fx = mats[0]
    fy = np.asfortranarray(mats[1].transpose(1, 0, 2))
    fz = np.asfortranarray(mats[2].transpose(1, 0, 2))
    cube = np.asfortranarray(1.0 + rng.uniform(0.0, 1.0, (n,) * 3))
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
    op = dage_amd.Coulomb3D(ctx, g, g, q, coda=False)
    op.set_matrices(fx, fy, fz)
    assert all(p > 0 for p in op.tiling()["pitch"][:2])
    ref = op.transform_cube(cube)
    got = op.transform_cube(cube, precision=BF16X6)
    op.close()
    monkeypatch.setenv("DAGE_APPLY_NB", "2,2,2")
    op = dage_amd.Coulomb3D(ctx, g, g, q, coda=False)
    op.set_matrices(fx, fy, fz)
    grid16 = op.transform_cube(cube, precision=BF16X6)
    op.close()
    assert np.array_equal(got, grid16)
    err = rel_maxdiff(got, ref)
    print(f"box-banded bf16x6 vs FP64: {err:.3e}")
    assert err < 2e-7
