"""Sub-cubes with leading dimensions at the C-ABI (the explicit `ld` that
replaces the reference's pitched CudaCube slices and sub-cubes,
src/bubbles/cube.cu:1042-1177): the operator apply and the bubbles inject on
boxes inside larger cubes equal the contiguous calls bit for bit, leave the
rest of the parent cubes untouched, and refuse leading dimensions below the
box.
"""
import numpy as np
import pytest

import dage_amd
import oracle as O
from _helpers import random_gaussians, slater_bubbles

pytestmark = pytest.mark.gpu


def test_apply_on_sub_cubes(ctx):
    ax = O.Axis.ncell_fixed(-5.0, 5.0, 7, 6)   # 37^3
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
    op = dage_amd.Coulomb3D(ctx, g, g, dage_amd.GaussQuad(12, 12))
    rng = np.random.default_rng(8)
    rho = O.gauss_density(*random_gaussians(rng, 3), ax.points, ax.points, ax.points)
    a, b = dage_amd.DeviceCube(ctx, rho.shape, rho), dage_amd.DeviceCube(ctx, rho.shape)
    op.transform_cuda_cube(a, b)
    ref = b.download()
    big_in = np.asfortranarray(rng.standard_normal((45, 41, 40)))
    big_in[3:40, 2:39, 1:38] = rho
    P = dage_amd.DeviceCube(ctx, big_in.shape, big_in)
    fill = np.asfortranarray(rng.standard_normal((39, 44, 41)))
    Q = dage_amd.DeviceCube(ctx, fill.shape, fill)
    # strided in and out, strided in only, strided out only
    op.transform_cuda_cube_strided(P.view((3, 2, 1), (37, 37, 37)), Q.view((1, 5, 2), (37, 37, 37)))
    got = Q.download()
    assert np.array_equal(got[1:38, 5:42, 2:39], ref)
    mask = np.ones(fill.shape, bool)
    mask[1:38, 5:42, 2:39] = False
    assert np.array_equal(got[mask], fill[mask])
    c = dage_amd.DeviceCube(ctx, rho.shape)
    op.transform_cuda_cube_strided(P.view((3, 2, 1), (37, 37, 37)), c)
    assert np.array_equal(c.download(), ref)
    Q.upload(fill)
    op.transform_cuda_cube_strided(a, Q.view((2, 0, 4), (37, 37, 37)))
    assert np.array_equal(Q.download()[2:39, 0:37, 4:41], ref)
    with pytest.raises(dage_amd.DageError):
        dage_amd._lib.call("dage_operator_apply_device_strided", op.handle, P.ptr, 36, 41,
                           Q.ptr, 39, 44, 0)
    with pytest.raises(ValueError):
        P.view((10, 0, 0), (37, 37, 37))
    for obj in (a, b, c, P, Q, op, g):
        obj.close()


def test_inject_into_sub_cube(ctx):
    ax = O.Axis.ncell_fixed(-4.0, 4.0, 7, 5)   # 31^3
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
    rng = np.random.default_rng(9)
    bset, Z = slater_bubbles(rng, 3, lmax=3, box=2.5)
    bub = dage_amd.Bubbles(ctx, 3, 3, 0, 7)
    for i in range(3):
        bub.set_bubble(i, bset.centers[i], Z[i], bset.steps[i], bset.bf[i])
    base = np.asfortranarray(rng.standard_normal((31, 31, 31)))
    d = dage_amd.DeviceCube(ctx, base.shape, base)
    bub.inject(g, d, lmin=1, multiplier=0.5)
    ref = d.download()
    big = np.asfortranarray(rng.standard_normal((36, 33, 35)))
    big[4:35, 1:32, 3:34] = base
    B = dage_amd.DeviceCube(ctx, big.shape, big)
    bub.inject_strided(g, B.view((4, 1, 3), (31, 31, 31)), lmin=1, multiplier=0.5)
    got = B.download()
    assert np.array_equal(got[4:35, 1:32, 3:34], ref)
    mask = np.ones(big.shape, bool)
    mask[4:35, 1:32, 3:34] = False
    assert np.array_equal(got[mask], big[mask])
    for obj in (bub, d, B, g):
        obj.close()
