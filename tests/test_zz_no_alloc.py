"""dage_operator_apply_device_noalloc: the FP64 apply with its intermediates in
the library's fixed per-GPU scratch array (no allocation), against
dage_operator_apply_device on two calls in a row, the second on a smaller
grid, and against the oracle.

Bars: with the same t-chunk (DAGE_APPLY_TCHUNK=1 for both) the two entry
points run the same kernels on the same intermediates, so they must agree bit
for bit; on their own t-chunking 1e-13 of max|V| (only the order of the
t-chunk sums differs); against the oracle the FP64 bar, 1e-10 of max|V|.
"""
import ctypes
import os

import numpy as np
import pytest

import dage_amd
import oracle as O
from _helpers import random_gaussians, rel_maxdiff
from dage_amd._lib import DageError, call

pytestmark = pytest.mark.gpu


def noalloc(op, a, b):
    call("dage_operator_apply_device_noalloc", op.handle, ctypes.c_void_p(a.ptr),
         ctypes.c_void_p(b.ptr))


def device_apply(op, a, b):
    call("dage_operator_apply_device", op.handle, ctypes.c_void_p(a.ptr), ctypes.c_void_p(b.ptr),
         dage_amd.DAGE_PREC_FP64)


def with_tchunk(value, fn):
    old = os.environ.get("DAGE_APPLY_TCHUNK")
    os.environ["DAGE_APPLY_TCHUNK"] = value
    try:
        return fn()
    finally:
        if old is None:
            os.environ.pop("DAGE_APPLY_TCHUNK", None)
        else:
            os.environ["DAGE_APPLY_TCHUNK"] = old


def one_case(ctx, ncell, nlip, energy, seed):
    ax = O.Axis.ncell_fixed(-6.0, 6.0, nlip, ncell)
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, nlip)
    q = dage_amd.GaussQuad(nlin=20, nlog=20, tlin=2.0, tlog=500.0)
    if energy is None:
        op = dage_amd.Coulomb3D(ctx, g, g, q)
        w = O.coulomb_weights(q.tw)
    else:
        op = dage_amd.Helmholtz3D(ctx, g, g, q, energy=0.0)
        op.set_energy(energy)
        w = O.helmholtz_weights(q.tp, q.tw, energy)
    rng = np.random.default_rng(seed)
    rho = O.gauss_density(*random_gaussians(rng, 4), ax.points, ax.points, ax.points)
    n = ax.n
    a = dage_amd.DeviceCube(ctx, (n, n, n), rho)
    b1 = dage_amd.DeviceCube(ctx, (n, n, n))
    b2 = dage_amd.DeviceCube(ctx, (n, n, n))
    # same t-chunk: bit for bit
    with_tchunk("1", lambda: (noalloc(op, a, b1), device_apply(op, a, b2)))
    got1, got2 = b1.download(), b2.download()
    print(f"{n}^3 E={energy}: t-chunk 1, {np.count_nonzero(got1 != got2)} points differ")
    assert np.array_equal(got1, got2)
    # own t-chunking
    noalloc(op, a, b1)
    device_apply(op, a, b2)
    got1, got2 = b1.download(), b2.download()
    d = rel_maxdiff(got1, got2)
    fx, fy, fz = O.coulomb3d_matrices([ax] * 3, [ax] * 3, q.tp)
    ref = O.transform_cube(fx, fy, fz, w, O.coda(q.tlog), rho)
    e = rel_maxdiff(got1, ref)
    print(f"{n}^3 E={energy}: vs apply_device {d:.3e}, vs oracle {e:.3e}")
    assert d < 1e-13
    assert e < 1e-10
    for obj in (a, b1, b2, op, g):
        obj.close()


def test_noalloc_apply_two_calls_second_smaller(ctx):
    limit = ctypes.c_size_t()
    call("dage_operator_noalloc_limit", ctypes.byref(limit))
    assert limit.value == 128 << 20
    one_case(ctx, 12, 7, None, 71)   # 73^3 Coulomb
    one_case(ctx, 8, 5, -0.7, 72)    # 33^3 Helmholtz, smaller


def test_noalloc_apply_rejects_too_large(ctx):
    # 257^3: U of one t-point alone (135 MB) exceeds the 128 MiB scratch
    ax = O.Axis.ncell_fixed(-8.0, 8.0, 5, 64)
    assert ax.n == 257
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 5)
    op = dage_amd.Coulomb3D(ctx, g, g, dage_amd.GaussQuad(nlin=4, nlog=4, tlin=2.0, tlog=500.0))
    a = dage_amd.DeviceCube(ctx, (ax.n,) * 3)
    b = dage_amd.DeviceCube(ctx, (ax.n,) * 3)
    with pytest.raises(DageError) as e:
        noalloc(op, a, b)
    assert "workspace" in str(e.value)
    for obj in (a, b, op, g):
        obj.close()
