"""The opt-in fused x/y pass (xy_fused_kernel, DAGE_APPLY_FUSED=1) against
the separate x and y passes and the oracle.

With the same t-chunk (DAGE_APPLY_TCHUNK) both paths hand the z pass the same
intermediate U, so their outputs must agree bit for bit; on its own chunking
the fused path meets the FP64 bar, 1e-10 of max|V| against the oracle.
"""
import os

import numpy as np
import pytest

import dage_amd
import oracle as O
from _helpers import random_gaussians, rel_maxdiff

pytestmark = pytest.mark.gpu


def with_env(env, fn):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return fn()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def lib_grid(ctx, axes, nlip):
    return dage_amd.Grid3D(ctx, [a.qmin for a in axes], [a.steps for a in axes], nlip)


def quad():
    return dage_amd.GaussQuad(nlin=20, nlog=20, tlin=2.0, tlog=500.0)


CASES = [
    # input axes, output axes (None: same), nlip, Helmholtz energy (None: Coulomb), t-chunk
    ([(-4.0, 4.0, 9), (-3.0, 5.0, 6), (-2.0, 2.0, 5)], None, 7, None, 7),
    ([(-8.0, 8.0, 16)] * 3, None, 7, None, 40),             # 97^3, config-1 shape
    ([(-6.0, 6.0, 10)] * 3, None, 5, -0.8, 3),
    ([(-9.6, 9.6, 32)] * 3, None, 5, -0.3, 11),             # 129^3 Helmholtz
    ([(-6.0, 6.0, 10)] * 3, [(-2.4, 1.2, 3), (-6.0, 0.0, 5), (0.0, 3.6, 3)], 7, None, 13),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_fused_equals_separate_passes(ctx, case):
    ain, aout, nlip, energy, tchunk = CASES[case]
    ain = [O.Axis.ncell_fixed(a, b, nlip, n) for a, b, n in ain]
    aout = ain if aout is None else [O.Axis.ncell_fixed(a, b, nlip, n) for a, b, n in aout]
    same = aout is ain
    rng = np.random.default_rng(40 + case)
    q_, alpha, c = random_gaussians(rng, 4)
    rho = O.gauss_density(q_, alpha, c, *(a.points for a in ain))
    q = quad()
    gin = lib_grid(ctx, ain, nlip)
    gout = gin if same else lib_grid(ctx, aout, nlip)
    # the fused path has neither the low-rank small-t points nor the banded
    # large-t passes, so the separate passes are compared on an operator
    # without them
    if energy is None:
        op = with_env({"DAGE_APPLY_BAND": "0", "DAGE_APPLY_LOWRANK": "0"},
                      lambda: dage_amd.Coulomb3D(ctx, gin, gout, q, coda=same))
        w = O.coulomb_weights(q.tw)
    else:
        op = with_env({"DAGE_APPLY_BAND": "0", "DAGE_APPLY_LOWRANK": "0"},
                      lambda: dage_amd.Helmholtz3D(ctx, gin, gout, q, energy=0.0, coda=same))
        op.set_energy(energy)
        w = O.helmholtz_weights(q.tp, q.tw, energy)
    assert op.tiling()["band_t0"] == q.get_npoints() and op.tiling()["lowrank_t1"] == 0
    fused = with_env({"DAGE_APPLY_TCHUNK": str(tchunk), "DAGE_APPLY_FUSED": "1"},
                     lambda: op.transform_cube(rho))
    sep = with_env({"DAGE_APPLY_TCHUNK": str(tchunk), "DAGE_APPLY_FUSED": "0"},
                   lambda: op.transform_cube(rho))
    ndiff = int(np.count_nonzero(fused != sep))
    print(f"case {case}: shape {fused.shape}, t-chunk {tchunk}: {ndiff} points differ")
    assert ndiff == 0
    got = with_env({"DAGE_APPLY_FUSED": "1"}, lambda: op.transform_cube(rho))  # own t-chunking
    fx = O.exponentials(ain[0], aout[0], q.tp, False)
    fy = O.exponentials(ain[1], aout[1], q.tp, True)
    fz = O.exponentials(ain[2], aout[2], q.tp, True)
    coda = O.coda(q.tlog) if same else 0.0
    ref = O.transform_cube(fx, fy, fz, w, coda, rho, add_coda=same)
    err = rel_maxdiff(got, ref)
    print(f"case {case}: fused path vs oracle {err:.3e}")
    assert err < 1e-10
    op.close()
