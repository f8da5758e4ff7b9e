"""Batched Helmholtz applies (one hipGraph launch for many orbitals with
different energies and cube sizes) against per-orbital applies, the oracle
and the Yukawa closed form.

Bars: FP64 batch vs per-orbital FP64 applies 1e-13 (the split-K partition of
the z pass depends on the batch size, so sums round differently); vs the
oracle 1e-10; split-BF16 (bf16x6) batch vs FP64 2e-7 on these toy cubes
(1e-7 at the config shapes, tests/test_gpu_mixed.py and bench.py); graph
replays equal to the eager launch bit for bit; tuned tiles give the same
numbers.
"""
import numpy as np
import pytest

import dage_amd
import oracle as O
from _helpers import random_gaussians, rel_maxdiff

pytestmark = pytest.mark.gpu


def centred_axis(n, h=0.15, nlip=5):
    ncell = (n - 1) // (nlip - 1)
    half = ncell * (nlip - 1) * h / 2
    return O.Axis(-half, np.full(ncell, h), nlip)


def make_batch(ctx, sizes, energies, precision=dage_amd.DAGE_PREC_FP64, seed=0, coda=True):
    rng = np.random.default_rng(seed)
    axes = {n: centred_axis(n) for n in set(sizes)}
    grids = {n: dage_amd.Grid3D(ctx, [axes[n].qmin] * 3, [axes[n].steps] * 3, 5) for n in axes}
    gq = dage_amd.GaussQuad(10, 10, tlin=2.0, tlog=500.0)
    rhos = []
    for n in sizes:
        ax = axes[n]
        q_, a, c = random_gaussians(rng, 3, box=0.6 * ax.points[-1])
        rhos.append(O.gauss_density(q_, a, c, ax.points, ax.points, ax.points))
    batch = dage_amd.HelmholtzBatch(ctx, [grids[n] for n in sizes], energies, gq,
                                    precision=precision, coda=coda)
    for i, r in enumerate(rhos):
        batch.upload(i, r)
    return batch, grids, axes, gq, rhos


def test_batch_matches_single_applies_and_oracle(ctx):
    sizes = [17, 25, 17, 33, 25, 17]
    energies = [-0.3, -1.2, -0.05, -2.0, 0.0, -0.8]
    batch, grids, axes, gq, rhos = make_batch(ctx, sizes, energies)
    assert batch.ngroups == 3
    batch.apply()
    for i, (n, e) in enumerate(zip(sizes, energies)):
        got = batch.download(i)
        op = dage_amd.Helmholtz3D(ctx, grids[n], grids[n], gq, energy=e)
        single = op.transform_cube(rhos[i])
        assert rel_maxdiff(got, single) < 1e-13, i
        ax = [axes[n]] * 3
        fx, fy, fz = O.coulomb3d_matrices(ax, ax, gq.tp)
        ref = O.transform_cube(fx, fy, fz, O.helmholtz_weights(gq.tp, gq.tw, e), O.coda(500.0),
                               rhos[i])
        assert rel_maxdiff(got, ref) < 1e-10, i


def test_graph_replays_match_eager(ctx):
    sizes = [25, 17, 25]
    batch, grids, axes, gq, rhos = make_batch(ctx, sizes, [-0.5, -1.0, -0.1], seed=3)
    batch.apply_eager()
    eager = [batch.download(i) for i in range(len(sizes))]
    for rep in range(4):
        # unrelated eager work on the same stream between replays
        d = dage_amd.DeviceCube(ctx, rhos[0].shape, rhos[0])
        d.set_to_zero()
        batch.apply()
        for i in range(len(sizes)):
            assert np.array_equal(batch.download(i), eager[i]), (rep, i)


def test_set_energies_and_no_coda(ctx):
    sizes = [17, 25]
    batch, grids, axes, gq, rhos = make_batch(ctx, sizes, [-0.5, -1.0], seed=5, coda=False)
    batch.set_energies([-2.5, -0.01])
    batch.apply()
    for i, (n, e) in enumerate(zip(sizes, [-2.5, -0.01])):
        op = dage_amd.Helmholtz3D(ctx, grids[n], grids[n], gq, energy=e, coda=False)
        assert rel_maxdiff(batch.download(i), op.transform_cube(rhos[i])) < 1e-13


def test_mixed_batch_and_tuning(ctx):
    sizes = [33, 25, 33, 17]
    energies = [-0.4, -1.5, -3.0, -0.2]
    b64, grids, axes, gq, rhos = make_batch(ctx, sizes, energies, seed=7)
    b64.apply()
    ref = [b64.download(i) for i in range(len(sizes))]
    tiles = b64.tune(reps=1)
    assert tiles.shape == (3, 3) and tiles.min() >= 0 and tiles.max() <= 8  # 0: own tiling
    for i in range(len(sizes)):
        b64.upload(i, rhos[i])  # tuning overwrote the outputs, not the inputs
    b64.apply()
    for i in range(len(sizes)):
        assert np.array_equal(b64.download(i), ref[i]), i
    b6, _, _, _, _ = make_batch(ctx, sizes, energies, precision=dage_amd.DAGE_PREC_BF16X6, seed=7)
    b6.tune(reps=1)
    b6.apply()
    for i in range(len(sizes)):
        # 3-term splits carry 24 bits per operand (2^-24 = 6e-8 per pass); on
        # these toy cubes (17^3-33^3, densities cut by the box) the three
        # passes reach ~1.1e-7 of max|V|; the config shapes stay below 1e-7
        assert rel_maxdiff(b6.download(i), ref[i]) < 2e-7, i


def test_batch_against_yukawa_closed_form(ctx):
    """65^3 at h = 0.15 (nlip 5), k^2 = 1.7, T = 40 quadrature of config 3."""
    ax = centred_axis(65)
    rng = np.random.default_rng(0xDA6E0003)
    q_, a, c = random_gaussians(rng, 4, box=2.0)
    rho = O.gauss_density(q_, a, c, ax.points, ax.points, ax.points)
    k2 = 1.7
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 5)
    gq = dage_amd.GaussQuad(20, 20, tlin=2.0, tlog=500.0)
    batch = dage_amd.HelmholtzBatch(ctx, [g], [-k2 / 2], gq, coda=False)
    batch.upload(0, rho)
    batch.apply()
    got = batch.download(0)
    exact = O.yukawa_gauss_potential(q_, a, c, np.sqrt(k2), ax.points, ax.points, ax.points)
    err = rel_maxdiff(got, exact)
    print(f"batched Helmholtz vs Yukawa closed form: {err:.3e}")
    assert err < 2e-3


def test_batch_errors(ctx):
    ax = centred_axis(17)
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 5)
    gq = dage_amd.GaussQuad(5, 5)
    with pytest.raises(dage_amd.DageError):
        dage_amd.HelmholtzBatch(ctx, [g], [np.nan], gq)
    with pytest.raises(dage_amd.DageError):
        dage_amd.HelmholtzBatch(ctx, [g], [-1.0], gq, precision=7)
    b = dage_amd.HelmholtzBatch(ctx, [g], [-1.0], gq)
    host = np.zeros(17 ** 3)
    with pytest.raises(dage_amd.DageError):  # the C-ABI checks the orbital index
        dage_amd._lib.call("dage_batch_download", b.handle, 1,
                           host.ctypes.data_as(dage_amd._lib._dp))


@pytest.mark.parametrize("precision,tol", [(dage_amd.DAGE_PREC_FP64, 1e-13),
                                           (dage_amd.DAGE_PREC_BF16X6, 2e-7)])
def test_batch_slices_under_workspace_cap(ctx, monkeypatch, precision, tol):
    """A group whose workspace would exceed the cap runs in slices of
    right-hand sides inside the same graph; the result matches the unsliced
    batch (the z pass's K split depends on the slice size, so only to
    rounding)."""
    sizes = [25] * 7
    energies = [-0.3, -1.2, -0.05, -2.0, -0.5, -0.8, -1.5]
    whole, _, _, _, _ = make_batch(ctx, sizes, energies, precision=precision, seed=4)
    whole.apply()
    ref = [whole.download(i) for i in range(len(sizes))]
    whole.close()
    monkeypatch.setenv("DAGE_BATCH_WORKSPACE_MB", "1")
    sliced, _, _, _, _ = make_batch(ctx, sizes, energies, precision=precision, seed=4)
    for _ in range(2):  # replays of the captured graph
        sliced.apply()
        for i in range(len(sizes)):
            assert rel_maxdiff(sliced.download(i), ref[i]) < tol, i
    sliced.close()
