"""Pins of the oracle's bubbles evaluation (CPU only).

Racah solid harmonics (harmonics.F90:586-659): explicit low-order forms and
the addition theorem sum_m S_lm^2 = r^(2l).  Radial grid (grid.F90:569-603):
cell ends on the closed-form map.  Cell lookup (grid.F90:1212-1246): brute
force.  Bubbles_eval_3dgrid (bubbles.F90:2017-2211): exact for radial data
that is a polynomial of degree <= nlip-1 per cell, r^k and lmin handling, and
the Slater-type radial functions against their analytic values.
"""
import math

import numpy as np
import pytest

import oracle as O
from _helpers import slater_bubbles


def test_solid_harmonics_low_orders():
    x, y, z = 0.3, -0.7, 1.1
    S = O.solid_harmonics(3, x, y, z)
    r2 = x * x + y * y + z * z
    s3 = math.sqrt(3.0)
    expect = [1.0, y, z, x, s3 * x * y, s3 * y * z, z * z - 0.5 * (x * x + y * y), s3 * x * z,
              0.5 * s3 * (x * x - y * y)]
    assert np.allclose(S[:9], expect, rtol=0, atol=1e-15)
    # S_30 = z (5 z^2 - 3 r^2) / 2 (Racah)
    assert abs(S[12] - 0.5 * z * (5 * z * z - 3 * r2)) < 1e-14


def test_solid_harmonics_addition_theorem():
    rng = np.random.default_rng(0)
    for _ in range(20):
        x, y, z = rng.normal(size=3)
        S = O.solid_harmonics(8, x, y, z)
        r2 = x * x + y * y + z * z
        for l in range(9):
            assert abs(np.sum(S[l * l:(l + 1) ** 2] ** 2) - r2 ** l) < 1e-12 * max(1.0, r2 ** l)


def test_radial_grid_and_cell_lookup():
    st = O.radial_steps(6.0, 33, 7, 20.0)
    ax = O.Axis(0.0, st, 7)
    assert ax.n == 199 and abs(ax.points[-1] - 20.0) < 1e-12
    c = 8.0 * 6.0 ** -1.5
    dx = 20.0 / 33
    for i in (1, 5, 17, 32):
        xi = c / ((20.0 + c) / (i * dx) - 1.0)
        assert abs(ax.points[6 * i] - xi) < 1e-13 * max(1.0, xi)
    rng = np.random.default_rng(1)
    for r in np.concatenate([rng.uniform(0, 20, 200), ax.celld, [0.0, 20.0, 20.1, -1e-9]]):
        ic = O.get_icell(ax, r)
        if r < 0 or r > ax.points[-1]:
            assert ic == 0
        else:
            ref = np.searchsorted(ax.celld, r, side="right")
            assert ic == ref, (r, ic, ref)


def _direct(bub, fr, x, y, z, lmin=0):
    """sum_b sum_lm fr(b, l, m, r) S_lm / r^l evaluated directly."""
    out = np.zeros((len(x), len(y), len(z)))
    X, Y, Zg = np.meshgrid(x, y, z, indexing="ij")
    for b, cen in enumerate(bub.centers):
        dx, dy, dz = X - cen[0], Y - cen[1], Zg - cen[2]
        r = np.sqrt(dx * dx + dy * dy + dz * dz)
        inside = r <= bub.axis(b).points[-1]
        for idx in np.ndindex(r.shape):
            if not inside[idx]:
                continue
            S = O.solid_harmonics(bub.lmax, dx[idx], dy[idx], dz[idx])
            rr = r[idx]
            for l in range(lmin, bub.lmax + 1):
                for m in range(-l, l + 1):
                    den = rr ** l if rr > 2.2e-16 else 1.0
                    out[idx] += fr(b, l, m, rr) * S[l * l + l + m] / den
    return out


def test_eval_grid_exact_for_polynomial_radial_data():
    rng = np.random.default_rng(2)
    lmax, nlip = 3, 7
    coef = rng.uniform(-1, 1, (2, (lmax + 1) ** 2, nlip))  # degree <= 6 in r
    centers = np.array([[0.3, -0.2, 0.1], [-1.0, 0.5, 0.4]])
    steps = [O.radial_steps(z, 10, nlip, 4.0) for z in (1.0, 3.0)]
    bfs = []
    for b in range(2):
        r = O.Axis(0.0, steps[b], nlip).points
        bfs.append(np.asfortranarray(np.stack([np.polyval(coef[b, lm], r)
                                               for lm in range((lmax + 1) ** 2)], axis=1)))
    bub = O.BubbleSet(centers, steps, bfs, lmax, 0, nlip)
    x = np.linspace(-2, 2, 9)
    got = O.bubbles_eval_grid(bub, x, x, x)
    ref = _direct(bub, lambda b, l, m, r: np.polyval(coef[b, l * l + l + m], r), x, x, x)
    assert np.max(np.abs(got - ref)) < 1e-11 * np.max(np.abs(ref))
    # lmin drops the low-l terms; k multiplies the radial part by r^k
    got2 = O.bubbles_eval_grid(bub, x, x, x, lmin=2)
    ref2 = _direct(bub, lambda b, l, m, r: np.polyval(coef[b, l * l + l + m], r), x, x, x, lmin=2)
    assert np.max(np.abs(got2 - ref2)) < 1e-11 * np.max(np.abs(ref2))
    bub1 = O.BubbleSet(centers, steps, bfs, lmax, -1, nlip)
    x2 = np.linspace(-1.7, 1.9, 7)  # avoid grid points at a bubble centre
    got3 = O.bubbles_eval_grid(bub1, x2, x2, x2)
    assert np.all(np.isfinite(got3))


def test_slater_bubbles_vs_analytic():
    rng = np.random.default_rng(4)
    bub, Z = slater_bubbles(rng, 2, lmax=4, zset=(1.0, 2.0), box=1.0)
    x = np.linspace(-3, 3, 7) + 0.013
    got = O.bubbles_eval_grid(bub, x, x, x)
    def coeff(b, lm, l):  # c_lm recovered from the second radial grid point
        r1 = bub.axis(b).points[1]
        return bub.bf[b][1, lm] / (r1 ** l * math.exp(-Z[b] * r1))

    ref = _direct(bub, lambda b, l, m, r: coeff(b, l * l + l + m, l) * r ** l * math.exp(-Z[b] * r),
                  x, x, x)
    # LIP interpolation of r^l exp(-Zr) on the 199-point grid
    assert np.max(np.abs(got - ref)) < 1e-6 * np.max(np.abs(ref))
