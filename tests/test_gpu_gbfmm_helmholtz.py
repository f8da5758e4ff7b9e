"""GBFMM-Helmholtz (row f4): dage_gbfmm_set_energy turns the GBFMM operator
into GBFMMHelmholtz3D (gbfmm_helmholtz3d.F90:241-964), near field with the
Helmholtz t-weights and the far field of exp(-kappa r)/r by Yukawa
multipoles, M2M/M2L/L2L and local expansions.

Against the oracle (the same near field, and the far field summed directly
over the points of the non-adjacent boxes, which is what the expansions
approximate) on a 33^3 grid with 4x4x4 and 8x8x8 boxes (the latter with the
upward and downward passes): the difference is the expansions' truncation,
shrinking with lmax (kappa 1.1: 2e-4 at lmax 6, 1.2-1.5e-6 at lmax 12 of
max|V|; bar 5e-6).  Linear in the density;
setting the energy back to 0 gives the Coulomb operator bit for bit.
"""
import numpy as np
import pytest

import dage_amd
import oracle as O
from _helpers import rel_maxdiff

pytestmark = pytest.mark.gpu


def setup(ctx, maxlevel):
    ax = O.Axis.ncell_fixed(-6.0, 6.0, 5, 8)  # 33 points, h = 0.375
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 5)
    rng = np.random.default_rng(20 + maxlevel)
    al = rng.uniform(0.8, 3.0, 5)
    R = rng.uniform(-3.5, 3.5, (5, 3))
    q = rng.uniform(-1.0, 1.5, 5)
    rho = O.gauss_density(q, al, R, ax.points, ax.points, ax.points)
    return ax, g, rho


@pytest.mark.parametrize("maxlevel,energy", [(2, -0.6), (3, -0.6), (3, -3.0)])
def test_gbfmm_helmholtz_vs_oracle(ctx, maxlevel, energy):
    ax, g, rho = setup(ctx, maxlevel)
    gq = dage_amd.GaussQuad(25, 25)
    ref = O.gbfmm_helmholtz_potential([ax] * 3, rho, maxlevel, energy, gq.tp, gq.tw)
    errs = {}
    for lmax in (6, 12):
        f = dage_amd.GBFMMCoulomb3D(ctx, g, maxlevel=maxlevel, lmax=lmax, gauss=gq)
        f.set_energy(energy)
        pot = f.apply(rho)
        errs[lmax] = rel_maxdiff(pot, ref)
        f.close()
    print(f"maxlevel {maxlevel} E {energy}: GPU GBFMM-Helmholtz vs oracle: lmax 6 "
          f"{errs[6]:.2e}, lmax 12 {errs[12]:.2e}")
    assert errs[12] < 5e-6
    assert errs[12] < errs[6]
    g.close()


def test_gbfmm_helmholtz_linear_and_back_to_coulomb(ctx):
    ax, g, rho = setup(ctx, 3)
    rng = np.random.default_rng(5)
    rho2 = np.asfortranarray(rng.standard_normal(rho.shape) * 1e-2 + rho[::-1, :, :])
    gq = dage_amd.GaussQuad(25, 25)
    f = dage_amd.GBFMMCoulomb3D(ctx, g, maxlevel=3, lmax=10, gauss=gq)
    coulomb = f.apply(rho)
    f.set_energy(-1.1)
    p1, p2 = f.apply(rho), f.apply(rho2)
    p12 = f.apply(np.asfortranarray(2.0 * rho - 0.5 * rho2))
    lin = rel_maxdiff(p12, 2.0 * p1 - 0.5 * p2)
    print(f"linearity {lin:.2e}")
    assert lin < 1e-13
    assert rel_maxdiff(p1, coulomb) > 1e-2  # the screening changes the potential
    f.set_energy(0.0)
    assert np.array_equal(f.apply(rho), coulomb)
    f.set_energy(-1.1)
    assert np.array_equal(f.apply(rho), p1)
    f.close()
    g.close()


def test_gbfmm_lmax22_far_field(ctx):
    """lmax 22, the reference's GBFMM-Helmholtz far-field order
    (gbfmm_helmholtz3d.F90:66-71): with 529 coefficients per box the
    expansions' truncation falls below the 1e-6 of lmax 12, for the screened
    kernel (vs the oracle's direct far-field sum) and for 1/r (the oracle at
    E = 0); boxes of unequal shape stay limited to lmax 15."""
    ax, g, rho = setup(ctx, 3)
    gq = dage_amd.GaussQuad(25, 25)
    for energy in (-0.6, 0.0):
        ref = O.gbfmm_helmholtz_potential([ax] * 3, rho, 3, energy, gq.tp, gq.tw)
        errs = {}
        for lmax in (12, 22):
            f = dage_amd.GBFMMCoulomb3D(ctx, g, maxlevel=3, lmax=lmax, gauss=gq)
            if energy != 0.0:
                f.set_energy(energy)
            errs[lmax] = rel_maxdiff(f.apply(rho), ref)
            f.close()
        print(f"E {energy}: GBFMM vs oracle's direct far field, lmax 12 {errs[12]:.2e}, "
              f"lmax 22 {errs[22]:.2e}")
        assert errs[22] < 0.1 * errs[12]
        assert errs[22] < 2e-7
    g.close()
    # cells of two lengths: boxes of unequal shape at maxlevel 2
    steps = np.array([0.3, 0.3, 0.4, 0.4, 0.3, 0.3, 0.4, 0.4])
    g2 = dage_amd.Grid3D(ctx, [-5.6] * 3, [steps] * 3, 5)
    with pytest.raises(dage_amd.DageError):
        dage_amd.GBFMMCoulomb3D(ctx, g2, maxlevel=2, lmax=16, gauss=gq)
    g2.close()
