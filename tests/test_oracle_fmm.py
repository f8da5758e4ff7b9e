"""Oracle FMM matrices (multipole.F90:33-560) pinned by the reference's own
known answers, and the library's host-built matrices against the oracle.

Pins:
  * displaced_monopoles_test goldens (tests/multipole_test.F90:198-290):
    |qi^T T(boxj-boxi) qj - 1/|ri-rj|| for lmax 0..15, reproduced to 1e-5
    relative (the goldens are differences of numbers ~0.09, so the last
    digits carry rounding noise);
  * translation_test (:137-183): a unit monopole at `old` expanded about
    `new` has multipoles S(old - new), and translating back restores it;
  * interaction_test (:54-135): dipole-dipole energy from T to 1e-14;
  * mathematics: translations compose, W(b-c) W(a-b) = W(a-c), and the M2L
    of point charges converges to the direct 1/r sum.
"""
import numpy as np
import pytest

import dage_amd
import oracle as O

RJ = np.array([-12.31523266, -2.77677722, -13.87351588])
RI = np.array([-13.80289944, -12.66934534, -8.92620243])
BOXJ = np.array([-10.625, -2.125, -14.875])
BOXI = np.array([-14.875, -10.625, -10.625])
ERR_REF = [6.4549767744542658e-003, 3.5239702059782635e-003, 3.6472898611333848e-005,
           3.5551144335181639e-004, 1.2985977061118203e-004, 3.0889425959704231e-005,
           1.8391214269969369e-006, 2.0053722605550828e-006, 1.0603102051365276e-006,
           2.9594595833681847e-007, 3.4396637579869882e-008, 1.2534625434823532e-008,
           9.0637385791181302e-009, 2.9670064183306977e-009, 4.8909977456990106e-010,
           6.3554106422003542e-011]


def monopole_about(lmax, at, centre):
    e = np.zeros((lmax + 1) ** 2)
    e[0] = 1.0
    return O.fmm_translate(lmax, O.fmm_translation_packed(lmax, at - centre), e)


def test_displaced_monopoles_golden():
    lmax = 15
    T = O.fmm_interaction(lmax, BOXJ - BOXI)
    qj = monopole_about(lmax, RJ, BOXJ)
    qi = monopole_about(lmax, RI, BOXI)
    U = 1.0 / np.linalg.norm(RI - RJ)
    for l in range(lmax + 1):
        n = (l + 1) ** 2
        err = abs(qi[:n] @ (T[:n, :n] @ qj[:n]) - U)
        assert err == pytest.approx(ERR_REF[l], rel=1e-5), l


@pytest.mark.parametrize("lmax", [0, 3, 12])
def test_translation_kat(lmax):
    rng = np.random.default_rng(83453)
    old, new = rng.uniform(0, 1, 3), rng.uniform(0, 1, 3)
    q = monopole_about(lmax, old, new)
    S = O.solid_harmonics(lmax, *(old - new))
    assert np.max(np.abs(q - S)) < 1e-15
    back = O.fmm_translate(lmax, O.fmm_translation_packed(lmax, new - old), q)
    e = np.zeros_like(back)
    e[0] = 1.0
    assert np.max(np.abs(back - e)) < 1e-12  # cancellation in the l = 12 terms


def test_interaction_dipoles():
    lmax = 4
    rng = np.random.default_rng(86456)
    mu = rng.uniform(0, 1, 3)
    q = np.zeros((lmax + 1) ** 2)
    q[1:4] = [mu[1], mu[2], mu[0]]  # S_1 = (y, z, x)
    d = 3.0
    for theta in np.linspace(0, 2 * np.pi, 40):
        dr = np.array([d * np.cos(theta), d * np.sin(theta), 0.7])
        T = O.fmm_interaction(lmax, dr)
        r = np.linalg.norm(dr)
        ref = mu @ mu / r ** 3 - 3 * (mu @ dr) ** 2 / r ** 5
        assert abs(q @ T @ q - ref) < 1e-14


def test_translations_compose():
    lmax = 8
    a, b, c = np.array([0.3, -0.2, 0.5]), np.array([-0.4, 0.1, 0.2]), np.array([0.2, 0.6, -0.3])
    Wab, Wbc, Wac = (O.fmm_translation_full(lmax, x) for x in (a - b, b - c, a - c))
    assert np.max(np.abs(Wbc @ Wab - Wac)) < 1e-13


def test_m2l_converges_to_direct():
    rng = np.random.default_rng(5)
    src_c, tgt_c = np.zeros(3), np.array([4.0, 1.0, -2.0])
    charges = rng.uniform(-1, 1, 6)
    pos = src_c + rng.uniform(-0.6, 0.6, (6, 3))
    probe = tgt_c + np.array([0.4, -0.3, 0.5])
    direct = sum(qk / np.linalg.norm(probe - p) for qk, p in zip(charges, pos))
    errs = []
    for lmax in (2, 6, 12):
        q = sum(qk * monopole_about(lmax, p, src_c) for qk, p in zip(charges, pos))
        L = O.fmm_interaction(lmax, src_c - tgt_c) @ q
        V = L @ O.solid_harmonics(lmax, *(probe - tgt_c))
        errs.append(abs(V - direct))
    assert errs[2] < 1e-8 * abs(direct) and errs[0] > errs[1] > errs[2]


@pytest.mark.parametrize("lmax", [0, 1, 5, 12, 15])
def test_library_matrices_match_oracle(lmax):
    """The library builds T and W by its own polynomial projections; both must
    equal the reference's coefficient formulas (host code, no GPU needed)."""
    rng = np.random.default_rng(lmax)
    for _ in range(3):
        d = rng.uniform(-6, 6, 3)
        T, To = dage_amd.interaction_matrix(lmax, d), O.fmm_interaction(lmax, d)
        W, Wo = dage_amd.translation_matrix(lmax, d), O.fmm_translation_full(lmax, d)
        assert np.max(np.abs(T - To)) <= 1e-13 * np.max(np.abs(To))
        assert np.max(np.abs(W - Wo)) <= 1e-13 * np.max(np.abs(Wo))
    with pytest.raises(dage_amd.DageError):
        dage_amd.interaction_matrix(lmax, [0.0, 0.0, 0.0])


# ------------------------------------------------------------------------
# The oracle's GBFMM pieces, pinned on the CPU without the HIP library.

def test_box_multipoles_of_gaussians_are_solid_harmonics_at_centres():
    """A spherically symmetric density about R integrates any harmonic
    polynomial P to q P(R) (mean-value property), so the box multipoles
    (integrator.cu:1122-1196) of a sum of normalised Gaussians are
    sum_g q_g S_lm(R_g - c) exactly, up to the grid quadrature (here < 1e-7
    on a 145^3 box grid with 6-point-wide cells of 0.25 bohr)."""
    ax = O.Axis.ncell_fixed(-3.0, 3.0, 7, 24)
    R = np.array([[0.2, -0.3, 0.15], [-0.25, 0.1, 0.3], [0.1, 0.35, -0.45]])
    q = np.array([1.0, -0.7, 0.4])
    rho = O.gauss_density(q, np.array([4.0, 6.0, 5.0]), R, ax.points, ax.points, ax.points)
    c = np.array([0.1, 0.05, -0.2])
    lmax = 12
    got = O.box_multipoles(lmax, rho, [ax] * 3, c)
    ref = sum(qg * O.solid_harmonics(lmax, *(Rg - c)) for qg, Rg in zip(q, R))
    err = np.max(np.abs(got - ref))
    print(f"box multipoles vs sum q S(R - c): {err:.2e}")
    assert err < 1e-7
    # a plausible slip (a dropped weight axis) is far outside that bar
    bad = O.box_multipoles(lmax, rho, [ax, ax, O.Axis(ax.qmin, ax.steps * 1.01, 7)], c)
    assert np.max(np.abs(bad - ref)) > 1e-4


def test_local_expansion_of_a_point_charge_is_its_potential():
    """The local expansion of a unit charge (its multipoles S(R - c_s) about
    a far box, moved by the pinned interaction matrix) evaluated by the
    oracle at points of the target box gives 1/|r - R| to the lmax-12
    truncation, and equals the direct sum_lm L_lm S_lm(r - c_t)."""
    lmax = 12
    ct = np.array([0.3, -0.1, 0.2])
    cs = np.array([4.0, 3.0, -3.5])
    Rp = cs + np.array([0.3, -0.2, 0.4])
    N = (lmax + 1) ** 2
    e0 = np.zeros(N)
    e0[0] = 1.0
    qs = O.fmm_translate(lmax, O.fmm_translation_packed(lmax, Rp - cs), e0)
    L = O.fmm_interaction(lmax, cs - ct) @ qs
    x = ct[0] + np.linspace(-0.5, 0.5, 7)
    y = ct[1] + np.linspace(-0.5, 0.5, 6)
    z = ct[2] + np.linspace(-0.5, 0.5, 5)
    V = O.eval_local_expansion(lmax, L, x, y, z, ct)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    exact = 1.0 / np.sqrt((X - Rp[0]) ** 2 + (Y - Rp[1]) ** 2 + (Z - Rp[2]) ** 2)
    assert np.max(np.abs(V - exact)) < 1e-10 * np.max(exact)
    direct = np.array([[[L @ O.solid_harmonics(lmax, xi - ct[0], yi - ct[1], zi - ct[2])
                         for zi in z] for yi in y] for xi in x])
    assert np.max(np.abs(V - direct)) < 1e-14 * np.max(np.abs(direct))


def test_gbfmm_oracle_matches_direct_apply_and_analytic():
    """The oracle's whole GBFMM (near field on sub-grids, multipoles, M2L,
    local expansions; gbfmm_coulomb3d.F90:650-821) against the oracle's
    direct whole-grid apply without coda (pinned by the reference's
    gauss_test goldens) on a 49^3 grid with 4x4x4 boxes: they differ only by
    the far-field multipole approximation, which shrinks with lmax (1e-3 at
    lmax 4, 1.1e-5 at lmax 12, where the multipole quadrature on the coarse
    grid sets the floor); both sit at the same grid error from the analytic
    potential."""
    ax = O.Axis.ncell_fixed(-6.0, 6.0, 7, 8)
    rng = np.random.default_rng(3)
    al = rng.uniform(1.0, 4.0, 5)
    R = rng.uniform(-3.5, 3.5, (5, 3))
    q = rng.uniform(0.5, 1.5, 5)
    rho = O.gauss_density(q, al, R, ax.points, ax.points, ax.points)
    tp, tw = O.gauss_quad(25, 25)
    fx, fy, fz = O.coulomb3d_matrices([ax] * 3, [ax] * 3, tp)
    direct = O.transform_cube(fx, fy, fz, O.coulomb_weights(tw), 0.0, rho, add_coda=False)
    exact = O.gauss_potential(q, al, R, ax.points, ax.points, ax.points)
    errs = {}
    for lmax in (4, 12):
        pot, _, _ = O.gbfmm_potential([ax] * 3, rho, 2, lmax, tp, tw)
        errs[lmax] = np.max(np.abs(pot - direct)) / np.max(np.abs(direct))
    print(f"oracle GBFMM vs direct apply: lmax 4 {errs[4]:.2e}, lmax 12 {errs[12]:.2e}")
    assert errs[12] < 3e-5
    assert errs[4] > 30 * errs[12]
    e_an = np.max(np.abs(pot - exact)) / np.max(exact)
    e_dir = np.max(np.abs(direct - exact)) / np.max(exact)
    assert e_an < 2e-3 and abs(e_an - e_dir) < 5e-5


def test_gbfmm_helmholtz_oracle_pinned():
    """The oracle's GBFMM-Helmholtz (near field with Helmholtz t-weights plus
    the far field summed directly over the non-adjacent boxes' points,
    gbfmm_helmholtz3d.F90:517-964) pinned on a 49^3 grid with 4x4x4 boxes:
    * at energy 0 the direct far field is the 1/r one: it agrees with the
      oracle's multipole GBFMM (lmax 12) to the multipole truncation;
    * at energy -0.6 (kappa 1.095) it agrees with the whole-grid Helmholtz
      apply (no coda; t-quadrature of exp(-kappa r)/r on the cell
      interpolants) to the same level (6e-6), and both sit at the same
      distance from the closed-form Yukawa potential of the Gaussians (2.9e-3
      on this coarse grid)."""
    ax = O.Axis.ncell_fixed(-6.0, 6.0, 7, 8)
    rng = np.random.default_rng(3)
    al = rng.uniform(1.0, 4.0, 5)
    R = rng.uniform(-3.5, 3.5, (5, 3))
    q = rng.uniform(0.5, 1.5, 5)
    rho = O.gauss_density(q, al, R, ax.points, ax.points, ax.points)
    tp, tw = O.gauss_quad(25, 25)
    # 1/r: direct far field against the multipole one
    direct0 = O.gbfmm_helmholtz_potential([ax] * 3, rho, 2, 0.0, tp, tw)
    mp, _, _ = O.gbfmm_potential([ax] * 3, rho, 2, 12, tp, tw)
    e0 = np.max(np.abs(direct0 - mp)) / np.max(np.abs(mp))
    print(f"energy 0: direct far field vs multipoles (lmax 12): {e0:.2e}")
    assert e0 < 3e-5
    # Yukawa
    E = -0.6
    kappa = np.sqrt(-2 * E)
    yk = O.gbfmm_helmholtz_potential([ax] * 3, rho, 2, E, tp, tw)
    fx, fy, fz = O.coulomb3d_matrices([ax] * 3, [ax] * 3, tp)
    whole = O.transform_cube(fx, fy, fz, O.helmholtz_weights(tp, tw, E), 0.0, rho,
                             add_coda=False)
    exact = O.yukawa_gauss_potential(q, al, R, kappa, ax.points, ax.points, ax.points)
    e_whole = np.max(np.abs(yk - whole)) / np.max(np.abs(whole))
    e_an = np.max(np.abs(yk - exact)) / np.max(np.abs(exact))
    e_an_whole = np.max(np.abs(whole - exact)) / np.max(np.abs(exact))
    print(f"energy {E}: vs whole-grid Helmholtz apply {e_whole:.2e}; vs closed form "
          f"{e_an:.2e} (whole-grid apply {e_an_whole:.2e})")
    assert e_whole < 3e-5
    assert e_an < 5e-3 and abs(e_an - e_an_whole) < 5e-5
