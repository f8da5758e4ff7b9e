"""Helmholtz-kernel SCF on the cube (rows A27-A28).

Mirrors the reference's HelmholtzSCFCycle / PowerMethod loop
(src/bubbles/scf_cycle.F90:2927-3054, src/bubbles/scf.F90:99-300):

  per orbital, in energy order:  V psi_i (level-shifted by -(eps_i + 0.1)
  when eps_i > 0, energy set to -0.1), Helmholtz3D with that energy,
  times -1/(2 pi), update Delta_i = G psi_i - psi_i, psi_i += w Delta_i;
  then the orbitals are rotated to the Ritz vectors of the Hamiltonian in
  their span (calculate_hamiltonian_matrix + diagonalize +
  update_orbitals_linear_combination), and the run stops when
  iteration > 1, |E_k - E_{k-1}| < 1e-6 and every |eps_i,k - eps_i,k-1| < 1e-6
  (SCFOptimizer_is_converged, scf.F90:99-136).

The model Hamiltonian (the reference's orbitals carry bubbles; this module
has the cube part only): H = -1/2 laplacian + V_nuc + J[rho] + v_x[rho] with
nuclei smeared as normalised Gaussians of exponent `beta` (so the orbitals
need no cusp), J the Coulomb3D potential of rho = 2 sum |psi_i|^2 and v_x the
Slater (LDA) exchange; closed shell, occupation 2.  All cube work runs on the
GPU through the C-ABI: the Helmholtz applies of all orbitals are ONE batched
hipGraph launch (FP64 here), J one Coulomb apply, the Ritz matrices
(overlap, kinetic from the cells' exact stiffness integrals, potential)
deterministic FP64 Gram
reductions; only the (nocc x nocc) eigenproblem runs on the host.

`update` selects how Delta becomes the proposal: None -> the plain
iteration psi + w Delta; a callable (the learned preconditioner) maps the
list of (psi_i, Delta_i) device cubes to proposals, and the plain proposal
stays the fallback: of the two, the one with the lower occupied Ritz
eigenvalue sum is kept.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import DAGE_PREC_FP64, call
from .operators import Context, Coulomb3D, DeviceCube, GaussQuad, Grid3D, HelmholtzBatch


def _vp(x):
    return ctypes.c_void_p(x if isinstance(x, int) else x.ptr)


def _ptr_array(cubes):
    arr = (ctypes.c_void_p * len(cubes))(*[c if isinstance(c, int) else c.ptr for c in cubes])
    return arr


@dataclass
class Molecule:
    Z: np.ndarray        # nuclear charges
    R: np.ndarray        # (natom, 3) positions, bohr

    def nuclear_repulsion(self) -> float:
        e = 0.0
        for a in range(len(self.Z)):
            for b in range(a):
                e += self.Z[a] * self.Z[b] / float(np.linalg.norm(self.R[a] - self.R[b]))
        return e


def random_molecule(rng, natom=10, zset=(1, 6, 7, 8), radius=6.0, dmin=1.8) -> Molecule:
    """Config-4 generator: Z uniform in zset, positions uniform in a sphere of
    `radius` bohr with all pairwise distances >= dmin (rejection sampling)."""
    R = []
    while len(R) < natom:
        p = rng.uniform(-radius, radius, 3)
        if np.linalg.norm(p) > radius:
            continue
        if all(np.linalg.norm(p - q) >= dmin for q in R):
            R.append(p)
    return Molecule(rng.choice(np.asarray(zset, float), natom), np.array(R))


def is_converged(energies, eigen_values, e_tol=1e-6, eig_tol=1e-6) -> bool:
    """SCFOptimizer_is_converged (scf.F90:99-136): iteration > 1, |dE| < e_tol
    and every |d eps_i| < eig_tol (the reference leaves the eigenvalue flag
    unset when no eigenvalue fails the test; the evident intent is kept)."""
    if len(energies) < 2:
        return False
    if abs(energies[-1] - energies[-2]) >= e_tol:
        return False
    d = np.abs(np.asarray(eigen_values[-1]) - np.asarray(eigen_values[-2]))
    return bool(np.all(d <= eig_tol))


def backoff_skips(rejected: int, backoff) -> bool:
    """Whether the learned update is skipped this iteration: after n = backoff[0]
    proposals in a row lost to the plain step, only every backoff[1]-th
    iteration evaluates it (counting from the n-th loss) until one wins."""
    if backoff is None or rejected < backoff[0]:
        return False
    return (rejected - backoff[0]) % backoff[1] != 0


@dataclass
class SCFResult:
    energies: list = field(default_factory=list)
    eigen_values: list = field(default_factory=list)
    times: list = field(default_factory=list)       # wall seconds at each iteration end
    learned_used: list = field(default_factory=list)
    converged: bool = False
    iterations: int = 0
    wall_s: float = 0.0


class HelmholtzSCF:
    def __init__(self, ctx: Context, grid: Grid3D, molecule: Molecule, nocc: int,
                 gauss: GaussQuad | None = None, beta: float = 4.0, exchange: bool = True,
                 update_weight: float = 1.0, external_potential: np.ndarray | None = None,
                 hartree: bool = True, tune: bool = False):
        """tune: run the batch autotuner once (GEMM column blocks per pass) on
        zeroed inputs before the SCF starts."""
        from scipy.special import erf
        self.ctx, self.grid, self.mol, self.nocc = ctx, grid, molecule, nocc
        self.gauss = gauss if gauss is not None else GaussQuad(20, 20, tlin=2.0, tlog=500.0)
        self.exchange, self.hartree, self.w = exchange, hartree, float(update_weight)
        self.shape = grid.get_shape()
        self.n = int(np.prod(self.shape))
        self.box_volume = float(np.prod([np.sum(grid.get_ints(a)) for a in range(3)]))
        x, y, z = (grid.get_coord(a) for a in range(3))
        X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
        if external_potential is not None:
            vnuc = np.asarray(external_potential, float)
        else:
            vnuc = np.zeros(self.shape)
            sb = math.sqrt(beta)
            for Za, Ra in zip(molecule.Z, molecule.R):
                r = np.sqrt((X - Ra[0]) ** 2 + (Y - Ra[1]) ** 2 + (Z - Ra[2]) ** 2)
                rs = np.where(r < 1e-12, 1.0, r)
                vnuc -= Za * np.where(r < 1e-12, 2 * sb / math.sqrt(math.pi), erf(sb * rs) / rs)
        self._X = (X, Y, Z)
        mk = lambda h=None: DeviceCube(ctx, self.shape, h)
        self.vnuc = mk(np.asfortranarray(vnuc))
        self.rho, self.J, self.vx, self.ex, self.V = mk(), mk(), mk(), mk(), mk()
        self.psi = [mk() for _ in range(nocc)]
        self.phi = [mk() for _ in range(nocc)]       # proposals
        self.alt = [mk() for _ in range(nocc)]       # second proposal / scratch
        self.delta = [mk() for _ in range(nocc)]
        self.work = [mk() for _ in range(nocc)]      # V phi, gradients
        self.grad = [[mk() for _ in range(nocc)] for _ in range(3)]
        self.coul = Coulomb3D(ctx, grid, grid, self.gauss)
        self.batch = HelmholtzBatch(ctx, [grid] * nocc, [-0.5] * nocc, self.gauss,
                                    precision=DAGE_PREC_FP64)
        if tune:
            zero = np.zeros(self.shape, order="F")
            for i in range(nocc):
                self.batch.upload(i, zero)
            self.batch.tune(reps=1)
        self.eps = np.zeros(nocc)
        self.kin_diag = np.zeros(nocc)

    # ---------------------------------------------------------------- helpers
    def _axpby(self, a, x, b, y):
        call("dage_vec_axpby", self.ctx.handle, self.n, float(a), _vp(x), float(b), _vp(y))

    def _mul(self, x, y, out):
        call("dage_vec_mul", self.ctx.handle, self.n, _vp(x), _vp(y), _vp(out))

    def _gram(self, f, g):
        S = np.zeros(len(f) * len(g))
        call("dage_cube_gram", self.grid.handle, len(f), _ptr_array(f), len(g), _ptr_array(g),
             S.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return S.reshape((len(f), len(g)), order="F")

    def _lincomb(self, src, coef, dst):
        c = np.asfortranarray(coef, dtype=np.float64)
        call("dage_vec_lincomb", self.ctx.handle, self.n, len(src), _ptr_array(src),
             c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(dst), _ptr_array(dst))

    def _gradient(self, axis, f, df):
        call("dage_cube_gradient", self.grid.handle, axis, _vp(f), _vp(df))

    def integrate(self, f, g=None) -> float:
        return self.grid.integrate(f, g)

    # ----------------------------------------------------------- Hamiltonian
    def build_potential(self):
        """rho, J[rho], v_x[rho] and V = V_nuc + J + v_x of the current orbitals."""
        self.rho.set_to_zero()
        for p in self.psi:
            call("dage_vec_square_acc", self.ctx.handle, self.n, 2.0, _vp(p), _vp(self.rho))
        self.V.copy_from(self.vnuc)
        if self.hartree:
            self.coul.transform_cuda_cube(self.rho, self.J)
            self._axpby(1.0, self.J, 1.0, self.V)
        if self.exchange:
            call("dage_vec_lda_x", self.ctx.handle, self.n, _vp(self.rho), _vp(self.vx),
                 _vp(self.ex))
            self._axpby(1.0, self.vx, 1.0, self.V)

    def total_energy(self) -> float:
        """E = sum_i 2 <psi_i|T|psi_i> + int rho V_nuc + 1/2 int rho J + E_x + E_nn
        for the current orbitals (the potential must be built from them)."""
        e = 2.0 * float(np.sum(self.kin_diag))
        e += self.integrate(self.rho, self.vnuc)
        if self.hartree:
            e += 0.5 * self.integrate(self.rho, self.J)
        if self.exchange:
            e += self.integrate(self.ex)
        return e + self.mol.nuclear_repulsion()

    def _scratch(self, k):
        """Gradient and work cubes for k functions (grown on demand)."""
        while len(self.work) < k:
            self.work.append(DeviceCube(self.ctx, self.shape))
            for g in self.grad:
                g.append(DeviceCube(self.ctx, self.shape))

    def ritz(self, phi):
        """Ritz vectors of H (with the current V) in span(phi): returns
        (eigenvalues, coefficients (k x m), kinetic matrix), m <= k.  The
        overlap is inverted by canonical orthogonalisation: directions with
        overlap eigenvalue below 1e-10 of the largest are dropped."""
        self._scratch(len(phi))
        S = self._gram(phi, phi)
        # exact kinetic integrals of the cell interpolants: the three axes'
        # stiffness results are summed first (the Gram is linear in them)
        k = len(phi)
        for i, f in enumerate(phi):
            for a in range(3):
                call("dage_cube_stiffness", self.grid.handle, a, _vp(f), _vp(self.grad[a][i]))
            self._axpby(1.0, self.grad[1][i], 1.0, self.grad[0][i])
            self._axpby(1.0, self.grad[2][i], 1.0, self.grad[0][i])
        T = 0.5 * self._gram(phi, self.grad[0][:k])
        for i, f in enumerate(phi):
            self._mul(self.V, f, self.work[i])
        Vm = self._gram(phi, self.work[:len(phi)])
        H = T + Vm
        H = 0.5 * (H + H.T)
        S = 0.5 * (S + S.T)
        s, U = np.linalg.eigh(S)
        keep = s > 1e-10 * s[-1]
        X = U[:, keep] / np.sqrt(s[keep])
        e, C = np.linalg.eigh(X.T @ H @ X)
        return e, X @ C, T

    def _adopt(self, phi, e, C, T):
        """psi <- the lowest nocc Ritz vectors phi C (orthonormal, signs kept
        close to phi)."""
        k = self.nocc
        C = C[:, :k].copy()
        for j in range(k):
            if C[j, j] < 0:
                C[:, j] = -C[:, j]
        self._lincomb(phi, C, self.psi)
        self.eps = e[:k].copy()
        self.kin_diag = np.diag(C.T @ T @ C).copy()

    # --------------------------------------------------------------- guesses
    def initial_guess(self, exponents=None):
        """Atom-centred s and p Gaussians (s for every atom, then p_x, p_y,
        p_z on the heaviest atoms) until nocc functions, rotated to the Ritz
        vectors of H[V_nuc] (no electron-electron terms yet)."""
        X, Y, Z = self._X
        funcs = []
        order = np.argsort(-self.mol.Z)
        for a in order:
            Ra = self.mol.R[a]
            al = 0.5 * (exponents[a] if exponents is not None else 1.0)
            funcs.append(np.exp(-al * ((X - Ra[0]) ** 2 + (Y - Ra[1]) ** 2 + (Z - Ra[2]) ** 2)))
        for a in order:
            Ra = self.mol.R[a]
            g = np.exp(-0.5 * ((X - Ra[0]) ** 2 + (Y - Ra[1]) ** 2 + (Z - Ra[2]) ** 2))
            for d, D in enumerate((X, Y, Z)):
                funcs.append((D - Ra[d]) * g)
        funcs = funcs[:self.nocc]
        if len(funcs) < self.nocc:
            raise ValueError("not enough guess functions")
        for f, c in zip(funcs, self.phi):
            c.upload(np.asfortranarray(f))
        self.V.copy_from(self.vnuc)
        e, C, T = self.ritz(self.phi)
        self._adopt(self.phi, e, C, T)

    # ------------------------------------------------------------- iteration
    def helmholtz_step(self):
        """Delta_i = -1/(2 pi) G(eps_i)[V psi_i] - psi_i for every orbital:
        one batched graph launch for all Helmholtz applies."""
        energies = []
        for i, p in enumerate(self.psi):
            dst = self.batch.input_ptr(i)
            self._mul(self.V, p, dst)
            e = float(self.eps[i])
            if e > 0.0:  # level shift, scf_cycle.F90:2972-2978
                self._axpby(-(e + 0.1), p, 1.0, dst)
                e = -0.1
            energies.append(e)
        self.batch.set_energies(energies)
        self.batch.apply()
        for i, p in enumerate(self.psi):
            out = self.batch.output_ptr(i)
            self.delta[i].copy_from(self.psi[i])
            # delta = -1/(2 pi) G - psi
            self._axpby(-1.0 / (2.0 * math.pi), out, -1.0, self.delta[i])

    def run(self, max_iter=100, e_tol=1e-6, eig_tol=1e-6, update=None, log=None,
            combine="choose", backoff=None) -> SCFResult:
        """`combine` says how an `update` proposal meets the plain one:
        "choose" keeps whichever set has the lower occupied Ritz eigenvalue
        sum; "span" takes the lowest nocc Ritz vectors of the span of both
        (never above either, by the variational principle).
        `backoff` = (n, period), "choose" only: after n proposals in a row
        lost to the plain step, the update is evaluated only every period-th
        iteration until one wins again (the plain step is taken meanwhile)."""
        res = SCFResult()
        t0 = time.perf_counter()
        rejected = 0
        for it in range(1, max_iter + 1):
            self.build_potential()
            E = self.total_energy()
            res.energies.append(E)
            res.eigen_values.append(self.eps.copy())
            self.ctx.synchronize()
            res.times.append(time.perf_counter() - t0)
            res.iterations = it
            if log:
                log(it, E, self.eps)
            if is_converged(res.energies, res.eigen_values, e_tol, eig_tol):
                res.converged = True
                break
            self.helmholtz_step()
            # plain proposal psi + w Delta
            for i in range(self.nocc):
                self.phi[i].copy_from(self.psi[i])
                self._axpby(self.w, self.delta[i], 1.0, self.phi[i])
            used = False
            if update is not None and combine == "span":
                update(self.psi, self.delta, self.alt)
                both = self.phi + self.alt
                e2, C2, T2 = self.ritz(both)
                chosen = (both, e2, C2, T2)
                used = True
            else:
                e, C, T = self.ritz(self.phi)
                chosen = (self.phi, e, C, T)
                if update is not None and backoff_skips(rejected, backoff):
                    rejected += 1
                elif update is not None:
                    update(self.psi, self.delta, self.alt)
                    e2, C2, T2 = self.ritz(self.alt)
                    if np.sum(e2[:self.nocc]) < np.sum(e[:self.nocc]):
                        chosen = (self.alt, e2, C2, T2)
                        used = True
                        rejected = 0
                    else:
                        rejected += 1
            res.learned_used.append(used)
            self._adopt(*chosen)
        res.wall_s = time.perf_counter() - t0
        return res

    def orbitals(self):
        return [p.download() for p in self.psi]

    def close(self):
        for obj in ([self.vnuc, self.rho, self.J, self.vx, self.ex, self.V] + self.psi + self.phi
                    + self.alt + self.delta + self.work + [c for g in self.grad for c in g]
                    + [self.coul, self.batch]):
            obj.close()


def helmholtz_update_functions(op, orbitals, potentials, eigen_values, scaling_factor=1.0,
                               lmax_keep=13):
    """HelmholtzSCFCycle_update_orbitals_worker (scf_cycle.F90:2927-3054) on
    orbitals that carry bubbles (Function3D: cube + bubbles), in energy order:

      input   = V psi_i (potentials[i]), level-shifted by -(eps_i + 0.1) psi_i
                and the energy set to -0.1 when eps_i > 0
      G       = Helmholtz3D(eps_i) operate_on input: Bessel-convolved bubbles
                and the cube transform, then x -1/(2 pi)
      update  = G with bubbles above l = 13 dropped and those above the
                orbital's lmax injected into the cube, minus psi_i, times
                scaling_factor
      psi_i  += update, normalised with the Function3D dot product (cube
                integral of the product plus the bubbles' radial integrals)

    op: a Helmholtz3D on the orbitals' grid; orbitals / potentials: lists of
    dage_amd.function3d.Function3D; the orbitals are updated in place."""
    from .function3d import operate_on
    order = np.argsort(np.asarray(eigen_values, dtype=float), kind="stable")
    for i in order:
        psi = orbitals[i]
        inp = potentials[i].copy()
        e = float(eigen_values[i])
        if e > 0.0:
            inp.axpby(-(e + 0.1), psi, 1.0)
            e = -0.1
        op.set_energy(e)
        G = operate_on(op, inp)
        inp.close()
        G.scale(-1.0 / (2.0 * math.pi))
        G.dispose_extra_bubbles(lmax_keep)
        if psi.bubbles is not None:
            G.inject_extra_bubbles(psi.bubbles.lmax)
        G.axpby(-1.0, psi, 1.0)
        G.scale(scaling_factor)
        psi.axpby(1.0, G, 1.0)
        G.close()
        psi.scale(1.0 / math.sqrt(psi.dot(psi)))
    return orbitals


class BubbleHartreeFock:
    """Closed-shell Hartree-Fock on orbitals that carry bubbles: the
    reference's RHF Helmholtz cycle (HelmholtzSCFCycle, scf_cycle.F90:
    2389-2395 Lowdin, 1434-1470 diagonalize, 2007-2070 Hamiltonian matrix,
    2927-3054 update) with every orbital a Function3D: cube + lmax-`lmax`
    bubbles on every atom's radial grid.

    Operators, all through the C-ABI on the GPU:
      V_nuc    bubbles -Z/r (k = -1) on every atom, zero cube
      rho      2 sum_j psi_j psi_j (Function3D products, bubbles to 2 lmax)
      J        Coulomb3D operate_on rho (radial Poisson + cube transform)
      K psi_i  sum_j psi_j Coulomb3D[psi_i psi_j]   (exchange=True)
      V psi_i  (V_nuc + J) psi_i - K psi_i          (bubbles to 3 lmax, k = -1)
    One iteration: Hamiltonian matrix in span(psi), rotation to its
    eigenvectors (eigenvalues eps), the Helmholtz update of every orbital
    (helmholtz_update_functions' steps), Lowdin orthonormalisation, and the
    potentials of the new orbitals.

    The kinetic matrix needs no Laplacian: the update psi~ = -1/(2 pi)
    G(e)[V psi - s psi] solves (T - e) psi~ = -(V psi - s psi), so
    T psi~ = e psi~ - (V psi - s psi) is formed from functions already at
    hand and carried through every linear step applied to the orbitals
    (mixing, normalisation, Lowdin, rotation); <psi_i|T psi_j> is then the
    kinetic matrix of the same Green's function the update uses, so its
    fixed points are the Ritz pairs exactly.  The guess (single-exponential
    s functions e^(-zeta r) on the atoms) has T psi in closed form:
    -1/2 laplacian e^(-zeta r) = r^-1 (zeta - zeta^2 r / 2) e^(-zeta r).

    Energy: E = sum_i (h_ii + eps_i) + E_nn with sum_i h_ii = tr T +
    1/2 int V_nuc rho (occupation 2).

    Products of the cusp-type functions carry the Taylor-series terms
    (Function3D.multiply), which keep -Z/r times the cube and the other
    centres' bubbles out of the cube.  taylor=None uses them when there is
    more than one atom: with one atom every function lies in the bubbles and
    the cube holds only rounding noise, whose interpolant derivatives at the
    nucleus (noise / h^2) the Taylor series would feed back into the cube
    (measured: the noise grows 1.4-2.8x per iteration at h = 0.2-0.4)."""

    def __init__(self, ctx: Context, grid: Grid3D, molecule: Molecule, nocc: int,
                 gauss: GaussQuad | None = None, lmax: int = 4, ncell: int = 33,
                 r_max: float = 20.0, exchange: bool = True, interaction: bool = True,
                 update_weight: float = 1.0, taylor: bool | None = None):
        """interaction=False drops J and K (one-electron problem in V_nuc)."""
        from .function3d import F3D_TYPE_CUSP, Function3D
        from .operators import Bubbles, Helmholtz3D, radial_grid_steps
        self.ctx, self.grid, self.mol, self.nocc = ctx, grid, molecule, nocc
        self.gauss = gauss if gauss is not None else GaussQuad(20, 20, tlin=2.0, tlog=500.0)
        self.lmax, self.exchange, self.interaction = lmax, exchange, interaction
        self.w = float(update_weight)
        self._F3, self._B, self._cusp = Function3D, Bubbles, F3D_TYPE_CUSP
        self.steps = [radial_grid_steps(float(z), ncell, 7, r_max) for z in molecule.Z]
        self.npts = [len(s) * 6 + 1 for s in self.steps]
        self.nat = len(molecule.Z)
        self.taylor = (self.nat > 1) if taylor is None else bool(taylor)
        vb = Bubbles(ctx, self.nat, 0, -1, 7)
        for a in range(self.nat):
            vb.set_bubble(a, molecule.R[a], float(molecule.Z[a]), self.steps[a],
                          np.full((self.npts[a], 1), -float(molecule.Z[a]), order="F"))
        self.vnuc = Function3D(grid, DeviceCube(ctx, grid.get_shape()), vb, type=F3D_TYPE_CUSP,
                               label="nuclear")
        self.vnuc.cube.set_to_zero()
        self.coul = Coulomb3D(ctx, grid, grid, self.gauss)
        self.helm = Helmholtz3D(ctx, grid, grid, self.gauss)
        self.psi, self.tpsi, self.vpsi = [], [], []
        self.eps = np.zeros(nocc)
        self.rho = None

    # ------------------------------------------------------------ helpers
    def _r(self, a):
        """Radial points of atom a's bubble grid: 7 equidistant LIP nodes per
        cell, steps[c] the node spacing of cell c."""
        st = self.steps[a]
        starts = np.concatenate([[0.0], np.cumsum(6.0 * st)[:-1]])
        r = (starts[:, None] + st[:, None] * np.arange(6)[None, :]).ravel()
        return np.concatenate([r, [6.0 * np.sum(st)]])

    def _new(self, label="f"):
        f = self._F3(self.grid, DeviceCube(self.ctx, self.grid.get_shape()), None,
                     type=self._cusp, label=label)
        f.cube.set_to_zero()
        return f

    def _lincomb(self, funcs, C):
        """out_j = sum_i C[i, j] funcs_i (cubes in one launch, bubbles by axpby);
        all inputs share lmax and k."""
        n = int(np.prod(self.grid.get_shape()))
        out = [self._new(funcs[0].label) for _ in range(C.shape[1])]
        c = np.asfortranarray(C, dtype=np.float64)
        call("dage_vec_lincomb", self.ctx.handle, n, len(funcs),
             _ptr_array([f.cube for f in funcs]), c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
             len(out), _ptr_array([f.cube for f in out]))
        src = funcs[0].bubbles
        if src is not None:
            for j, f in enumerate(out):
                f.bubbles = self._B.like(src, k=src.k)
                for i, g in enumerate(funcs):
                    if C[i, j] != 0.0:
                        f.bubbles.axpby(float(C[i, j]), g.bubbles, 1.0)
        return out

    def _replace(self, name, new):
        for f in getattr(self, name):
            f.close()
        setattr(self, name, new)

    def _mul(self, a, b):
        return a.multiply(b, taylor=None if self.taylor else False)

    def _dot(self, a, b):
        return a.dot(b, taylor=None if self.taylor else False)

    def _gram(self, f, g, symmetric=False):
        M = np.zeros((len(f), len(g)))
        for i in range(len(f)):
            for j in range(len(g)):
                if symmetric and j < i:
                    M[i, j] = M[j, i]
                else:
                    M[i, j] = self._dot(f[i], g[j])
        return M

    # ------------------------------------------------------------- guess
    def initial_guess(self, exponents=None):
        """nocc s functions e^(-zeta r): zeta = Z on every atom (heaviest
        first), then zeta = Z/2, Z/3, ... in further rounds; with T psi in
        closed form, then orthonormalised and rotated like any iterate."""
        order = np.argsort(-self.mol.Z, kind="stable")
        funcs, tfuncs = [], []
        rnd = 1
        while len(funcs) < self.nocc:
            for a in order:
                if len(funcs) == self.nocc:
                    break
                zeta = float(self.mol.Z[a]) / rnd if exponents is None else exponents[a] / rnd
                f, t = self._new("orbital"), self._new("T orbital")
                f.bubbles = self._B(self.ctx, self.nat, self.lmax, 0, 7)
                t.bubbles = self._B(self.ctx, self.nat, self.lmax, -1, 7)
                nlm = (self.lmax + 1) ** 2
                for b in range(self.nat):
                    bf = np.zeros((self.npts[b], nlm), order="F")
                    bt = np.zeros((self.npts[b], nlm), order="F")
                    if b == a:
                        r = self._r(b)
                        e = np.exp(-zeta * r)
                        bf[:, 0] = e
                        bt[:, 0] = (zeta - 0.5 * zeta * zeta * r) * e
                    args = (b, self.mol.R[b], float(self.mol.Z[b]), self.steps[b])
                    f.bubbles.set_bubble(*args, bf)
                    t.bubbles.set_bubble(*args, bt)
                funcs.append(f)
                tfuncs.append(t)
            rnd += 1
        self._replace("psi", funcs)
        self._replace("tpsi", tfuncs)
        self.eps = np.full(self.nocc, -0.5)
        self.lowdin()
        self.orbital_potentials()
        return self.diagonalize()

    # ------------------------------------------------------------ steps
    def lowdin(self):
        """HelmholtzSCFCycle_lowdin_orthogonalize_orbitals: psi <- psi S^-1/2."""
        S = self._gram(self.psi, self.psi, symmetric=True)
        s, U = np.linalg.eigh(0.5 * (S + S.T))
        X = (U / np.sqrt(s)) @ U.T
        self._replace("psi", self._lincomb(self.psi, X))
        self._replace("tpsi", self._lincomb(self.tpsi, X))
        return S

    def orbital_potentials(self):
        """V psi_i for the current (orthonormal) orbitals; also rho."""
        F3 = self._F3
        if self.rho is not None:
            self.rho.close()
        rho = self._new("density")
        for p in self.psi:
            sq = self._mul(p, p)
            rho.axpby(2.0, sq, 1.0)
            sq.close()
        self.rho = rho
        # (V_nuc + J): k = -1 bubbles of J's lmax
        from .function3d import operate_on
        if self.interaction:
            J = operate_on(self.coul, rho)
            vj = F3(self.grid, J.cube, self._B.like(J.bubbles, k=-1), type=self._cusp,
                    label="V+J")
            vj.bubbles.axpby(1.0, J.bubbles, 0.0)
            vj.bubbles.axpby(1.0, self.vnuc.bubbles, 1.0)
            J.bubbles.close()
        else:
            vj = self.vnuc.copy()
        vpsi = [self._mul(vj, p) for p in self.psi]
        vj.close()
        if self.interaction and self.exchange:
            n = len(self.psi)
            for i in range(n):
                for j in range(i, n):
                    pair = self._mul(self.psi[i], self.psi[j])
                    pot = operate_on(self.coul, pair)
                    pair.close()
                    kij = self._mul(self.psi[j], pot)
                    vpsi[i].axpby(-1.0, kij, 1.0)
                    kij.close()
                    if j != i:
                        kji = self._mul(self.psi[i], pot)
                        vpsi[j].axpby(-1.0, kji, 1.0)
                        kji.close()
                    pot.close()
        self._replace("vpsi", vpsi)

    def hamiltonian(self):
        """(H, tr T) in span(psi): H_ij = <psi_i | T psi_j + V psi_j>
        symmetrised, and sum_i <psi_i | T psi_i>.  T psi_j + V psi_j is summed
        into a k = -1 function whose bubbles above the orbitals' lmax are
        injected into its cube, so that the products stay within 2 lmax."""
        n = len(self.psi)
        fpsi = []
        for j in range(n):
            f = self.vpsi[j].copy()
            f.axpby(1.0, self.tpsi[j], 1.0)
            f.inject_extra_bubbles(self.lmax)
            fpsi.append(f)
        H = self._gram(self.psi, fpsi)
        for f in fpsi:
            f.close()
        trT = sum(self._dot(p, t) for p, t in zip(self.psi, self.tpsi))
        return 0.5 * (H + H.T), float(trT)

    def diagonalize(self):
        """SCFCycle_diagonalize_worker on orthonormal orbitals: eigenvectors
        of H, orbitals (and T psi, V psi) rotated to them; returns the
        total energy."""
        H, trT = self.hamiltonian()
        e, C = np.linalg.eigh(H)
        for j in range(C.shape[1]):
            if C[np.argmax(np.abs(C[:, j])), j] < 0:
                C[:, j] = -C[:, j]
        self._replace("psi", self._lincomb(self.psi, C))
        self._replace("tpsi", self._lincomb(self.tpsi, C))
        self._replace("vpsi", self._lincomb(self.vpsi, C))
        self.eps = e.copy()
        en = self._dot(self.vnuc, self.rho)
        self.energy = trT + 0.5 * en + float(np.sum(e)) + self.mol.nuclear_repulsion()
        self.kinetic = 2.0 * trT
        return self.energy

    def update(self):
        """HelmholtzSCFCycle_update_orbitals_worker, keeping T psi: for each
        orbital (energy order) the level-shifted input, the Helmholtz
        operate_on, x -1/(2 pi), bubbles above l = 13 dropped and above the
        orbital's lmax injected, psi += w (psi~ - psi), normalisation."""
        from .function3d import operate_on
        order = np.argsort(self.eps, kind="stable")
        for i in order:
            psi, tpsi = self.psi[i], self.tpsi[i]
            inp = self.vpsi[i].copy()
            e = float(self.eps[i])
            if e > 0.0:
                inp.axpby(-(e + 0.1), psi, 1.0)
                e = -0.1
            self.helm.set_energy(e)
            G = operate_on(self.helm, inp)
            G.scale(-1.0 / (2.0 * math.pi))
            G.dispose_extra_bubbles(13)
            # T G = e G - inp, reduced to the orbitals' lmax like G
            tg = G.copy()
            tg_b = self._B.like(inp.bubbles, k=inp.bubbles.k)
            tg_b.axpby(1.0, G.bubbles, 0.0)
            tg.bubbles.close()
            tg.bubbles = tg_b
            tg.axpby(-1.0, inp, e)       # cube: e G - inp; bubbles: e G - inp
            inp.close()
            G.inject_extra_bubbles(self.lmax)
            tg.inject_extra_bubbles(self.lmax)
            # psi <- (1 - w) psi + w G and the same for T psi
            G.axpby(1.0 - self.w, psi, self.w)
            tg.axpby(1.0 - self.w, tpsi, self.w)
            G.type = tg.type = self._cusp
            norm = 1.0 / math.sqrt(self._dot(G, G))
            G.scale(norm)
            tg.scale(norm)
            psi.close()
            tpsi.close()
            self.psi[i], self.tpsi[i] = G, tg

    def iterate(self):
        """One SCF iteration after the guess: update, Lowdin, potentials,
        diagonalisation; returns the energy."""
        self.update()
        self.lowdin()
        self.orbital_potentials()
        return self.diagonalize()

    def run(self, max_iter=50, e_tol=1e-6, eig_tol=1e-6, log=None) -> SCFResult:
        res = SCFResult()
        t0 = time.perf_counter()
        res.energies.append(self.initial_guess())
        res.eigen_values.append(self.eps.copy())
        res.times.append(time.perf_counter() - t0)
        for it in range(max_iter):
            res.energies.append(self.iterate())
            res.eigen_values.append(self.eps.copy())
            res.times.append(time.perf_counter() - t0)
            res.iterations = it + 1
            if log is not None:
                log(it + 1, res.energies[-1], self.eps)
            if is_converged(res.energies, res.eigen_values, e_tol, eig_tol):
                res.converged = True
                break
        res.wall_s = time.perf_counter() - t0
        return res

    def close(self):
        for name in ("psi", "tpsi", "vpsi"):
            self._replace(name, [])
        for obj in (self.rho, self.vnuc, self.coul, self.helm):
            if obj is not None:
                obj.close()
        self.rho = None
