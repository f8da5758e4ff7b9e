/*
 * dage_oracle.h -- CPU FP64 restatement of DAGE's operator-application hot
 * path (tensor-quadrature Coulomb3D/Helmholtz3D apply).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * library in dage_amd/ is compared against.  Only tests/, the smoke entry
 * point and the CPU-baseline leg of bench.py may load it; the product path
 * never calls it and has no CPU fallback.
 *
 * Every function cites the reference routine it restates (paths relative to
 * the reference repository root).  Arrays are Fortran column-major, 0-based
 * in C.
 */
#ifndef DAGE_ORACLE_H
#define DAGE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Constants as src/bubbles/globals.F90:54-61 defines them. */
double dage_oracle_pi(void);

/* gauleg: src/bubbles/gauss.F90:337-384 (Newton on P_n, EPS 3e-15). */
void dage_oracle_gauleg(int ngp, double *xabsc, double *weig);

/* GaussQuadInit + gauleg_int: src/bubbles/gauss.F90:81-122, 307-329.
 * tp, tw have nlin+nlog entries. */
void dage_oracle_gauss_quad(int nlin, int nlog, double tstart, double tlin,
                            double tlog, double *tp, double *tw);

/* gauss_lobatto_grid: src/bubbles/gauss.F90:409-446 (n = 6, 7 only).
 * Returns 0 on success, -1 for an unsupported n. */
int dage_oracle_gauss_lobatto_grid(int n, double left, double right,
                                   double *tp);

/* LIPBasisInit first/last: src/bubbles/lip.F90:118-142. */
void dage_oracle_lip_first_last(int nlip, int grid_type, double *first,
                                double *last);

/* LIPBasis_coeffs(0): src/bubbles/lip.F90:285-344.  coeffs is nlip x nlip
 * column-major; coeffs[j + nlip*k] is the coefficient of x^(nlip-1-k) of
 * basis polynomial j.  Returns -1 for an unsupported grid type. */
int dage_oracle_lip_coeffs(int nlip, int grid_type, double *coeffs);

/* LIPBasis_integrals: src/bubbles/lip.F90:367-415. */
int dage_oracle_lip_integrals(int nlip, int grid_type, double *integrals);

/* Grid1D_init_step: src/bubbles/grid.F90:446-537.  points has
 * ncell*(nlip-1)+1 entries, celld has ncell entries. */
int dage_oracle_grid1d(double qmin, int ncell, int nlip, const double *steps,
                       int grid_type, double *points, double *celld);

/* Grid1D_init_equidistant cell count and step: src/bubbles/grid.F90:539-567
 * (also Grid3D_init_equidistant :704-751). */
void dage_oracle_equidistant_cells(double r0, double r1, int nlip,
                                   double stepmax, int *ncell, double *step);

/* Grid1D_get_ints: src/bubbles/grid.F90:1398-1424 (whole axis). */
int dage_oracle_grid1d_ints(int ncell, int nlip, const double *steps,
                            int grid_type, double *ints);

/* integrals(a,b,t,p,n): src/bubbles/exponentials.F90:647-666.
 * values[n-1-i] = int_a^b x^i exp(-t^2 (x-p)^2) dx. */
void dage_oracle_integrals(double a, double b, double t, double p, int n,
                           double *values);

/* expo_0 .. expo_6 closed forms: src/bubbles/exponentials.F90:695-1297. */
double dage_oracle_expo(int n, double a, double b, double t, double p);

/* gauss_expo_n: src/bubbles/exponentials.F90:2287-2429. */
double dage_oracle_gauss_expo(int n, double a, double b, double t, double p);

/* exponentials(): src/bubbles/exponentials.F90:178-346 for one axis, no
 * periodic images.  Input axis: (qmin_in, ncell_in, steps_in); output axis
 * points (points_out, nout).  When transpose == 0 the result is
 * (nout, nin, T) column-major (the X matrix); otherwise (nin, nout, T) (the
 * Y and Z matrices, exponentials.F90:629-634). */
int dage_oracle_exponentials(double qmin_in, int ncell_in,
                             const double *steps_in, int nlip, int grid_type,
                             const double *points_out, int nout,
                             const double *tp, int ntp, int transpose,
                             double *out);

/* Operator3D_transform_cube: src/bubbles/function3d.F90:4230-4349, CPU
 * branch (loop :4301-4314), then out += coda*in (:4327-4332) when coda != 0
 * (same input/output grid, where cube_project is the identity copy,
 * :4143-4146).
 * fx: (nxo,nxi,T), fy: (nyi,nyo,T), fz: (nzi,nzo,T); in: (nxi,nyi,nzi);
 * out: (nxo,nyo,nzo), overwritten.  Only t-points [t0, t1) are summed
 * (t0=0, t1=T for the full operator); the coda is added only when
 * add_coda != 0. */
int dage_oracle_transform_cube(const int dims_in[3], const int dims_out[3],
                               int ntp, const double *fx, const double *fy,
                               const double *fz, const double *w, double coda,
                               const double *in, double *out, int t0, int t1,
                               int add_coda);

/* Helmholtz3D_set_energy weights: src/bubbles/helmholtz3d.F90:411-444 and
 * Coulomb3D w = (2/sqrt(pi))*tw, coulomb.F90:150. */
void dage_oracle_helmholtz_weights(const double *tp, const double *tw, int ntp,
                                   double energy, double *w);

/* GaussGenerator_gen_cubef: src/bubbles/gauss_generator.F90:291-318. */
void dage_oracle_gauss_density(int ngau, const double *q, const double *alpha,
                               const double *centers, const double *x, int nx,
                               const double *y, int ny, const double *z,
                               int nz, double *cube);

/* GaussPotGenerator_gen_cubef: src/bubbles/gauss_generator.F90:384-429. */
void dage_oracle_gauss_potential(int ngau, const double *q,
                                 const double *alpha, const double *centers,
                                 const double *x, int nx, const double *y,
                                 int ny, const double *z, int nz,
                                 double *cube);

/* Cube integral sum_{ijk} wx_i wy_j wz_k f_ijk (reduce3d / integrate_cube,
 * src/bubbles/integrator.cu:320-373, parallel_info.F90 integrate_cube). */
double dage_oracle_cube_integral(const double *wx, int nx, const double *wy,
                                 int ny, const double *wz, int nz,
                                 const double *f, const double *g);

/* Grid1D_init_radial cell steps (src/bubbles/grid.F90:569-603) for a given
 * cell count: x_i = c/((r_max+c)/(i dx) - 1), c = 8 z^-1.5, dx = r_max/ncell,
 * step_i = (x_i - x_{i-1})/(nlip-1).  (The reference takes ncell =
 * int(n0 z^0.25).) */
void dage_oracle_radial_steps(double z, int ncell, int nlip, double r_max,
                              double *steps);

/* Grid1D_get_icell (src/bubbles/grid.F90:1212-1246): 1-based cell index of
 * x by bisection over the cell starts, 0 outside [qmin, qmax]. */
int dage_oracle_get_icell(const double *celld, int ncell, double qmin,
                          double qmax, double x);

/* Racah-normalised real regular solid harmonics S_lm(x,y,z), l = 0..lmax,
 * S[l*l + l + m] (src/bubbles/harmonics.F90:586-659 recursion). */
void dage_oracle_solid_harmonics(int lmax, double x, double y, double z,
                                 double *S);

/* Bubbles_eval_3dgrid CPU path (src/bubbles/bubbles.F90:2148-2211 ->
 * Bubbles_eval_all :2017-2073 -> Bubbles_simple_eval_pick :2075-2146):
 * cube[i,j,k] += sum_b sum_{l>=lmin} interp(r^k f_lm)(r) * S_lm/r^l
 * with SimpleInterpolator1D_eval (src/bubbles/evaluators.F90:508-550).
 * Bubble b has radial grid (ncell[b] cells of nlip points from r = 0, steps
 * steps[b]) and radial functions bf[b] (npts x (lmax+1)^2, column-major). */
int dage_oracle_bubbles_eval_grid(int nbub, const double *centers, int lmax,
                                  int k, int lmin, const int *ncell, int nlip,
                                  const double *const *steps,
                                  const double *const *bf, const double *x,
                                  int nx, const double *y, int ny,
                                  const double *z, int nz, double *cube);

/* ------------------------------------------------------------------- FMM */
/* InteractionMatrix (multipole.F90:104-183, 268-312): T (N x N, column-major,
 * N = (lmax+1)^2), L_target += T(dR) q_source with dR = source - target. */
int dage_oracle_fmm_interaction(int lmax, const double dR[3], double *T);
/* TranslationMatrix (multipole.F90:349-470): packed unit-lower W(dR). */
int dage_oracle_fmm_translation(int lmax, const double dR[3], double *W);
/* TranslationMatrix_apply (multipole.F90:475-492): q := W q (transpose 0) or
 * W^T q (transpose 1). */
void dage_oracle_fmm_translate(int lmax, const double *W, int transpose, double *q);
/* Box multipoles (integrator.cu:1122-1196), lmax <= 24. */
void dage_oracle_box_multipoles(int lmax, const double *rho, int ld_x, int ld_y,
                                const double *x, const double *wx, int nx,
                                const double *y, const double *wy, int ny,
                                const double *z, const double *wz, int nz,
                                const double centre[3], double *q);
/* Local-expansion evaluation (gbfmm_coulomb3d.cu:81-280). */
void dage_oracle_eval_local_expansion(int lmax, const double *L, const double *x, int nx,
                                      const double *y, int ny, const double *z, int nz,
                                      const double centre[3], double *out, int ld_x, int ld_y);

/* ------------------------------------------------ bubble radial transforms */
/* LIPBasis_inward_integrals / _outward_integrals (lip.F90:364-485): nlip x
 * (nlip-1) column-major each. */
int dage_oracle_lip_inout_integrals(int nlip, int grid_type, double *ints_in,
                                    double *ints_out);
/* InOutIntegrator1D_inwards / _outwards (evaluators.F90:1263-1320) on a
 * radial grid from r = 0 (either output may be NULL). */
int dage_oracle_integrate_inout(int ncell, int nlip, const double *steps, const double *f,
                                double *inw, double *outw);
/* FirstModSphBesselCollection_evaluate / SecondModSphBesselCollection_evaluate
 * (helmholtz3d.F90:879-1000, 1133-1210) at one point, l = 0..lmax, computed in
 * __float128 as the reference's REAL128. */
void dage_oracle_bessel_i(int lmax, double x, double *out);
void dage_oracle_bessel_k(int lmax, double x, double *out);
/* Coulomb3D_evaluate_bubbles_potential (coulomb.F90:222-336) for one bubble:
 * bf, out: npts x (lmax+1)^2 column-major; k of the input bubbles. */
int dage_oracle_bubble_coulomb(int lmax, int k, int ncell, int nlip, const double *steps,
                               const double *bf, double *out);
/* Helmholtz3D_transform_bubbles_sub (helmholtz3d.F90:519-657) for one bubble
 * (equidistant cells), kappa = sqrt(-2 energy). */
void dage_oracle_exp_inout(int nlip, double mu, double *ii, double *io);
int dage_oracle_bubble_helmholtz(int lmax, int k, int ncell, int nlip, const double *steps,
                                 double energy, const double *bf, double *out);

/* --------------------------------------------------------------- products */
/* YProduct (harmonic.F90:884-1000): dense product coefficients of Racah real
 * spherical harmonics, coef[i + n1*(j + n2*o)], L = l1max + l2max. */
int dage_oracle_yproduct(int l1max, int l2max, double *coef);
/* BubblesMultiplier_multiply without Taylor terms (bubbles.F90:3703-3756),
 * one bubble: out (npts x (l1max+l2max+1)^2). */
int dage_oracle_bubble_product(int l1max, int l2max, int npts, const double *f1,
                               const double *f2, double *out);

/* OpenMP threads of transform_cube and bubbles_eval_grid (their parallel
 * loops give the serial loops' results bit for bit); returns the count in
 * use (1 without OpenMP). */
int dage_oracle_set_threads(int n);
/* GBFMM far field summed directly over the points of the non-adjacent
 * boxes, kernel exp(-kappa r)/r (kappa = 0: 1/r); pot +=. */
void dage_oracle_gbfmm_far_direct(int nb, const int *lo, const int *n, const int *cut,
                                  const double *w, int wld, const double *x, const double *y,
                                  const double *z, int ldx, int ldy, const double *rho,
                                  double kappa, double *pot);

#ifdef __cplusplus
}
#endif
#endif
