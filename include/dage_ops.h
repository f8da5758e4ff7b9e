/*
 * dage_ops.h -- C-ABI of the MI355X (gfx950) operator-application library.
 *
 * Drop-in boundary for DAGE's Operator3D GPU path.  Each entry point names
 * the reference routine or bind(C) interface it replaces (paths relative to
 * the reference repository root).  A Fortran ISO_C_BINDING module that
 * mirrors the reference's Operator3D API on top of these calls is in
 * dage_amd/fortran/dage_ops_mod.F90.
 *
 * Conventions
 *  - Every function returns an int status: DAGE_OK (0) or a DAGE_ERR_* code.
 *    dage_last_error() returns the message of the last failure on the
 *    calling host thread.  (The reference prints and calls exit(-1),
 *    src/bubbles/streamcontainer.h:67-95.)
 *  - Cubes are Fortran column-major real(8) arrays, x fastest:
 *    index = x + nx*(y + ny*z).  Indices are 0-based in C.
 *  - Host pointers are owned by the caller; device memory allocated by the
 *    library is owned by the library.  *_device entry points take device
 *    pointers (hipMalloc'd, caller-owned) and enqueue on the context stream
 *    without synchronising.
 *  - A context is one HIP device plus one stream.  Handles may be used from
 *    several host threads when each thread uses its own context (the
 *    reference calls the CUDA path from 3 OpenMP threads,
 *    src/gbfmm/gbfmm_coulomb3d.F90:767-785).
 */
#ifndef DAGE_OPS_H
#define DAGE_OPS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DAGE_ABI_VERSION 1

enum {
    DAGE_OK = 0,
    DAGE_ERR_ARG = 1,         /* invalid argument / shape */
    DAGE_ERR_HIP = 2,         /* HIP runtime error */
    DAGE_ERR_ALLOC = 3,       /* device or host allocation failed */
    DAGE_ERR_STATE = 4,       /* handle in the wrong state */
    DAGE_ERR_UNSUPPORTED = 5  /* feature outside the implemented scope */
};

/* Precision modes of the apply. */
enum {
    DAGE_PREC_FP64 = 0,       /* FP64 MFMA throughout */
    DAGE_PREC_BF16X6 = 1,     /* split-BF16, 3 terms per operand, 6 MFMA
                                 products, FP32 sums of 32 products promoted
                                 to FP64; FP32 intermediates (~5e-8 of
                                 max|V| at config 1) */
    DAGE_PREC_BF16X3 = 2      /* split-BF16, 2 terms, 3 products (~2e-6) */
};

/* Operator creation flags. */
enum {
    DAGE_OP_NO_CODA = 1       /* omit coda*rho, as the reference CUDA path
                                 does (src/bubbles/function3d.F90:4473-4479) */
};

typedef struct dage_ctx_s *dage_ctx;
typedef struct dage_grid3d_s *dage_grid3d;
typedef struct dage_operator_s *dage_operator;
typedef struct dage_cube_s *dage_cube;
typedef struct dage_bubbles_s *dage_bubbles;
typedef struct dage_gbfmm_s *dage_gbfmm;
typedef struct dage_batch_s *dage_batch;

/* ----------------------------------------------------------------- library */
int dage_abi_version(void);
const char *dage_last_error(void);

/* ----------------------------------------------------------------- context
 * Replaces streamcontainer_init(StreamContainer**, int)
 * (src/bubbles/streamcontainer.cpp:377) for one device. */
int dage_ctx_create(int device, dage_ctx *ctx);
int dage_ctx_destroy(dage_ctx ctx);
/* Use a caller-provided hipStream_t (NULL restores the context's own).
 * Synchronises the outgoing stream first, so the caller may destroy it
 * afterwards. */
int dage_ctx_set_stream(dage_ctx ctx, void *hip_stream);
int dage_ctx_get_stream(dage_ctx ctx, void **hip_stream);
/* Replaces cuda_sync_all_ (src/bubbles/cublas_iface.cpp) for this stream. */
int dage_ctx_synchronize(dage_ctx ctx);
/* Per-kernel timing with HIP events on the context stream (the reference's
 * bigben/NVTX ranges, src/bubbles/chrono.F90:85-200,
 * src/bubbles/cuda_profiling.cu:58-90).  enable != 0 starts recording and
 * clears the counters. */
int dage_ctx_profile(dage_ctx ctx, int enable);
/* Total milliseconds and launch count of kernel `name` since profiling was
 * enabled (synchronises the stream). */
int dage_ctx_profile_query(dage_ctx ctx, const char *name, double *total_ms,
                           int *launches);

/* Event timer on the context stream: begin records a start event; end
 * records a stop event, synchronises and returns the elapsed milliseconds. */
int dage_ctx_time_begin(dage_ctx ctx);
int dage_ctx_time_end(dage_ctx ctx, double *ms);

/* -------------------------------------------------------------- device cube
 * Device-resident cube, x fastest, replacing the reference's CudaCube
 * (src/bubbles/cube.cu:1182-1269; Fortran side src/bubbles/cuda.F90:70-200):
 * cudacube_init_cuda, cudacube_upload_cuda, cudacube_download_cuda,
 * cudacube_set_to_zero_cuda, cudacube_destroy_cuda.  The memory is one
 * contiguous allocation (no pitch, no multi-device split); a new cube holds
 * zeros (filled on the context stream). */
int dage_cube_create(dage_ctx ctx, const int shape[3], dage_cube *cube);
int dage_cube_destroy(dage_cube cube);
/* Synchronises the context stream before returning (the host buffer may be
 * reused immediately). */
int dage_cube_upload(dage_cube cube, const double *host);
/* Synchronises the context stream before returning. */
int dage_cube_download(dage_cube cube, double *host);
int dage_cube_set_to_zero(dage_cube cube);
/* Device address of the cube, for the *_device entry points. */
int dage_cube_device_pointer(dage_cube cube, double **dev);
/* dst = src on the device (same shape), with a streaming copy kernel
 * ("cube_copy_kernel"); asynchronous on the context stream. */
int dage_cube_copy(dage_cube dst, dage_cube src);

/* -------------------------------------------------------------- quadrature
 * GaussQuadInit (src/bubbles/gauss.F90:81-122, gauleg_int :307-329,
 * gauleg :337-384).  tp, tw receive nlin+nlog values. */
int dage_gauss_quad(int nlin, int nlog, double tstart, double tlin,
                    double tlog, double *tp, double *tw);

/* -------------------------------------------------------------------- grid
 * Grid3D_init_step (src/bubbles/grid.F90:652-702) with Grid1D_init_step
 * (:446-537).  steps_x/y/z hold ncell[axis] cell steps.  grid_type 1 is
 * equidistant cells, 2 Gauss-Lobatto cells (nlip 6 or 7).  Replaces
 * grid3d_init_cuda / grid1d_init_cuda (src/bubbles/grid.cu:498-521). */
int dage_grid3d_create(dage_ctx ctx, const double qmin[3], const int ncell[3],
                       int nlip, const double *steps_x, const double *steps_y,
                       const double *steps_z, int grid_type,
                       dage_grid3d *grid);
/* Grid3D_init_equidistant (src/bubbles/grid.F90:704-751); ranges is the
 * Fortran (2,3) array {xmin,xmax,ymin,ymax,zmin,zmax}. */
int dage_grid3d_create_equidistant(dage_ctx ctx, const double ranges[6],
                                   int nlip, double stepmax, int grid_type,
                                   dage_grid3d *grid);
int dage_grid3d_destroy(dage_grid3d grid);
int dage_grid3d_shape(dage_grid3d grid, int shape[3]);
/* Grid1D_get_coord (grid.F90) of one axis (0,1,2) into host memory. */
int dage_grid3d_coord(dage_grid3d grid, int axis, double *points);
/* Grid1D_get_ints (grid.F90:1398-1424): LIP integration weights. */
int dage_grid3d_weights(dage_grid3d grid, int axis, double *weights);

/* ---------------------------------------------------------------- operator
 * Coulomb3D_init_explicit_sub (src/bubbles/coulomb.F90:109-176): w =
 * (2/sqrt(pi))*tw, coda = pi/tlog^2, and the 1D operator matrices of
 * make_Coulomb3D_matrices (src/bubbles/exponentials.F90:602-639), generated
 * on the GPU in FP64 from the reference's erf closed forms; then
 * Operator3D_cuda_init (src/bubbles/function3d.F90:4172-4209).
 * gridin/gridout must share nlip and grid_type. */
int dage_coulomb3d_create(dage_ctx ctx, dage_grid3d gridin,
                          dage_grid3d gridout, const double *tp,
                          const double *tw, int ntp, double tlog, int flags,
                          dage_operator *op);
/* Operator3D_set_transformation_weights (function3d.F90:320-336). */
int dage_operator_set_weights(dage_operator op, const double *w);
int dage_operator_get_weights(dage_operator op, double *w);
/* Helmholtz3D_set_energy (src/bubbles/helmholtz3d.F90:411-444): w_t =
 * (2/sqrt(pi)) tw_t exp(2E/(4 t^2)) for E < -0.005, else Coulomb weights. */
int dage_operator_set_energy(dage_operator op, double energy);
/* Upload externally built matrices, as Operator3D_cuda_init uploads
 * f(X_), f(Y_), f(Z_) (function3d.F90:4200-4206).  fx: (nxo,nxi,T), fy:
 * (nyi,nyo,T), fz: (nzi,nzo,T), host memory.  As for the generated
 * matrices, entries at most 1e-20 times the largest |entry| of their row
 * (one input point, one t) are stored as zeros (environment variable
 * DAGE_FLUSH_CUTOFF: another factor, 0 to keep them): the passes skip the
 * panels such Gaussian tails fill, and a pass's output moves by at most
 * n * 1e-20 * max|F_t| * max|input|. */
int dage_operator_set_matrices(dage_operator op, const double *fx,
                               const double *fy, const double *fz);
/* Download the matrices as applied (same layouts) to host memory. */
int dage_operator_get_matrices(dage_operator op, double *fx, double *fy,
                               double *fz);
/* Operator3D_transform_cube (function3d.F90:4230-4349): cubeout =
 * sum_t w_t (Fx_t x Fy_t x Fz_t) cubein + coda*cubein.  Host buffers: the
 * cube is uploaded, transformed and downloaded, then the stream is
 * synchronised. */
int dage_operator_apply(dage_operator op, const double *cubein,
                        double *cubeout, int precision);
/* Operator3D_transform_cuda_cube (function3d.F90:4354-4480) on device
 * buffers; asynchronous on the context stream. */
int dage_operator_apply_device(dage_operator op, const double *cubein_dev,
                               double *cubeout_dev, int precision);
/* The device apply on sub-cubes: the input box (the input grid's shape) at
 * in with element (i, j, k) at in[i + ldx_in (j + ldy_in k)], the output box
 * likewise (ld >= the grid's own extent; equal: dage_operator_apply_device).
 * Replaces the pitched CudaCube slices and sub-cubes the reference applies
 * to (cudacube_get_subcube_cuda / _get_slice_cuda, src/bubbles/cube.cu:
 * 1042-1177, 1182-1269).  A strided box is packed into the context's
 * auxiliary buffer (grown on demand) before the apply and unpacked after it,
 * so the result equals the contiguous apply's bit for bit; the boxes may
 * overlap when either is strided.  Asynchronous on the context stream. */
int dage_operator_apply_device_strided(dage_operator op, const double *in, int ldx_in,
                                       int ldy_in, double *out, int ldx_out, int ldy_out,
                                       int precision);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the FP64 apply of dage_operator_apply_device, but with
 * its intermediates in a fixed per-GPU scratch array of
 * dage_operator_noalloc_limit() bytes (128 MiB) that the library holds for
 * the life of the process.  The t-chunk shrinks to fit; when the
 * intermediates of one t-point (8*nxo*nyo*nzi bytes with the fused x/y pass,
 * plus the z-pass partials) do not fit, DAGE_ERR_ARG is returned and nothing
 * is launched.  Calls no allocator.  Every call on one GPU shares that GPU's
 * array, so calls that overlap on one GPU (other streams, other contexts,
 * other host threads) race; calls on one stream are ordered. */
int dage_operator_apply_device_noalloc(dage_operator op, const double *cubein_dev,
                                       double *cubeout_dev);
int dage_operator_noalloc_limit(size_t *bytes);
/* Bytes of device workspace the operator holds for its intermediates. */
int dage_operator_workspace_bytes(dage_operator op, size_t *bytes);
/* No reference counterpart (diagnostic): the output-column tiling the FP64
 * passes x, y, z use, chosen from the matrices' zero blocks.  nb[p]: 16-wide
 * column blocks per tile (0: picked from the output width at each apply);
 * pitch[p]: tile pitch in points, 0 for tiles of 16*nb[p] columns;
 * skip[p]: 1 when all-zero K panels are skipped. */
int dage_operator_get_tiling(dage_operator op, int nb[3], int pitch[3], int skip[3]);
/* No reference counterpart (diagnostic): the first t-point whose passes run
 * banded on the FP64 VALU (every nonzero entry of Fx_t, Fy_t and Fz_t within
 * 8 points of the diagonal, for this and all later t); *t0 = T when none (or
 * with DAGE_APPLY_BAND=0 at operator set-up).  An apply takes them when it
 * is large enough to fill the GPU (DAGE_APPLY_BAND=2: always). */
int dage_operator_get_band(dage_operator op, int *t0);
/* No reference counterpart (diagnostic): the t-points 0..t1-1 applied
 * through low-rank factors of their 1D matrices (column-pivoted
 * Gram-Schmidt, truncated where the largest residual column norm falls to
 * 1e-14 of the largest column norm; every axis's rank at most half its
 * length), and the sum of their x ranks.  t1 = 0: none (or with
 * DAGE_APPLY_LOWRANK=0 at operator set-up). */
int dage_operator_get_lowrank(dage_operator op, int *t1, int *kx_sum);
int dage_operator_destroy(dage_operator op);

/* ----------------------------------------------------------------- bubbles
 * Atom-centred radial functions times real solid harmonics
 * (src/bubbles/bubbles.F90:90-230).  Replaces bubbles_init_cuda,
 * bubble_init_cuda, bubble_upload_all_cuda and bubbles_inject_cuda
 * (src/bubbles/bubbles_cuda.cu:3715-3776).  lmax (0..12) and k are shared
 * by all bubbles (r^k f_lm(r) is the radial part); radial grids start at
 * r = 0 with equidistant-cell LIP basis of nlip points. */
int dage_bubbles_create(dage_ctx ctx, int nbub, int lmax, int k, int nlip,
                        dage_bubbles *bubbles);
/* Bubble ibub (0-based): centre, nuclear charge (informational), radial grid
 * (ncell cells with the given steps) and radial functions bf, an
 * npts x (lmax+1)^2 column-major array (npts = ncell*(nlip-1)+1, lm index
 * l*l+l+m), as Fortran holds bf(ibub)%p. */
int dage_bubbles_set(dage_bubbles bubbles, int ibub, const double center[3],
                     double charge, int ncell, const double *steps,
                     const double *bf);
/* Bubbles_eval_3dgrid / bubbles_inject_cuda: cube_dev += multiplier *
 * sum_b sum_{l>=lmin} interp(r^k f_lm)(r) S_lm(r-c_b)/r^l on the points of
 * grid (asynchronous on the context stream). */
int dage_bubbles_inject_device(dage_bubbles bubbles, dage_grid3d grid, int lmin,
                               double multiplier, double *cube_dev);
/* dage_bubbles_inject_device into a sub-cube: the grid's points are the box
 * at cube with element (i, j, k) at cube[i + ldx (j + ldy k)] (bubbles_inject
 * on a CudaCube sub-cube, src/bubbles/bubbles_cuda.cu:3780-3800, cube.cu:
 * 1042-1177).  Packed through the context's auxiliary buffer; equal to the
 * contiguous inject bit for bit.  Asynchronous. */
int dage_bubbles_inject_device_strided(dage_bubbles bubbles, dage_grid3d grid, int lmin,
                                       double multiplier, double *cube, int ldx, int ldy);
/* Bubbles_eval_3dgrid (src/bubbles/bubbles.F90:2148-2211) on a host cube:
 * cube += multiplier * bubbles (l >= lmin) on the points of grid.  The cube
 * is uploaded, injected on the GPU and downloaded; returns after the
 * context stream has finished. */
int dage_bubbles_inject(dage_bubbles bubbles, dage_grid3d grid, int lmin, double multiplier,
                        double *cube);
int dage_bubbles_destroy(dage_bubbles bubbles);
/* Bubbles(src, copy_content=.false., lmax=lmax, k=k) (Bubbles_init_copy,
 * src/bubbles/bubbles.F90:802-830):
 * the same centres, charges and radial grids as src, radial values zero. */
int dage_bubbles_create_like(dage_bubbles src, int lmax, int k, dage_bubbles *out);
int dage_bubbles_info(dage_bubbles bubbles, int *nbub, int *lmax, int *k, int *nlip);
/* Radial points of bubble ibub (ncell*(nlip-1)+1). */
int dage_bubbles_npoints(dage_bubbles bubbles, int ibub, int *npts);
/* Bubble ibub's centre[3], charge, cell steps (ncell) and radial values f
 * (npts x (lmax+1)^2, the values given to dage_bubbles_set, without r^k);
 * any output may be NULL; synchronises. */
int dage_bubbles_get(dage_bubbles bubbles, int ibub, double *center, double *charge,
                     double *steps, double *bf);
/* y = a*x + beta*y on bubbles with the same grids (Bubbles_add and
 * Bubbles_times_REAL64, src/bubbles/bubbles.F90:2499-2535, 3311-3330): x's
 * k may exceed y's (x enters as r^(kx-ky) f_x in y's representation), not
 * fall below it; l above x's lmax: y = beta*y.  Asynchronous. */
int dage_bubbles_axpby(double a, dage_bubbles x, double beta, dage_bubbles y);
/* Coulomb3D_evaluate_bubbles_potential (src/bubbles/coulomb.F90:222-336):
 * out = radial Poisson potential of in (out: same grids and lmax, k = 0),
 * V_lm(r) = 4pi/(2l+1) [r^-(l+1) int_0^r s^(l+2+k) f_lm + r^l int_r s^(1-l+k) f_lm].
 * Asynchronous on the context stream. */
int dage_bubbles_coulomb_transform(dage_bubbles in, dage_bubbles out);
/* Helmholtz3D_transform_bubbles_sub (src/bubbles/helmholtz3d.F90:519-657):
 * out = radial convolution of in with exp(-kappa r)/r, kappa = sqrt(-2E),
 * E < 0 (out: same grids and lmax, k = 0; 7-point radial cells).
 * Asynchronous on the context stream. */
int dage_bubbles_helmholtz_transform(dage_bubbles in, double energy, dage_bubbles out);
/* Bubbles(src, copy_content=.true., lmax=lmax) (Bubbles_init_copy,
 * bubbles.F90:802-830): a new set with lmax (values above src's lmax zero,
 * l above lmax dropped); dispose_extra_bubbles
 * (src/bubbles/function3d.F90:2225-2240) keeps it as the function's
 * bubbles. */
int dage_bubbles_copy_lmax(dage_bubbles src, int lmax, dage_bubbles *out);
/* Function3D_inject_extra_bubbles (src/bubbles/function3d.F90:2198-2222):
 * cube_dev += the l > lmax part of `in` on grid, and *kept = the l <=
 * min(lmax, in's lmax) part of `in` (a new bubbles handle). */
int dage_function3d_inject_extra_bubbles(dage_grid3d grid, double *cube_dev,
                                         dage_bubbles in, int lmax, dage_bubbles *kept);
/* Operator3D_operate_on_Function3D (src/bubbles/function3d.F90:3727-3792)
 * and Helmholtz3D_operate (src/bubbles/helmholtz3d.F90:446-481) on a function
 * of a device cube (operator's input grid) and bubbles: the bubble transform
 * (Coulomb: radial Poisson; after dage_operator_set_energy: Helmholtz with
 * the energy set, which must be negative) into bub_out (same grids and lmax,
 * k = 0), then the FP64 cube transform into cube_out.  only_cube != 0 (or
 * bub_in NULL) skips the bubbles.  A Helmholtz operator is reset to energy 0
 * afterwards, as the reference does (helmholtz3d.F90:479).  Asynchronous on
 * the context stream. */
int dage_operator_operate_on(dage_operator op, const double *cube_in, dage_bubbles bub_in,
                             double *cube_out, dage_bubbles bub_out, int only_cube);
/* Bubbles_integrate_all (src/bubbles/bubbles.F90:3146-3220): sum over
 * bubbles of 4 pi int r^(2+k) f_00(r) dr with the radial LIP weights (only
 * the s part survives the angular integral); synchronises. */
int dage_bubbles_integrate(dage_bubbles bubbles, double *result);
/* BubblesMultiplier_multiply, simple product (src/bubbles/bubbles.F90:
 * 3553-3760): out = a * b centre by centre, g_LM = sum c(l1m1,l2m2->LM) f1 f2
 * with the product coefficients of the Racah harmonics (YProduct,
 * src/bubbles/harmonic.F90:884-1000).  out: same grids, lmax = la + lb,
 * k = ka + kb.  The Taylor-series terms the reference adds for cusp-type
 * factors are dage_function3d_multiply_taylor's.  Asynchronous. */
int dage_bubbles_multiply(dage_bubbles a, dage_bubbles b, dage_bubbles out);
/* Function3D_multiply_sub (src/bubbles/function3d.F90:1097-1308, the product
 * without Taylor-series terms, as the reference forms it when a factor is not
 * cusp-type): bub_out = b1 * b2 (when both given) and
 * cube_out = (cube2 + B2)(cube1 + B1) - B_out, B = bubbles injected on grid.
 * Device cubes on grid; cube_out may be cube2 but not cube1.  Asynchronous. */
int dage_function3d_multiply(dage_grid3d grid, const double *cube1, dage_bubbles b1,
                             const double *cube2, dage_bubbles b2, double *cube_out,
                             dage_bubbles bub_out);
/* Taylor-series "contaminants" of f = cube + bubbles at every centre of b
 * (Function3D_get_cube_contaminants, src/bubbles/function3d.F90:1570-1622,
 * plus Bubbles_get_contaminants, src/bubbles/bubbles.F90:2844-2903): the
 * Cartesian derivatives d^kappa (cube + bubbles of the other centres) at the
 * centre, |kappa| <= taylor_order (0..2), in CartIter order (x exponent
 * outermost, z innermost); out is (taylor_order+1)(taylor_order+2)
 * (taylor_order+3)/6 x nbub, column per centre.  cube may be NULL (bubbles
 * only); a centre outside the grid gets no cube part.  Synchronous. */
int dage_function3d_contaminants(dage_grid3d grid, const double *cube, dage_bubbles b,
                                 int taylor_order, double *out);
/* Function3D_get_taylor_series_bubbles (src/bubbles/function3d.F90:1338-1404,
 * Bubbles_make_taylor, src/bubbles/bubbles.F90:1435-1533): new bubbles on
 * b's centres and grids, k of b, lmax max(lmax_b, taylor_order), holding
 * erfc(r - 2/r)/2 * sum_kappa d_kappa / |kappa|! (x - R)^kappa with the
 * contaminants d above.  Synchronous; *out is the caller's to destroy. */
int dage_function3d_taylor_bubbles(dage_grid3d grid, const double *cube, dage_bubbles b,
                                   int taylor_order, dage_bubbles *out);
/* Function3D_multiply_sub of two cusp-type functions, with the Taylor-series
 * bubbles T1, T2 of both factors (Function3D_multiply_bubbles,
 * src/bubbles/function3d.F90:1406-1470; BubblesMultiplier_multiply,
 * src/bubbles/bubbles.F90:3567-3740): bub_out = (b1 + T1)(b2 + T2) - T1 T2,
 * cube_out = (cube1 + B1)(cube2 + B2) - B_out on the grid.  bub_out: b1's
 * grids, lmax max(l1, t) + max(l2, t), k = k1 + k2.  cube_out may be cube2
 * but not cube1.  Synchronous (the contaminants are read on the host). */
int dage_function3d_multiply_taylor(dage_grid3d grid, const double *cube1, dage_bubbles b1,
                                    const double *cube2, dage_bubbles b2, int taylor_order,
                                    double *cube_out, dage_bubbles bub_out);
/* Grid1D_init_radial (src/bubbles/grid.F90:569-603) cell steps for ncell
 * cells: x_i = c/((r_max+c)/(i*dx)-1), c = 8 z^-1.5, dx = r_max/ncell. */
int dage_radial_grid_steps(double z, int ncell, int nlip, double r_max,
                           double *steps);

/* ------------------------------------------------------------ binary files
 * DAGE's Function3D binary format (.fun: Function3D_dump / Function3D_load,
 * src/bubbles/function3d.F90:812-935; bubbles part Bubbles_write /
 * Bubbles_read_and_init, src/bubbles/bubbles.F90:712-756, 1251-1276) and
 * grid format (.g3d: Grid3D_dump / Grid3D_init_read,
 * src/bubbles/grid.F90:1072-1145).  Host buffers; no GPU work.
 * dage_fun_write: shape NULL stores no cube; nbub < 0 stores no bubbles,
 * otherwise charges[nbub], centers[3*nbub], npts[nbub] and bf[i] (npts[i] x
 * (lmax+1)^2, without r^k) per bubble.  The label is stored in 20 blank-padded
 * characters. */
int dage_fun_write(const char *path, const char *label, int type, const int shape[3],
                   const double *cube, int nbub, int lmax, int k, const double *charges,
                   const double *centers, const int *npts, const double *const *bf);
/* Header of a .fun file: label (21 bytes incl. terminator), type, cube shape
 * (-1s without a cube), bubble count (-1 without bubbles), lmax, k.  Any
 * output may be NULL. */
int dage_fun_read_info(const char *path, char *label, int *type, int shape[3], int *nbub,
                       int *lmax, int *k);
/* Contents of a .fun file (outputs may be NULL); the radial grids are not in
 * the file, so npts (per bubble) comes from the caller's mould. */
int dage_fun_read(const char *path, double *cube, double *charges, double *centers,
                  const int *npts, double *const *bf);
int dage_g3d_write(const char *path, int nlip, int grid_type, const int ncell[3],
                   const double qmin[3], const double *const steps[3]);
/* Reads a .g3d file: nlip, grid_type, ncell[3], qmin[3], and the cell steps
 * into steps[a] (ncell[a] values; steps or steps[a] NULL skips them). */
int dage_g3d_read(const char *path, int *nlip, int *grid_type, int ncell[3], double qmin[3],
                  double *const steps[3]);

/* ---------------------------------------------------------------- cube sums
 * Integrator3D reduce3d (src/bubbles/integrator.cu:320-373) /
 * SerialInfo_integrate_cube (src/bubbles/parallel_info.F90:512-530), and the
 * cube part of Function3D_dot_product (function3d.F90:1497-1530):
 * result = sum_ijk wx_i wy_j wz_k f_ijk g_ijk (g may be NULL).  Device
 * buffers; synchronises and writes *result on the host. */
int dage_cube_integrate_device(dage_grid3d grid, const double *f_dev,
                               const double *g_dev, double *result);

/* -------------------------------------------------------------------- GBFMM
 * Grid-based fast multipole Coulomb potential of a cube
 * (src/gbfmm/gbfmm_coulomb3d.F90:306-1330; GPU side gbfmm_coulomb3d.cu:
 * 550-676, whose gbfmmcoulomb3d_init_cuda / _calculate_multipole_moments_cuda
 * / _evaluate_potential_le_cuda / _destroy_cuda these replace).
 * dage_gbfmm_create (GBFMMCoulomb3D_init + init_nearfield +
 * init_interaction_matrices): the grid's cells split into 2^maxlevel boxes
 * per axis (every axis's cell count must divide), multipoles up to lmax
 * (0..22, 22 being the reference's GBFMM-Helmholtz far-field order; boxes of
 * unequal shape, which take the per-box multipole kernel, up to 15),
 * near-field operators built on the sub-grids with the t-quadrature
 * (tp, tw, ntp) and Coulomb weights 2/sqrt(pi)*tw, no coda. */
int dage_gbfmm_create(dage_ctx ctx, dage_grid3d grid, int maxlevel, int lmax,
                      const double *tp, const double *tw, int ntp,
                      dage_gbfmm *gbfmm);
/* GBFMMCoulomb3D_calculate_potential, cube part (gbfmm_coulomb3d.F90:650-821):
 * pot = near field + local expansions, on the grid's points (device buffers,
 * asynchronous on the context stream).  precision selects the near-field
 * arithmetic (DAGE_PREC_*); multipoles, translations and local expansions
 * are always FP64.  rho and pot must not overlap (DAGE_ERR_ARG): the far
 * field reads rho on a second stream while pot is written. */
int dage_gbfmm_apply_device(dage_gbfmm gbfmm, const double *rho_dev,
                            double *pot_dev, int precision);
/* Host-buffer variant (uploads, applies, downloads, synchronises). */
int dage_gbfmm_apply(dage_gbfmm gbfmm, const double *rho, double *pot,
                     int precision);
/* Number of boxes over levels 2..maxlevel (the columns of the arrays below). */
/* GBFMMHelmholtz3D_set_energy (src/gbfmm/gbfmm_helmholtz3d.F90:478-515) and
 * the GBFMMHelmholtz3D far field (:730-964, gbfmm_helmholtz3d.cu:300-420):
 * energy < -0.005 turns the operator into the Helmholtz one, near field with
 * the t-weights times exp(2E/(4t^2)) and far field of the screened kernel
 * exp(-kappa r)/r, kappa = sqrt(-2E) (multipoles and local expansions in
 * i_l(kappa r) C_lm, translations from dage_fmm_yukawa_*_matrix); energy >=
 * -0.005 restores the 1/r operator.  Needs boxes of equal point counts and
 * spacing (DAGE_ERR_UNSUPPORTED otherwise).  Rebuilds the far-field
 * matrices (host sampling and a GPU projection GEMM) and synchronises the
 * context stream.  The potential of dage_gbfmm_apply is then
 * G rho = int exp(-kappa |r - r'|)/|r - r'| rho(r') dr' (without the
 * -1/(2 pi) of the Helmholtz update). */
int dage_gbfmm_set_energy(dage_gbfmm gbfmm, double energy);
int dage_gbfmm_box_count(dage_gbfmm gbfmm, int *nbox_total);
/* No reference counterpart (diagnostic): dage_operator_get_band of the
 * near-field operator. */
int dage_gbfmm_get_band(dage_gbfmm gbfmm, int *t0);
/* No reference counterpart (diagnostic): dage_operator_get_lowrank of the
 * near-field operator. */
int dage_gbfmm_get_lowrank(dage_gbfmm gbfmm, int *t1, int *kx_sum);
/* After an apply: multipole moments and local expansions, (lmax+1)^2 x
 * nbox_total column-major, boxes in the reference's order (level maxlevel
 * first, children of a box consecutive), as multipole_moments(:, :) and
 * farfield_potential(:, :) of gbfmm_coulomb3d.F90. */
int dage_gbfmm_get_multipoles(dage_gbfmm gbfmm, double *q);
int dage_gbfmm_get_local_expansions(dage_gbfmm gbfmm, double *L);
int dage_gbfmm_destroy(dage_gbfmm gbfmm);
/* InteractionMatrix%eval (src/bubbles/multipole.F90:268-312): T, (lmax+1)^2
 * square column-major, with L_target += T q_source for dR = source centre -
 * target centre.  Host computation. */
int dage_fmm_interaction_matrix(int lmax, const double dR[3], double *T);
/* TranslationMatrix%eval (multipole.F90:444-470) unpacked: W, (lmax+1)^2
 * square column-major, unit lower triangular, q(new) = W(old - new) q(old). */
int dage_fmm_translation_matrix(int lmax, const double dR[3], double *W);
/* Modified spherical Bessel functions i_l(x), k_l(x), l = 0..lmax (<= 16),
 * 0 < x <= 600, k_0(x) = (pi/2) exp(-x)/x (the reference's first and second
 * modified spherical Bessel functions, src/bubbles/modified_bessel.F90,
 * bessels.cu:183-300).  Host computation. */
int dage_bessel_modified_spherical(int lmax, double x, double *i_l, double *k_l);
/* Screened-Coulomb (Yukawa, exp(-kappa r)/r) counterparts of the two
 * matrices above, for the GBFMM-Helmholtz far field
 * (HelmholtzMultipoleConverter / HelmholtzMultipoleTranslator,
 * src/bubbles/helmholtz_multipole_tools.F90:338-426).  Multipoles
 * q_lm = sum rho w i_l(kappa r) C_lm, local expansions
 * phi = sum L_lm i_l(kappa r) C_lm (C_lm = S_lm / r^l).  Interaction: L_target
 * += T q_source for dR = source centre - target centre, exact on the target
 * sphere of radius rho (|dR| > 2 rho) up to the lmax truncation.
 * Translation: A with q(parent) = A q(child) and L(child) = A^T L(parent)
 * for d = child centre - parent centre, for points within rho of the child
 * centre.  Built by projection on a sphere; host computation. */
int dage_fmm_yukawa_interaction_matrix(int lmax, double kappa, const double dR[3], double rho,
                                       double *T);
int dage_fmm_yukawa_translation_matrix(int lmax, double kappa, const double d[3], double rho,
                                       double *A);

/* ---------------------------------------------------------- batched applies
 * Helmholtz applies of many orbitals at once: the per-orbital loop of
 * HelmholtzSCFCycle_update_orbitals_worker (src/bubbles/scf_cycle.F90:
 * 2927-3054) that sets each orbital's energy (Helmholtz3D_set_energy,
 * helmholtz3d.F90:411-444) and applies the operator once per orbital.
 * Right-hand side i has its own grid (in = out grid; cube sizes may differ)
 * and energy E_i; orbitals on identical grids share matrices.  The batch owns
 * one device input and output cube per right-hand side (dage_batch_input /
 * _output) and captures all of its work into one hipGraph, replayed by
 * dage_batch_apply with a single graph launch on the context stream.
 * precision: DAGE_PREC_*; flags: DAGE_OP_NO_CODA. */
int dage_helmholtz_batch_create(dage_ctx ctx, int nrhs, const dage_grid3d *grids,
                                const double *energies, const double *tp,
                                const double *tw, int ntp, double tlog,
                                int flags, int precision, dage_batch *batch);
int dage_batch_input(dage_batch batch, int i, double **cube_dev);
int dage_batch_output(dage_batch batch, int i, double **cube_dev);
/* Synchronous host copies into input cube i / out of output cube i. */
int dage_batch_upload(dage_batch batch, int i, const double *host);
int dage_batch_download(dage_batch batch, int i, double *host);
/* New energies (device weights are rewritten; the graph is kept). */
int dage_batch_set_energies(dage_batch batch, const double *energies);
/* out_i = G(E_i) in_i for every i: one hipGraphLaunch (asynchronous). */
int dage_batch_apply(dage_batch batch);
/* The same work launched kernel by kernel (for profiling). */
int dage_batch_apply_eager(dage_batch batch);
/* Autotuner: times every GEMM column-block count 1..8 and the operator's own
 * tiling (0) for each pass of each grid group (reps timed runs each), keeps
 * the fastest and re-captures the graph.  tiles (optional, 3 per group)
 * receives the choices.  Overwrites outputs. */
int dage_batch_tune(dage_batch batch, int reps, int *tiles);
int dage_batch_group_count(dage_batch batch, int *ngroups);
int dage_batch_destroy(dage_batch batch);

/* -------------------------------------------------------------- cube algebra
 * The cube part of the Function3D arithmetic of the Helmholtz SCF cycle
 * (src/bubbles/scf_cycle.F90:2927-3054: product_in_place_REAL64,
 * add_in_place, subtract_in_place; electron density :2333-2388; overlap
 * matrix :2399-2424; orbital rotations :936-968).  Device buffers of n
 * doubles; asynchronous on the context stream unless stated. */
/* y = alpha*x + beta*y */
int dage_vec_axpby(dage_ctx ctx, long n, double alpha, const double *x,
                   double beta, double *y);
/* out = x*y (elementwise; out may alias x or y) */
int dage_vec_mul(dage_ctx ctx, long n, const double *x, const double *y,
                 double *out);
/* rho += occ*x^2 */
int dage_vec_square_acc(dage_ctx ctx, long n, double occ, const double *x,
                        double *rho);
/* Slater exchange of the density: vx = -(3 rho/pi)^(1/3) and (ex may be NULL)
 * ex = -(3/4)(3/pi)^(1/3) rho^(4/3) (negative rho counts as 0). */
int dage_vec_lda_x(dage_ctx ctx, long n, const double *rho, double *vx,
                   double *ex);
/* out[j] = sum_i coef[i + nin*j] in[i], j < nout (no output may alias an
 * input); synchronises. */
int dage_vec_lincomb(dage_ctx ctx, long n, int nin, const double *const *in,
                     const double *coef, int nout, double *const *out);
/* df = d f / d(axis) on the grid's points (axis 0..2) from the cells' LIP
 * interpolants (mean of both cells at shared cell ends); equidistant cell
 * nodes (grid_type 1); asynchronous on the context stream. */
int dage_cube_gradient(dage_grid3d grid, int axis, const double *f,
                       double *df);
/* out = (K f)/w along axis, K the assembled stiffness matrix of the cells'
 * Lagrange interpolants (int L_a' L_b', exact) and w the LIP weights, so that
 * the Gram of g with out is sum_lines w_perp int g' f' along that axis (the
 * interpolants' exact kinetic integral); asynchronous on the context
 * stream. */
int dage_cube_stiffness(dage_grid3d grid, int axis, const double *f,
                        double *out);
/* S[i + nf*j] = sum_p wx wy wz f_i(p) g_j(p) with the grid's LIP weights
 * (overlap matrix); deterministic; synchronises and writes S on the host. */
int dage_cube_gram(dage_grid3d grid, int nf, const double *const *f, int ng,
                   const double *const *g, double *S);

/* ----------------------------------------- learned preconditioner layers
 * No reference counterpart (the reference has no learned orbital-update
 * preconditioner).  3x3x3 convolutions with zero padding: activations FP32
 * [c][z][y][x], weights BF16 (uint16 bit patterns) [cout][cin][27] with tap
 * (dx+1) + 3(dy+1) + 9(dz+1), FP32 accumulation; 1 <= cin, cout <= 16
 * (forward: cout 1 or 16).  Device buffers; asynchronous unless stated. */
int dage_conv3d_forward(dage_ctx ctx, int nx, int ny, int nz, int cin, int cout,
                        const float *in, const unsigned short *w_bf16,
                        const float *bias, int relu, float *out);
int dage_conv3d_backward_data(dage_ctx ctx, int nx, int ny, int nz, int cin,
                              int cout, const float *grad_out,
                              const unsigned short *w_bf16, float *grad_in);
/* Deterministic filter and bias gradients (fixed-order two-stage sums). */
int dage_conv3d_backward_filter(dage_ctx ctx, int nx, int ny, int nz, int cin,
                                int cout, const float *in, const float *grad_out,
                                float *grad_w, float *grad_b);
/* grad = 0 where out <= 0 */
int dage_relu_backward(dage_ctx ctx, long n, const float *out, float *grad);
/* Adam on FP32 master weights (bias-corrected with `step` >= 1), BF16 copy
 * refreshed. */
int dage_adam_step(dage_ctx ctx, int n, float *w, const float *grad, float *m,
                   float *v, float lr, float beta1, float beta2, float eps,
                   int step, unsigned short *w_bf16);
int dage_weights_to_bf16(dage_ctx ctx, int n, const float *w,
                         unsigned short *w_bf16);
/* grad = 2 (pred - target)/n, *loss = mean (pred - target)^2; synchronises. */
int dage_mse_loss(dage_ctx ctx, long n, const float *pred, const float *target,
                  float *grad, double *loss);
/* y = (float)(scale x) ; y += a x */
int dage_cvt_f64_f32(dage_ctx ctx, long n, const double *x, double scale,
                     float *y);
int dage_axpy_f32_f64(dage_ctx ctx, long n, double a, const float *x,
                      double *y);
/* Raw device memory for callers without their own allocator (the Python
 * mirror): allocate, free, and synchronous copies (kind 0 host->device,
 * 1 device->host, 2 device->device). */
int dage_device_alloc(dage_ctx ctx, size_t bytes, void **ptr);
int dage_device_free(dage_ctx ctx, void *ptr);
int dage_memcpy(dage_ctx ctx, void *dst, const void *src, size_t bytes,
                int kind);

#ifdef __cplusplus
}
#endif
#endif /* DAGE_OPS_H */
