// Internal declarations of the gfx950 operator-application library.
// Not part of the C-ABI: the public surface is include/dage_ops.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/dage_ops.h"

namespace dage {

// ------------------------------------------------------------------ errors
void set_error(const char *fmt, ...);

#define DAGE_HIP_CHECK(expr)                                                  \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) {                                               \
            ::dage::set_error("%s failed: %s (%s:%d)", #expr,                 \
                              hipGetErrorString(e_), __FILE__, __LINE__);     \
            return DAGE_ERR_HIP;                                              \
        }                                                                     \
    } while (0)

#define DAGE_REQUIRE(cond, code, ...)                                         \
    do {                                                                      \
        if (!(cond)) {                                                        \
            ::dage::set_error(__VA_ARGS__);                                   \
            return (code);                                                    \
        }                                                                     \
    } while (0)

#define DAGE_TRY(expr)                                                        \
    do {                                                                      \
        int rc_ = (expr);                                                     \
        if (rc_ != DAGE_OK) return rc_;                                       \
    } while (0)

// ------------------------------------------------------------ host helpers
// Restatements used by the library itself (quadrature, LIP basis, grids).
// Their reference routines are cited where they are defined (host_math.cpp).
void gauleg(int ngp, double *x, double *w);
void gauss_quad(int nlin, int nlog, double tstart, double tlin, double tlog,
                double *tp, double *tw);
int gauss_lobatto_grid(int n, double left, double right, double *tp);
void lip_first_last(int nlip, double *first, double *last);
int lip_coeffs(int nlip, int grid_type, double *c);  // c[j + nlip*k]
int lip_integrals(int nlip, int grid_type, double *v);
int lip_inout_integrals(int nlip, double *ii, double *io);  // (nlip, nlip-1) each
int lip_inout_integrals_exp(int nlip, double mu, double *ii, double *io);
double two_over_sqrtpi();
double pi();

constexpr int kMaxNlip = 8;
// Banded y/z passes (band.hip): largest half-width |o - i| of a nonzero
// matrix entry they handle, and the row length of their coefficient tables.
constexpr int kBandH = 8;
constexpr int kBandW = 2 * kBandH + 1;

struct Axis {
    double qmin = 0.0;
    int ncell = 0;
    int nlip = 0;
    int grid_type = 1;
    std::vector<double> steps;   // ncell
    std::vector<double> points;  // ncell*(nlip-1)+1
    std::vector<double> celld;   // ncell
    int n() const { return ncell * (nlip - 1) + 1; }
};
int make_axis(double qmin, int ncell, int nlip, const double *steps,
              int grid_type, Axis *ax);

// --------------------------------------------------------------- profiling
struct ProfileEntry {
    double total_ms = 0.0;
    int launches = 0;
};

}  // namespace dage

struct dage_ctx_s {
    // Grids and operators hold a reference; the context is freed when the
    // last holder (including the creator's dage_ctx_destroy) lets go.
    int refs = 1;
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    bool profiling = false;
    // pending (start, stop, name) event triples, resolved on query
    struct Pending {
        hipEvent_t start, stop;
        std::string name;
    };
    std::vector<Pending> pending;
    std::map<std::string, dage::ProfileEntry> profile;
    hipEvent_t timer_start = nullptr, timer_stop = nullptr;
    // grow-only device scratch for per-call temporaries (dage::ctx_scratch)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    void *aux = nullptr;      // dage::ctx_aux
    size_t aux_bytes = 0;
    hipStream_t scratch_stream = nullptr;
    // this device's copy of the allocation-free apply's fixed scratch array
    // (abi.hip), looked up at context creation
    void *noalloc_scratch = nullptr;
};

struct dage_cube_s {
    dage_ctx ctx = nullptr;
    int shape[3] = {0, 0, 0};
    double *d = nullptr;
    size_t count() const { return (size_t)shape[0] * shape[1] * shape[2]; }
};

struct dage_grid3d_s {
    dage_ctx ctx = nullptr;
    dage::Axis axis[3];
    double *d_weights[3] = {nullptr, nullptr, nullptr};  // LIP weights (device)
    double *d_points[3] = {nullptr, nullptr, nullptr};   // grid points (device)
    // per-axis cell tables built on first use: LIP derivative matrix at the
    // nodes and exact stiffness matrix (nlip x nlip each)
    double *d_deriv[3] = {nullptr, nullptr, nullptr};
    double *d_stiff[3] = {nullptr, nullptr, nullptr};
    int shape(int a) const { return axis[a].n(); }
};

namespace dage {
constexpr int kMaxBubbleLmax = 12;

// Device view of one bubble (read by the inject kernel).
struct BubbleDev {
    double cx, cy, cz;
    double qmax;           // last radial grid point
    const double *celld;   // ncell cell starts
    const double *steps;   // ncell cell steps
    const double *f;       // npts x (lmax+1)^2, r^k folded in, point fastest
    int ncell, npts;
    // cell-lookup table: ctab[k] = cell (1-based) holding r = k*qmax/kCellTab
    const int *ctab;
    double tscale;         // kCellTab / qmax
};
constexpr int kCellTab = 1024;

struct BubbleHost {
    double center[3] = {0.0, 0.0, 0.0};
    double charge = 0.0;
    int ncell = 0, npts = 0;
    double qmax = 0.0;
    // celld | steps | r^k f | f (the last block only when k != 0; with k = 0
    // the raw values are the r^k-folded ones)
    double *d_buf = nullptr;
    int *d_tab = nullptr;     // kCellTab cell-lookup entries
    std::vector<double> h_celld, h_steps;  // host copies of the radial grid
    bool set = false;
    double *celld() const { return d_buf; }
    double *steps() const { return d_buf + ncell; }
    double *fk() const { return d_buf + 2 * ncell; }
};
}  // namespace dage

struct dage_bubbles_s {
    dage_ctx ctx = nullptr;
    int nbub = 0, lmax = 0, k = 0, nlip = 7;
    std::vector<dage::BubbleHost> host;
    dage::BubbleDev *d_desc = nullptr;
    bool dirty = true;
    int nlm() const { return (lmax + 1) * (lmax + 1); }
    // raw radial values f (npts x nlm, point fastest) of bubble i
    double *f(int i) const {
        const auto &h = host[i];
        return h.fk() + (k != 0 ? (size_t)h.npts * nlm() : 0);
    }
    size_t buf_doubles(int ncell, int npts) const {
        return 2 * (size_t)ncell + (size_t)npts * nlm() * (k != 0 ? 2 : 1);
    }
};

// Layout of the split-BF16 matrix planes (apply_mixed.hip): per mode, column
// count padded to the block width and K padded to a multiple of 32.
struct MixedLayout {
    int nbx = 0, nby = 0, nbz = 0;
    int npx = 0, npy = 0, npz = 0;
    int kpx = 0, kpy = 0, kpz = 0;
    size_t off_x = 0, off_y = 0, off_z = 0;  // element offsets into the plane buffer
};

// Factors of the low-rank t-points (lowrank.hip), element offsets into one
// device buffer `base`: RxAll (Kx x nxi) and QxAll (nxo x Kx) with the x
// factors of t in rows / columns koff[t].., and per t R_y^T, R_z^T
// (n_in x k) and Q_y^T, Q_z^T (k x n_out), column-major.
struct LowRank {
    int t1 = 0;    // t-points 0..t1-1
    int mode = 1;  // 1: where the apply is large enough, 2: always
    int Kx = 0;
    std::vector<int> k[3], koff;
    size_t off_rx = 0, off_qx = 0;
    std::vector<size_t> off_rt[2], off_qt[2];  // [0]: y, [1]: z
    const double *base = nullptr;
};

struct dage_operator_s {
    dage_ctx ctx = nullptr;
    dage_grid3d gin = nullptr, gout = nullptr;  // set during creation only
    int ntp = 0;
    // Helmholtz3D once dage_operator_set_energy has been called (its bubble
    // transform then convolves with exp(-kappa r)/r, kappa = sqrt(-2 energy))
    bool helmholtz = false;
    double energy = 0.0;
    std::vector<double> tp, tw, w;  // host copies
    double coda = 0.0;
    int flags = 0;
    int din[3] = {0, 0, 0}, dout[3] = {0, 0, 0};
    double *d_fx = nullptr, *d_fy = nullptr, *d_fz = nullptr;  // reference layouts
    double *d_gz = nullptr;  // Fz relaid as ((zi + nzi*t), zo)
    double *d_w = nullptr;   // weights (device)
    double *d_v = nullptr;   // x-mode intermediate
    double *d_u = nullptr;   // y-mode intermediate, ((xo,yo), (zi,t))
    double *d_in = nullptr, *d_out = nullptr;  // staging for host applies
    size_t ws_bytes = 0;
    uint64_t *d_mask = nullptr;  // B-row masks of the three passes (build_masks)
    int nb_pick[3] = {0, 0, 0};  // column-block counts chosen from the masks
    bool mask_used[3] = {false, false, false};  // masks cut the pass's work enough
    // FP64 passes may instead tile the output columns at a pitch that follows
    // the matrices' zero blocks (pitch[p] > 0, column blocks nbp[p]); d_tmask
    // then holds per row one bit per such tile, same row order as d_mask
    uint64_t *d_tmask = nullptr;
    int pitch[3] = {0, 0, 0}, nbp[3] = {0, 0, 0};
    // the same choice for the split-BF16 passes, which run every t (the FP64
    // ones are chosen for the t-points between the low-rank head and the
    // banded tail); mix_pitch[p] is 0 or pitch[p]
    int mix_nb_pick[3] = {0, 0, 0}, mix_pitch[3] = {0, 0, 0}, mix_nbp[3] = {0, 0, 0};
    bool mix_mask_used[3] = {false, false, false};
    // Nonzero bands for the fused x/y pass: d_xband[t*nxt + j] = [first,
    // last+1) k-step (4 input x points) of Fx_t rows reaching output x tile j
    // (16 points); d_yband[t*nyc + c] = [first, last+1) 16-point output y tile
    // reached by Fy_t rows of input y chunk c (16 points)
    int2 *d_xband = nullptr, *d_yband = nullptr;
    int nxt = 0, nyc = 0;
    void *d_planes = nullptr;   // split-BF16 matrix planes (built on first mixed apply)
    bool planes_valid = false;  // cleared when the matrices change
    MixedLayout mixed;
    // Banded passes for t >= band_t0 (band_t0 == ntp: none): band_h[t] =
    // largest |o - i| of a nonzero entry of Fx_t, Fy_t or Fz_t; d_band holds
    // the coefficient tables of Fy, Fz, then Fx for t = band_t0..ntp-1
    // (band.hip)
    double *d_lr = nullptr;  // low-rank factors (lowrank.hip), described by lr
    LowRank lr;
    double *d_band = nullptr;
    size_t band_off[2] = {0, 0};  // element offsets of the Fz and Fx tables
    int band_t0 = 0;
    int band_mode = 1;  // 1: where the launch fills the GPU, 2: always
    std::vector<int> band_h;
};

namespace dage {

// Context reference counting (grids and operators keep their context alive).
void ctx_retain(dage_ctx ctx);
void ctx_release(dage_ctx ctx);
// Device scratch of at least `bytes` for temporaries of one library call.
// Uses are ordered by the context stream (copies into it must be enqueued
// on ctx->stream); growing it, or a change of the context stream since the
// last use, first waits for the previous users.
int ctx_scratch(dage_ctx ctx, size_t bytes, void **p);

// Event-bracketed launch helpers (record only when profiling is on).
void prof_begin(dage_ctx ctx, const char *name, hipEvent_t *start);
void prof_end(dage_ctx ctx, const char *name, hipEvent_t start);
int prof_resolve(dage_ctx ctx);

// ----------------------------------------------------------------- kernels
// Operator matrix generation (matrices.hip).
int generate_matrices(dage_operator op);
// FP64 apply (apply_f64.hip).
int apply_f64(dage_operator op, const double *in_dev, double *out_dev);
// The same apply with its intermediates in a caller-given device buffer of
// ws_bytes: the t-chunk and K split shrink to fit, nothing is allocated, and
// DAGE_ERR_ARG is returned (nothing launched) when one t-point does not fit.
int apply_f64_fixed(dage_operator op, const double *in_dev, double *out_dev, void *ws,
                    size_t ws_bytes);
// One operator's matrices in the library's device layouts (Fx (nxo,nxi,T),
// Fy (nyi,nyo,T), Gz ((zi + nzi*t), zo)) and weights.
// mask (optional): per B row of each pass, bit c set when the row has a
// nonzero in output columns [16c, 16c + 16): x rows (t, xi) at t*nxi + xi,
// then y rows (t, yi), then z rows zi + nzi*t; null for a pass whose output
// axis exceeds 1024 points.
struct ApplyMats {
    const double *fx = nullptr, *fy = nullptr, *gz = nullptr, *w = nullptr;
    int din[3] = {0, 0, 0}, dout[3] = {0, 0, 0};
    int ntp = 0;
    const uint64_t *mask[3] = {nullptr, nullptr, nullptr};
    int nb_pick[3] = {0, 0, 0};
    // pitch tiling (pitch[p] > 0): tile masks, same row order, bit j = tile j
    const uint64_t *tmask[3] = {nullptr, nullptr, nullptr};
    int pitch[3] = {0, 0, 0}, nbp[3] = {0, 0, 0};
    // nonzero bands of Fx and Fy (dage_operator_s::d_xband/d_yband; null: none)
    const int2 *xband = nullptr, *yband = nullptr;
    int nxt = 0, nyc = 0;
    // banded y/z passes for t >= band_t0 (see dage_operator_s)
    const double *band_y = nullptr, *band_z = nullptr, *band_x = nullptr;
    int band_t0 = 0, band_mode = 1;
    const int *band_h = nullptr;
    // low-rank t-points 0..lr->t1-1 (null: none)
    const LowRank *lr = nullptr;
};
// Device scratch owned by the caller, grown on demand.
struct Workspace {
    double *p = nullptr;
    size_t bytes = 0;
};
ApplyMats operator_mats(dage_operator op);
// A batch of right-hand sides on one operator's grids: `in` and `out` hold B
// cubes back to back; w (B x T, row b at w + b*T) replaces the operator's
// weights when given; coda != 0 adds coda*in (in and out on the same grid);
// nb[pass] overrides the GEMM column-block count (0: heuristic); no_alloc
// makes a too-small workspace an error instead of growing it.
struct BatchSpec {
    int B = 1;
    const double *w = nullptr;
    double coda = 0.0;
    int nb[3] = {0, 0, 0};
    bool no_alloc = false;
    // > 0: the workspace (V, U and the z-pass partials) must fit this many
    // bytes; t-chunk and K split shrink to fit (DAGE_ERR_ARG when even one
    // t-point without split does not)
    size_t ws_cap = 0;
};
int apply_f64_batch(dage_ctx ctx, const ApplyMats &m, const double *in, double *out,
                    const BatchSpec &bs, Workspace *ws);
int apply_f64_batch_workspace(const ApplyMats &m, int B, size_t *bytes);
int apply_f64_workspace(dage_operator op, size_t *bytes);
// Allows the fused x/y pass its dynamic LDS on the current device (called
// once per context, outside any stream capture).
int prepare_fused_kernels();
// Mixed-precision apply (apply_mixed.hip): npair = 6 (bf16x6) or 3 (bf16x3).
int apply_mixed(dage_operator op, const double *in_dev, double *out_dev, int npair);
// Batch of right-hand sides on one operator (same conventions as apply_f64_batch).
int apply_mixed_batch(dage_operator op, const double *in, double *out, int npair,
                      const BatchSpec &bs, Workspace *ws);
int apply_mixed_batch_workspace(dage_operator op, int B, size_t *bytes);
// Builds the operator's split-BF16 matrix planes now (they are otherwise built
// on the first mixed apply).
int prepare_mixed(dage_operator op);
// FMM matrices (fmm_host.cpp).  Sparse term lists per matrix entry (row +
// N*col): interaction entries combine I_{L,n}(D), translation entries S_{L,n}(d).
struct FmmTables {
    int lmax = 0;
    std::vector<int> t_start, t_idx;
    std::vector<double> t_val;
    std::vector<int> w_start, w_idx;
    std::vector<double> w_val;
};
const FmmTables &fmm_tables(int lmax);
// T (N x N, column-major): L_target += T q_source, D = source - target centre.
void fmm_interaction(const FmmTables &tab, const double D[3], double *T);
// W (N x N, column-major, unit lower triangular): q(new) = W(old - new) q(old).
void fmm_translation(const FmmTables &tab, const double d[3], double *W);
void solid_harmonics_host(int lmax, double x, double y, double z, double *S);

// Yukawa (exp(-kappa r)/r) far-field math (yukawa_host.cpp).  Matrices are
// column-major N x N, N = (lmax+1)^2, lm = l*l + l + m.
constexpr int kMaxFmmBasisL = 24;
void bessel_i_scaled(int lmax, double x, double *g);  // i_l(x)/x^l, 0 <= x <= 600
void bessel_k(int lmax, double x, double *k);         // k_l(x), k_0 = (pi/2) e^-x / x
void yukawa_regular(int lmax, double kappa, const double v[3], double *R);
void yukawa_singular(int lmax, double kappa, const double v[3], double *K);
// Product rule on the unit sphere: Q nodes, weights, C_lm at the nodes [lm + N q].
struct YukawaSphere {
    int lmax = 0, Q = 0;
    std::vector<double> nodes, w, C;
};
YukawaSphere yukawa_sphere(int lmax);
// P[lm + N q]: regular coefficients of a function sampled at rho W_q
void yukawa_projector(const YukawaSphere &s, double kappa, double rho, double *P);
// Kq[q + Q lm]: (2 kappa/pi)(2l+1) K_lm(rho W_q - D), the far field at the
// target sphere of unit source multipoles at offset D (source - target)
void yukawa_m2l_samples(const YukawaSphere &s, double kappa, const double D[3], double rho,
                        double *Kq);
// M2L matrix T = P Kq (target local coefficients from source multipoles)
void yukawa_m2l(const YukawaSphere &s, double kappa, const double D[3], double rho, double *T);
// A(lm, l'm'): R_lm(z + d) = sum A R_l'm'(z) for |z| <= rho (M2M: q_parent = A
// q_child, L2L: L_child = A^T L_parent, d = child centre - parent centre)
void yukawa_shift(const YukawaSphere &s, double kappa, const double d[3], double rho, double *A);
// Cube integral (reduce.hip).
int cube_integrate(dage_grid3d g, const double *f, const double *h,
                   double *result);
// Relayout Fz (nzi, nzo, T) -> Gz ((zi + nzi*t), zo).
int relayout_fz(dage_operator op);
// C = A B with the FP64 MFMA GEMM, column-major: A(m,k) = A[m + k*lda],
// B(k,n) = B[k + n*ldb], C(m,n) = C[m + n*M].  splits > 1 splits K into
// that many partial products (part: splits*M*N doubles) summed in order.
int gemm_colmajor(dage_ctx ctx, const char *name, int M, int N, int K, const double *A, long lda,
                  const double *B, long ldb, double *C, int splits, double *part);
// batch independent products C_z = A_z B_z (z < batch), same layouts, with
// A_z = A + z*strideA etc.
int gemm_colmajor_batched(dage_ctx ctx, const char *name, int M, int N, int K, const double *A,
                          long lda, long strideA, const double *B, long ldb, long strideB,
                          double *C, long strideC, int batch);
// Row masks of the operator's matrices and the column-block counts they make
// cheapest; run after every change of the matrices (synchronises).
int build_masks(dage_operator op);
int flush_matrices(dage_operator op);
// Banded passes (band.hip): half-widths and coefficient tables from the row
// spans build_masks computes; the x pass V = Fx_t . in, the y pass
// U = w * (V . Fy_t) and the z pass
// C (+)= sum_t U_t . Fz_t over t = t0..t0+tc-1 (all >= band_t0), for B
// right-hand sides, in the layouts of apply_f64.hip's passes 2 and 3.
int build_bands(dage_operator op, const std::vector<int> &span);
// Whether an apply of B right-hand sides takes the banded passes.
bool band_pays(const ApplyMats &m, int B);
// Column-major FP64 MFMA GEMM, batched: for batch index z (z_lo = z % Z0,
// z_hi = z / Z0 when Z0 > 0, else z_lo = z, z_hi = 0)
//   C_z (M x N, column stride sCn) (+)= alpha * A_z (M x K, column stride sAk)
//                                              . B_z (K x N, column stride sBn)
// with X_z = X + z_lo*bX + z_hi*bX1 and alpha = alpha[z_lo + aS1*z_hi] (null: 1).
struct ColGemm {
    const double *A = nullptr, *B = nullptr;
    double *C = nullptr;
    int M = 0, N = 0, K = 0;
    long sAk = 0, sBn = 0, sCn = 0;
    long bA = 0, bB = 0, bC = 0;
    int Z0 = 0;
    long bA1 = 0, bB1 = 0, bC1 = 0;
    const double *alpha = nullptr;
    long aS1 = 0;
    int accumulate = 0;
};
int gemm_colmajor_general(dage_ctx ctx, const char *name, const ColGemm &c, int batch);
// Low-rank t-points (lowrank.hip): factors from the operator's matrices
// (after build_masks), workspace (doubles: first region, second region) and
// the apply of t = 0..lr->t1-1 for B right-hand sides into C_b = C + b*cS
// (overwritten).
int build_lowrank(dage_operator op);
bool lowrank_pays(const ApplyMats &m, int B);
void lowrank_workspace(const ApplyMats &m, int B, size_t *first, size_t *second);
int lowrank_apply(dage_ctx ctx, const ApplyMats &m, const double *in, double *C, long cS, int B,
                  const double *W, long wS, double *ws1, double *ws2, const char *name);
int band_pass_x(dage_ctx ctx, const ApplyMats &m, const double *in, double *V, int t0, int tc,
                int B, const char *name);
int band_pass_y(dage_ctx ctx, const ApplyMats &m, const double *V, double *U, const double *w,
                long wS, int t0, int tc, int B, const char *name);
// The banded x and y passes in one kernel (no V in memory; the same sums in
// the same order as band_pass_x then band_pass_y); off with DAGE_APPLY_BANDXY=0.
int band_pass_xy(dage_ctx ctx, const ApplyMats &m, const double *in, double *U, const double *w,
                 long wS, int t0, int tc, int B, const char *name);
bool band_xy_fused();
int band_pass_z(dage_ctx ctx, const ApplyMats &m, const double *U, double *C, long cS,
                bool accumulate, int t0, int tc, int B, const char *name);
// Bubbles (inject.hip, abi.hip).
int bubbles_sync_device(dage_bubbles bub);
// Bubble ibub's buffers on the radial grid ax (cell-lookup table built from
// it), radial values zeroed.
int bubble_alloc_host(dage_bubbles b, int ibub, const double center[3], double charge,
                      const Axis &ax);
// Bubble ibub on the radial grid of bubble sbub of src (device copies),
// radial values zeroed.
int bubble_alloc_like(dage_bubbles b, int ibub, dage_bubbles src, int sbub);
// r^k f -> the inject block after the raw values changed on the device
// (k != 0 only; asynchronous on the context stream).
int bubbles_fold_rk(dage_bubbles b);
// out_i(:, LM) += sum c(lm1, lm2 -> LM) f1_i(:, lm1) f2_i(:, lm2) for nb
// bubbles on npts_i radial points (raw values, l1max/l2max the factors'
// lmax; out holds at least (l1max+l2max+1)^2 columns); asynchronous.
int bubble_product_acc(dage_ctx ctx, int nb, const double *const *f1, int l1max,
                       const double *const *f2, int l2max, double *const *out, const int *npts);
// a second grow-only device buffer for temporaries that must outlive the
// calls that use ctx_scratch within one library call
int ctx_aux(dage_ctx ctx, size_t bytes, void **p);
int bubbles_inject(dage_bubbles bub, dage_grid3d grid, int lmin, double multiplier,
                   double *cube);
// Zero n doubles with a fill kernel on the context stream (graph-safe, unlike
// a memset node).
int fill_zero(dage_ctx ctx, double *p, size_t n);
// dst[i] = src[i] with a streaming copy kernel.
int copy_device(dage_ctx ctx, double *dst, const double *src, size_t n);

}  // namespace dage
