// Grid-based fast multipole method for the Coulomb potential of a cube
// (GBFMMCoulomb3D, src/gbfmm/gbfmm_coulomb3d.F90:306-1330, its CUDA side
// gbfmm_coulomb3d.cu and the box geometry of gbfmm_parallel_info.F90).
//
// The cube's cells are split into 2^maxlevel boxes per axis.  For every box:
//   near field  exact Coulomb operator (no coda) from the box and its nearest
//               neighbours to the box's points, i.e. the operator built on
//               the near-field sub-grid -> box sub-grid (Coulomb3D_get_suboperator,
//               coulomb.F90:517-635).  The near-field region of every point is
//               the tensor product of per-axis ranges, so the near fields of
//               all boxes together are ONE separable operator on the whole
//               cube whose 1D matrices are banded: row xo holds the sub-
//               operator row of the box owning xo, placed at that box's
//               near-field columns.  The near field is therefore one
//               whole-cube apply (FP64 MFMA, or split-BF16) with those
//               assembled matrices instead of one apply per box;
//   far field   box multipoles q_lm (lmax, Racah solid harmonics, box sub-grid
//               weights), translated up the tree (M2M), converted to local
//               expansions from each interaction list (M2L), translated down
//               (L2L) and evaluated on the box's points (LE).
// Every output point belongs to exactly one box: the last point of a box is
// left to the next box unless it is the last point of the cube
// (cut_end_borders, gbfmm_coulomb3d.F90:470-473).
//
// Boxes are numbered as in the reference: at level 1 index x + 2y + 4z of
// the 2x2x2 split, and each box's 8 children are numbered consecutively
// (parent*8 + octant, octant = ox + 2 oy + 4 oz).  Multipole and local-
// expansion arrays hold level maxlevel first, then maxlevel-1, ..., level 2.
//
// Device layout: q and L are (N, nbox_total) FP64 with N = (lmax+1)^2; the
// per-level M2L matrices are full N x N column-major blocks keyed by the
// integer box offset of the source; M2M/L2L use 8 N x N blocks per level.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "dage_internal.hpp"

struct dage_gbfmm_s {
    dage_ctx ctx = nullptr;
    int maxlevel = 0, start_level = 2, lmax = 0, N = 0;
    int nb = 0;  // boxes per axis at maxlevel
    int npt[3] = {0, 0, 0};
    int cells_per_box[3] = {0, 0, 0};
    int nlip = 0;
    double qmin[3] = {0, 0, 0}, qmax[3] = {0, 0, 0};
    std::vector<int> level_offset;  // by level
    int nbox_total = 0;
    // per axis and box coordinate: box and near-field point ranges
    std::vector<int> box_lo[3], box_n[3], box_cut[3], nf_lo[3], nf_n[3];
    dage_operator nf_op = nullptr;    // whole-cube operator with the banded near-field matrices
    // the far field up to the local expansions runs on its own stream,
    // beside the near field (both only read the density)
    hipStream_t s_far = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::vector<int> boxvec;          // 3 per maxlevel box, by box index
    // far field
    double *d_q = nullptr, *d_L = nullptr, *d_part = nullptr;
    int zsplit = 1;
    int *d_boxvec = nullptr;
    int *d_geo = nullptr;             // [axis][box_lo, box_n, box_cut][nb]
    double *d_wbox = nullptr;         // [axis][nb][maxn] box sub-grid weights
    int wbox_ld = 0;
    double *d_points = nullptr;       // [axis][maxpt] global points
    int points_ld = 0;
    double *d_centers = nullptr;      // 3 per maxlevel box
    double *d_coef = nullptr;         // a[lm], b[lm], f[l] of the solid-harmonic recursion
    std::vector<double *> d_T;        // per level: M2L matrices
    std::vector<int *> d_list;        // per level: CSR start of each target's pairs (nbox+1)
    std::vector<int> list_len;
    std::vector<double *> d_W;        // per child level: 8 translation matrices
    // M2L as one batched GEMM per level: source columns grouped by box
    // offset (matrix), slots padded to the largest group
    std::vector<int> m2l_nmat, m2l_smax;
    std::vector<int *> d_m2l_gsrc;    // per level: [matrix][slot] source column (-1: none)
    std::vector<int *> d_m2l_pos;     // per level: per pair (CSR order) its [matrix][slot]
    double *d_m2l_G = nullptr, *d_m2l_out = nullptr;
    // Uniform boxes (equal point counts, relative positions and sub-grid
    // weights): multipoles and local expansions as FP64 MFMA GEMMs with one
    // shared matrix of solid harmonics over a box's points.
    bool uniform = false;
    int pb[3] = {0, 0, 0};            // points per box and axis
    long P = 0;                       // points per box
    double *d_S = nullptr;            // [lm][p]  S_lm(point p - centre)
    double *d_Sw = nullptr;           // [p][lm]  S_lm * wx wy wz
    double *d_R = nullptr;            // [box][p] density at the box's points
    double *d_Pbox = nullptr;         // [box][p] far-field potential at the box's points
    double *d_gpart = nullptr;        // split-K partials of the multipole GEMM
    int msplits = 1;
    // Far-field kernel: 1/r, or exp(-kappa r)/r (kappa = sqrt(-2 E)) after
    // dage_gbfmm_set_energy with E < -0.005 (uniform boxes only).  The
    // inputs below are kept from construction to rebuild the far-field
    // matrices and the box basis when the kernel changes.
    bool yukawa = false;
    double kappa = 0.0;
    std::vector<double> box_rel, box_w;  // [axis][box_ld] box-relative coordinates, weights
    int box_ld = 0;
    std::vector<std::vector<int>> m2l_keys;  // per level and matrix: box-offset key
};

namespace dage {
namespace {

constexpr int kMaxFmmLmax = 22;  // the reference's GBFMM-Helmholtz default
constexpr int kMaxFmmN = (kMaxFmmLmax + 1) * (kMaxFmmLmax + 1);
constexpr int kMaxBoxKernelLmax = 15;  // box_multipoles_kernel<LMAX> instances

int box_index(int vx, int vy, int vz, int level) {
    int idx = 0;
    for (int l = level - 1; l >= 0; --l)
        idx = idx * 8 + ((vx >> l) & 1) + 2 * ((vy >> l) & 1) + 4 * ((vz >> l) & 1);
    return idx;
}

void box_vector(int idx, int level, int *v) {
    v[0] = v[1] = v[2] = 0;
    for (int l = 0; l < level; ++l) {
        const int o = (idx >> (3 * l)) & 7;
        v[0] |= (o & 1) << l;
        v[1] |= ((o >> 1) & 1) << l;
        v[2] |= ((o >> 2) & 1) << l;
    }
}

// Racah solid harmonics S_lm(x,y,z), l <= lmax, produced column by column
// (fixed m, increasing l) so that only two previous values are live; each
// value is handed to f(lm, value).  Same recursion and association as the
// full-array form (harmonics.F90:586-659): coef holds a[lm], b[lm], f[l].
template <typename F>
__device__ inline void solid_harmonics_by_column(int lmax, double x, double y, double z,
                                                 const double *coef, F &&f) {
    const int N = (lmax + 1) * (lmax + 1);
    const double *ca = coef, *cb = coef + N, *cf = coef + 2 * N;
    const double r2 = x * x + y * y + z * z;
    double dp = 1.0, dm = 0.0;  // S_{mm,mm}, S_{mm,-mm}
    for (int mm = 0; mm <= lmax; ++mm) {
        if (mm == 1) {
            dp = x;
            dm = y;
        } else if (mm > 1) {
            const double np = cf[mm] * (x * dp - y * dm);
            const double nm = cf[mm] * (y * dp + x * dm);
            dp = np;
            dm = nm;
        }
        for (int sg = 0; sg < (mm ? 2 : 1); ++sg) {
            const int m = sg ? -mm : mm;
            double p2 = 0.0, p1 = sg ? dm : dp;
            f(mm * mm + mm + m, p1);
            for (int l = mm + 1; l <= lmax; ++l) {
                const int lm = l * l + l + m;
                double v = ca[lm] * (z * p1);
                if (l - mm > 1) v = v - cb[lm] * (r2 * p2);
                f(lm, v);
                p2 = p1;
                p1 = v;
            }
        }
    }
}

struct MultArgs {
    const double *rho;
    long ldx, ldy;
    const int *boxvec, *geo;
    int nb;
    const double *wbox;
    int wbox_ld;
    const double *points;
    int points_ld;
    const double *centers;
    const double *coef;
    int zsplit, nbox;
    double *part;  // [zsplit][nbox][N]
};

// The (l,m) of one lmax split into groups of whole m-columns (|m| = mm, both
// signs) holding at most 64 values each, in the order the column recursion
// produces them: group[mm] and the group sizes.
template <int LMAX>
struct MpPlan {
    int group[LMAX + 1];
    int size[8];
    int ngroups;
    constexpr MpPlan() : group(), size(), ngroups(0) {
        int g = 0, fill = 0;
        for (int mm = 0; mm <= LMAX; ++mm) {
            const int c = (mm ? 2 : 1) * (LMAX + 1 - mm);
            if (fill + c > 64) {
                ++g;
                fill = 0;
            }
            group[mm] = g;
            fill += c;
            size[g] = fill;
        }
        ngroups = g + 1;
    }
};

// Accumulates val * S_lm(x,y,z) for the lm of group G into acc[0..size).
// Written as template recursion so every accumulator index is a compile-
// time constant (the accumulators stay in registers); the m-columns of
// other groups are skipped and the diagonal chain S_{mm,+-mm} that seeds
// each column is always advanced.  Same recursion and association as
// solid_harmonics_by_column.  `coef` is read through an offset that is 0 but
// changes per point batch, so the reads stay LDS reads inside the loop
// instead of being hoisted into registers.
template <int LMAX, int M, int L, int K>
__device__ __forceinline__ void mp_column(double z, double r2, double val, double p1, double p2,
                                          const double *coef, double (&acc)[64]) {
    if constexpr (L <= LMAX) {
        constexpr int N = (LMAX + 1) * (LMAX + 1);
        constexpr int AM = M < 0 ? -M : M;
        constexpr int lm = L * L + L + M;
        double v = coef[lm] * (z * p1);
        if constexpr (L - AM > 1) v = v - coef[N + lm] * (r2 * p2);
        acc[K] += v * val;
        mp_column<LMAX, M, L + 1, K + 1>(z, r2, val, v, p1, coef, acc);
    }
}

template <int LMAX, int G, int MM, int K>
__device__ __forceinline__ void mp_columns(double x, double y, double z, double r2, double val,
                                           double dp, double dm, const double *coef,
                                           double (&acc)[64]) {
    constexpr MpPlan<LMAX> plan{};
    if constexpr (MM <= LMAX) {
        constexpr int N = (LMAX + 1) * (LMAX + 1);
        if constexpr (plan.group[MM] <= G) {
            double ndp = dp, ndm = dm;
            if constexpr (MM == 1) {
                ndp = x;
                ndm = y;
            } else if constexpr (MM > 1) {
                const double f = coef[2 * N + MM];
                ndp = f * (x * dp - y * dm);
                ndm = f * (y * dp + x * dm);
            }
            if constexpr (plan.group[MM] == G) {
                acc[K] += ndp * val;
                mp_column<LMAX, MM, MM + 1, K + 1>(z, r2, val, ndp, 0.0, coef, acc);
                if constexpr (MM > 0) {
                    constexpr int K2 = K + (LMAX + 1 - MM);
                    acc[K2] += ndm * val;
                    mp_column<LMAX, -MM, MM + 1, K2 + 1>(z, r2, val, ndm, 0.0, coef, acc);
                }
                constexpr int C = (MM ? 2 : 1) * (LMAX + 1 - MM);
                mp_columns<LMAX, G, MM + 1, K + C>(x, y, z, r2, val, ndp, ndm, coef, acc);
            } else {
                mp_columns<LMAX, G, MM + 1, K>(x, y, z, r2, val, ndp, ndm, coef, acc);
            }
        }
    }
}

template <int LMAX, int G>
__device__ __forceinline__ void multipole_point(double x, double y, double z, double val,
                                                const double *coef, double (&acc)[64]) {
    const double r2 = x * x + y * y + z * z;
    mp_columns<LMAX, G, 0, 0>(x, y, z, r2, val, 1.0, 0.0, coef, acc);
}

template <int LMAX, int G>
__device__ void box_multipoles_group(const MultArgs &A, const double *coef, double (*red)[64]) {
    constexpr MpPlan<LMAX> plan{};
    constexpr int N = (LMAX + 1) * (LMAX + 1);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int box = blockIdx.x;
    const int bx = A.boxvec[3 * box], by = A.boxvec[3 * box + 1], bz = A.boxvec[3 * box + 2];
    const int nb = A.nb;
    const int x0 = A.geo[0 * 3 * nb + bx], nx = A.geo[0 * 3 * nb + nb + bx];
    const int y0 = A.geo[1 * 3 * nb + by], ny = A.geo[1 * 3 * nb + nb + by];
    const int z0 = A.geo[2 * 3 * nb + bz], nz = A.geo[2 * 3 * nb + nb + bz];
    const int zper = (nz + A.zsplit - 1) / A.zsplit;
    const int zb = blockIdx.y * zper, ze = min(nz, zb + zper);
    const double *wx = A.wbox + (long)(0 * nb + bx) * A.wbox_ld;
    const double *wy = A.wbox + (long)(1 * nb + by) * A.wbox_ld;
    const double *wz = A.wbox + (long)(2 * nb + bz) * A.wbox_ld;
    const double *px = A.points, *py = A.points + A.points_ld, *pz = A.points + 2 * A.points_ld;
    const double cx = A.centers[3 * box], cy = A.centers[3 * box + 1], cz = A.centers[3 * box + 2];
    double acc[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) acc[i] = 0.0;
    const long total = (long)nx * ny * max(0, ze - zb);
    for (long base = 0; base < total; base += 256) {
        const long e = base + tid;
        if (e >= total) break;
        const int i = (int)(e % nx);
        const long r = e / nx;
        const int j = (int)(r % ny);
        const int kz = zb + (int)(r / ny);
        const double val = A.rho[(x0 + i) + A.ldx * ((y0 + j) + A.ldy * (long)(z0 + kz))] *
                           wx[i] * wy[j] * wz[kz];
        multipole_point<LMAX, G>(px[x0 + i] - cx, py[y0 + j] - cy, pz[z0 + kz] - cz, val,
                                 coef + (int)(base >> 62), acc);
    }
    // wavefront sums in a fixed butterfly order, then the 4 wavefronts in order
#pragma unroll
    for (int i = 0; i < plan.size[G]; ++i) {
        double v = acc[i];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) red[wave][i] = v;
    }
    __syncthreads();
    if (tid < plan.size[G]) {
        // lm of the tid-th value of group G (the packing of MpPlan, at run time)
        int g = 0, fill = 0, lm = 0;
        for (int mm = 0, k = 0; mm <= LMAX; ++mm) {
            const int c = (mm ? 2 : 1) * (LMAX + 1 - mm);
            if (fill + c > 64) {
                ++g;
                fill = 0;
            }
            fill += c;
            if (g != G) continue;
            for (int sg = 0; sg < (mm ? 2 : 1); ++sg)
                for (int l = mm; l <= LMAX; ++l, ++k)
                    if (k == tid) lm = l * l + l + (sg ? -mm : mm);
        }
        const double s = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
        A.part[((long)blockIdx.y * A.nbox + box) * N + lm] = s;
    }
}

// q_lm(box) = sum over the box sub-grid of rho * wx wy wz * S_lm(r - centre)
// (GBFMMCoulomb3DMultipoleEvaluator, integrator.cu:1122-1196).  Block
// (box, z-slab, lm group): every thread keeps its group's <= 64 sums in
// registers over its points and the sums are combined once at the end.
template <int LMAX>
__global__ void __launch_bounds__(256) box_multipoles_kernel(MultArgs A) {
    constexpr MpPlan<LMAX> plan{};
    constexpr int N = (LMAX + 1) * (LMAX + 1);
    __shared__ double coef[2 * N + LMAX + 1];
    __shared__ double red[4][64];
    for (int i = threadIdx.x; i < 2 * N + LMAX + 1; i += 256) coef[i] = A.coef[i];
    __syncthreads();
    switch (blockIdx.z) {
    case 0: box_multipoles_group<LMAX, 0>(A, coef, red); break;
    case 1: if constexpr (plan.ngroups > 1) box_multipoles_group<LMAX, 1>(A, coef, red); break;
    case 2: if constexpr (plan.ngroups > 2) box_multipoles_group<LMAX, 2>(A, coef, red); break;
    case 3: if constexpr (plan.ngroups > 3) box_multipoles_group<LMAX, 3>(A, coef, red); break;
    case 4: if constexpr (plan.ngroups > 4) box_multipoles_group<LMAX, 4>(A, coef, red); break;
    default: break;
    }
}

template <int LMAX>
int launch_multipoles(const MultArgs &A, int nbox, hipStream_t s) {
    constexpr MpPlan<LMAX> plan{};
    static_assert(plan.ngroups <= 5, "multipole groups");
    hipLaunchKernelGGL(box_multipoles_kernel<LMAX>, dim3(nbox, A.zsplit, plan.ngroups),
                       dim3(256), 0, s, A);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int multipoles(int lmax, const MultArgs &A, int nbox, hipStream_t s) {
    switch (lmax) {
    case 0: return launch_multipoles<0>(A, nbox, s);
    case 1: return launch_multipoles<1>(A, nbox, s);
    case 2: return launch_multipoles<2>(A, nbox, s);
    case 3: return launch_multipoles<3>(A, nbox, s);
    case 4: return launch_multipoles<4>(A, nbox, s);
    case 5: return launch_multipoles<5>(A, nbox, s);
    case 6: return launch_multipoles<6>(A, nbox, s);
    case 7: return launch_multipoles<7>(A, nbox, s);
    case 8: return launch_multipoles<8>(A, nbox, s);
    case 9: return launch_multipoles<9>(A, nbox, s);
    case 10: return launch_multipoles<10>(A, nbox, s);
    case 11: return launch_multipoles<11>(A, nbox, s);
    case 12: return launch_multipoles<12>(A, nbox, s);
    case 13: return launch_multipoles<13>(A, nbox, s);
    case 14: return launch_multipoles<14>(A, nbox, s);
    case 15: return launch_multipoles<15>(A, nbox, s);
    default: set_error("lmax %d unsupported", lmax); return DAGE_ERR_UNSUPPORTED;
    }
}

__global__ void __launch_bounds__(256) sum_parts_kernel(const double *part, int zsplit, long len,
                                                        double *q) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < len;
         i += (long)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int k = 0; k < zsplit; ++k) s += part[(long)k * len + i];
        q[i] = s;
    }
}

// q(parent) = sum_children W_octant q(child)  (translate_multipoles,
// gbfmm_coulomb3d.F90:1502-1558); W unit lower triangular for 1/r (only its
// lower triangle is read), full for the Yukawa kernel.
__global__ void __launch_bounds__(256) m2m_kernel(const double *W, const double *qchild,
                                                  double *qparent, int N, int full) {
    const int P = blockIdx.x;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {  // one lm row per thread pass
        double acc = 0.0;
        const int send = full ? N : i + 1;
        for (int o = 0; o < 8; ++o) {
            const double *Wo = W + (long)o * N * N;
            const double *qc = qchild + (long)(8 * P + o) * N;
            for (int s = 0; s < send; ++s) acc += Wo[i + (long)N * s] * qc[s];
        }
        qparent[(long)P * N + i] = acc;
    }
}

// L(target) += sum_list T(offset) q(source)  (calculate_local_farfield_potential,
// gbfmm_coulomb3d.F90:1461-1500), as one batched GEMM over the box offsets:
// gather the sources' multipoles per offset, multiply, add back per target.
// G[col][lm] = q[gsrc[col]][lm] (0 for empty slots): the M2L GEMM operand
__global__ void __launch_bounds__(256) m2l_gather_kernel(const double *q, const int *gsrc,
                                                         long ncol, int N, double *G) {
    const long total = ncol * N;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const long col = e / N;
        const int src = gsrc[col];
        G[e] = src >= 0 ? q[(long)src * N + (e - col * N)] : 0.0;
    }
}

// L[t] += the target's GEMM columns, in interaction-list order (deterministic)
__global__ void __launch_bounds__(256) m2l_scatter_kernel(const double *out, const int *start,
                                                          const int *pos, int N, double *L) {
    const int t = blockIdx.x;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        double acc = 0.0;
        for (int e = start[t]; e < start[t + 1]; ++e) acc += out[(long)pos[e] * N + i];
        L[(long)t * N + i] += acc;
    }
}

// L(child) += W_octant^T L(parent)  (translate_potential, gbfmm_coulomb3d.F90:1562-1622).
__global__ void __launch_bounds__(256) l2l_kernel(const double *W, const double *Lparent,
                                                  double *Lchild, int N, int full) {
    const int c = blockIdx.x;
    const double *Wo = W + (long)(c & 7) * N * N;
    const double *Lp = Lparent + (long)(c >> 3) * N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        double acc = 0.0;
        for (int s = full ? 0 : i; s < N; ++s) acc += Wo[s + (long)N * i] * Lp[s];
        Lchild[(long)c * N + i] += acc;
    }
}

struct LeArgs {
    double *out;
    long ldx, ldy;
    const int *boxvec, *geo;
    int nb;
    const double *points;
    int points_ld;
    const double *centers;
    const double *coef;
    const double *L;  // maxlevel section
    int lmax, N;
};

// out(p) += sum_lm L_lm(box) S_lm(p - centre) over the box's own points
// (GBFMMCoulomb3D_evaluate_le_grid, gbfmm_coulomb3d.cu:233-280).  Block
// (box, z-plane of the box).
__global__ void __launch_bounds__(256) local_expansion_kernel(LeArgs A) {
    __shared__ double coef[2 * kMaxFmmN + kMaxFmmLmax + 1];
    __shared__ double Ls[kMaxFmmN];
    const int tid = threadIdx.x;
    const int box = blockIdx.x;
    const int nb = A.nb;
    const int bx = A.boxvec[3 * box], by = A.boxvec[3 * box + 1], bz = A.boxvec[3 * box + 2];
    const int nzc = A.geo[2 * 3 * nb + 2 * nb + bz];
    const int kz = blockIdx.y;
    if (kz >= nzc) return;  // uniform per block
    const int ncoef = 2 * A.N + A.lmax + 1;
    for (int i = tid; i < ncoef; i += 256) coef[i] = A.coef[i];
    for (int i = tid; i < A.N; i += 256) Ls[i] = A.L[(long)box * A.N + i];
    __syncthreads();
    const int x0 = A.geo[0 * 3 * nb + bx], nxc = A.geo[0 * 3 * nb + 2 * nb + bx];
    const int y0 = A.geo[1 * 3 * nb + by], nyc = A.geo[1 * 3 * nb + 2 * nb + by];
    const int z = A.geo[2 * 3 * nb + bz] + kz;
    const double cx = A.centers[3 * box], cy = A.centers[3 * box + 1], cz = A.centers[3 * box + 2];
    const double dz = A.points[2 * A.points_ld + z] - cz;
    for (int e = tid; e < nxc * nyc; e += 256) {
        const int i = e % nxc, j = e / nxc;
        const double dx = A.points[x0 + i] - cx;
        const double dy = A.points[A.points_ld + y0 + j] - cy;
        double res = 0.0;
        solid_harmonics_by_column(A.lmax, dx, dy, dz, coef,
                                  [&](int lm, double s) { res += s * Ls[lm]; });
        double *dst = A.out + (x0 + i) + A.ldx * ((y0 + j) + A.ldy * (long)z);
        *dst = *dst + res;
    }
}

int upload(void **dst, const void *src, size_t bytes) {
    DAGE_HIP_CHECK(hipMalloc(dst, bytes ? bytes : 8));
    if (bytes) DAGE_HIP_CHECK(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return DAGE_OK;
}

// S[lm*P + p] = S_lm(rel(p)), Sw[p*N + lm] = S_lm(rel(p)) w(p) for the points
// p = i + pb0 (j + pb1 k) of a box (relative positions and weights per axis).
__global__ void __launch_bounds__(256) box_harmonics_kernel(const double *rel, const double *w,
                                                            int pb0, int pb1, int pb2, int ld,
                                                            int lmax, const double *coef,
                                                            double *S, double *Sw) {
    const long P = (long)pb0 * pb1 * pb2;
    const int N = (lmax + 1) * (lmax + 1);
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < P;
         p += (long)gridDim.x * blockDim.x) {
        const int i = (int)(p % pb0), j = (int)((p / pb0) % pb1), k = (int)(p / ((long)pb0 * pb1));
        const double wp = w[i] * w[ld + j] * w[2 * ld + k];
        solid_harmonics_by_column(lmax, rel[i], rel[ld + j], rel[2 * ld + k], coef,
                                  [&](int lm, double v) {
                                      S[(long)lm * P + p] = v;
                                      Sw[p * N + lm] = v * wp;
                                  });
    }
}

// R[b*P + p] = rho at point p of box b (gather into the GEMM operand)
__global__ void __launch_bounds__(256) box_gather_kernel(const double *rho, long ldx, long ldy,
                                                         const int *boxvec, const int *geo, int nb,
                                                         int pb0, int pb1, int pb2, int nbox,
                                                         double *R) {
    const long P = (long)pb0 * pb1 * pb2, total = P * nbox;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int b = (int)(e / P);
        const long p = e - (long)b * P;
        const int i = (int)(p % pb0), j = (int)((p / pb0) % pb1), k = (int)(p / ((long)pb0 * pb1));
        const int x = geo[boxvec[3 * b]] + i;
        const int y = geo[3 * nb + boxvec[3 * b + 1]] + j;
        const int z = geo[6 * nb + boxvec[3 * b + 2]] + k;
        R[e] = rho[x + ldx * (y + ldy * (long)z)];
    }
}

// pot += Pbox over each box's own points (the cut rule: box_cut per axis)
__global__ void __launch_bounds__(256) box_scatter_kernel(const double *Pbox, long ldx, long ldy,
                                                          const int *boxvec, const int *geo,
                                                          int nb, int pb0, int pb1, int pb2,
                                                          int nbox, double *pot) {
    const long P = (long)pb0 * pb1 * pb2, total = P * nbox;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int b = (int)(e / P);
        const long p = e - (long)b * P;
        const int i = (int)(p % pb0), j = (int)((p / pb0) % pb1), k = (int)(p / ((long)pb0 * pb1));
        const int bx = boxvec[3 * b], by = boxvec[3 * b + 1], bz = boxvec[3 * b + 2];
        if (i >= geo[2 * nb + bx] || j >= geo[3 * nb + 2 * nb + by] ||
            k >= geo[6 * nb + 2 * nb + bz])
            continue;
        const int x = geo[bx] + i, y = geo[3 * nb + by] + j, z = geo[6 * nb + bz] + k;
        double *dst = pot + x + ldx * (y + ldy * (long)z);
        *dst = *dst + Pbox[e];
    }
}

void free_gbfmm(dage_gbfmm f) {
    if (f->nf_op) dage_operator_destroy(f->nf_op);
    void *bufs[] = {f->d_q,      f->d_L,      f->d_part,   f->d_boxvec,
                    f->d_geo,    f->d_wbox,   f->d_points, f->d_centers,  f->d_coef,
                    f->d_S,      f->d_Sw,     f->d_R,      f->d_Pbox,     f->d_gpart};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    for (double *p : f->d_T)
        if (p) (void)hipFree(p);
    for (int *p : f->d_list)
        if (p) (void)hipFree(p);
    for (double *p : f->d_W)
        if (p) (void)hipFree(p);
    for (int *p : f->d_m2l_gsrc)
        if (p) (void)hipFree(p);
    for (int *p : f->d_m2l_pos)
        if (p) (void)hipFree(p);
    if (f->d_m2l_G) (void)hipFree(f->d_m2l_G);
    if (f->d_m2l_out) (void)hipFree(f->d_m2l_out);
    if (f->ev_fork) (void)hipEventDestroy(f->ev_fork);
    if (f->ev_join) (void)hipEventDestroy(f->ev_join);
    if (f->s_far) (void)hipStreamDestroy(f->s_far);
    if (f->ctx) ctx_release(f->ctx);
    delete f;
}

// Size of a box at `level` along each axis.
void box_size(const dage_gbfmm f, int level, double size[3]) {
    for (int a = 0; a < 3; ++a) size[a] = (f->qmax[a] - f->qmin[a]) / (double)(1 << level);
}

// The shared basis of the uniform-box GEMMs: S[p + P lm] and Sw[lm + N p] =
// S w at the box points p = i + pb0 (j + pb1 k), with S_lm(point - centre)
// for 1/r (on the GPU) or the Yukawa regular basis i_l(kappa r) C_lm.
int build_box_basis(dage_gbfmm f) {
    const int N = f->N, lmax = f->lmax, ld = f->box_ld;
    const long P = f->P;
    if (!f->yukawa) {
        double *d_rel = nullptr, *d_w = nullptr;
        DAGE_TRY(upload((void **)&d_rel, f->box_rel.data(), 8 * f->box_rel.size()));
        int rc = upload((void **)&d_w, f->box_w.data(), 8 * f->box_w.size());
        if (rc == DAGE_OK) {
            const int grid_sz = (int)std::min<long>((P + 255) / 256, 65536);
            hipLaunchKernelGGL(box_harmonics_kernel, dim3(grid_sz), dim3(256), 0, f->ctx->stream,
                               d_rel, d_w, f->pb[0], f->pb[1], f->pb[2], ld, lmax, f->d_coef,
                               f->d_S, f->d_Sw);
            if (hipGetLastError() != hipSuccess ||
                hipStreamSynchronize(f->ctx->stream) != hipSuccess) {
                set_error("box harmonics failed");
                rc = DAGE_ERR_HIP;
            }
        }
        if (d_rel) (void)hipFree(d_rel);
        if (d_w) (void)hipFree(d_w);
        return rc;
    }
    std::vector<double> S((size_t)P * N), Sw((size_t)P * N), R(N);
    for (long p = 0; p < P; ++p) {
        const int i = (int)(p % f->pb[0]), j = (int)((p / f->pb[0]) % f->pb[1]),
                  k = (int)(p / ((long)f->pb[0] * f->pb[1]));
        const double v[3] = {f->box_rel[i], f->box_rel[ld + j], f->box_rel[2 * ld + k]};
        const double wp = f->box_w[i] * f->box_w[ld + j] * f->box_w[2 * ld + k];
        yukawa_regular(lmax, f->kappa, v, R.data());
        for (int lm = 0; lm < N; ++lm) {
            S[(size_t)p + (size_t)P * lm] = R[lm];
            Sw[(size_t)lm + (size_t)N * p] = R[lm] * wp;
        }
    }
    DAGE_HIP_CHECK(hipMemcpy(f->d_S, S.data(), 8 * S.size(), hipMemcpyHostToDevice));
    DAGE_HIP_CHECK(hipMemcpy(f->d_Sw, Sw.data(), 8 * Sw.size(), hipMemcpyHostToDevice));
    return DAGE_OK;
}

// M2L matrices per level and box offset (L_target += T q_source) and the
// M2M/L2L matrices per child level and octant, for the current kernel.
// 1/r: the closed forms of fmm_host.cpp on the host.  Yukawa: projection on
// the target sphere (yukawa_host.cpp), T = P K for all offsets of a level as
// one batched GEMM on the GPU with the shared projector P, and A = R P^T for
// the 8 octants likewise (R: the regular basis shifted by the octant vector).
int build_far_matrices(dage_gbfmm f) {
    const int L = f->maxlevel, N = f->N, lmax = f->lmax;
    dage_ctx ctx = f->ctx;
    hipStream_t s = ctx->stream;
    const FmmTables *tab = f->yukawa ? nullptr : &fmm_tables(lmax);
    YukawaSphere sph;
    if (f->yukawa) sph = yukawa_sphere(lmax);
    const int Q = sph.Q;
    constexpr int kChunk = 32;  // M2L matrices sampled per upload
    double *d_P = nullptr, *d_K = nullptr;
    auto cleanup = [&]() {
        if (d_P) (void)hipFree(d_P);
        if (d_K) (void)hipFree(d_K);
        d_P = d_K = nullptr;
    };
    if (f->yukawa) {
        if (hipMalloc((void **)&d_P, 8 * (size_t)N * Q) != hipSuccess ||
            hipMalloc((void **)&d_K, 8 * (size_t)Q * N * std::max(kChunk, 8)) != hipSuccess) {
            cleanup();
            set_error("Yukawa projection buffers: allocation failed");
            return DAGE_ERR_HIP;
        }
    }
    int rc = DAGE_OK;
    for (int l = f->start_level; l <= L && rc == DAGE_OK; ++l) {
        double size[3];
        box_size(f, l, size);
        const std::vector<int> &keys = f->m2l_keys[l];
        const int nmat = f->m2l_nmat[l];
        if (!f->d_T[l] && hipMalloc((void **)&f->d_T[l], 8 * (size_t)N * N * nmat) != hipSuccess) {
            set_error("M2L matrices: allocation failed");
            rc = DAGE_ERR_HIP;
            break;
        }
        if (l > f->start_level && !f->d_W[l] &&
            hipMalloc((void **)&f->d_W[l], 8 * (size_t)8 * N * N) != hipSuccess) {
            set_error("translation matrices: allocation failed");
            rc = DAGE_ERR_HIP;
            break;
        }
        auto offset = [&](int key, double D[3]) {
            const int ox = key % 17 - 8, oy = (key / 17) % 17 - 8, oz = key / 289 - 8;
            D[0] = ox * size[0];
            D[1] = oy * size[1];
            D[2] = oz * size[2];
        };
        auto octant = [&](int o, double d[3]) {  // child centre - parent centre
            d[0] = ((o & 1) - 0.5) * size[0];
            d[1] = (((o >> 1) & 1) - 0.5) * size[1];
            d[2] = (((o >> 2) & 1) - 0.5) * size[2];
        };
        if (!f->yukawa) {
            std::vector<double> Tall((size_t)N * N * nmat, 0.0);
            for (size_t id = 0; id < keys.size(); ++id) {
                double D[3];
                offset(keys[id], D);
                fmm_interaction(*tab, D, &Tall[id * N * N]);
            }
            if (hipMemcpy(f->d_T[l], Tall.data(), 8 * Tall.size(), hipMemcpyHostToDevice) !=
                hipSuccess) {
                set_error("M2L matrices: upload failed");
                rc = DAGE_ERR_HIP;
            }
            if (rc == DAGE_OK && l > f->start_level) {
                std::vector<double> Wall((size_t)8 * N * N);
                for (int o = 0; o < 8; ++o) {
                    double d[3];
                    octant(o, d);
                    fmm_translation(*tab, d, &Wall[(size_t)o * N * N]);
                }
                if (hipMemcpy(f->d_W[l], Wall.data(), 8 * Wall.size(), hipMemcpyHostToDevice) !=
                    hipSuccess) {
                    set_error("translation matrices: upload failed");
                    rc = DAGE_ERR_HIP;
                }
            }
            continue;
        }
        // Yukawa: target sphere of the level's box (its half diagonal)
        const double rho = 0.5 * std::sqrt(size[0] * size[0] + size[1] * size[1] +
                                           size[2] * size[2]);
        if (!(f->kappa * rho <= 600.0)) {
            set_error("kappa %g too large for boxes of half diagonal %g (kappa rho > 600)",
                      f->kappa, rho);
            rc = DAGE_ERR_ARG;
            break;
        }
        std::vector<double> P((size_t)N * Q), K((size_t)Q * N * kChunk);
        yukawa_projector(sph, f->kappa, rho, P.data());
        if (hipMemcpy(d_P, P.data(), 8 * P.size(), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("projector upload failed");
            rc = DAGE_ERR_HIP;
            break;
        }
        for (size_t id0 = 0; id0 < keys.size() && rc == DAGE_OK; id0 += kChunk) {
            const int cnt = (int)std::min<size_t>(kChunk, keys.size() - id0);
            for (int j = 0; j < cnt; ++j) {
                double D[3];
                offset(keys[id0 + j], D);
                yukawa_m2l_samples(sph, f->kappa, D, rho, &K[(size_t)j * Q * N]);
            }
            if (hipMemcpy(d_K, K.data(), 8 * (size_t)Q * N * cnt, hipMemcpyHostToDevice) !=
                hipSuccess) {
                set_error("M2L samples upload failed");
                rc = DAGE_ERR_HIP;
                break;
            }
            rc = gemm_colmajor_batched(ctx, "gemm_f64_yukawa_m2l", N, N, Q, d_P, N, 0, d_K, Q,
                                       (long)Q * N, f->d_T[l] + id0 * N * N, (long)N * N, cnt);
            if (rc == DAGE_OK && hipStreamSynchronize(s) != hipSuccess) {
                set_error("M2L projection failed");
                rc = DAGE_ERR_HIP;
            }
        }
        if (rc != DAGE_OK || l == f->start_level) continue;
        // A(lm, l'm') = sum_q R_lm(rho_c W_q + d) P_c(l'm', q) on the child sphere
        const double rc_ = 0.5 * rho;
        std::vector<double> Pc((size_t)N * Q), PcT((size_t)Q * N), Rq((size_t)8 * N * Q), R(N);
        yukawa_projector(sph, f->kappa, rc_, Pc.data());
        for (int q = 0; q < Q; ++q)
            for (int c = 0; c < N; ++c) PcT[(size_t)q + (size_t)Q * c] = Pc[(size_t)c + (size_t)N * q];
        for (int o = 0; o < 8; ++o) {
            double d[3];
            octant(o, d);
            for (int q = 0; q < Q; ++q) {
                const double v[3] = {rc_ * sph.nodes[3 * (size_t)q] + d[0],
                                     rc_ * sph.nodes[3 * (size_t)q + 1] + d[1],
                                     rc_ * sph.nodes[3 * (size_t)q + 2] + d[2]};
                yukawa_regular(lmax, f->kappa, v, R.data());
                std::copy(R.begin(), R.end(), Rq.begin() + ((size_t)o * Q + q) * N);
            }
        }
        if (hipMemcpy(d_P, PcT.data(), 8 * PcT.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d_K, Rq.data(), 8 * Rq.size(), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("translation samples upload failed");
            rc = DAGE_ERR_HIP;
            break;
        }
        rc = gemm_colmajor_batched(ctx, "gemm_f64_yukawa_shift", N, N, Q, d_K, N, (long)N * Q,
                                   d_P, Q, 0, f->d_W[l], (long)N * N, 8);
        if (rc == DAGE_OK && hipStreamSynchronize(s) != hipSuccess) {
            set_error("translation projection failed");
            rc = DAGE_ERR_HIP;
        }
    }
    cleanup();
    return rc;
}

int build_gbfmm(dage_gbfmm f, dage_grid3d grid, const double *tp, const double *tw, int ntp) {
    const int L = f->maxlevel, nb = f->nb, lmax = f->lmax, N = f->N;
    // ---- geometry per axis (get_box_cell_index_limits, get_nearfield_cell_index_limits)
    int maxn = 0;
    for (int a = 0; a < 3; ++a) {
        const Axis &ax = grid->axis[a];
        const int s = ax.ncell / nb, np1 = ax.nlip - 1;
        f->cells_per_box[a] = s;
        f->npt[a] = ax.n();
        f->qmin[a] = ax.points.front();
        f->qmax[a] = ax.points.back();
        for (int b = 0; b < nb; ++b) {
            f->box_lo[a].push_back(b * s * np1);
            f->box_n[a].push_back(s * np1 + 1);
            f->box_cut[a].push_back(b < nb - 1 ? s * np1 : s * np1 + 1);
            const int c0 = std::max(0, b - 1) * s, c1 = std::min(nb, b + 2) * s;
            f->nf_lo[a].push_back(c0 * np1);
            f->nf_n[a].push_back((c1 - c0) * np1 + 1);
        }
        maxn = std::max(maxn, s * np1 + 1);
    }
    // ---- near field: the sub-operator of each box coordinate (near-field
    // sub-grid -> box sub-grid), assembled into banded whole-cube matrices:
    // output row o (owned by box b under the cut rule) takes row o - box_lo
    // of box b's sub-operator at columns nf_lo(b) + j.
    {
        const int T = ntp;
        const int n0 = f->npt[0], n1 = f->npt[1], n2 = f->npt[2];
        std::vector<double> FX((size_t)n0 * n0 * T, 0.0), FY((size_t)n1 * n1 * T, 0.0),
            FZ((size_t)n2 * n2 * T, 0.0);
        for (int b = 0; b < nb; ++b) {
            double q_nf[3], q_box[3];
            int n_nf[3], n_box[3];
            std::vector<double> st_nf[3], st_box[3];
            for (int a = 0; a < 3; ++a) {
                const Axis &ax = grid->axis[a];
                const int s = f->cells_per_box[a];
                const int c0 = std::max(0, b - 1) * s, c1 = std::min(nb, b + 2) * s;
                q_nf[a] = ax.celld[c0];
                n_nf[a] = c1 - c0;
                st_nf[a].assign(ax.steps.begin() + c0, ax.steps.begin() + c1);
                q_box[a] = ax.celld[b * s];
                n_box[a] = s;
                st_box[a].assign(ax.steps.begin() + b * s, ax.steps.begin() + (b + 1) * s);
            }
            const int bn[3] = {f->box_n[0][b], f->box_n[1][b], f->box_n[2][b]};
            const int nn[3] = {f->nf_n[0][b], f->nf_n[1][b], f->nf_n[2][b]};
            std::vector<double> fx((size_t)bn[0] * nn[0] * T), fy((size_t)nn[1] * bn[1] * T),
                fz((size_t)nn[2] * bn[2] * T);
            dage_grid3d gnf = nullptr, gbox = nullptr;
            dage_operator op = nullptr;
            int rc = dage_grid3d_create(f->ctx, q_nf, n_nf, f->nlip, st_nf[0].data(),
                                        st_nf[1].data(), st_nf[2].data(), grid->axis[0].grid_type,
                                        &gnf);
            if (rc == DAGE_OK)
                rc = dage_grid3d_create(f->ctx, q_box, n_box, f->nlip, st_box[0].data(),
                                        st_box[1].data(), st_box[2].data(),
                                        grid->axis[0].grid_type, &gbox);
            if (rc == DAGE_OK)
                rc = dage_coulomb3d_create(f->ctx, gnf, gbox, tp, tw, ntp, 1.0, DAGE_OP_NO_CODA,
                                           &op);
            if (rc == DAGE_OK) rc = dage_operator_get_matrices(op, fx.data(), fy.data(), fz.data());
            if (op) dage_operator_destroy(op);
            if (gbox) dage_grid3d_destroy(gbox);
            if (gnf) dage_grid3d_destroy(gnf);
            if (rc != DAGE_OK) return rc;
            const int cx = f->box_cut[0][b], cy = f->box_cut[1][b], cz = f->box_cut[2][b];
            for (int t = 0; t < T; ++t) {
                for (int xi = 0; xi < nn[0]; ++xi)
                    for (int xo = 0; xo < cx; ++xo)
                        FX[(size_t)(f->box_lo[0][b] + xo) +
                           (size_t)n0 * ((f->nf_lo[0][b] + xi) + (size_t)n0 * t)] =
                            fx[(size_t)xo + (size_t)bn[0] * (xi + (size_t)nn[0] * t)];
                for (int yo = 0; yo < cy; ++yo)
                    for (int yi = 0; yi < nn[1]; ++yi)
                        FY[(size_t)(f->nf_lo[1][b] + yi) +
                           (size_t)n1 * ((f->box_lo[1][b] + yo) + (size_t)n1 * t)] =
                            fy[(size_t)yi + (size_t)nn[1] * (yo + (size_t)bn[1] * t)];
                for (int zo = 0; zo < cz; ++zo)
                    for (int zi = 0; zi < nn[2]; ++zi)
                        FZ[(size_t)(f->nf_lo[2][b] + zi) +
                           (size_t)n2 * ((f->box_lo[2][b] + zo) + (size_t)n2 * t)] =
                            fz[(size_t)zi + (size_t)nn[2] * (zo + (size_t)bn[2] * t)];
            }
        }
        DAGE_TRY(dage_coulomb3d_create(f->ctx, grid, grid, tp, tw, ntp, 1.0, DAGE_OP_NO_CODA,
                                       &f->nf_op));
        DAGE_TRY(dage_operator_set_matrices(f->nf_op, FX.data(), FY.data(), FZ.data()));
    }

    // ---- box numbering and centres (get_box_vector, get_box_center)
    const int nbox_max = nb * nb * nb;
    f->boxvec.resize(3 * nbox_max);
    std::vector<double> centers(3 * nbox_max);
    for (int idx = 0; idx < nbox_max; ++idx) {
        int v[3];
        box_vector(idx, L, v);
        for (int a = 0; a < 3; ++a) {
            f->boxvec[3 * idx + a] = v[a];
            const double size = (f->qmax[a] - f->qmin[a]) / (double)nb;
            centers[3 * idx + a] = f->qmin[a] + (v[a] + 0.5) * size;
        }
    }
    f->level_offset.assign(L + 1, 0);
    for (int l = L - 1; l >= f->start_level; --l) f->level_offset[l] = f->level_offset[l + 1] + (1 << (3 * (l + 1)));
    f->nbox_total = 0;
    for (int l = f->start_level; l <= L; ++l) f->nbox_total += 1 << (3 * l);

    std::vector<int> geo(9 * nb);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < nb; ++b) {
            geo[a * 3 * nb + b] = f->box_lo[a][b];
            geo[a * 3 * nb + nb + b] = f->box_n[a][b];
            geo[a * 3 * nb + 2 * nb + b] = f->box_cut[a][b];
        }
    // box sub-grid weights (Grid1D_get_ints of the box sub-axis)
    f->wbox_ld = maxn;
    std::vector<double> wbox((size_t)3 * nb * maxn, 0.0);
    double base[kMaxNlip];
    lip_integrals(f->nlip, grid->axis[0].grid_type, base);
    for (int a = 0; a < 3; ++a) {
        const Axis &ax = grid->axis[a];
        const int s = f->cells_per_box[a], np1 = ax.nlip - 1;
        for (int b = 0; b < nb; ++b) {
            double *w = &wbox[(size_t)(a * nb + b) * maxn];
            for (int c = 0; c < s; ++c)
                for (int j = 0; j < ax.nlip; ++j) w[c * np1 + j] += base[j] * ax.steps[b * s + c];
        }
    }
    f->points_ld = std::max(f->npt[0], std::max(f->npt[1], f->npt[2]));
    std::vector<double> pts((size_t)3 * f->points_ld, 0.0);
    for (int a = 0; a < 3; ++a)
        std::copy(grid->axis[a].points.begin(), grid->axis[a].points.end(),
                  pts.begin() + (size_t)a * f->points_ld);
    // recursion coefficients a[lm], b[lm], f[l]
    std::vector<double> coef(2 * N + lmax + 1, 0.0);
    for (int l = 0; l <= lmax; ++l) {
        coef[2 * N + l] = std::sqrt((2 * l - 1.0) / (2.0 * l));
        for (int m = -l; m <= l; ++m) {
            const int lm = l * l + l + m;
            if (std::abs(m) < l) {
                coef[lm] = (2 * l - 1.0) / std::sqrt(1.0 * (l + m) * (l - m));
                if (l - std::abs(m) > 1)
                    coef[N + lm] = std::sqrt((l + m - 1.0) * (l - m - 1.0) / ((l + m) * (l - m)));
            }
        }
    }
    DAGE_TRY(upload((void **)&f->d_boxvec, f->boxvec.data(), 4 * f->boxvec.size()));
    DAGE_TRY(upload((void **)&f->d_geo, geo.data(), 4 * geo.size()));
    DAGE_TRY(upload((void **)&f->d_wbox, wbox.data(), 8 * wbox.size()));
    DAGE_TRY(upload((void **)&f->d_points, pts.data(), 8 * pts.size()));
    DAGE_TRY(upload((void **)&f->d_centers, centers.data(), 8 * centers.size()));
    DAGE_TRY(upload((void **)&f->d_coef, coef.data(), 8 * coef.size()));
    DAGE_HIP_CHECK(hipMalloc((void **)&f->d_q, 8 * (size_t)N * f->nbox_total));
    DAGE_HIP_CHECK(hipMalloc((void **)&f->d_L, 8 * (size_t)N * f->nbox_total));
    f->zsplit = std::max(1, std::min(maxn, (2048 + nbox_max - 1) / nbox_max));
    DAGE_HIP_CHECK(hipMalloc((void **)&f->d_part, 8 * (size_t)N * nbox_max * f->zsplit));

    // ---- uniform boxes: the shared solid-harmonic matrix of the GEMM route
    {
        bool uni = true;
        std::vector<double> rel((size_t)3 * maxn, 0.0), w0((size_t)3 * maxn, 0.0);
        for (int a = 0; a < 3 && uni; ++a) {
            const int n0 = f->box_n[a][0];
            const double ext = f->qmax[a] - f->qmin[a];
            // centre of box coordinate b along axis a (boxes of one slab share it)
            auto centre = [&](int b) {
                int v[3] = {0, 0, 0};
                v[a] = b;
                return centers[3 * box_index(v[0], v[1], v[2], L) + a];
            };
            for (int i = 0; i < n0; ++i) {
                rel[(size_t)a * maxn + i] = grid->axis[a].points[f->box_lo[a][0] + i] - centre(0);
                w0[(size_t)a * maxn + i] = wbox[(size_t)(a * nb) * maxn + i];
            }
            for (int b = 1; b < nb && uni; ++b) {
                if (f->box_n[a][b] != n0) uni = false;
                for (int i = 0; i < n0 && uni; ++i) {
                    const double r = grid->axis[a].points[f->box_lo[a][b] + i] - centre(b);
                    const double wv = wbox[(size_t)(a * nb + b) * maxn + i];
                    if (std::fabs(r - rel[(size_t)a * maxn + i]) > 1e-12 * ext ||
                        std::fabs(wv - w0[(size_t)a * maxn + i]) >
                            1e-13 * std::fabs(w0[(size_t)a * maxn + i]))
                        uni = false;
                }
            }
            f->pb[a] = n0;
        }
        f->uniform = uni;
        DAGE_REQUIRE(uni || f->lmax <= kMaxBoxKernelLmax, DAGE_ERR_UNSUPPORTED,
                     "boxes of unequal shape need lmax <= %d (got %d): the per-box multipole "
                     "kernel is instantiated up to that order", kMaxBoxKernelLmax, f->lmax);
        if (uni) {
            f->P = (long)f->pb[0] * f->pb[1] * f->pb[2];
            f->box_rel = rel;
            f->box_w = w0;
            f->box_ld = maxn;
            int rc = DAGE_OK;
            const size_t ps = (size_t)f->P;
            if (rc == DAGE_OK && (hipMalloc((void **)&f->d_S, 8 * ps * N) != hipSuccess ||
                                  hipMalloc((void **)&f->d_Sw, 8 * ps * N) != hipSuccess ||
                                  hipMalloc((void **)&f->d_R, 8 * ps * nbox_max) != hipSuccess ||
                                  hipMalloc((void **)&f->d_Pbox, 8 * ps * nbox_max) != hipSuccess)) {
                set_error("GBFMM box-harmonics allocation failed");
                rc = DAGE_ERR_HIP;
            }
            // multipole GEMM: K = P split so that ~1024 workgroups run
            const int mt = (N + 63) / 64, nt = (nbox_max + 63) / 64;
            f->msplits = (int)std::max<long>(1, std::min<long>(f->P / 64, 1024 / (mt * nt)));
            if (rc == DAGE_OK && hipMalloc((void **)&f->d_gpart,
                                           8 * (size_t)f->msplits * N * nbox_max) != hipSuccess) {
                set_error("GBFMM partials allocation failed");
                rc = DAGE_ERR_HIP;
            }
            if (rc == DAGE_OK) rc = build_box_basis(f);
            DAGE_TRY(rc);
        }
    }

    // ---- interaction lists and far-field matrices (init_interaction_matrices,
    // get_local_farfield_indices)
    f->m2l_keys.assign(L + 1, {});
    f->d_T.assign(L + 1, nullptr);
    f->d_list.assign(L + 1, nullptr);
    f->list_len.assign(L + 1, 0);
    f->d_W.assign(L + 1, nullptr);
    f->m2l_nmat.assign(L + 1, 0);
    f->m2l_smax.assign(L + 1, 0);
    f->d_m2l_gsrc.assign(L + 1, nullptr);
    f->d_m2l_pos.assign(L + 1, nullptr);
    size_t m2l_cols = 1;
    for (int l = f->start_level; l <= L; ++l) {
        const int n = 1 << l, nbox = n * n * n;
        std::map<int, int> mat_id;
        std::vector<int> start(nbox + 1, 0), src, mat;
        for (int t = 0; t < nbox; ++t) {
            int v[3], p[3];
            box_vector(t, l, v);
            for (int a = 0; a < 3; ++a) p[a] = v[a] >> 1;
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (!dx && !dy && !dz) continue;
                        const int q[3] = {p[0] + dx, p[1] + dy, p[2] + dz};
                        if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= n / 2 || q[1] >= n / 2 ||
                            q[2] >= n / 2)
                            continue;
                        for (int o = 0; o < 8; ++o) {
                            const int c[3] = {2 * q[0] + (o & 1), 2 * q[1] + ((o >> 1) & 1),
                                              2 * q[2] + ((o >> 2) & 1)};
                            const int ox = c[0] - v[0], oy = c[1] - v[1], oz = c[2] - v[2];
                            if (std::abs(ox) <= 1 && std::abs(oy) <= 1 && std::abs(oz) <= 1)
                                continue;  // nearest neighbour: near field
                            const int key = (ox + 8) + 17 * ((oy + 8) + 17 * (oz + 8));
                            auto it = mat_id.find(key);
                            const int id = (it == mat_id.end()) ? (int)mat_id.size() : it->second;
                            if (it == mat_id.end()) mat_id[key] = id;
                            src.push_back(f->level_offset[l] + box_index(c[0], c[1], c[2], l));
                            mat.push_back(id);
                        }
                    }
            start[t + 1] = (int)src.size();
        }
        f->m2l_keys[l].assign(mat_id.size(), 0);
        for (auto &kv : mat_id) f->m2l_keys[l][kv.second] = kv.first;
        f->list_len[l] = (int)src.size();
        DAGE_TRY(upload((void **)&f->d_list[l], start.data(), 4 * start.size()));
        {   // group the pairs by matrix: slot s of matrix o = its s-th pair in list order
            const int nmat = std::max<int>(1, (int)mat_id.size());
            std::vector<int> cnt(nmat, 0);
            for (int m : mat) ++cnt[m];
            const int smax = std::max(1, *std::max_element(cnt.begin(), cnt.end()));
            std::vector<int> gsrc((size_t)nmat * smax, -1), pos(std::max<size_t>(1, src.size()));
            std::vector<int> next(nmat, 0);
            for (size_t e = 0; e < src.size(); ++e) {
                const int o = mat[e], slot = next[o]++;
                gsrc[(size_t)o * smax + slot] = src[e];
                pos[e] = o * smax + slot;
            }
            f->m2l_nmat[l] = nmat;
            f->m2l_smax[l] = smax;
            DAGE_TRY(upload((void **)&f->d_m2l_gsrc[l], gsrc.data(), 4 * gsrc.size()));
            DAGE_TRY(upload((void **)&f->d_m2l_pos[l], pos.data(), 4 * pos.size()));
            m2l_cols = std::max(m2l_cols, (size_t)nmat * smax);
        }
    }
    DAGE_HIP_CHECK(hipMalloc((void **)&f->d_m2l_G, 8 * (size_t)N * m2l_cols));
    DAGE_HIP_CHECK(hipMalloc((void **)&f->d_m2l_out, 8 * (size_t)N * m2l_cols));
    return build_far_matrices(f);
}

// Runs the context's launches on another stream while in scope.
struct StreamSwap {
    dage_ctx ctx;
    hipStream_t saved;
    StreamSwap(dage_ctx c, hipStream_t s) : ctx(c), saved(c->stream) { c->stream = s; }
    ~StreamSwap() { ctx->stream = saved; }
};

// Far field of the finest boxes up to their local-expansion coefficients
// d_L: multipoles, M2M, M2L, L2L (reads only rho), and for equal boxes the
// expansions evaluated at every box's points (d_Pbox).
int gbfmm_far_coefficients(dage_gbfmm f, const double *rho);

int gbfmm_apply(dage_gbfmm f, const double *rho, double *pot, int precision) {
    dage_ctx ctx = f->ctx;
    hipStream_t s = ctx->stream;
    const int nb = f->nb, N = f->N;
    const long ldx = f->npt[0], ldy = f->npt[1];
    hipEvent_t ev;
    const int nbox_max = nb * nb * nb;
    // near field of every box: one whole-cube apply with the banded matrices
    // (context stream); the far-field coefficients beside it on s_far
    DAGE_HIP_CHECK(hipEventRecord(f->ev_fork, s));
    DAGE_HIP_CHECK(hipStreamWaitEvent(f->s_far, f->ev_fork, 0));
    if (precision == DAGE_PREC_FP64) DAGE_TRY(apply_f64(f->nf_op, rho, pot));
    else DAGE_TRY(apply_mixed(f->nf_op, rho, pot, precision == DAGE_PREC_BF16X6 ? 6 : 3));
    {
        StreamSwap swap(ctx, f->s_far);
        DAGE_TRY(gbfmm_far_coefficients(f, rho));
    }
    DAGE_HIP_CHECK(hipEventRecord(f->ev_join, f->s_far));
    DAGE_HIP_CHECK(hipStreamWaitEvent(s, f->ev_join, 0));
    // local expansions on the finest boxes' points (after the near field)
    LeArgs la;
    la.out = pot; la.ldx = ldx; la.ldy = ldy;
    la.boxvec = f->d_boxvec; la.geo = f->d_geo; la.nb = nb;
    la.points = f->d_points; la.points_ld = f->points_ld;
    la.centers = f->d_centers; la.coef = f->d_coef;
    la.L = f->d_L;  // level maxlevel has offset 0
    la.lmax = f->lmax; la.N = N;
    if (f->uniform) {
        // potential at the box's points = S^T L (computed on s_far), added
        // over the points each box owns
        prof_begin(ctx, "local_expansion_gemm", &ev);
        const long tot = f->P * nbox_max;
        const int gs = (int)std::min<long>((tot + 255) / 256, 65536);
        hipLaunchKernelGGL(box_scatter_kernel, dim3(gs), dim3(256), 0, s, f->d_Pbox, ldx, ldy,
                           f->d_boxvec, f->d_geo, nb, f->pb[0], f->pb[1], f->pb[2], nbox_max,
                           pot);
        DAGE_HIP_CHECK(hipGetLastError());
        prof_end(ctx, "local_expansion_gemm", ev);
        return DAGE_OK;
    }
    int maxz = 0;
    for (int b = 0; b < nb; ++b) maxz = std::max(maxz, f->box_cut[2][b]);
    prof_begin(ctx, "local_expansion_kernel", &ev);
    hipLaunchKernelGGL(local_expansion_kernel, dim3(nbox_max, maxz), dim3(256), 0, s, la);
    DAGE_HIP_CHECK(hipGetLastError());
    prof_end(ctx, "local_expansion_kernel", ev);
    return DAGE_OK;
}

int gbfmm_far_coefficients(dage_gbfmm f, const double *rho) {
    dage_ctx ctx = f->ctx;
    hipStream_t s = ctx->stream;
    const int nb = f->nb, N = f->N, L = f->maxlevel;
    const long ldx = f->npt[0], ldy = f->npt[1];
    hipEvent_t ev;
    // multipoles of the finest boxes
    const int nbox_max = nb * nb * nb;
    DAGE_TRY(fill_zero(ctx, f->d_L, (size_t)N * f->nbox_total));
    MultArgs ma;
    ma.rho = rho; ma.ldx = ldx; ma.ldy = ldy;
    ma.boxvec = f->d_boxvec; ma.geo = f->d_geo; ma.nb = nb;
    ma.wbox = f->d_wbox; ma.wbox_ld = f->wbox_ld;
    ma.points = f->d_points; ma.points_ld = f->points_ld;
    ma.centers = f->d_centers; ma.coef = f->d_coef;
    ma.zsplit = f->zsplit; ma.nbox = nbox_max;
    ma.part = f->d_part;
    if (f->uniform) {
        // q(:, box) = Sw^T-layout GEMM over the box's points: q = (S w) R
        const long tot = f->P * nbox_max;
        const int gs = (int)std::min<long>((tot + 255) / 256, 65536);
        hipEvent_t ev2;
        prof_begin(ctx, "box_multipoles_gemm", &ev2);
        hipLaunchKernelGGL(box_gather_kernel, dim3(gs), dim3(256), 0, s, rho, ldx, ldy,
                           f->d_boxvec, f->d_geo, nb, f->pb[0], f->pb[1], f->pb[2], nbox_max,
                           f->d_R);
        DAGE_HIP_CHECK(hipGetLastError());
        DAGE_TRY(gemm_colmajor(ctx, "gemm_f64_multipoles", N, nbox_max, (int)f->P, f->d_Sw, N,
                               f->d_R, f->P, f->d_q, f->msplits, f->d_gpart));
        prof_end(ctx, "box_multipoles_gemm", ev2);
    } else {
        prof_begin(ctx, "box_multipoles_kernel", &ev);
        DAGE_TRY(multipoles(f->lmax, ma, nbox_max, s));
        prof_end(ctx, "box_multipoles_kernel", ev);
        const long qlen = (long)N * nbox_max;
        hipLaunchKernelGGL(sum_parts_kernel, dim3((int)std::min<long>((qlen + 255) / 256, 65536)),
                           dim3(256), 0, s, f->d_part, f->zsplit, qlen, f->d_q);
        DAGE_HIP_CHECK(hipGetLastError());
    }
    // upward pass
    for (int l = L - 1; l >= f->start_level; --l) {
        prof_begin(ctx, "m2m_kernel", &ev);
        hipLaunchKernelGGL(m2m_kernel, dim3(1 << (3 * l)), dim3(256), 0, s, f->d_W[l + 1],
                           f->d_q + (size_t)N * f->level_offset[l + 1],
                           f->d_q + (size_t)N * f->level_offset[l], N, f->yukawa ? 1 : 0);
        DAGE_HIP_CHECK(hipGetLastError());
        prof_end(ctx, "m2m_kernel", ev);
    }
    // far field per level, translated down
    for (int l = f->start_level; l <= L; ++l) {
        const int nbox = 1 << (3 * l);
        double *Ll = f->d_L + (size_t)N * f->level_offset[l];
        if (l > f->start_level) {
            prof_begin(ctx, "l2l_kernel", &ev);
            hipLaunchKernelGGL(l2l_kernel, dim3(nbox), dim3(256), 0, s, f->d_W[l],
                               f->d_L + (size_t)N * f->level_offset[l - 1], Ll, N,
                               f->yukawa ? 1 : 0);
            DAGE_HIP_CHECK(hipGetLastError());
            prof_end(ctx, "l2l_kernel", ev);
        }
        if (f->list_len[l] == 0) continue;
        // one GEMM per box offset: T(offset) times the multipoles of every
        // source at that offset, then each target adds its columns in list order
        const int nmat = f->m2l_nmat[l], smax = f->m2l_smax[l];
        const long ncol = (long)nmat * smax;
        prof_begin(ctx, "m2l_gemm", &ev);
        hipLaunchKernelGGL(m2l_gather_kernel,
                           dim3((int)std::min<long>((ncol * N + 255) / 256, 65536)), dim3(256), 0,
                           s, f->d_q, f->d_m2l_gsrc[l], ncol, N, f->d_m2l_G);
        DAGE_HIP_CHECK(hipGetLastError());
        DAGE_TRY(gemm_colmajor_batched(ctx, "gemm_f64_m2l", N, smax, N, f->d_T[l], N,
                                       (long)N * N, f->d_m2l_G, N, (long)N * smax,
                                       f->d_m2l_out, (long)N * smax, nmat));
        hipLaunchKernelGGL(m2l_scatter_kernel, dim3(nbox), dim3(256), 0, s, f->d_m2l_out,
                           f->d_list[l], f->d_m2l_pos[l], N, Ll);
        DAGE_HIP_CHECK(hipGetLastError());
        prof_end(ctx, "m2l_gemm", ev);
    }
    if (f->uniform) {
        // the local expansions at every box's points, P_box = S^T L (needs
        // only L; the scatter into the potential waits for the near field)
        prof_begin(ctx, "local_expansion_gemm", &ev);
        DAGE_TRY(gemm_colmajor(ctx, "gemm_f64_local_expansion", (int)f->P, nbox_max, N, f->d_S,
                               f->P, f->d_L, N, f->d_Pbox, 1, nullptr));
        prof_end(ctx, "local_expansion_gemm", ev);
    }
    return DAGE_OK;
}

int use_device(dage_ctx ctx) {
    DAGE_REQUIRE(ctx != nullptr, DAGE_ERR_ARG, "null context handle");
    DAGE_HIP_CHECK(hipSetDevice(ctx->device));
    return DAGE_OK;
}

}  // namespace
}  // namespace dage

using namespace dage;

extern "C" {

int dage_gbfmm_create(dage_ctx ctx, dage_grid3d grid, int maxlevel, int lmax, const double *tp,
                      const double *tw, int ntp, dage_gbfmm *out) {
    DAGE_REQUIRE(out != nullptr, DAGE_ERR_ARG, "null output pointer");
    *out = nullptr;
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(grid && tp && tw && ntp > 0, DAGE_ERR_ARG, "null or empty argument");
    DAGE_REQUIRE(grid->ctx == ctx, DAGE_ERR_ARG, "grid belongs to another context");
    DAGE_REQUIRE(maxlevel >= 2 && maxlevel <= 6, DAGE_ERR_UNSUPPORTED,
                 "maxlevel %d outside 2..6 (boxes start at level 2)", maxlevel);
    DAGE_REQUIRE(lmax >= 0 && lmax <= kMaxFmmLmax, DAGE_ERR_UNSUPPORTED, "lmax %d outside 0..%d",
                 lmax, kMaxFmmLmax);
    const int nb = 1 << maxlevel;
    for (int a = 0; a < 3; ++a)
        DAGE_REQUIRE(grid->axis[a].ncell % nb == 0, DAGE_ERR_ARG,
                     "axis %d: %d cells do not split into %d boxes", a, grid->axis[a].ncell, nb);
    auto *f = new (std::nothrow) dage_gbfmm_s();
    DAGE_REQUIRE(f != nullptr, DAGE_ERR_ALLOC, "out of host memory");
    f->ctx = ctx;
    ctx_retain(ctx);
    f->maxlevel = maxlevel;
    f->lmax = lmax;
    f->N = (lmax + 1) * (lmax + 1);
    f->nb = nb;
    f->nlip = grid->axis[0].nlip;
    int rc = DAGE_OK;
    if (hipStreamCreateWithFlags(&f->s_far, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&f->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&f->ev_join, hipEventDisableTiming) != hipSuccess) {
        set_error("GBFMM stream or event creation failed");
        rc = DAGE_ERR_HIP;
    }
    if (rc == DAGE_OK) rc = build_gbfmm(f, grid, tp, tw, ntp);
    if (rc != DAGE_OK) {
        free_gbfmm(f);
        return rc;
    }
    *out = f;
    return DAGE_OK;
}

int dage_gbfmm_apply_device(dage_gbfmm f, const double *rho, double *pot, int precision) {
    DAGE_REQUIRE(f && rho && pot, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE(precision == DAGE_PREC_FP64 || precision == DAGE_PREC_BF16X6 ||
                     precision == DAGE_PREC_BF16X3,
                 DAGE_ERR_ARG, "unknown precision mode %d", precision);
    // the far field reads rho on a second stream while the near field writes
    // pot on the context stream, so the two buffers must not overlap
    const size_t n = (size_t)f->npt[0] * f->npt[1] * f->npt[2];
    DAGE_REQUIRE(rho + n <= pot || pot + n <= rho, DAGE_ERR_ARG,
                 "rho and pot overlap (the far field reads rho while pot is written)");
    DAGE_TRY(use_device(f->ctx));
    return gbfmm_apply(f, rho, pot, precision);
}

int dage_gbfmm_apply(dage_gbfmm f, const double *rho, double *pot, int precision) {
    DAGE_REQUIRE(f && rho && pot, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(use_device(f->ctx));
    const size_t n = (size_t)f->npt[0] * f->npt[1] * f->npt[2];
    double *d_in = nullptr, *d_out = nullptr;
    DAGE_HIP_CHECK(hipMalloc((void **)&d_in, 8 * n));
    if (hipMalloc((void **)&d_out, 8 * n) != hipSuccess) {
        (void)hipFree(d_in);
        set_error("device allocation of %zu bytes failed", 8 * n);
        return DAGE_ERR_HIP;
    }
    int rc = DAGE_OK;
    if (hipMemcpyAsync(d_in, rho, 8 * n, hipMemcpyHostToDevice, f->ctx->stream) != hipSuccess) {
        set_error("upload failed");
        rc = DAGE_ERR_HIP;
    }
    if (rc == DAGE_OK) rc = dage_gbfmm_apply_device(f, d_in, d_out, precision);
    if (rc == DAGE_OK &&
        hipMemcpyAsync(pot, d_out, 8 * n, hipMemcpyDeviceToHost, f->ctx->stream) != hipSuccess) {
        set_error("download failed");
        rc = DAGE_ERR_HIP;
    }
    if (hipStreamSynchronize(f->ctx->stream) != hipSuccess && rc == DAGE_OK) {
        set_error("stream synchronisation failed");
        rc = DAGE_ERR_HIP;
    }
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    return rc;
}

int dage_gbfmm_set_energy(dage_gbfmm f, double energy) {
    DAGE_REQUIRE(f != nullptr, DAGE_ERR_ARG, "null GBFMM handle");
    DAGE_REQUIRE(std::isfinite(energy), DAGE_ERR_ARG, "energy is not finite");
    DAGE_TRY(use_device(f->ctx));
    const bool yk = energy < -0.005;
    DAGE_REQUIRE(!yk || f->uniform, DAGE_ERR_UNSUPPORTED,
                 "the Helmholtz far field needs boxes of equal point counts and spacing");
    DAGE_HIP_CHECK(hipStreamSynchronize(f->ctx->stream));
    DAGE_HIP_CHECK(hipStreamSynchronize(f->s_far));
    // near field: the banded operator's weights (Helmholtz3D_set_energy)
    DAGE_TRY(dage_operator_set_energy(f->nf_op, energy));
    const bool was = f->yukawa;
    const double kappa_was = f->kappa;
    f->yukawa = yk;
    f->kappa = yk ? std::sqrt(-2.0 * energy) : 0.0;
    if (was == yk && kappa_was == f->kappa) return DAGE_OK;
    int rc = build_far_matrices(f);
    if (rc == DAGE_OK) rc = build_box_basis(f);
    if (rc != DAGE_OK) {  // back to the 1/r far field, which always builds
        f->yukawa = false;
        f->kappa = 0.0;
        (void)dage_operator_set_energy(f->nf_op, 0.0);
        (void)build_far_matrices(f);
        (void)build_box_basis(f);
    }
    return rc;
}

int dage_gbfmm_box_count(dage_gbfmm f, int *nbox_total) {
    DAGE_REQUIRE(f && nbox_total, DAGE_ERR_ARG, "null argument");
    *nbox_total = f->nbox_total;
    return DAGE_OK;
}

int dage_gbfmm_get_band(dage_gbfmm f, int *t0) {
    DAGE_REQUIRE(f && t0, DAGE_ERR_ARG, "null argument");
    return dage_operator_get_band(f->nf_op, t0);
}

int dage_gbfmm_get_lowrank(dage_gbfmm f, int *t1, int *kx_sum) {
    DAGE_REQUIRE(f && t1 && kx_sum, DAGE_ERR_ARG, "null argument");
    return dage_operator_get_lowrank(f->nf_op, t1, kx_sum);
}

int dage_gbfmm_get_multipoles(dage_gbfmm f, double *q) {
    DAGE_REQUIRE(f && q, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(use_device(f->ctx));
    DAGE_HIP_CHECK(hipStreamSynchronize(f->ctx->stream));
    DAGE_HIP_CHECK(hipMemcpy(q, f->d_q, 8 * (size_t)f->N * f->nbox_total, hipMemcpyDeviceToHost));
    return DAGE_OK;
}

int dage_gbfmm_get_local_expansions(dage_gbfmm f, double *L) {
    DAGE_REQUIRE(f && L, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(use_device(f->ctx));
    DAGE_HIP_CHECK(hipStreamSynchronize(f->ctx->stream));
    DAGE_HIP_CHECK(hipMemcpy(L, f->d_L, 8 * (size_t)f->N * f->nbox_total, hipMemcpyDeviceToHost));
    return DAGE_OK;
}

int dage_gbfmm_destroy(dage_gbfmm f) {
    if (!f) return DAGE_OK;
    (void)hipSetDevice(f->ctx->device);
    (void)hipStreamSynchronize(f->ctx->stream);
    free_gbfmm(f);
    return DAGE_OK;
}

int dage_fmm_interaction_matrix(int lmax, const double dR[3], double *T) {
    DAGE_REQUIRE(dR && T, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE(lmax >= 0 && lmax <= kMaxFmmLmax, DAGE_ERR_UNSUPPORTED, "lmax %d outside 0..%d",
                 lmax, kMaxFmmLmax);
    DAGE_REQUIRE(dR[0] != 0.0 || dR[1] != 0.0 || dR[2] != 0.0, DAGE_ERR_ARG,
                 "the interaction matrix is undefined at zero separation");
    fmm_interaction(fmm_tables(lmax), dR, T);
    return DAGE_OK;
}

int dage_fmm_translation_matrix(int lmax, const double dR[3], double *W) {
    DAGE_REQUIRE(dR && W, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE(lmax >= 0 && lmax <= kMaxFmmLmax, DAGE_ERR_UNSUPPORTED, "lmax %d outside 0..%d",
                 lmax, kMaxFmmLmax);
    fmm_translation(fmm_tables(lmax), dR, W);
    return DAGE_OK;
}

int dage_bessel_modified_spherical(int lmax, double x, double *i_l, double *k_l) {
    DAGE_REQUIRE(i_l && k_l, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE(lmax >= 0 && lmax <= kMaxFmmBasisL, DAGE_ERR_UNSUPPORTED,
                 "lmax %d outside 0..%d", lmax, kMaxFmmBasisL);
    DAGE_REQUIRE(x > 0.0 && x <= 600.0, DAGE_ERR_ARG, "x = %g outside (0, 600]", x);
    double g[kMaxFmmBasisL + 1];
    bessel_i_scaled(lmax, x, g);
    double xl = 1.0;
    for (int l = 0; l <= lmax; ++l) {
        i_l[l] = g[l] * xl;
        xl *= x;
    }
    bessel_k(lmax, x, k_l);
    return DAGE_OK;
}

int dage_fmm_yukawa_interaction_matrix(int lmax, double kappa, const double dR[3], double rho,
                                       double *T) {
    DAGE_REQUIRE(dR && T, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE(lmax >= 0 && lmax <= kMaxFmmLmax, DAGE_ERR_UNSUPPORTED, "lmax %d outside 0..%d",
                 lmax, kMaxFmmLmax);
    const double d = std::sqrt(dR[0] * dR[0] + dR[1] * dR[1] + dR[2] * dR[2]);
    DAGE_REQUIRE(kappa > 0.0 && rho > 0.0 && d > 2.0 * rho && kappa * rho <= 600.0, DAGE_ERR_ARG,
                 "needs kappa > 0, 0 < rho < |dR|/2 and kappa rho <= 600");
    yukawa_m2l(yukawa_sphere(lmax), kappa, dR, rho, T);
    return DAGE_OK;
}

int dage_fmm_yukawa_translation_matrix(int lmax, double kappa, const double d[3], double rho,
                                       double *A) {
    DAGE_REQUIRE(d && A, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE(lmax >= 0 && lmax <= kMaxFmmLmax, DAGE_ERR_UNSUPPORTED, "lmax %d outside 0..%d",
                 lmax, kMaxFmmLmax);
    DAGE_REQUIRE(kappa > 0.0 && rho > 0.0 && kappa * rho <= 600.0, DAGE_ERR_ARG,
                 "needs kappa > 0, rho > 0 and kappa rho <= 600");
    yukawa_shift(yukawa_sphere(lmax), kappa, d, rho, A);
    return DAGE_OK;
}

}  // extern "C"
