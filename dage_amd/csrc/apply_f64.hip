// FP64 tensor-quadrature apply on gfx950 matrix cores.
//
// out(xo,yo,zo) = sum_t w_t sum_{xi,yi,zi} Fx_t(xo,xi) Fy_t(yi,yo) Fz_t(zi,zo)
//                 in(xi,yi,zi)  + coda*in
// (Operator3D_transform_cube, src/bubbles/function3d.F90:4230-4349, and the
// cuBLAS path it replaces, function3d.F90:4354-4480 / cuda_blas.cu:141-330).
//
// Three passes of one tiled FP64 MFMA GEMM (v_mfma_f64_16x16x4f64, LDS-staged
// operand panels, register-prefetched next panel):
//   1. x mode, batched over t:  V_t(xo,yi,zi) = sum_xi Fx_t(xo,xi) in(xi,yi,zi)
//   2. y mode, batched over t:  U((xo,yo),(zi,t)) = w_t sum_yi V_t(xo,yi,zi) Fy_t(yi,yo)
//   3. z mode with the t-sum folded into K = (zi,t):
//        out((xo,yo),zo) = sum_(zi,t) U((xo,yo),(zi,t)) Gz((zi,t),zo) + coda*in
//      split over K for occupancy, then a deterministic reduction adds the
//      splits and the coda term.
// t is processed in chunks so that V and U stay within a workspace budget.
#include "dage_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace dage {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;  // 4 wavefronts of 64
constexpr int kBK = 16;        // K panel depth staged in LDS

struct Gemm {
    const double *A;
    const double *B;
    double *C;
    int M, N, K;
    int M0;                   // m = m_lo + M0*m_hi (two-level row index)
    long sAm0, sAm1, sAk;     // A(m,k)
    long sBk, sBn;            // B(k,n)
    long sCm0, sCm1, sCn;     // C(m,n)
    long bA, bB, bC;          // strides per batch index z_lo
    int Z0;                   // >0: blockIdx.z = z_lo + Z0*z_hi (z_hi: right-hand side)
    long bA1, bB1, bC1;       // strides per z_hi
    const double *alpha;      // alpha[z_lo + aS1*z_hi] (null: 1)
    long aS1;
    int Ksplit;               // >0: z_lo selects a K range of this size
    int accumulate;           // C += result instead of C = result
    // Optional row masks of B: bit c of mask[k] is set when B(k, n) != 0 for
    // some n in [16c, 16c + 16).  Rows of batch index z_lo start at
    // mask + z_lo*mS (mS = 0 with Ksplit: the K range start is used).
    const uint64_t *mask;
    long mS;
    // >0: column tile j is [j*pitch, (j+1)*pitch), the last one running to N
    // (each at most 16*NB wide), and bit j of a mask row stands for tile j
    int pitch;
    // Tile grid (row tiles, column tiles, batch indices), set at launch; the
    // launch itself is one-dimensional (see block_coords).
    int gx, gy, gz;
    int remap;  // 0: plain row-tile-fastest order (for comparison runs)
    int batch_inner;  // remap order walks z_lo before the row tile (x pass)
    // PASS 4 (x pass, t-groups): a block covers 4 t-points of tcount (z_lo
    // = t-group) with the same A; column group j of a tile is t = j / (NB/4),
    // columns (j % (NB/4))*16.., at B + t*tB and C + t*tC
    long tB, tC;
    int tcount;
};

// Block coordinates of a one-dimensional launch.  Hardware hands block L to
// XCD L % 8, so L is first mapped to a logical tile index that gives each XCD
// one contiguous range of the tile order; the order walks column tiles
// fastest, so the blocks that share an A row tile run together on one XCD and
// read it through that XCD's L2 once instead of once per column tile.
// batch_inner (the x pass, whose A is the input cube for every t) walks the
// batch index z_lo before the row tile for the same reason.
__device__ inline void block_coords(const Gemm &g, int *bx, int *by, int *bz) {
    unsigned L = blockIdx.x;
    if (g.remap) {
        const unsigned total = gridDim.x;
        const unsigned q = total >> 3, r = total & 7u, x = L & 7u;
        L = x * q + min(x, r) + (L >> 3);
        *by = (int)(L % (unsigned)g.gy);
        L /= (unsigned)g.gy;
        if (g.batch_inner && g.Z0 > 0) {
            const unsigned zl = L % (unsigned)g.Z0;
            L /= (unsigned)g.Z0;
            *bx = (int)(L % (unsigned)g.gx);
            *bz = (int)(zl + (unsigned)g.Z0 * (L / (unsigned)g.gx));
        } else {
            *bx = (int)(L % (unsigned)g.gx);
            *bz = (int)(L / (unsigned)g.gx);
        }
    } else {
        *bx = (int)(L % (unsigned)g.gx);
        L /= (unsigned)g.gx;
        *by = (int)(L % (unsigned)g.gy);
        *bz = (int)(L / (unsigned)g.gy);
    }
}

constexpr int kMaxPanels = 2048;  // K panels a block can list (K <= 32768)

// C tile of (64*MW) x (16*NB) per block; wavefront w owns rows
// [w*16*MW, (w+1)*16*MW) and all 16*NB columns.  MFMA f64 16x16x4 operand
// layout on gfx950 (measured with a one-hot probe): lane l holds A(l%16, l/16)
// and B(l/16, l%16); result register r of lane l is D(4*r + l/16, l%16).
//
// Operand panels of KB k-values are staged in LDS; the next panel's global
// loads are issued into registers before the current panel's MFMAs, from
// clamped addresses so no load sits behind a branch.  Values outside the
// matrix are loaded as zeros, so padded k-steps add nothing.  PASS 0-2 are
// the three apply passes (one LDS buffer), PASS 3 the column-major GEMMs
// (two buffers).
template <int MW, int NB, bool A_MCONTIG, bool B_KCONTIG, int PASS, int KB = kBK>
__global__ void __launch_bounds__(kThreads) gemm_f64_kernel(Gemm g) {
    constexpr int BM = 64 * MW;
    constexpr int BN = 16 * NB;
    constexpr int A_PER = BM * KB / kThreads;  // elements per thread
    constexpr int B_PER = BN * KB / kThreads;
    constexpr int LDA = BM + 1;  // +1 double breaks power-of-two strides
    constexpr int LDB = BN + 1;
    // The mode passes stage one panel at a time (next panel prefetched into
    // registers, two barriers per panel): 21 KB of LDS lets five blocks share
    // a CU, and independent blocks hide more latency than a second buffer
    // does (x and z passes 5 % faster at 241^3 than double buffering at four
    // blocks per CU).  The column-major GEMMs keep two buffers.
    constexpr int NBUF = (PASS <= 2 || PASS == 4) ? 1 : 2;
    // PASS 4: the x pass with TG t-points per block sharing the A panels
    constexpr bool TGRP = PASS == 4;
    constexpr int TG = TGRP ? 4 : 1;
    constexpr int CG = TGRP ? (NB >= 4 ? NB / 4 : 1) : NB;  // column groups per t
    __shared__ double As[NBUF][KB * LDA];
    __shared__ double Bs[NBUF][KB * LDB];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    int bx, by, bz;
    block_coords(g, &bx, &by, &bz);
    const int m0 = bx * BM;
    const int n0 = g.pitch > 0 ? by * g.pitch : by * 16 * CG;
    const int nend = g.pitch > 0 && by + 1 < g.gy ? n0 + g.pitch : g.N;

    const double *A = g.A;
    const double *B = g.B;
    double *C = g.C;
    double alpha = 1.0;
    int kbeg = 0, kend = g.K;
    const int zlo = g.Z0 > 0 ? bz % g.Z0 : bz;
    const int zhi = g.Z0 > 0 ? bz / g.Z0 : 0;
    A += (long)zhi * g.bA1;
    B += (long)zhi * g.bB1;
    C += (long)zlo * g.bC + (long)zhi * g.bC1;
    if (g.Ksplit > 0) {
        kbeg = zlo * g.Ksplit;
        kend = min(g.K, kbeg + g.Ksplit);
        A += (long)kbeg * g.sAk;
        B += (long)kbeg * g.sBk;
    } else {
        A += (long)zlo * g.bA;
        B += (long)zlo * g.bB;
        if (g.alpha) alpha = g.alpha[zlo + g.aS1 * zhi];
    }
    const int K = kend - kbeg;
    const int tval = TGRP ? min(TG, g.tcount - zlo * TG) : 1;  // t-points of this block
    const uint64_t *mrow = nullptr;
    if (g.mask) mrow = g.mask + (g.Ksplit > 0 ? (long)kbeg : (long)zlo * TG * g.mS);

    // Loader geometry: every thread loads A_PER elements of the A panel and NB
    // of the B panel at fixed (m, k) / (k, n) steps from one base position, so
    // a slot's address is base + slot * compile-time multiple of a stride.
    //   A m-contiguous: m = tid % BM, k = tid / BM + slot * (256 / BM)
    //   A k-contiguous: k = tid % 16, m = tid / 16 + slot * 16 (one-level M)
    //   B k-contiguous: k = tid % 16, n = tid / 16 + slot * 16
    //   B n-contiguous: n = tid % 16 + slot * 16, k = tid / 16
    constexpr int A_DK = A_MCONTIG ? kThreads / BM : 0;
    constexpr int A_DM = A_MCONTIG ? 0 : kThreads / KB;
    const int a_m = A_MCONTIG ? tid % BM : tid / KB;
    const int a_k = A_MCONTIG ? tid / BM : tid % KB;
    long a_base;
    bool a_row_ok;  // m-contiguous: the single row of this thread is inside M
    {
        const int gm = m0 + a_m;
        a_row_ok = gm < g.M;
        const int mlo = a_row_ok ? gm % g.M0 : 0, mhi = a_row_ok ? gm / g.M0 : 0;
        a_base = (long)mlo * g.sAm0 + (long)mhi * g.sAm1 + (long)a_k * g.sAk;
    }
    const long a_step = A_MCONTIG ? (long)A_DK * g.sAk : (long)A_DM * g.sAm0;
    const int a_sm0 = a_k * LDA + a_m;
    constexpr int a_sstep = A_MCONTIG ? A_DK * LDA : A_DM;

    // B slot r sits (b_dk(r), b_dn(r)) from the thread's base (k, n):
    //   k-contiguous: k = tid % KB, n = tid / KB + r * (256 / KB)
    //   n-contiguous: n = tid % 16 + (r % NB) * 16, k = tid / 16 + (r / NB) * 16
    const int bk = B_KCONTIG ? tid % KB : tid / 16;
    const int bn = B_KCONTIG ? tid / KB : tid % 16;
    const long b_base = (long)bk * g.sBk + (long)(n0 + bn) * g.sBn;
    const int b_sm0 = bk * LDB + bn;
    auto b_dk = [](int r) { return B_KCONTIG ? 0 : (r / NB) * 16; };
    auto b_dn = [](int r) { return B_KCONTIG ? r * (kThreads / KB) : (r % NB) * 16; };
    // global position of slot r: column offset within the tile and t-point
    // (PASS 4), the LDS position (b_dn) keeps the group order
    auto b_gn = [](int r) { return B_KCONTIG ? r * (kThreads / KB) : ((r % NB) % CG) * 16; };
    auto b_gt = [](int r) { return TGRP ? (r % NB) / CG : 0; };

    // The loaded values stay raw in registers and their validity in a bit
    // mask: the zero for an out-of-range slot is selected when the panel is
    // stored, so no instruction consumes a prefetched value before the
    // current panel's MFMAs (a select right after each load made the wave
    // wait for the next panel's loads before computing on this one).
    double ra[A_PER], rb[B_PER];
    unsigned amask = 0, bmask = 0;
    auto load_panel = [&](int kp) {
        const int lim = K - kp;  // valid k within this panel: [0, lim)
        const long koff = (long)kp * g.sAk;
        amask = 0;
        bmask = 0;
#pragma unroll
        for (int r = 0; r < A_PER; ++r) {
            const bool ok = A_MCONTIG ? (a_row_ok && a_k + r * A_DK < lim)
                                      : (a_k < lim && m0 + a_m + r * A_DM < g.M);
            ra[r] = A[ok ? a_base + r * a_step + koff : 0];
            amask |= (ok ? 1u : 0u) << r;
        }
        const long kboff = (long)kp * g.sBk;
#pragma unroll
        for (int r = 0; r < B_PER; ++r) {
            const bool ok = bk + b_dk(r) < lim && n0 + bn + b_gn(r) < nend && b_gt(r) < tval;
            rb[r] = B[ok ? b_base + (long)b_dk(r) * g.sBk + (long)b_gn(r) * g.sBn +
                               (long)b_gt(r) * g.tB + kboff
                         : 0];
            bmask |= (ok ? 1u : 0u) << r;
        }
    };
    auto store_panel = [&](int buf) {
#pragma unroll
        for (int r = 0; r < A_PER; ++r)
            As[buf][a_sm0 + r * a_sstep] = ((amask >> r) & 1u) ? ra[r] : 0.0;
#pragma unroll
        for (int r = 0; r < B_PER; ++r)
            Bs[buf][b_sm0 + b_dk(r) * LDB + b_dn(r)] = ((bmask >> r) & 1u) ? rb[r] : 0.0;
    };

    d4 acc[MW][NB];
#pragma unroll
    for (int i = 0; i < MW; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

    const int lrow = lane & 15;
    const int lk = lane >> 4;
    const int wrow = wave * 16 * MW;

    auto kstep = [&](const double *as, const double *bs, int ks) {
        const int kk = ks * 4 + lk;
        double a[MW], b[NB];
#pragma unroll
        for (int i = 0; i < MW; ++i) a[i] = as[kk * LDA + wrow + i * 16 + lrow];
#pragma unroll
        for (int j = 0; j < NB; ++j) b[j] = bs[kk * LDB + j * 16 + lrow];
#pragma unroll
        for (int i = 0; i < MW; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    };

    // K panels whose B rows are zero on all of this block's columns add
    // exactly zero and are skipped.  Wavefront 0 lists the others in
    // ascending order (ballot + prefix count), so the summation order is the
    // same as without masks.
    __shared__ unsigned short plist[kMaxPanels];
    __shared__ int plist_n;
    int npanels = (K + KB - 1) / KB;
    const bool listed = mrow != nullptr && npanels <= kMaxPanels;
    if (listed) {
        if (wave == 0) {
            const uint64_t cols = g.pitch > 0 ? 1ull << by
                                              : ((1ull << CG) - 1ull) << (n0 / 16);
            int cnt = 0;
            for (int base = 0; base < npanels; base += 64) {
                const int p = base + lane;
                bool need = false;
                if (p < npanels) {
                    const int k0 = p * KB;
                    const int kn = min(KB, K - k0);
                    uint64_t any = 0;
                    for (int tl = 0; tl < tval; ++tl)  // PASS 4: any t of the block
                        for (int j = 0; j < kn; ++j) any |= mrow[(long)tl * g.mS + k0 + j];
                    need = (any & cols) != 0;
                }
                const uint64_t bal = __ballot(need);
                if (need) plist[cnt + __popcll(bal & ((1ull << lane) - 1ull))] = (unsigned short)p;
                cnt += __popcll(bal);
            }
            if (lane == 0) plist_n = cnt;
        }
        __syncthreads();
        npanels = plist_n;
    }
    auto panel_k = [&](int p) { return (listed ? (int)plist[p] : p) * KB; };

    if (npanels > 0) {
        load_panel(panel_k(0));
        store_panel(0);
        __syncthreads();
    }
    for (int p = 0; p < npanels; ++p) {
        const int buf = NBUF == 2 ? (p & 1) : 0;
        const bool more = p + 1 < npanels;
        if (more) load_panel(panel_k(p + 1));
        const double *as = As[buf];
        const double *bs = Bs[buf];
        // uniform bound: a panel at the end of K skips its all-padding k-steps
        // (running them as one branch-free block pipelines the LDS reads but
        // holds every k-step's operands live: 3 instead of 4 waves per SIMD,
        // 10 % slower at 385^3)
        const int ksteps = min(KB / 4, (K - panel_k(p) + 3) / 4);
#pragma unroll
        for (int ks = 0; ks < KB / 4; ++ks)
            if (ks < ksteps) kstep(as, bs, ks);
        if constexpr (NBUF == 1) {
            __syncthreads();  // every wavefront is done with the panel
            if (more) store_panel(0);
        } else {
            if (more) store_panel(buf ^ 1);
        }
        __syncthreads();
    }

    // Epilogue.  Register r of lane l holds D(4*r + l/16, l%16).
    if constexpr (PASS == 1 && MW == 1 && NB <= 4) {
        // y pass (its output rows are the contiguous axis and it writes the
        // largest array, U): each wavefront's 16 x 16*NB tile goes through
        // LDS (the operand buffers, free after the loop's last barrier; one
        // tile fits each, so wavefronts 0-1 go first, then 2-3) so that
        // stores run along rows, 16 consecutive rows (128 B) per column
        // instead of 4 (32 B).  The x pass's output is contiguous along
        // columns already.
        static_assert(16 * LDB <= KB * LDA && 16 * LDB <= KB * LDB,
                      "a wavefront tile fits each operand buffer");
        double *R = (wave & 1) ? &Bs[0][0] : &As[0][0];
        const int gm = m0 + wrow + lrow;
        const long crow =
            gm < g.M ? (long)(gm % g.M0) * g.sCm0 + (long)(gm / g.M0) * g.sCm1 : 0;
        for (int round = 0; round < 2; ++round) {
            const bool mine = (wave >> 1) == round;
            if (round > 0) __syncthreads();  // the first round's reads are done
            if (mine) {
#pragma unroll
                for (int j = 0; j < NB; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        R[(4 * r + lk) * LDB + j * 16 + lrow] = acc[0][j][r];
            }
            __syncthreads();
            if (mine && gm < g.M) {
#pragma unroll
                for (int i = 0; i < 4 * NB; ++i) {
                    const int nl = lk + 4 * i;
                    const int gn = n0 + nl;
                    if (gn >= nend) continue;
                    double *dst = C + crow + (long)gn * g.sCn;
                    const double v = alpha * R[lrow * LDB + nl];
                    *dst = g.accumulate ? *dst + v : v;
                }
            }
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < MW; ++i) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gm = m0 + wrow + i * 16 + 4 * r + lk;
            if (gm >= g.M) continue;
            const long crow = (long)(gm % g.M0) * g.sCm0 + (long)(gm / g.M0) * g.sCm1;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const int jt = TGRP ? j / CG : 0, jc = TGRP ? j % CG : j;
                if (jt >= tval) continue;
                const int gn = n0 + jc * 16 + lrow;
                if (gn >= nend) continue;
                double *dst = C + (long)jt * g.tC + crow + (long)gn * g.sCn;
                const double v = alpha * acc[i][j][r];
                *dst = g.accumulate ? *dst + v : v;
            }
        }
    }
}

// Fused x and y passes: for one (right-hand side b, t, input z-plane zi) and
// 64 output x points (16 per wavefront),
//   U(xo, yo) = w_t sum_yi [sum_xi Fx_t(xo,xi) in(xi,yi,zi)] Fy_t(yi,yo),
// without the x-pass intermediate V reaching memory.  The input y axis is
// walked in chunks of 16 points.  Per chunk a wavefront first forms the
// 16 x 16 tile V^T(yi, xo) = in^T(yi, xi) Fx_t^T(xi, xo) on the MFMA (A = the
// input chunk from LDS, B = Fx_t from global memory); result register r of
// lane l then holds V(xo = l%16, yi = 4r + l/16), which is exactly the
// A-operand element of k-step r of the y product, so the tile feeds
//   U(xo, yo) += V(xo, yi-chunk) Fy_t(yi-chunk, yo)
// straight from registers (B = the Fy_t chunk from LDS), one accumulator per
// 16-point output y tile.  The k-grouping, operand roles and k order are those
// of the separate x and y passes, and only products with a zero matrix entry
// outside the bands are skipped, so U is bit-identical to theirs.
//
// The input chunk (16 columns of nxi points, contiguous in memory) and the
// Fy_t chunk (16 rows) are staged in LDS, the next chunk's loads issued into
// registers before the current chunk's MFMAs.  NT >= the output y tiles and
// NT >= nxi / 16 (it also sizes the loader's registers).
struct XYF {
    const double *in;   // in_b(xi, yi, zi) at in + b*nin + xi + nxi*(yi + nyi*zi)
    const double *fx;   // Fx_t(xo, xi) at fx + xo + nxo*(xi + nxi*t)
    const double *fy;   // Fy_t(yi, yo) at fy + yi + nyi*(yo + nyo*t)
    double *U;          // U + xo + nxo*(yo + nyo*zi) + Ut*(tl + tc*b)
    const double *w;    // alpha = w[t0 + tl + wS*b]
    long wS;
    int nxi, nyi, nxo, nyo, nzi;
    int t0, tc;
    long nin, Ut;
    const int2 *xband, *yband;
    int nxt, nyc;
    int gx;      // 64-point output x blocks
    int ldy;     // LDS row count of the Fy chunk (16 * output y tiles)
    int remap;
};

template <int NT>
__global__ void __launch_bounds__(kThreads) xy_fused_kernel(XYF g) {
    extern __shared__ double smem[];
    double *Is = smem;                   // input chunk: (xi, yl) at yl*nxi + xi
    double *Ys = smem + 16 * g.nxi;      // Fy chunk: (yl, yo) at yo*17 + yl
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int l16 = lane & 15, lk = lane >> 4;

    // block -> (b, tl, zi, xb), xb fastest; XCD-aware as in block_coords
    unsigned L = blockIdx.x;
    if (g.remap) {
        const unsigned total = gridDim.x;
        const unsigned q = total >> 3, r = total & 7u, x = L & 7u;
        L = x * q + min(x, r) + (L >> 3);
    }
    const int xb = (int)(L % (unsigned)g.gx);
    L /= (unsigned)g.gx;
    const int zi = (int)(L % (unsigned)g.nzi);
    L /= (unsigned)g.nzi;
    const int tl = (int)(L % (unsigned)g.tc);
    const int b = (int)(L / (unsigned)g.tc);
    const int t = g.t0 + tl;

    const double *in = g.in + (long)b * g.nin + (long)zi * g.nxi * g.nyi;
    const double *fx = g.fx + (long)t * g.nxo * g.nxi;
    const double *fy = g.fy + (long)t * g.nyi * g.nyo;
    const int xt = xb * 4 + wave;  // this wavefront's 16-point output x tile
    const bool active = xt < g.nxt;
    const int xo = xt * 16 + l16;
    const bool xo_ok = xo < g.nxo;
    const int2 xr = active ? g.xband[(long)t * g.nxt + xt] : make_int2(0, 0);
    const double *fxl = fx + (xo_ok ? xo : 0);

    // loaders: input chunk element e = tid + 256*s (contiguous in memory),
    // Fy chunk element e: yl = e % 16, yo = e / 16
    const int ichunk = 16 * g.nxi;
    double ri[NT], ry[NT];
    auto load_chunk = [&](int c) {
        const int yc = 16 * c;
        const int ilim = min(16, g.nyi - yc) * g.nxi;  // valid elements
        const double *src = in + (long)yc * g.nxi;
#pragma unroll
        for (int s = 0; s < NT; ++s) {
            const int e = tid + kThreads * s;
            const bool ok = e < ilim;
            const double v = src[ok ? e : 0];
            ri[s] = ok ? v : 0.0;
        }
#pragma unroll
        for (int s = 0; s < NT; ++s) {
            const int e = tid + kThreads * s;
            const int yl = e & 15, yo = e >> 4;
            const bool ok = yo < g.nyo && yc + yl < g.nyi;
            const double v = fy[ok ? (long)(yc + yl) + (long)g.nyi * yo : 0];
            ry[s] = ok ? v : 0.0;
        }
    };
    auto store_chunk = [&]() {
#pragma unroll
        for (int s = 0; s < NT; ++s) {
            const int e = tid + kThreads * s;
            if (e < ichunk) Is[e] = ri[s];
        }
#pragma unroll
        for (int s = 0; s < NT; ++s) {
            const int e = tid + kThreads * s;
            const int yl = e & 15, yo = e >> 4;
            if (yo < g.ldy) Ys[yo * 17 + yl] = ry[s];
        }
    };

    d4 acc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[j] = d4{0.0, 0.0, 0.0, 0.0};

    const int nchunks = (g.nyi + 15) / 16;
    load_chunk(0);
    for (int c = 0; c < nchunks; ++c) {
        store_chunk();
        __syncthreads();
        if (c + 1 < nchunks) load_chunk(c + 1);
        const int2 yr = g.yband[(long)t * g.nyc + c];
        if (active && yr.x < yr.y && xr.x < xr.y) {
            // x part: V^T(yi, xo) over the k-steps of this tile's band
            d4 v = d4{0.0, 0.0, 0.0, 0.0};
            const double *ia = Is + l16 * g.nxi;
            for (int s0 = xr.x; s0 < xr.y; s0 += 8) {
                double fa[8], fb[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int xi = 4 * (s0 + u) + lk;
                    const bool ok = s0 + u < xr.y && xi < g.nxi;
                    const double bv = fxl[ok && xo_ok ? (long)g.nxo * xi : 0];
                    fb[u] = ok && xo_ok ? bv : 0.0;
                    fa[u] = ok ? ia[ok ? xi : 0] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (s0 + u < xr.y) v = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[u], fb[u], v, 0, 0, 0);
            }
            // y part: k-steps with some input y point inside the grid
            const int nks = min(4, (g.nyi - 16 * c + 3) / 4);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                if (s < nks) {
#pragma unroll
                    for (int j = 0; j < NT; ++j)
                        if (j >= yr.x && j < yr.y) {
                            const double bv = Ys[(16 * j + l16) * 17 + 4 * s + lk];
                            acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(v[s], bv, acc[j], 0, 0, 0);
                        }
                }
            }
        }
        __syncthreads();
    }
    if (!active) return;
    // Register r of lane l holds U(xo = xt*16 + 4r + l/16, yo = 16j + l%16).
    const double alpha = g.w[g.t0 + tl + g.wS * b];
    double *U = g.U + (long)zi * g.nxo * g.nyo + g.Ut * ((long)tl + (long)g.tc * b);
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int yo = 16 * j + l16;
        if (yo >= g.nyo) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int x = xt * 16 + 4 * r + lk;
            if (x < g.nxo) U[x + (long)g.nxo * yo] = alpha * acc[j][r];
        }
    }
}

// Fz (nzi, nzo, T) -> Gz((zi + nzi*t), zo)
__global__ void __launch_bounds__(256) relayout_fz_kernel(const double *fz, double *gz, int nzi,
                                                          int nzo, int T) {
    const long total = (long)nzi * nzo * T;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int zi = (int)(e % nzi);
        const long r = e / nzi;
        const int zo = (int)(r % nzo);
        const int t = (int)(r / nzo);
        gz[(long)zi + (long)nzi * t + (long)nzi * T * zo] = fz[e];
    }
}

// Entries at most `cutoff` times their row's largest |entry| are set to
// zero; one thread per row, element (row r, column n) at b[r*rs + n*cs].
__global__ void __launch_bounds__(256) flush_small_kernel(double *b, long rows, int n, long rs,
                                                          long cs, double cutoff) {
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < rows;
         r += (long)gridDim.x * blockDim.x) {
        double big = 0.0;
        for (int c = 0; c < n; ++c) big = fmax(big, fabs(b[r * rs + (long)c * cs]));
        const double thr = cutoff * big;
        for (int c = 0; c < n; ++c) {
            double &v = b[r * rs + (long)c * cs];
            if (v != 0.0 && fabs(v) <= thr) v = 0.0;
        }
    }
}

// Row masks: one thread per B row of a pass; bit c when the row has a nonzero
// entry among output columns [16c, 16c + 16).  Element (row r, column n) of a
// pass's B matrix is at b[r*rs + n*cs] (r local to one t for x and y).
// span[2r], span[2r+1]: first and last nonzero column of the row (n, -1 if
// none).
__global__ void __launch_bounds__(256) row_mask_kernel(const double *b, long rows, int n, long rs,
                                                       long cs, uint64_t *mask, int *span) {
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < rows;
         r += (long)gridDim.x * blockDim.x) {
        uint64_t m = 0;
        int lo = n, hi = -1;
        for (int c = 0; c < n; ++c)
            if (b[r * rs + (long)c * cs] != 0.0) {
                m |= 1ull << (c >> 4);
                if (lo == n) lo = c;
                hi = c;
            }
        mask[r] = m;
        span[2 * r] = lo;
        span[2 * r + 1] = hi;
    }
}

// Column tiles at pitch p: n / p of them, the last one also taking the n % p
// leftover columns.
int pitch_tiles(int n, int p) { return std::max(1, n / p); }

// Pick the column-block count NB (<= 8) that wastes the least MFMA width on
// an output dimension of n points.
int pick_nb(int n) {
    int best = 8;
    double best_cost = 1e30;
    for (int nb = 8; nb >= 1; --nb) {
        const int bn = 16 * nb;
        const int tiles = (n + bn - 1) / bn;
        const double cost = (double)tiles * bn / n + 0.02 * tiles;  // waste + per-tile cost
        if (cost < best_cost - 1e-12) { best_cost = cost; best = nb; }
    }
    return best;
}

constexpr int kMW = 1;  // 64-row tiles: 2-3 waves per SIMD at NB = 7

template <int NB, bool AM, bool BK, int PASS>
void launch_t(const Gemm &g, dim3 grid, hipStream_t s) {
    hipLaunchKernelGGL((gemm_f64_kernel<kMW, NB, AM, BK, PASS>), grid, dim3(kThreads), 0, s, g);
}

// XCD-aware block order (block_coords) for GEMMs whose A operand (per batch
// index) exceeds half of the 256 MB MALL: there the re-reads of A by the
// column tiles go to HBM, and the remap, which runs the tiles sharing an A
// row block together on one XCD, makes the 385^3 apply 99.1 -> 89.8 ms;
// below it the re-reads hit the MALL and the plain order, which streams A
// rows in order, is 3-5 % faster (241^3).  DAGE_APPLY_REMAP=0 / 1 forces
// either order (comparison runs; read at every launch).
int remap_default(double a_bytes) {
    const char *e = std::getenv("DAGE_APPLY_REMAP");
    if (e && e[0] == '0') return 0;
    if (e && e[0] == '1') return 1;
    return a_bytes > 128.0 * (1 << 20) ? 1 : 0;
}

template <bool AM, bool BK, int PASS>
int launch_nb(const Gemm &g0, int nb, int zdim, hipStream_t s) {
    const int BM = 64 * kMW;
    // PASS 4 tiles are 16*nb/4 columns wide (nb/4 groups for each of 4 t)
    const int tw = PASS == 4 ? 16 * std::max(1, nb / 4) : 16 * nb;
    DAGE_REQUIRE(PASS != 4 || (g0.pitch <= 0 && (nb == 4 || nb == 8)), DAGE_ERR_STATE,
                 "t-grouped x pass needs 16-column tiles and 4 or 8 column groups");
    const int ntiles = g0.pitch > 0 ? pitch_tiles(g0.N, g0.pitch) : (g0.N + tw - 1) / tw;
    DAGE_REQUIRE(g0.pitch <= 0 || (ntiles <= 64 && g0.pitch <= 16 * nb &&
                                   g0.N - (ntiles - 1) * g0.pitch <= 16 * nb),
                 DAGE_ERR_STATE, "column pitch %d does not fit %d-column tiles", g0.pitch, 16 * nb);
    Gemm g = g0;
    g.gx = (g.M + BM - 1) / BM;
    g.gy = ntiles;
    g.gz = zdim;
    g.remap = remap_default(8.0 * g.M * (double)(g.Ksplit > 0 ? g.Ksplit : g.K));
    g.batch_inner = PASS == 0 || PASS == 4;  // the x pass's A (the input) is the same for every t
    const long blocks = (long)g.gx * g.gy * g.gz;
    DAGE_REQUIRE(g.Z0 <= 0 || zdim % g.Z0 == 0, DAGE_ERR_STATE,
                 "batch grid %d is not a multiple of %d", zdim, g.Z0);
    DAGE_REQUIRE(blocks > 0 && blocks < (1L << 31), DAGE_ERR_ARG,
                 "GEMM of %ld tiles does not fit one launch", blocks);
    const dim3 grid((unsigned)blocks);
    if constexpr (PASS == 4) {
        if (nb == 4) launch_t<4, AM, BK, PASS>(g, grid, s);
        else launch_t<8, AM, BK, PASS>(g, grid, s);
        DAGE_HIP_CHECK(hipGetLastError());
        return DAGE_OK;
    }
    switch (nb) {
    case 1: launch_t<1, AM, BK, PASS>(g, grid, s); break;
    case 2: launch_t<2, AM, BK, PASS>(g, grid, s); break;
    case 3: launch_t<3, AM, BK, PASS>(g, grid, s); break;
    case 4: launch_t<4, AM, BK, PASS>(g, grid, s); break;
    case 5: launch_t<5, AM, BK, PASS>(g, grid, s); break;
    case 6: launch_t<6, AM, BK, PASS>(g, grid, s); break;
    case 7: launch_t<7, AM, BK, PASS>(g, grid, s); break;
    default: launch_t<8, AM, BK, PASS>(g, grid, s); break;
    }
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

// The three passes have fixed operand layouts: x mode reads k-contiguous A and
// n-contiguous B; y and z modes read m-contiguous A and k-contiguous B.
template <int PASS>
int launch_gemm(dage_ctx ctx, const char *name, const Gemm &g, int zdim, int nb_override = 0) {
    // the k-contiguous A loader steps rows linearly: one-level row index only
    DAGE_REQUIRE((PASS != 0 && PASS != 4) || g.M0 >= g.M, DAGE_ERR_STATE,
                 "x-mode GEMM needs a one-level M");
    const int nb = (nb_override >= 1 && nb_override <= 8) ? nb_override : pick_nb(g.N);
    hipEvent_t ev;
    prof_begin(ctx, name, &ev);
    int rc;
    if constexpr (PASS == 0 || PASS == 4) rc = launch_nb<false, false, PASS>(g, nb, zdim, ctx->stream);
    else rc = launch_nb<true, true, PASS>(g, nb, zdim, ctx->stream);
    prof_end(ctx, name, ev);
    return rc;
}

// Fused x/y pass (xy_fused_kernel): opt-in with DAGE_APPLY_FUSED=1, when the
// bands are built and both the output y axis and the input x axis fit the
// largest register tile.  Measured on MI355X it is 1.5x (97^3) to 4.4x
// (385^3) slower than the separate x and y passes: per chunk each wavefront's
// x product is one dependent MFMA chain fed by global loads of Fx_t, at one
// wavefront per SIMD once the y accumulators pass 128 registers.  Returns
// the kernel's NT (0: separate passes).
constexpr int kMaxNT = 28;

// The x pass with 4 t-points per block (PASS 4) where the cost model picks
// 16- or 32-column tiles; DAGE_APPLY_TGROUP=0 keeps one t per block.
// Measured on the 385^3 sweep's middle t-points: 23.6 ms per launch against
// 24.1 ms for one t per block on 32-column tiles (the A panel of 64 input
// rows feeds 4 column groups instead of 2).
bool tgroup_x() {
    const char *e = std::getenv("DAGE_APPLY_TGROUP");
    return !(e && e[0] == '0');
}

int fused_nt(const ApplyMats &m) {
    if (!m.xband || !m.yband) return 0;
    const char *e = std::getenv("DAGE_APPLY_FUSED");
    if (!e || e[0] != '1') return 0;
    const int need = std::max((m.dout[1] + 15) / 16, (m.din[0] + 15) / 16);
    if (need > kMaxNT) return 0;
    return (need + 3) / 4 * 4;
}

// Largest dynamic LDS of the fused pass: 16 input columns of up to 16*kMaxNT
// points and 16*kMaxNT Fy rows of 17 doubles.
constexpr size_t kFusedMaxLds = 8 * (16 * 16 * kMaxNT + 17 * 16 * kMaxNT);

template <int NT>
int set_xy_attr() {
    DAGE_HIP_CHECK(hipFuncSetAttribute((const void *)xy_fused_kernel<NT>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)kFusedMaxLds));
    return DAGE_OK;
}

template <int NT>
int launch_xy_t(const XYF &g, unsigned blocks, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL(xy_fused_kernel<NT>, dim3(blocks), dim3(kThreads), lds, s, g);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int launch_xy(dage_ctx ctx, int nt, const XYF &g, int B, const char *name) {
    const long blocks = (long)g.gx * g.nzi * g.tc * B;
    DAGE_REQUIRE(blocks > 0 && blocks < (1L << 31), DAGE_ERR_ARG,
                 "fused pass of %ld blocks does not fit one launch", blocks);
    DAGE_REQUIRE(16 * nt >= g.ldy && 16 * nt >= g.nxi, DAGE_ERR_STATE,
                 "fused pass tile %d too small", nt);
    const size_t lds = 8 * (16 * (size_t)g.nxi + 17 * (size_t)g.ldy);
    hipEvent_t ev;
    prof_begin(ctx, name, &ev);
    int rc;
    const unsigned nb = (unsigned)blocks;
    switch (nt) {
    case 4: rc = launch_xy_t<4>(g, nb, lds, ctx->stream); break;
    case 8: rc = launch_xy_t<8>(g, nb, lds, ctx->stream); break;
    case 12: rc = launch_xy_t<12>(g, nb, lds, ctx->stream); break;
    case 16: rc = launch_xy_t<16>(g, nb, lds, ctx->stream); break;
    case 20: rc = launch_xy_t<20>(g, nb, lds, ctx->stream); break;
    case 24: rc = launch_xy_t<24>(g, nb, lds, ctx->stream); break;
    case 28: rc = launch_xy_t<28>(g, nb, lds, ctx->stream); break;
    default: DAGE_REQUIRE(false, DAGE_ERR_STATE, "no fused pass for tile %d", nt);
    }
    prof_end(ctx, name, ev);
    return rc;
}

// Workspace layout: V | U | P (split-K partials); no V with the fused pass
struct Plan {
    int tchunk;
    int splits;
    size_t v_elems, u_elems, p_elems;
};

// cap > 0: the whole workspace must fit cap bytes (the largest t-chunk, then
// the largest K split, that does); tchunk = 0 when nothing fits.
Plan make_plan(const ApplyMats &m, int B = 1, size_t cap = 0) {
    const bool fused = fused_nt(m) > 0;
    const long nyi = m.din[1], nzi = m.din[2];
    const long nxo = m.dout[0], nyo = m.dout[1], nzo = m.dout[2];
    size_t budget = 4ull << 30;  // bytes for V + U per right-hand side
    if (const char *e = std::getenv("DAGE_APPLY_WORKSPACE_MB")) budget = (size_t)atol(e) << 20;
    const size_t total_cap = 64ull << 30;  // and for all B right-hand sides together
    const size_t per_t = 8ull * (size_t)((fused ? 0 : nxo * nyi * nzi) + nxo * nyo * nzi);
    const long M3 = nxo * nyo;
    const int mtiles = (int)((M3 + 64 * kMW - 1) / (64 * kMW));
    const int nb = pick_nb((int)nzo);
    const int ntiles = (int)((nzo + 16 * nb - 1) / (16 * nb));
    auto fill = [&](int tchunk) {
        Plan p;
        p.tchunk = tchunk;
        const long K3 = nzi * p.tchunk;
        int splits = std::max(1, (1024 + mtiles * ntiles * B - 1) / (mtiles * ntiles * B));
        splits = (int)std::min<long>(splits, std::max<long>(1, K3 / 64));
        p.splits = splits;
        p.v_elems = fused ? 0 : (size_t)nxo * nyi * nzi * p.tchunk * B;
        p.u_elems = (size_t)nxo * nyo * nzi * p.tchunk * B;
        p.p_elems = (size_t)M3 * nzo * splits * B;
        return p;
    };
    auto bytes = [](const Plan &p) { return 8 * (p.v_elems + p.u_elems + p.p_elems); };
    if (cap > 0) {
        int tmax = m.ntp;
        if (const char *e = std::getenv("DAGE_APPLY_TCHUNK"))
            if (atoi(e) > 0) tmax = std::min(tmax, atoi(e));
        for (int tc = tmax; tc >= 1; --tc) {
            Plan p = fill(tc);
            while (bytes(p) > cap && p.splits > 1) {
                --p.splits;
                p.p_elems = (size_t)M3 * nzo * p.splits * B;
            }
            if (bytes(p) <= cap) return p;
        }
        Plan none = fill(1);
        none.tchunk = 0;
        return none;
    }
    // DAGE_APPLY_TCHUNK: a fixed t-chunk (comparison runs; the z pass then adds
    // the same partial sums in the same order with or without the fused pass)
    if (const char *e = std::getenv("DAGE_APPLY_TCHUNK"))
        if (atoi(e) > 0) return fill(std::min(m.ntp, atoi(e)));
    return fill((int)std::max<size_t>(
        1, std::min<size_t>({(size_t)m.ntp, budget / per_t, total_cap / (per_t * (size_t)B)})));
}

// out[b*len + i] = sum_s part[(b*splits + s)*len + i] + coda*in[b*len + i]
__global__ void __launch_bounds__(256) splitk_reduce_batch_kernel(const double *part, int splits,
                                                                  long len, int B,
                                                                  const double *in, double coda,
                                                                  double *out) {
    const long total = len * B;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const long b = e / len, i = e - b * len;
        double s = 0.0;
        for (int k = 0; k < splits; ++k) s += part[(b * splits + k) * len + i];
        if (in) s += coda * in[e];
        out[e] = s;
    }
}

}  // namespace

namespace {

// Entries of the 1D matrices at most kFlushCutoff times the largest |entry|
// of their row (one input point, one t) are stored as zeros, so that the
// passes skip the K panels they fill.  The operator matrices' far entries are
// Gaussian tails exp(-t^2 d^2) that reach exact zero only at underflow;
// between 1e-300 and 1e-20 of the row maximum they fill about a third of the
// 16x16 blocks of a 385-point axis at T = 40.  Dropping them changes an
// output of a pass by at most n * kFlushCutoff * max|F_t| * max|g|, about
// 1e-4 of the rounding-error bound of the dense FP64 sum itself
// (n * 2^-53 * max|F_t| * max|g|, n <= 1024).  Every tiling then skips the
// same entries, so results do not depend on it.  DAGE_FLUSH_CUTOFF
// overrides the cutoff (0: keep every nonzero entry).
constexpr double kFlushCutoff = 1e-20;

double flush_cutoff() {
    if (const char *e = std::getenv("DAGE_FLUSH_CUTOFF")) return std::max(0.0, atof(e));
    return kFlushCutoff;
}

}  // namespace

int flush_matrices(dage_operator op) {
    const double cut = flush_cutoff();
    if (cut == 0.0) return DAGE_OK;
    const long nxi = op->din[0], nyi = op->din[1], nzi = op->din[2];
    const long nxo = op->dout[0], nyo = op->dout[1], nzo = op->dout[2];
    const long T = op->ntp;
    hipStream_t s = op->ctx->stream;
    auto grid = [](long r) { return (int)std::min<long>((r + 255) / 256, 4096); };
    // Fx (nxo, nxi, T): row (xi, t) is contiguous over xo
    hipLaunchKernelGGL(flush_small_kernel, dim3(grid(T * nxi)), dim3(256), 0, s, op->d_fx,
                       T * nxi, (int)nxo, nxo, 1L, cut);
    DAGE_HIP_CHECK(hipGetLastError());
    // Fy (nyi, nyo, T), Fz (nzi, nzo, T): row (i, t) at i + ni*no*t, stride ni
    for (long t = 0; t < T; ++t) {
        hipLaunchKernelGGL(flush_small_kernel, dim3(grid(nyi)), dim3(256), 0, s,
                           op->d_fy + t * nyi * nyo, nyi, (int)nyo, 1L, nyi, cut);
        DAGE_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(flush_small_kernel, dim3(grid(nzi)), dim3(256), 0, s,
                           op->d_fz + t * nzi * nzo, nzi, (int)nzo, 1L, nzi, cut);
        DAGE_HIP_CHECK(hipGetLastError());
    }
    return DAGE_OK;
}

int relayout_fz(dage_operator op) {
    const int nzi = op->din[2], nzo = op->dout[2], T = op->ntp;
    const long total = (long)nzi * nzo * T;
    const int grid = (int)std::min<long>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(relayout_fz_kernel, dim3(grid), dim3(256), 0, op->ctx->stream, op->d_fz,
                       op->d_gz, nzi, nzo, T);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

namespace {

// Relative cost of one K panel of one block: NB column groups of MFMA work
// plus a fixed share (operand loads, barrier) worth kPanelCost groups.
constexpr double kPanelCost = 1.0;

// Estimated work of a pass whose B rows (nseg segments of klen rows, panels
// aligned to each segment) have the given masks, output axis n points, at
// column-block count nb; all panels count when dense.
double pass_cost(const std::vector<uint64_t> &m, long nseg, long klen, int n, int nb, bool dense) {
    double cost = 0.0;
    for (int n0 = 0; n0 < n; n0 += 16 * nb) {
        const uint64_t cols = ((1ull << nb) - 1ull) << (n0 / 16);
        long need = 0;
        for (long sgi = 0; sgi < nseg; ++sgi)
            for (long k0 = 0; k0 < klen; k0 += 16) {
                uint64_t any = 0;
                for (long k = k0; k < std::min(klen, k0 + 16); ++k) any |= m[sgi * klen + k];
                need += dense || (any & cols) != 0;
            }
        cost += (double)need * (nb + kPanelCost);
    }
    return cost;
}

// Column-block count for a pass (least estimated work) and whether the
// masks pay: listing a block's panels costs one dependent load before its
// first panel, so they are used only when they cut the work by 7 % or more.
int pick_nb_masked(const std::vector<uint64_t> &m, long nseg, long klen, int n, bool *use) {
    int best = 1;
    double best_cost = 1e300, best_dense = 1e300;
    for (int nb = 8; nb >= 1; --nb) {
        const double c = pass_cost(m, nseg, klen, n, nb, false);
        if (c < best_cost * (1.0 - 1e-9)) { best_cost = c; best = nb; }
        best_dense = std::min(best_dense, pass_cost(m, nseg, klen, n, nb, true));
    }
    *use = best_cost < 0.93 * best_dense;
    return *use ? best : pick_nb(n);
}

// The same estimate for tiles at column pitch P (pitch_tiles), from each
// row's first/last nonzero column: a panel is needed by a tile when one of
// its rows' nonzero span meets the tile's columns.
double pitch_cost(const std::vector<int> &span, long nseg, long klen, int n, int nb, int P) {
    const int nt = pitch_tiles(n, P);
    double cost = 0.0;
    for (int j = 0; j < nt; ++j) {
        const int a = j * P, b = (j + 1 < nt) ? a + P : n;  // columns [a, b)
        long need = 0;
        for (long sgi = 0; sgi < nseg; ++sgi)
            for (long k0 = 0; k0 < klen; k0 += 16) {
                bool any = false;
                for (long k = k0; k < std::min(klen, k0 + 16) && !any; ++k) {
                    const long r = sgi * klen + k;
                    any = span[2 * r] < b && span[2 * r + 1] >= a;
                }
                need += any;
            }
        cost += (double)need * (nb + kPanelCost);
    }
    return cost;
}

// Best (column blocks, pitch) over pitches within 15 columns below each
// 16*nb width whose last tile still fits; pitch 0 when none is valid.
void pick_pitch(const std::vector<int> &span, long nseg, long klen, int n, int *nb_out,
                int *p_out, double *cost_out) {
    *nb_out = 0;
    *p_out = 0;
    *cost_out = 1e300;
    for (int nb = 8; nb >= 1; --nb)
        for (int P = 16 * nb; P >= std::max(1, 16 * nb - 15); --P) {
            const int nt = pitch_tiles(n, P);
            if (nt > 64 || n - (nt - 1) * P > 16 * nb) continue;
            const double c = pitch_cost(span, nseg, klen, n, nb, P);
            if (c < *cost_out * (1.0 - 1e-9)) {
                *cost_out = c;
                *nb_out = nb;
                *p_out = P;
            }
        }
}

}  // namespace

int build_masks(dage_operator op) {
    const long nxi = op->din[0], nyi = op->din[1], nzi = op->din[2];
    const long nxo = op->dout[0], nyo = op->dout[1], nzo = op->dout[2];
    const long T = op->ntp;
    const long rx = T * nxi, ry = T * nyi, rz = nzi * T;
    const long rows = rx + ry + rz;
    if (!op->d_mask) DAGE_HIP_CHECK(hipMalloc((void **)&op->d_mask, 8 * rows));
    if (!op->d_tmask) DAGE_HIP_CHECK(hipMalloc((void **)&op->d_tmask, 8 * rows));
    int *d_span = nullptr;
    DAGE_TRY(ctx_scratch(op->ctx, 8 * (size_t)rows, (void **)&d_span));
    hipStream_t s = op->ctx->stream;
    auto grid = [](long r) { return (int)std::min<long>((r + 255) / 256, 4096); };
    // x: Fx (nxo, nxi, T), row (t, xi) = t*nxi + xi, column xo
    hipLaunchKernelGGL(row_mask_kernel, dim3(grid(rx)), dim3(256), 0, s, op->d_fx, rx, (int)nxo,
                       nxo, 1L, op->d_mask, d_span);
    DAGE_HIP_CHECK(hipGetLastError());
    // y: Fy (nyi, nyo, T); row t*nyi + yi sits at yi + nyi*nyo*t, column stride nyi
    for (long t = 0; t < T; ++t) {
        hipLaunchKernelGGL(row_mask_kernel, dim3(grid(nyi)), dim3(256), 0, s,
                           op->d_fy + t * nyi * nyo, nyi, (int)nyo, 1L, nyi,
                           op->d_mask + rx + t * nyi, d_span + 2 * (rx + t * nyi));
        DAGE_HIP_CHECK(hipGetLastError());
    }
    // z: Gz ((zi + nzi*t), zo), column stride nzi*T
    hipLaunchKernelGGL(row_mask_kernel, dim3(grid(rz)), dim3(256), 0, s, op->d_gz, rz, (int)nzo,
                       1L, rz, op->d_mask + rx + ry, d_span + 2 * (rx + ry));
    DAGE_HIP_CHECK(hipGetLastError());
    std::vector<uint64_t> h(rows);
    std::vector<int> hs(2 * rows);
    DAGE_HIP_CHECK(hipMemcpyAsync(h.data(), op->d_mask, 8 * h.size(), hipMemcpyDeviceToHost, s));
    DAGE_HIP_CHECK(hipMemcpyAsync(hs.data(), d_span, 4 * hs.size(), hipMemcpyDeviceToHost, s));
    DAGE_HIP_CHECK(hipStreamSynchronize(s));
    // banded large-t passes and low-rank small-t points (both need the
    // spans and the final matrices)
    DAGE_TRY(build_bands(op, hs));
    DAGE_HIP_CHECK(hipStreamSynchronize(s));
    DAGE_TRY(build_lowrank(op));
    // The tiles are chosen for the t-points the MFMA passes run at one
    // right-hand side: between the low-rank head and the banded tail when
    // those paths are taken at that size.
    int tlo = 0, thi = (int)T;
    {
        const ApplyMats m1 = operator_mats(op);
        if (lowrank_pays(m1, 1)) tlo = op->lr.t1;
        if (band_pays(m1, 1)) thi = op->band_t0;
        if (tlo >= thi) tlo = 0, thi = (int)T;
    }
    const long n[3] = {nxo, nyo, nzo};
    const long klen[3] = {nxi, nyi, rz}, nseg[3] = {T, T, 1}, off[3] = {0, rx, rx + ry};
    // the rows of those t-points: t*klen.. for x and y, (zi + nzi*t) for z
    const long rlo[3] = {tlo * nxi, tlo * nyi, nzi * tlo}, rhi[3] = {thi * nxi, thi * nyi, nzi * thi};
    const long sseg[3] = {thi - tlo, thi - tlo, 1};
    const long slen[3] = {nxi, nyi, nzi * (thi - tlo)};
    // tiles for the rows [r0, r1) of pass p: column blocks, whether masks
    // pay, and a pitch tiling that follows the zero blocks when it cuts the
    // dense estimate by 7 % and beats the 16-column grid by 5 % (pitch 0: none)
    struct Choice {
        int nb = 0, pitch = 0, nbp = 0;
        bool use = false;
    };
    auto choose = [&](int p, long r0, long r1, long ns, long kl) {
        Choice c;
        std::vector<uint64_t> mp(h.begin() + off[p] + r0, h.begin() + off[p] + r1);
        c.nb = pick_nb_masked(mp, ns, kl, (int)n[p], &c.use);
        std::vector<int> sp(hs.begin() + 2 * (off[p] + r0), hs.begin() + 2 * (off[p] + r1));
        int nbq = 0, P = 0;
        double pc = 0.0;
        pick_pitch(sp, ns, kl, (int)n[p], &nbq, &P, &pc);
        double dense = 1e300;
        for (int nb = 1; nb <= 8; ++nb)
            dense = std::min(dense, pass_cost(mp, ns, kl, (int)n[p], nb, true));
        const double gc = c.use ? pass_cost(mp, ns, kl, (int)n[p], c.nb, false) : dense;
        if (P > 0 && pc < 0.93 * dense && pc < 0.95 * gc) {
            c.pitch = P;
            c.nbp = nbq;
        }
        return c;
    };
    std::vector<uint64_t> ht(rows, 0);
    for (int p = 0; p < 3; ++p) {
        op->nb_pick[p] = 0;
        op->mask_used[p] = false;
        op->pitch[p] = 0;
        op->nbp[p] = 0;
        op->mix_nb_pick[p] = 0;
        op->mix_mask_used[p] = false;
        op->mix_pitch[p] = 0;
        op->mix_nbp[p] = 0;
        if (n[p] > 1024) continue;
        // FP64 passes: their t-range; split-BF16 passes: every t (the
        // pitch only where it equals the FP64 one: one tile-mask array)
        const Choice cf = choose(p, rlo[p], rhi[p], sseg[p], slen[p]);
        const Choice cm = choose(p, 0, nseg[p] * klen[p], nseg[p], klen[p]);
        op->nb_pick[p] = cf.nb;
        op->mask_used[p] = cf.use;
        op->mix_nb_pick[p] = cm.nb;
        op->mix_mask_used[p] = cm.use;
        if (cm.pitch > 0 && cm.pitch == cf.pitch) {
            op->mix_pitch[p] = cm.pitch;
            op->mix_nbp[p] = cm.nbp;
        }
        if (cf.pitch > 0) {
            const int P = cf.pitch;
            op->pitch[p] = P;
            op->nbp[p] = cf.nbp;
            const int nt = pitch_tiles((int)n[p], P);
            const int *spf = hs.data() + 2 * off[p];  // every row's span (all t)
            for (long r = 0; r < nseg[p] * klen[p]; ++r) {
                const int lo = spf[2 * r], hi = spf[2 * r + 1];
                uint64_t m = 0;
                for (int j = 0; j < nt; ++j) {
                    const int a = j * P, b = (j + 1 < nt) ? a + P : (int)n[p];
                    if (lo < b && hi >= a) m |= 1ull << j;
                }
                ht[off[p] + r] = m;
            }
        }
    }
    DAGE_HIP_CHECK(hipMemcpyAsync(op->d_tmask, ht.data(), 8 * ht.size(), hipMemcpyHostToDevice, s));
    // Bands of the fused x/y pass from the row spans: output x tile j (16
    // points) takes input points xi whose Fx_t row reaches it, input y chunk
    // c (16 points) reaches the output y tiles its Fy_t rows span.  Zero
    // entries inside a band are computed (they add exact zeros).
    const int nxt = (int)((nxo + 15) / 16), nyc = (int)((nyi + 15) / 16);
    std::vector<int2> xb((size_t)T * nxt), yb((size_t)T * nyc);
    for (long t = 0; t < T; ++t) {
        for (int j = 0; j < nxt; ++j) {
            const int a = 16 * j, b = (int)std::min<long>(a + 16, nxo);
            long lo = -1, hi = -1;
            for (long xi = 0; xi < nxi; ++xi) {
                const long r = t * nxi + xi;
                if (hs[2 * r] < b && hs[2 * r + 1] >= a) {
                    if (lo < 0) lo = xi;
                    hi = xi;
                }
            }
            xb[t * nxt + j] = lo < 0 ? make_int2(0, 0) : make_int2((int)(lo / 4), (int)(hi / 4 + 1));
        }
        for (int c = 0; c < nyc; ++c) {
            int lo = 1 << 30, hi = -1;
            for (long yi = 16L * c; yi < std::min<long>(16L * c + 16, nyi); ++yi) {
                const long r = rx + t * nyi + yi;
                if (hs[2 * r + 1] >= 0) {
                    lo = std::min(lo, hs[2 * r] / 16);
                    hi = std::max(hi, hs[2 * r + 1] / 16);
                }
            }
            yb[t * nyc + c] = hi < 0 ? make_int2(0, 0) : make_int2(lo, hi + 1);
        }
    }
    if (op->nxt != nxt || op->nyc != nyc || !op->d_xband) {
        if (op->d_xband) DAGE_HIP_CHECK(hipFree(op->d_xband));
        if (op->d_yband) DAGE_HIP_CHECK(hipFree(op->d_yband));
        op->d_xband = op->d_yband = nullptr;
        DAGE_HIP_CHECK(hipMalloc((void **)&op->d_xband, sizeof(int2) * xb.size()));
        DAGE_HIP_CHECK(hipMalloc((void **)&op->d_yband, sizeof(int2) * yb.size()));
        op->nxt = nxt;
        op->nyc = nyc;
    }
    DAGE_HIP_CHECK(hipMemcpyAsync(op->d_xband, xb.data(), sizeof(int2) * xb.size(),
                                  hipMemcpyHostToDevice, s));
    DAGE_HIP_CHECK(hipMemcpyAsync(op->d_yband, yb.data(), sizeof(int2) * yb.size(),
                                  hipMemcpyHostToDevice, s));
    DAGE_HIP_CHECK(hipStreamSynchronize(s));
    // tuning knob "x,y,z": fixed column-block counts, masks on wherever defined
    if (const char *e = std::getenv("DAGE_APPLY_NB")) {
        int v[3] = {0, 0, 0};
        if (std::sscanf(e, "%d,%d,%d", &v[0], &v[1], &v[2]) == 3)
            for (int p = 0; p < 3; ++p)
                if (v[p] >= 1 && v[p] <= 8) {
                    op->nb_pick[p] = op->mix_nb_pick[p] = v[p];
                    op->mask_used[p] = op->mix_mask_used[p] = n[p] <= 1024;
                    op->pitch[p] = op->mix_pitch[p] = 0;
                }
    }
    return DAGE_OK;
}

ApplyMats operator_mats(dage_operator op) {
    ApplyMats m;
    m.fx = op->d_fx;
    m.fy = op->d_fy;
    m.gz = op->d_gz;
    m.w = op->d_w;
    for (int i = 0; i < 3; ++i) {
        m.din[i] = op->din[i];
        m.dout[i] = op->dout[i];
    }
    m.ntp = op->ntp;
    m.xband = op->d_xband;
    m.yband = op->d_yband;
    m.nxt = op->nxt;
    m.nyc = op->nyc;
    m.band_t0 = op->ntp;
    if (op->d_band) {
        m.band_t0 = op->band_t0;
        m.band_y = op->d_band;
        m.band_z = op->d_band + op->band_off[0];
        m.band_x = op->d_band + op->band_off[1];
        m.band_h = op->band_h.data();
        m.band_mode = op->band_mode;
    }
    if (op->d_lr && op->lr.t1 > 0) m.lr = &op->lr;
    if (op->d_mask) {
        const long rx = (long)op->ntp * op->din[0], ry = (long)op->ntp * op->din[1];
        const uint64_t *base[3] = {op->d_mask, op->d_mask + rx, op->d_mask + rx + ry};
        for (int p = 0; p < 3; ++p) {
            m.mask[p] = op->mask_used[p] ? base[p] : nullptr;
            m.nb_pick[p] = op->nb_pick[p];
            if (op->pitch[p] > 0) {
                m.tmask[p] = op->d_tmask + (base[p] - op->d_mask);
                m.pitch[p] = op->pitch[p];
                m.nbp[p] = op->nbp[p];
            }
        }
    }
    return m;
}

int prepare_fused_kernels() {
    DAGE_TRY(set_xy_attr<4>());
    DAGE_TRY(set_xy_attr<8>());
    DAGE_TRY(set_xy_attr<12>());
    DAGE_TRY(set_xy_attr<16>());
    DAGE_TRY(set_xy_attr<20>());
    DAGE_TRY(set_xy_attr<24>());
    DAGE_TRY(set_xy_attr<28>());
    return DAGE_OK;
}

int apply_f64_workspace(dage_operator op, size_t *bytes) {
    const Plan p = make_plan(operator_mats(op));
    *bytes = 8 * (p.v_elems + p.u_elems + p.p_elems);
    return DAGE_OK;
}

int apply_f64_batch_workspace(const ApplyMats &m, int B, size_t *bytes) {
    const Plan p = make_plan(m, B);
    *bytes = 8 * (p.v_elems + p.u_elems + p.p_elems);
    return DAGE_OK;
}

int apply_f64_batch(dage_ctx ctx, const ApplyMats &m, const double *in, double *out,
                    const BatchSpec &bs, Workspace *ws) {
    const long nxi = m.din[0], nyi = m.din[1], nzi = m.din[2];
    const long nxo = m.dout[0], nyo = m.dout[1], nzo = m.dout[2];
    const int T = m.ntp, B = bs.B;
    const Plan plan = make_plan(m, B, bs.ws_cap);
    DAGE_REQUIRE(plan.tchunk > 0, DAGE_ERR_ARG,
                 "the apply needs more than the %zu bytes of workspace it may use (the "
                 "intermediates of one t-point do not fit)", bs.ws_cap);
    const size_t need = 8 * (plan.v_elems + plan.u_elems + plan.p_elems);
    if (need > ws->bytes) {
        DAGE_REQUIRE(!bs.no_alloc, DAGE_ERR_STATE, "workspace of %zu bytes not reserved", need);
        if (ws->p) DAGE_HIP_CHECK(hipFree(ws->p));
        ws->p = nullptr;
        ws->bytes = 0;
        DAGE_HIP_CHECK(hipMalloc((void **)&ws->p, need));
        ws->bytes = need;
    }
    double *V = ws->p;
    double *U = V + plan.v_elems;
    double *P = U + plan.u_elems;
    const double *W = bs.w ? bs.w : m.w;
    const long wS = bs.w ? T : 0;  // weights row per right-hand side
    const long nin = nxi * nyi * nzi, nout = nxo * nyo * nzo;
    const long M3 = nxo * nyo;
    const int tcmax = plan.tchunk;
    const long Vt = nxo * nyi * nzi, Ut = nxo * nyo * nzi;  // per (t, rhs)
    auto nb_for = [&](int p) { return bs.nb[p] > 0 ? bs.nb[p] : m.nb_pick[p]; };
    // a caller's column-block count keeps the 16-column tiles; otherwise a
    // pass with a pitch tiling uses it with its tile masks
    auto use_pitch = [&](int p) { return bs.nb[p] <= 0 && m.pitch[p] > 0; };
    auto set_tiles = [&](Gemm &g, int p, long row0, long mS, int *nb) {
        if (use_pitch(p)) {
            g.pitch = m.pitch[p];
            g.mask = m.tmask[p] + row0;
            *nb = m.nbp[p];
        } else {
            if (m.mask[p]) g.mask = m.mask[p] + row0;
            *nb = nb_for(p);
        }
        if (g.mask) g.mS = mS;
    };

    // One right-hand side, no K split and no coda: the z pass accumulates
    // straight into out (no partials, no reduction), unless out overlaps in,
    // which later t-chunks still read.
    const bool direct = B == 1 && plan.splits == 1 && bs.coda == 0.0 &&
                        (out + nout <= in || in + nin <= out);
    const int nt = fused_nt(m);
    // event names (profiling): batched right-hand sides with their own
    // weights (the batch graph) apart from single applies
    const bool bat = bs.w != nullptr;
    const char *nm_xy = bat ? "batch_xy_fused_kernel" : "xy_fused_kernel";
    const char *nm_x = bat ? "batch_f64_xmode" : "gemm_f64_xmode";
    const char *nm_y = bat ? "batch_f64_ymode" : "gemm_f64_ymode";
    const char *nm_z = bat ? "batch_f64_zmode" : "gemm_f64_zmode";
    const char *nm_r = bat ? "batch_splitk_reduce" : "splitk_reduce_kernel";
    const char *nm_bx = bat ? "batch_band_xmode" : "band_xmode";
    const char *nm_bxy = bat ? "batch_band_xymode" : "band_xymode";
    const char *nm_by = bat ? "batch_band_ymode" : "band_ymode";
    const char *nm_bz = bat ? "batch_band_zmode" : "band_zmode";
    // t-points from tb on (narrow bands) take the banded passes; the MFMA
    // chunks stop there
    const int tb = (nt == 0 && band_pays(m, B)) ? m.band_t0 : T;
    // the low-rank t-points (the smallest t) first, when their intermediates
    // fit the passes' workspace: partial 0 (or out) is overwritten, the other
    // partials cleared, and the chunks below accumulate
    int tlr = 0;
    if (nt == 0 && lowrank_pays(m, B)) {
        size_t w1 = 0, w2 = 0;
        lowrank_workspace(m, B, &w1, &w2);
        if (w1 <= plan.v_elems && w2 <= plan.u_elems) {
            const long slab = M3 * nzo;
            if (!direct)
                for (int b = 0; b < B && plan.splits > 1; ++b)
                    DAGE_TRY(fill_zero(ctx, P + ((size_t)b * plan.splits + 1) * slab,
                                       (size_t)(plan.splits - 1) * slab));
            DAGE_TRY(lowrank_apply(ctx, m, in, direct ? out : P, (long)plan.splits * slab, B, W,
                                   wS, V, U, bat ? "batch_lowrank" : "lowrank"));
            tlr = m.lr->t1;
        }
    }
    for (int t0 = tlr; t0 < T;) {
        const bool banded = t0 >= tb;
        const int tc = std::min(tcmax, (banded ? T : tb) - t0);
        if (nt > 0) {
            // passes 1 and 2 fused: U_b((xo,yo),(zi,t)) straight from in_b
            XYF f{};
            f.in = in;
            f.fx = m.fx;
            f.fy = m.fy;
            f.U = U;
            f.w = W;
            f.wS = wS;
            f.nxi = (int)nxi; f.nyi = (int)nyi; f.nxo = (int)nxo; f.nyo = (int)nyo;
            f.nzi = (int)nzi;
            f.t0 = t0; f.tc = tc;
            f.nin = nin; f.Ut = Ut;
            f.xband = m.xband; f.yband = m.yband;
            f.nxt = m.nxt; f.nyc = m.nyc;
            f.gx = (int)((nxo + 63) / 64);
            f.ldy = 16 * (int)((nyo + 15) / 16);
            f.remap = remap_default(8.0 * nxi * nyi * nzi);
            DAGE_TRY(launch_xy(ctx, nt, f, B, nm_xy));
        } else {
        // pass 1, batch (t, b): V_{b,t}((yi,zi), xo) = in_b((yi,zi), xi) . Fx_t^T(xi, xo)
        Gemm g1{};
        g1.A = in;
        g1.B = m.fx + (long)t0 * nxo * nxi;
        g1.C = V;
        g1.M = (int)(nyi * nzi); g1.N = (int)nxo; g1.K = (int)nxi;
        g1.M0 = g1.M; g1.sAm0 = nxi; g1.sAm1 = 0; g1.sAk = 1;
        g1.sBk = nxo; g1.sBn = 1;
        g1.sCm0 = nxo; g1.sCm1 = 0; g1.sCn = 1;
        g1.bA = 0; g1.bB = nxo * nxi; g1.bC = Vt;
        g1.Z0 = tc; g1.bA1 = nin; g1.bB1 = 0; g1.bC1 = Vt * tc;
        int nb1;
        set_tiles(g1, 0, (long)t0 * nxi, nxi, &nb1);
        if (banded && band_xy_fused()) {
            DAGE_TRY(band_pass_xy(ctx, m, in, U, W, wS, t0, tc, B, nm_bxy));
        } else if (banded) {
            DAGE_TRY(band_pass_x(ctx, m, in, V, t0, tc, B, nm_bx));
        } else if (tgroup_x() && tc >= 2 && g1.pitch <= 0 && (nb1 == 1 || nb1 == 2)) {
            // 4 t-points per block share each loaded input panel, on 16-column
            // tiles (4 column groups: 5 wavefronts per SIMD; 32-column tiles,
            // 8 groups at 3 per SIMD, were slower)
            Gemm gt = g1;
            gt.Z0 = (tc + 3) / 4;
            gt.bB = 4L * nxo * nxi;
            gt.bC = 4L * Vt;
            gt.tB = (long)nxo * nxi;
            gt.tC = Vt;
            gt.tcount = tc;
            DAGE_TRY(launch_gemm<4>(ctx, nm_x, gt, gt.Z0 * B, 4));
        } else {
            DAGE_TRY(launch_gemm<0>(ctx, nm_x, g1, tc * B, nb1));
        }

        if (banded) {
            if (!band_xy_fused()) DAGE_TRY(band_pass_y(ctx, m, V, U, W, wS, t0, tc, B, nm_by));
        } else {
        // pass 2, batch (t, b): U_b((xo,yo),(zi,t)) = w_{b,t} V_{b,t}((xo,zi), yi) . Fy_t(yi, yo)
        Gemm g2{};
        g2.A = V;
        g2.B = m.fy + (long)t0 * nyi * nyo;
        g2.C = U;
        g2.M = (int)(nxo * nzi); g2.N = (int)nyo; g2.K = (int)nyi;
        g2.M0 = (int)nxo; g2.sAm0 = 1; g2.sAm1 = nxo * nyi; g2.sAk = nxo;
        g2.sBk = 1; g2.sBn = nyi;
        g2.sCm0 = 1; g2.sCm1 = nxo * nyo; g2.sCn = nxo;
        g2.bA = Vt; g2.bB = nyi * nyo; g2.bC = Ut;
        g2.Z0 = tc; g2.bA1 = Vt * tc; g2.bB1 = 0; g2.bC1 = Ut * tc;
        g2.alpha = W + t0;
        g2.aS1 = wS;
        int nb2;
        set_tiles(g2, 1, (long)t0 * nyi, nyi, &nb2);
        DAGE_TRY(launch_gemm<1>(ctx, nm_y, g2, tc * B, nb2));
        }
        }

        if (banded) {
            // into partial 0 of each right-hand side (the others are zero when
            // no MFMA chunk ran before), or straight into out
            const long slab = M3 * nzo;
            if (t0 == 0 && !direct)
                for (int b = 0; b < B && plan.splits > 1; ++b)
                    DAGE_TRY(fill_zero(ctx, P + ((size_t)b * plan.splits + 1) * slab,
                                       (size_t)(plan.splits - 1) * slab));
            DAGE_TRY(band_pass_z(ctx, m, U, direct ? out : P, (long)plan.splits * slab, t0 > 0,
                                 t0, tc, B, nm_bz));
            t0 += tc;
            continue;
        }

        // pass 3, batch (split, b): P_{b,s}((xo,yo), zo) (+)= U_b((xo,yo), k) . Gz(k, zo)
        const int K3 = (int)(nzi * tc);
        const int ks = (K3 + plan.splits - 1) / plan.splits;
        const int splits = (K3 + ks - 1) / ks;
        Gemm g3{};
        g3.A = U;
        g3.B = m.gz + (long)nzi * t0;
        g3.C = direct ? out : P;
        g3.M = (int)M3; g3.N = (int)nzo; g3.K = K3;
        g3.M0 = g3.M; g3.sAm0 = 1; g3.sAm1 = 0; g3.sAk = M3;
        g3.sBk = 1; g3.sBn = (long)nzi * T;
        g3.sCm0 = 1; g3.sCm1 = 0; g3.sCn = M3;
        g3.bC = M3 * nzo;
        g3.Z0 = splits; g3.bA1 = Ut * tc; g3.bB1 = 0; g3.bC1 = M3 * nzo * plan.splits;
        g3.Ksplit = ks;
        g3.accumulate = (t0 > 0);
        int nb3;
        set_tiles(g3, 2, (long)nzi * t0, 0, &nb3);
        if (t0 == 0 && splits < plan.splits)
            for (int b = 0; b < B; ++b)
                DAGE_TRY(fill_zero(ctx, P + ((size_t)b * plan.splits + splits) * M3 * nzo,
                                   (size_t)(plan.splits - splits) * M3 * nzo));
        DAGE_TRY(launch_gemm<2>(ctx, nm_z, g3, splits * B, nb3));
        t0 += tc;
    }
    if (direct) return DAGE_OK;
    const long len = M3 * nzo;
    const long total = len * B;
    const int grid = (int)std::min<long>((total + 255) / 256, 65536);
    hipEvent_t ev;
    prof_begin(ctx, nm_r, &ev);
    hipLaunchKernelGGL(splitk_reduce_batch_kernel, dim3(grid), dim3(256), 0, ctx->stream, P,
                       plan.splits, len, B, bs.coda != 0.0 ? in : nullptr, bs.coda, out);
    DAGE_HIP_CHECK(hipGetLastError());
    prof_end(ctx, nm_r, ev);
    return DAGE_OK;
}

int gemm_colmajor(dage_ctx ctx, const char *name, int M, int N, int K, const double *A, long lda,
                  const double *B, long ldb, double *C, int splits, double *part) {
    Gemm g{};
    g.A = A;
    g.B = B;
    g.M = M; g.N = N; g.K = K;
    g.M0 = M; g.sAm0 = 1; g.sAm1 = 0; g.sAk = lda;
    g.sBk = 1; g.sBn = ldb;
    g.sCm0 = 1; g.sCm1 = 0; g.sCn = M;
    if (splits <= 1) {
        g.C = C;
        return launch_gemm<3>(ctx, name, g, 1);
    }
    const int ks = ((K + splits - 1) / splits + kBK - 1) / kBK * kBK;  // whole K panels
    const int nsplit = (K + ks - 1) / ks;
    g.C = part;
    g.bC = (long)M * N;
    g.Ksplit = ks;
    DAGE_TRY(launch_gemm<3>(ctx, name, g, nsplit));
    const long len = (long)M * N;
    const int grid = (int)std::min<long>((len + 255) / 256, 65536);
    hipLaunchKernelGGL(splitk_reduce_batch_kernel, dim3(grid), dim3(256), 0, ctx->stream, part,
                       nsplit, len, 1, (const double *)nullptr, 0.0, C);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int gemm_colmajor_general(dage_ctx ctx, const char *name, const ColGemm &c, int batch) {
    Gemm g{};
    g.A = c.A; g.B = c.B; g.C = c.C;
    g.M = c.M; g.N = c.N; g.K = c.K;
    g.M0 = c.M; g.sAm0 = 1; g.sAm1 = 0; g.sAk = c.sAk;
    g.sBk = 1; g.sBn = c.sBn;
    g.sCm0 = 1; g.sCm1 = 0; g.sCn = c.sCn;
    g.bA = c.bA; g.bB = c.bB; g.bC = c.bC;
    g.Z0 = c.Z0; g.bA1 = c.bA1; g.bB1 = c.bB1; g.bC1 = c.bC1;
    g.alpha = c.alpha; g.aS1 = c.aS1;
    g.accumulate = c.accumulate;
    return launch_gemm<3>(ctx, name, g, batch);
}

int gemm_colmajor_batched(dage_ctx ctx, const char *name, int M, int N, int K, const double *A,
                          long lda, long strideA, const double *B, long ldb, long strideB,
                          double *C, long strideC, int batch) {
    Gemm g{};
    g.A = A;
    g.B = B;
    g.C = C;
    g.M = M; g.N = N; g.K = K;
    g.M0 = M; g.sAm0 = 1; g.sAm1 = 0; g.sAk = lda;
    g.sBk = 1; g.sBn = ldb;
    g.sCm0 = 1; g.sCm1 = 0; g.sCn = M;
    g.bA = strideA; g.bB = strideB; g.bC = strideC;
    return launch_gemm<3>(ctx, name, g, batch);
}

int apply_f64(dage_operator op, const double *in, double *out) {
    BatchSpec bs;
    const bool coda = (op->coda != 0.0) && !(op->flags & DAGE_OP_NO_CODA);
    bs.coda = coda ? op->coda : 0.0;
    Workspace ws{op->d_v, op->ws_bytes};
    const int rc = apply_f64_batch(op->ctx, operator_mats(op), in, out, bs, &ws);
    op->d_v = ws.p;
    op->ws_bytes = ws.bytes;
    return rc;
}

int apply_f64_fixed(dage_operator op, const double *in, double *out, void *ws_ptr,
                    size_t ws_bytes) {
    BatchSpec bs;
    const bool coda = (op->coda != 0.0) && !(op->flags & DAGE_OP_NO_CODA);
    bs.coda = coda ? op->coda : 0.0;
    bs.no_alloc = true;
    bs.ws_cap = ws_bytes;
    Workspace ws{static_cast<double *>(ws_ptr), ws_bytes};
    return apply_f64_batch(op->ctx, operator_mats(op), in, out, bs, &ws);
}

}  // namespace dage
