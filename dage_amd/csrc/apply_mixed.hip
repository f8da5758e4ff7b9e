// Mixed-precision tensor-quadrature apply: split-BF16 on gfx950 MFMA with
// FP32 accumulation inside 32-deep K steps and FP64 accumulation across
// steps (and therefore across t, whose sum is folded into the z-pass K).
//
// Each operand element x is represented as x1 + x2 + x3 with x1 = bf16(x),
// x2 = bf16(x - x1), x3 = bf16(x - x1 - x2) (24 significant bits).  A product
// a*b is formed from the pairs with i + j <= 2 (6 MFMAs: "bf16x6") or with
// i + j <= 1 (3 MFMAs: "bf16x3", 16 bits).  bf16 x bf16 products are exact in
// FP32; one v_mfma_f32_16x16x32_bf16 sums 32 of them, and after every 32-deep
// K step the FP32 sums are added into FP64 accumulators, so rounding does not
// grow with K.  Intermediates between the mode passes are FP32.
//
// Data flow: the operator matrices are split once per operator into three
// bf16 planes laid out as MFMA B fragments ([plane][n][k], k padded to a
// multiple of 32 with zeros); a workgroup stages the planes of its column
// block in LDS with 16-byte copies (double-buffered) and its four
// wavefronts share them.  The data operand (cube, V or U) is loaded by each
// wavefront straight into registers, split there and fed to the MFMAs, so
// it never passes through LDS.  MFMA bf16 16x16x32 layout on gfx950
// (measured): lane l holds A(l%16, 8*(l/16)..+7) and B(8*(l/16)..+7, l%16);
// result register r of lane l is D(4*(l/16)+r, l%16).
//
// Passes (as in apply_f64.hip): x mode batched over t; y mode batched over t
// with w_t folded in; z mode with K = (zi,t), split for occupancy, then a
// deterministic FP64 reduction adding coda*in.
//
// Accuracy against the FP64 oracle at config 1 (97^3, T=40, Gaussian
// density): bf16x6 7.4e-8 of max|V|, bf16x3 1.4e-6 (tests/test_gpu_mixed.py).
#include "dage_internal.hpp"

#include <algorithm>
#include <cstdlib>

namespace dage {
namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kKS = 32;       // K depth of one bf16 MFMA = one K step
constexpr int kRow = kKS + 8; // LDS row of a B plane (bf16): 80 bytes

template <typename T>
__device__ inline void split3(T x, __bf16 &h, __bf16 &m, __bf16 &l) {
    h = (__bf16)(float)x;
    const T r1 = x - (T)(float)h;
    m = (__bf16)(float)r1;
    const T r2 = r1 - (T)(float)m;
    l = (__bf16)(float)r2;
}

// Matrix planes: out[q][t][n][k] for q < 3, t < T, n < Np, k < Kp; element
// (n,k) of batch t is src[n*s_n + k*s_k + t*s_t] inside (N,K), zero outside.
__global__ void __launch_bounds__(256) split_matrix_kernel(const double *src, long s_n, long s_k,
                                                           long s_t, int N, int K, int T, int Np,
                                                           int Kp, __bf16 *out) {
    const long plane = (long)T * Np * Kp;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < plane;
         e += (long)gridDim.x * blockDim.x) {
        const int k = (int)(e % Kp);
        const long r = e / Kp;
        const int n = (int)(r % Np);
        const int t = (int)(r / Np);
        const double v = (n < N && k < K) ? src[(long)n * s_n + (long)k * s_k + (long)t * s_t] : 0.0;
        __bf16 h, m, l;
        split3(v, h, m, l);
        out[e] = h;
        out[plane + e] = m;
        out[2 * plane + e] = l;
    }
}

struct SGemm {
    const void *A;            // A(m,k) = A[mlo*sAm0 + mhi*sAm1 + k*sAk], m = mlo + M0*mhi
    int M, K, M0;
    long sAm0, sAm1, sAk, bA;
    const __bf16 *Bp;         // B planes [3][.][Np][Kp]: plane stride sBq, batch stride bB
    int Np, Kp;
    long sBq, bB;
    void *C;                  // C(m,n) = C[mlo*sCm0 + mhi*sCm1 + n*sCn]
    int N;
    long sCm0, sCm1, sCn, bC;
    int Z0;                   // > 0: blockIdx.z = z_lo + Z0*z_hi (z_hi: right-hand side)
    long bA1, bC1;            // strides per z_hi (the B planes are shared)
    const double *alpha;      // alpha[z_lo + aS1*z_hi] (null: 1)
    long aS1;
    int Ksplit;               // > 0: z_lo selects K steps [z_lo*Ksplit, (z_lo+1)*Ksplit) of 32
    int accumulate;
    // Optional B row masks (bit c: a nonzero among columns [16c, 16c+16));
    // rows of batch index z_lo start at mask + z_lo*mS, or at the K range
    // start with Ksplit.  K steps that are zero on the block's columns are
    // skipped (they would add exactly zero).
    const uint64_t *mask;
    long mS;
    // > 0: column tile j is [j*pitch, (j+1)*pitch), the last one running to N,
    // and bit j of a mask row stands for tile j (the planes are padded past
    // N by at least 16*NB - pitch columns, so whole 16*NB tiles stay in them)
    int pitch;
};

constexpr int kMaxSteps = 2048;  // K steps a block can list (K <= 65536)

// Block: 4 wavefronts stacked along M, wavefront tile (16*MW) x (16*NB).
// MASKED: the block lists its nonzero K steps from g.mask (the unmasked
// instantiation carries neither the list nor its LDS).
template <int MW, int NB, int NPAIR, typename TA, typename TC, bool A_KCONTIG, int PASS,
          bool MASKED, int PIPE>
__global__ void __launch_bounds__(kThreads) gemm_split_kernel(SGemm g) {
    constexpr int BM = 64 * MW;
    constexpr int BN = 16 * NB;
    constexpr int NQB = (NPAIR == 6) ? 3 : 2;  // B planes in use
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2][NQB][BN][kRow];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int lrow = lane & 15;
    const int kg = lane >> 4;  // k group of 8
    const int m0 = blockIdx.x * BM;
    const int n0 = g.pitch > 0 ? blockIdx.y * g.pitch : blockIdx.y * BN;
    const int nend = g.pitch > 0 && blockIdx.y + 1 < gridDim.y ? n0 + g.pitch : g.N;

    const TA *A = static_cast<const TA *>(g.A);
    const __bf16 *Bp = g.Bp;
    TC *C = static_cast<TC *>(g.C);
    double alpha = 1.0;
    int kbeg = 0, kend = g.K;
    const int zlo = g.Z0 > 0 ? (int)blockIdx.z % g.Z0 : (int)blockIdx.z;
    const int zhi = g.Z0 > 0 ? (int)blockIdx.z / g.Z0 : 0;
    A += (long)zhi * g.bA1;
    C += (long)zlo * g.bC + (long)zhi * g.bC1;
    if (g.Ksplit > 0) {
        kbeg = zlo * g.Ksplit * kKS;
        kend = min(g.K, kbeg + g.Ksplit * kKS);
    } else {
        A += (long)zlo * g.bA;
        Bp += (long)zlo * g.bB;
        if (g.alpha) alpha = g.alpha[zlo + g.aS1 * zhi];
    }
    int nsteps = (kend - kbeg + kKS - 1) / kKS;
    __shared__ unsigned short slist[MASKED ? kMaxSteps : 1];
    __shared__ int slist_n;
    const uint64_t *mrow = nullptr;
    if (MASKED) mrow = g.mask + (g.Ksplit > 0 ? (long)kbeg : (long)zlo * g.mS);
    const bool listed = MASKED && nsteps <= kMaxSteps;
    if (listed) {
        if (wave == 0) {
            const uint64_t cols = g.pitch > 0 ? 1ull << blockIdx.y
                                              : ((1ull << NB) - 1ull) << (n0 / 16);
            const int klen = kend - kbeg;
            int cnt = 0;
            for (int base = 0; base < nsteps; base += 64) {
                const int st = base + lane;
                bool need = false;
                if (st < nsteps) {
                    const int k0 = st * kKS;
                    const int kn = min(kKS, klen - k0);
                    uint64_t any = 0;
                    for (int j = 0; j < kn; ++j) any |= mrow[k0 + j];
                    need = (any & cols) != 0;
                }
                const uint64_t bal = __ballot(need);
                if (need) slist[cnt + __popcll(bal & ((1ull << lane) - 1ull))] = (unsigned short)st;
                cnt += __popcll(bal);
            }
            if (lane == 0) slist_n = cnt;
        }
        __syncthreads();
        nsteps = slist_n;
    }
    auto step_of = [&](int s) { return listed ? (int)slist[s] : s; };

    long arow[MW];
    bool rok[MW];
#pragma unroll
    for (int i = 0; i < MW; ++i) {
        const int gm = m0 + wave * 16 * MW + i * 16 + lrow;
        rok[i] = gm < g.M;
        arow[i] = rok[i] ? (long)(gm % g.M0) * g.sAm0 + (long)(gm / g.M0) * g.sAm1 : 0;
    }

    TA ra[MW][8];
    auto load_a_to = [&](TA (&dst)[MW][8], int step) {
        const int kb = kbeg + step * kKS + kg * 8;
#pragma unroll
        for (int i = 0; i < MW; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = kb + j;
                const bool ok = rok[i] && k < kend;
                const TA v = A[ok ? arow[i] + (long)k * g.sAk : 0];
                dst[i][j] = ok ? v : (TA)0;
            }
    };
    auto load_a = [&](int step) { load_a_to(ra, step); };
    // 16-byte chunks of the block's B planes for one K step
    auto copy_b = [&](int step, int buf) {
        constexpr int chunks = NQB * BN * (kKS / 8);
        const long kofs = (long)kbeg + (long)step * kKS;
        for (int c = tid; c < chunks; c += kThreads) {
            const int q = c / (BN * 4);
            const int rem = c - q * (BN * 4);
            const int n = rem >> 2;
            const int part = rem & 3;
            const u4 v = *reinterpret_cast<const u4 *>(Bp + q * g.sBq + (long)(n0 + n) * g.Kp +
                                                       kofs + part * 8);
            *reinterpret_cast<u4 *>(&Bs[buf][q][n][part * 8]) = v;
        }
    };

    d4 acc[MW][NB];
#pragma unroll
    for (int i = 0; i < MW; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

    auto split_a = [&](const TA (&src)[MW][8], bf8 (&a)[3][MW]) {
#pragma unroll
        for (int i = 0; i < MW; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                __bf16 h, md, l;
                split3(src[i][j], h, md, l);
                a[0][i][j] = h;
                a[1][i][j] = md;
                a[2][i][j] = l;
            }
    };
    // The MFMAs of one K step on B buffer buf, the FP32 partials then added
    // into the FP64 accumulators.
    auto compute = [&](const bf8 (&a)[3][MW], int buf) {
        f4 c[MW][NB];
#pragma unroll
        for (int i = 0; i < MW; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) c[i][j] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int qb = 0; qb < NQB; ++qb) {
            bf8 b[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j)
                b[j] = *reinterpret_cast<const bf8 *>(&Bs[buf][qb][j * 16 + lrow][kg * 8]);
#pragma unroll
            for (int qa = 2; qa >= 0; --qa) {  // small terms first
                if (qa + qb > NQB - 1) continue;
#pragma unroll
                for (int i = 0; i < MW; ++i)
#pragma unroll
                    for (int j = 0; j < NB; ++j)
                        c[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[qa][i], b[j], c[i][j],
                                                                          0, 0, 0);
            }
        }
#pragma unroll
        for (int i = 0; i < MW; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[i][j][r] += (double)c[i][j][r];
    };

    if constexpr (PIPE == 2) {
        // Two-deep pipeline: the A values of step s + 2 are loaded (into the
        // register set of step s, free once split) while step s computes,
        // and the B planes of step s + 2 into registers, stored to LDS one
        // step later; every global load has one to two steps of MFMA work
        // to arrive in.  Same products, same order as the one-deep loop.
        constexpr int chunks = NQB * BN * (kKS / 8);
        constexpr int CH = (chunks + kThreads - 1) / kThreads;
        u4 rb[CH];
        TA ra1[MW][8];
        auto load_b_regs = [&](int step) {
            const long kofs = (long)kbeg + (long)step * kKS;
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                const int c = tid + i * kThreads;
                if (c < chunks) {
                    const int q = c / (BN * 4);
                    const int rem = c - q * (BN * 4);
                    rb[i] = *reinterpret_cast<const u4 *>(Bp + q * g.sBq +
                                                          (long)(n0 + (rem >> 2)) * g.Kp + kofs +
                                                          (rem & 3) * 8);
                }
            }
        };
        auto store_b_regs = [&](int buf) {
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                const int c = tid + i * kThreads;
                if (c < chunks) {
                    const int q = c / (BN * 4);
                    const int rem = c - q * (BN * 4);
                    *reinterpret_cast<u4 *>(&Bs[buf][q][rem >> 2][(rem & 3) * 8]) = rb[i];
                }
            }
        };
        auto body = [&](int s, TA (&rs)[MW][8], int buf) {
            bf8 a[3][MW];
            split_a(rs, a);
            if (s + 1 < nsteps) store_b_regs(buf ^ 1);
            if (s + 2 < nsteps) {
                load_a_to(rs, step_of(s + 2));
                load_b_regs(step_of(s + 2));
            }
            compute(a, buf);
            __syncthreads();
        };
        if (nsteps > 0) {
            copy_b(step_of(0), 0);
            load_a_to(ra, step_of(0));
        }
        if (nsteps > 1) {
            load_b_regs(step_of(1));
            load_a_to(ra1, step_of(1));
        }
        __syncthreads();
        for (int s = 0; s < nsteps; s += 2) {
            body(s, ra, 0);
            if (s + 1 < nsteps) body(s + 1, ra1, 1);
        }
    } else {
    if (nsteps > 0) {
        copy_b(step_of(0), 0);
        load_a(step_of(0));
        __syncthreads();
    }
    for (int s = 0; s < nsteps; ++s) {
        const int buf = s & 1;
        bf8 a[3][MW];
        split_a(ra, a);
        if (s + 1 < nsteps) {
            copy_b(step_of(s + 1), buf ^ 1);
            load_a(step_of(s + 1));
        }
        f4 c[MW][NB];
#pragma unroll
        for (int i = 0; i < MW; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) c[i][j] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int qb = 0; qb < NQB; ++qb) {
            bf8 b[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j)
                b[j] = *reinterpret_cast<const bf8 *>(&Bs[buf][qb][j * 16 + lrow][kg * 8]);
#pragma unroll
            for (int qa = 2; qa >= 0; --qa) {  // small terms first
                if (qa + qb > NQB - 1) continue;
#pragma unroll
                for (int i = 0; i < MW; ++i)
#pragma unroll
                    for (int j = 0; j < NB; ++j)
                        c[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[qa][i], b[j], c[i][j],
                                                                          0, 0, 0);
            }
        }
#pragma unroll
        for (int i = 0; i < MW; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[i][j][r] += (double)c[i][j][r];
        __syncthreads();
    }
    }

#pragma unroll
    for (int i = 0; i < MW; ++i) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gm = m0 + wave * 16 * MW + i * 16 + kg * 4 + r;
            if (gm >= g.M) continue;
            const long crow = (long)(gm % g.M0) * g.sCm0 + (long)(gm / g.M0) * g.sCm1;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const int gn = n0 + j * 16 + lrow;
                if (gn >= nend) continue;
                TC *dst = C + crow + (long)gn * g.sCn;
                const double v = alpha * acc[i][j][r];
                *dst = g.accumulate ? (TC)((double)*dst + v) : (TC)v;
            }
        }
    }
}

int pick_nb_mixed(int n) {
    int best = 8;
    double best_cost = 1e30;
    for (int nb = 8; nb >= 1; --nb) {
        const int bn = 16 * nb;
        const int tiles = (n + bn - 1) / bn;
        const double cost = (double)tiles * bn / n + 0.02 * tiles;
        if (cost < best_cost - 1e-12) {
            best_cost = cost;
            best = nb;
        }
    }
    return best;
}

int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Columns of a plane padded so that every NB in 1..8 reads inside it.
int pad_all_nb(int n) {
    int p = 0;
    for (int nb = 1; nb <= 8; ++nb) p = std::max(p, round_up(n, 16 * nb));
    return p;
}

// 64-row block tiles: 144-148 VGPRs, so two wavefronts per SIMD (128-row
// tiles need 256 VGPRs + 64 AGPRs, one wavefront, and ran 1.47x slower at 97^3)
constexpr int kMWs = 1;

// Pipeline depth of the split kernels: 2 (A two K steps ahead, B planes one
// step ahead in registers; 229 registers, two wavefronts per SIMD), or the
// one-deep loop (163 registers, three wavefronts per SIMD) with
// DAGE_MIXED_PIPE=1 (comparison runs; read at every launch).  Measured
// (bf16x6, same bits): 97^3 0.578 -> 0.564 ms, 241^3 14.5 -> 11.0 ms (z pass
// 4.97 -> 2.73 ms), 8-right-hand-side Helmholtz batch at 385^3 536 -> 470 ms.
int mixed_pipe() {
    const char *e = std::getenv("DAGE_MIXED_PIPE");
    return (e && e[0] == '1') ? 1 : 2;
}

template <int NB, int NPAIR, typename TA, typename TC, bool AK, int PASS>
void launch_s(const SGemm &g, dim3 grid, hipStream_t s) {
    const bool two = mixed_pipe() == 2;
    if (g.mask) {
        if (two)
            hipLaunchKernelGGL((gemm_split_kernel<kMWs, NB, NPAIR, TA, TC, AK, PASS, true, 2>),
                               grid, dim3(kThreads), 0, s, g);
        else
            hipLaunchKernelGGL((gemm_split_kernel<kMWs, NB, NPAIR, TA, TC, AK, PASS, true, 1>),
                               grid, dim3(kThreads), 0, s, g);
    } else {
        if (two)
            hipLaunchKernelGGL((gemm_split_kernel<kMWs, NB, NPAIR, TA, TC, AK, PASS, false, 2>),
                               grid, dim3(kThreads), 0, s, g);
        else
            hipLaunchKernelGGL((gemm_split_kernel<kMWs, NB, NPAIR, TA, TC, AK, PASS, false, 1>),
                               grid, dim3(kThreads), 0, s, g);
    }
}

template <int NPAIR, typename TA, typename TC, bool AK, int PASS>
int launch_split(dage_ctx ctx, const char *name, const SGemm &g, int nb, int zdim) {
    const int BM = 64 * kMWs;
    const int ntiles = g.pitch > 0 ? std::max(1, g.N / g.pitch) : (g.N + 16 * nb - 1) / (16 * nb);
    DAGE_REQUIRE(g.pitch <= 0 || (ntiles <= 64 && g.pitch <= 16 * nb && g.pitch >= 16 * nb - 15 &&
                                  g.N - (ntiles - 1) * g.pitch <= 16 * nb),
                 DAGE_ERR_STATE, "column pitch %d does not fit %d-column tiles", g.pitch, 16 * nb);
    const dim3 grid((g.M + BM - 1) / BM, ntiles, zdim);
    hipEvent_t ev;
    prof_begin(ctx, name, &ev);
    switch (nb) {
    case 1: launch_s<1, NPAIR, TA, TC, AK, PASS>(g, grid, ctx->stream); break;
    case 2: launch_s<2, NPAIR, TA, TC, AK, PASS>(g, grid, ctx->stream); break;
    case 3: launch_s<3, NPAIR, TA, TC, AK, PASS>(g, grid, ctx->stream); break;
    case 4: launch_s<4, NPAIR, TA, TC, AK, PASS>(g, grid, ctx->stream); break;
    case 5: launch_s<5, NPAIR, TA, TC, AK, PASS>(g, grid, ctx->stream); break;
    case 6: launch_s<6, NPAIR, TA, TC, AK, PASS>(g, grid, ctx->stream); break;
    case 7: launch_s<7, NPAIR, TA, TC, AK, PASS>(g, grid, ctx->stream); break;
    default: launch_s<8, NPAIR, TA, TC, AK, PASS>(g, grid, ctx->stream); break;
    }
    DAGE_HIP_CHECK(hipGetLastError());
    prof_end(ctx, name, ev);
    return DAGE_OK;
}

int launch_split_matrix(dage_ctx ctx, const double *src, long s_n, long s_k, long s_t, int N,
                        int K, int T, int Np, int Kp, __bf16 *out) {
    const long total = (long)T * Np * Kp;
    const int grid = (int)std::min<long>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(split_matrix_kernel, dim3(grid), dim3(256), 0, ctx->stream, src, s_n, s_k,
                       s_t, N, K, T, Np, Kp, out);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

// Split the operator matrices into bf16 planes (once per matrix set).
int prepare_planes(dage_operator op) {
    if (op->d_planes && op->planes_valid) return DAGE_OK;
    const int nxi = op->din[0], nyi = op->din[1], nzi = op->din[2];
    const int nxo = op->dout[0], nyo = op->dout[1], nzo = op->dout[2];
    const int T = op->ntp;
    MixedLayout &L = op->mixed;
    L.nbx = pick_nb_mixed(nxo);
    L.nby = pick_nb_mixed(nyo);
    L.nbz = pick_nb_mixed(nzo);
    L.npx = pad_all_nb(nxo);  // any column-block count may be launched on these planes
    L.npy = pad_all_nb(nyo);
    L.npz = pad_all_nb(nzo);
    L.kpx = round_up(nxi, kKS);
    L.kpy = round_up(nyi, kKS);
    // one extra zero K step: a t-chunk starts at a multiple of 8 rows, so its
    // last 32-row step can reach up to 24 rows past round_up(nzi*T, 32)
    L.kpz = round_up(nzi * T, kKS) + kKS;
    const size_t px = (size_t)T * L.npx * L.kpx, py = (size_t)T * L.npy * L.kpy,
                 pz = (size_t)L.npz * L.kpz;
    L.off_x = 0;
    L.off_y = 3 * px;
    L.off_z = 3 * (px + py);
    const size_t total = 3 * (px + py + pz);
    if (!op->d_planes) DAGE_HIP_CHECK(hipMalloc((void **)&op->d_planes, 2 * total));
    __bf16 *P = reinterpret_cast<__bf16 *>(op->d_planes);
    // x mode: B(k=xi, n=xo) = Fx[xo + nxo*(xi + nxi*t)]
    DAGE_TRY(launch_split_matrix(op->ctx, op->d_fx, 1, nxo, (long)nxo * nxi, nxo, nxi, T, L.npx,
                                 L.kpx, P + L.off_x));
    // y mode: B(k=yi, n=yo) = Fy[yi + nyi*(yo + nyo*t)]
    DAGE_TRY(launch_split_matrix(op->ctx, op->d_fy, nyi, 1, (long)nyi * nyo, nyo, nyi, T, L.npy,
                                 L.kpy, P + L.off_y));
    // z mode: B(k=zi+nzi*t, n=zo) = Gz[k + nzi*T*zo]
    DAGE_TRY(launch_split_matrix(op->ctx, op->d_gz, (long)nzi * T, 1, 0, nzo, nzi * T, 1, L.npz,
                                 L.kpz, P + L.off_z));
    op->planes_valid = true;
    return DAGE_OK;
}

struct MPlan {
    int tchunk, splits;
    size_t v_elems, u_elems, p_elems;  // FP32, FP32, FP64
};

MPlan make_mplan(dage_operator op, int B) {
    const long nyi = op->din[1], nzi = op->din[2];
    const long nxo = op->dout[0], nyo = op->dout[1], nzo = op->dout[2];
    size_t budget = 4ull << 30;  // bytes for V + U per right-hand side
    if (const char *e = std::getenv("DAGE_APPLY_WORKSPACE_MB")) budget = (size_t)atol(e) << 20;
    const size_t total_cap = 64ull << 30;  // and for all B right-hand sides together
    const size_t per_t = 4ull * (size_t)(nxo * nyi * nzi + nxo * nyo * nzi);
    MPlan p;
    int tc = (int)std::max<size_t>(
        1, std::min<size_t>({(size_t)op->ntp, budget / per_t, total_cap / (per_t * (size_t)B)}));
    // the z pass reads its B planes from row nzi*t0 in 16-byte (8 x bf16)
    // pieces, so nzi*tchunk % 8 == 0 unless one chunk covers all t (a K step
    // reaching past the chunk's rows multiplies them by zeroed A values)
    if (tc < op->ntp) {
        const int unit = 8 / (int)std::__gcd<long>(nzi, 8);
        tc = std::max(unit, tc / unit * unit);
        tc = std::min(tc, op->ntp);
    }
    p.tchunk = tc;
    const long M3 = nxo * nyo;
    const int mtiles = (int)((M3 + 64 * kMWs - 1) / (64 * kMWs));
    const int nb = pick_nb_mixed((int)nzo);
    const int ntiles = (int)((nzo + 16 * nb - 1) / (16 * nb));
    const long K3 = nzi * p.tchunk;
    int splits = std::max(1, (1024 + mtiles * ntiles * B - 1) / (mtiles * ntiles * B));
    splits = (int)std::min<long>(splits, std::max<long>(1, K3 / 128));
    p.splits = splits;
    p.v_elems = (size_t)nxo * nyi * nzi * p.tchunk * B;
    p.u_elems = (size_t)nxo * nyo * nzi * p.tchunk * B;
    p.p_elems = (size_t)M3 * nzo * splits * B;
    return p;
}

// out[b*len + i] = sum_s part[(b*splits + s)*len + i] + coda*in[b*len + i]
__global__ void __launch_bounds__(256) mixed_reduce_kernel(const double *part, int splits,
                                                           long len, int B, const double *in,
                                                           double coda, double *out) {
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

template <int NPAIR>
int apply_mixed_t(dage_operator op, const double *in, double *out, const BatchSpec &bs,
                  Workspace *ws) {
    const long nxi = op->din[0], nyi = op->din[1], nzi = op->din[2];
    const long nxo = op->dout[0], nyo = op->dout[1], nzo = op->dout[2];
    const int T = op->ntp, B = bs.B;
    DAGE_TRY(prepare_planes(op));
    const MixedLayout &L = op->mixed;
    // passes whose zero-panel masks pay use them, with the column-block count
    // the mask cost model chose
    auto nb_of = [&](int p, int dflt) {
        if (!bs.nb[p] && op->mix_pitch[p] > 0) return op->mix_nbp[p];
        return bs.nb[p] ? bs.nb[p] : (op->mix_mask_used[p] ? op->mix_nb_pick[p] : dflt);
    };
    const int nbx = nb_of(0, L.nbx), nby = nb_of(1, L.nby), nbz = nb_of(2, L.nbz);
    const long mrx = (long)T * nxi, mry = (long)T * nyi;
    // masks of pass p starting at row r: the pitch tiling's when it is used
    // (no caller column-block count), else the 16-column ones if they pay
    auto set_mask = [&](SGemm &g, int p, long r, long mS) {
        if (!bs.nb[p] && op->mix_pitch[p] > 0) {
            g.pitch = op->mix_pitch[p];
            g.mask = op->d_tmask + r;
        } else if (op->mix_mask_used[p]) {
            g.mask = op->d_mask + r;
        }
        if (g.mask) g.mS = mS;
    };
    const __bf16 *P16 = reinterpret_cast<const __bf16 *>(op->d_planes);
    const MPlan plan = make_mplan(op, B);
    const size_t need = 4 * (plan.v_elems + plan.u_elems) + 8 * plan.p_elems + 256;
    if (need > ws->bytes) {
        DAGE_REQUIRE(!bs.no_alloc, DAGE_ERR_STATE, "workspace of %zu bytes not reserved", need);
        if (ws->p) DAGE_HIP_CHECK(hipFree(ws->p));
        ws->p = nullptr;
        ws->bytes = 0;
        DAGE_HIP_CHECK(hipMalloc((void **)&ws->p, need));
        ws->bytes = need;
    }
    double *P = ws->p;  // FP64 partials first (8-byte aligned)
    float *V = reinterpret_cast<float *>(P + plan.p_elems);
    float *U = V + plan.v_elems;
    const double *W = bs.w ? bs.w : op->d_w;
    const long wS = bs.w ? T : 0;
    const long nin = nxi * nyi * nzi;
    const long M3 = nxo * nyo;
    const int tcmax = plan.tchunk;
    const long Vt = nxo * nyi * nzi, Ut = nxo * nyo * nzi;
    for (int t0 = 0; t0 < T; t0 += tcmax) {
        const int tc = std::min(tcmax, T - t0);
        // x mode, batch (t, b): V_{b,t}((yi,zi), xo) = in_b((yi,zi), xi) . Fx_t^T(xi, xo)
        SGemm g1{};
        g1.A = in;
        g1.M = (int)(nyi * nzi); g1.K = (int)nxi; g1.M0 = g1.M;
        g1.sAm0 = nxi; g1.sAm1 = 0; g1.sAk = 1; g1.bA = 0;
        g1.Bp = P16 + L.off_x + (size_t)t0 * L.npx * L.kpx;
        g1.Np = L.npx; g1.Kp = L.kpx;
        g1.sBq = (long)T * L.npx * L.kpx; g1.bB = (long)L.npx * L.kpx;
        g1.C = V; g1.N = (int)nxo;
        g1.sCm0 = nxo; g1.sCm1 = 0; g1.sCn = 1; g1.bC = Vt;
        g1.Z0 = tc; g1.bA1 = nin; g1.bC1 = Vt * tc;
        set_mask(g1, 0, (long)t0 * nxi, nxi);
        DAGE_TRY((launch_split<NPAIR, double, float, true, 0>(op->ctx, "gemm_mixed_xmode", g1,
                                                              nbx, tc * B)));
        // y mode, batch (t, b): U_b((xo,yo),(zi,t)) = w_{b,t} V_{b,t}((xo,zi), yi) . Fy_t(yi, yo)
        SGemm g2{};
        g2.A = V;
        g2.M = (int)(nxo * nzi); g2.K = (int)nyi; g2.M0 = (int)nxo;
        g2.sAm0 = 1; g2.sAm1 = nxo * nyi; g2.sAk = nxo; g2.bA = Vt;
        g2.Bp = P16 + L.off_y + (size_t)t0 * L.npy * L.kpy;
        g2.Np = L.npy; g2.Kp = L.kpy;
        g2.sBq = (long)T * L.npy * L.kpy; g2.bB = (long)L.npy * L.kpy;
        g2.C = U; g2.N = (int)nyo;
        g2.sCm0 = 1; g2.sCm1 = nxo * nyo; g2.sCn = nxo; g2.bC = Ut;
        g2.Z0 = tc; g2.bA1 = Vt * tc; g2.bC1 = Ut * tc;
        g2.alpha = W + t0;
        g2.aS1 = wS;
        set_mask(g2, 1, mrx + (long)t0 * nyi, nyi);
        DAGE_TRY((launch_split<NPAIR, float, float, false, 1>(op->ctx, "gemm_mixed_ymode", g2,
                                                              nby, tc * B)));
        // z mode, batch (split, b): P_{b,s}((xo,yo), zo) (+)= U_b((xo,yo), k) . Gz(k, zo)
        const int K3 = (int)(nzi * tc);
        const int steps = (K3 + kKS - 1) / kKS;
        const int per = (steps + plan.splits - 1) / plan.splits;
        const int splits = (steps + per - 1) / per;
        SGemm g3{};
        g3.A = U;
        g3.M = (int)M3; g3.K = K3; g3.M0 = g3.M;
        g3.sAm0 = 1; g3.sAm1 = 0; g3.sAk = M3; g3.bA = 0;
        g3.Bp = P16 + L.off_z + (size_t)nzi * t0;  // nzi*t0 is a multiple of 8 (see make_mplan)
        g3.Np = L.npz; g3.Kp = L.kpz;
        g3.sBq = (long)L.npz * L.kpz; g3.bB = 0;
        g3.C = P; g3.N = (int)nzo;
        g3.sCm0 = 1; g3.sCm1 = 0; g3.sCn = M3; g3.bC = M3 * nzo;
        g3.Z0 = splits; g3.bA1 = Ut * tc; g3.bC1 = M3 * nzo * plan.splits;
        g3.Ksplit = per;
        set_mask(g3, 2, mrx + mry + (long)nzi * t0, 0);
        g3.accumulate = (t0 > 0);
        if (t0 == 0 && splits < plan.splits)
            for (int b = 0; b < B; ++b)
                DAGE_TRY(fill_zero(op->ctx, P + ((size_t)b * plan.splits + splits) * M3 * nzo,
                                   (size_t)(plan.splits - splits) * M3 * nzo));
        DAGE_TRY((launch_split<NPAIR, float, double, false, 2>(op->ctx, "gemm_mixed_zmode", g3,
                                                               nbz, splits * B)));
    }
    const long len = M3 * nzo;
    const long total = len * B;
    const int grid = (int)std::min<long>((total + 255) / 256, 65536);
    hipEvent_t ev;
    prof_begin(op->ctx, "mixed_reduce_kernel", &ev);
    hipLaunchKernelGGL(mixed_reduce_kernel, dim3(grid), dim3(256), 0, op->ctx->stream, P,
                       plan.splits, len, B, bs.coda != 0.0 ? in : nullptr, bs.coda, out);
    DAGE_HIP_CHECK(hipGetLastError());
    prof_end(op->ctx, "mixed_reduce_kernel", ev);
    return DAGE_OK;
}

}  // namespace

int apply_mixed(dage_operator op, const double *in, double *out, int npair) {
    BatchSpec bs;
    const bool coda = (op->coda != 0.0) && !(op->flags & DAGE_OP_NO_CODA);
    bs.coda = coda ? op->coda : 0.0;
    Workspace ws{op->d_v, op->ws_bytes};
    const int rc = apply_mixed_batch(op, in, out, npair, bs, &ws);
    op->d_v = ws.p;
    op->ws_bytes = ws.bytes;
    return rc;
}

int apply_mixed_batch(dage_operator op, const double *in, double *out, int npair,
                      const BatchSpec &bs, Workspace *ws) {
    if (npair == 6) return apply_mixed_t<6>(op, in, out, bs, ws);
    if (npair == 3) return apply_mixed_t<3>(op, in, out, bs, ws);
    set_error("unsupported split-BF16 product count %d", npair);
    return DAGE_ERR_UNSUPPORTED;
}

int apply_mixed_batch_workspace(dage_operator op, int B, size_t *bytes) {
    const MPlan p = make_mplan(op, B);
    *bytes = 4 * (p.v_elems + p.u_elems) + 8 * p.p_elems + 256;
    return DAGE_OK;
}

int prepare_mixed(dage_operator op) { return prepare_planes(op); }

}  // namespace dage
