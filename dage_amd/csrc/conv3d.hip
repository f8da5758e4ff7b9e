// 3x3x3 convolution layers for the learned orbital-update preconditioner:
// BF16 weights, FP32 activations and FP32 accumulation, zero padding,
// forward / backward-data / backward-filter, plus the Adam step, the MSE
// loss and FP64 <-> FP32 conversions between SCF cubes and activations.
//
// Layout: activations [c][z][y][x] (x fastest: the cube layout, one cube per
// channel); weights [cout][cin][27] with tap = (dx+1) + 3 (dy+1) + 9 (dz+1);
// BF16 weights are the upper 16 bits of the FP32 master weights rounded to
// nearest even.  Every reduction (filter gradient, loss) is a fixed-order
// two-stage sum, so training is deterministic.
//
// The 16 -> 16 layers (forward, and backward-data as the same convolution
// with transposed, mirrored weights) run on bf16 MFMA: every FP32 activation
// is split into three bf16 terms (24 bits, so the products are as exact as
// FP32 ones) and multiplied with the bf16 weights, FP32 accumulation.  A
// block streams along z through its (16*NS) x 4 column of the volume: one
// input plane (with a one-point halo) sits in LDS, pre-split, and feeds the
// three output planes it touches, whose accumulators stay in registers; the
// weights stay in registers as MFMA operands for the whole launch.
// Algorithmic work: 2*16*16*27 = 13824 FLOPs per voxel (dense, as FP32);
// HBM roofline 128 bytes per voxel (16 channels in and out, FP32).
//
// Layers with 1 or 2 channels on a side run as FP32 VALU kernels; the
// backward filter stages, per output x-row and per dz, the 3 neighbouring
// input rows of every channel in LDS, and every thread owns one (cout, cin)
// pair's 27 taps.
#include <algorithm>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "dage_internal.hpp"

namespace dage {
namespace {

constexpr int kMaxC = 16;

int use_device(dage_ctx ctx) {
    DAGE_REQUIRE(ctx != nullptr, DAGE_ERR_ARG, "null context handle");
    DAGE_HIP_CHECK(hipSetDevice(ctx->device));
    return DAGE_OK;
}

int grid_of(long n) { return (int)std::min<long>((n + 255) / 256, 65536); }

__device__ inline float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ inline uint16_t f32_to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
    const uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);                    // nearest even
    return (uint16_t)(r >> 16);
}

// The VALU layers' weight table: w[(o*cin + c)*27 + t] (bf16) as FP32 at
// wf[(c*27 + t)*cout + o], so one (c, t) tap's cout weights are contiguous.
__global__ void __launch_bounds__(256) weights_f32_kernel(const uint16_t *w, int cout, int cin,
                                                          float *wf) {
    const int n = cout * cin * 27;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int o = i / (cin * 27), ct = i - o * (cin * 27);
        wf[ct * cout + o] = bf16_to_f32(w[i]);
    }
}

// The weights (FP32 copies of the bf16 values) are read at wavefront-uniform
// indices through scalar loads, so every multiply-add takes its weight from
// a scalar register; no LDS, no per-lane weight loads.
template <int COUT>
__global__ void __launch_bounds__(256) conv_fwd_kernel(const float *__restrict__ in, int cin,
                                                       const float *__restrict__ wf,
                                                       const float *__restrict__ bias, int relu,
                                                       float *__restrict__ out, int nx, int ny,
                                                       int nz) {
    typedef const float __attribute__((address_space(1))) *gfptr;
    const gfptr ws = (gfptr)wf;
    const gfptr bs = (gfptr)bias;
    const long vol = (long)nx * ny * nz;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < vol;
         p += (long)gridDim.x * blockDim.x) {
        const int x = (int)(p % nx);
        const long r = p / nx;
        const int y = (int)(r % ny);
        const int z = (int)(r / ny);
        float acc[COUT];
#pragma unroll
        for (int o = 0; o < COUT; ++o) acc[o] = bs[o];
        for (int c = 0; c < cin; ++c) {
            const float *ic = in + (long)c * vol;
#pragma unroll
            for (int t = 0; t < 27; ++t) {
                const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
                const int xx = x + dx, yy = y + dy, zz = z + dz;
                const bool ok = xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz;
                const float v = ok ? ic[xx + (long)nx * (yy + (long)ny * zz)] : 0.0f;
#pragma unroll
                for (int o = 0; o < COUT; ++o) acc[o] += ws[(c * 27 + t) * COUT + o] * v;
            }
        }
#pragma unroll
        for (int o = 0; o < COUT; ++o)
            out[(long)o * vol + p] = (relu && acc[o] < 0.0f) ? 0.0f : acc[o];
    }
}

// gin[c][p] = sum_o sum_t W[o][c][t] gout[o][p - off(t)]
__global__ void __launch_bounds__(256) conv_bwd_data_kernel(const float *gout, int cout, int cin,
                                                            const uint16_t *w, float *gin, int nx,
                                                            int ny, int nz) {
    __shared__ float ws[kMaxC * kMaxC * 27];
    const int nw = cout * cin * 27;
    for (int i = threadIdx.x; i < nw; i += 256) ws[i] = bf16_to_f32(w[i]);
    __syncthreads();
    const long vol = (long)nx * ny * nz;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < vol;
         p += (long)gridDim.x * blockDim.x) {
        const int x = (int)(p % nx);
        const long r = p / nx;
        const int y = (int)(r % ny);
        const int z = (int)(r / ny);
        float acc[kMaxC];
#pragma unroll
        for (int c = 0; c < kMaxC; ++c) acc[c] = 0.0f;
        for (int o = 0; o < cout; ++o) {
            const float *go = gout + (long)o * vol;
#pragma unroll
            for (int t = 0; t < 27; ++t) {
                const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
                const int xx = x - dx, yy = y - dy, zz = z - dz;
                const bool ok = xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz;
                const float g = ok ? go[xx + (long)nx * (yy + (long)ny * zz)] : 0.0f;
#pragma unroll
                for (int c = 0; c < kMaxC; ++c)
                    if (c < cin) acc[c] += ws[(o * cin + c) * 27 + t] * g;
            }
        }
        for (int c = 0; c < cin; ++c) gin[(long)c * vol + p] = acc[c];
    }
}

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kCY = 4;  // output y-rows per block: one per wavefront

__device__ inline __bf16 bf16_bits(uint16_t b) { return __builtin_bit_cast(__bf16, b); }

// 16 -> 16 channel 3x3x3 convolution on bf16 MFMA (v_mfma_f32_16x16x32_bf16;
// lane l holds A(l%16, 8(l/16)..+7) and B(8(l/16)..+7, l%16); result
// register r of lane l is D(4(l/16)+r, l%16)).  D rows are 16 x-points of
// one output row, columns the 16 output channels; K = 32 runs over two taps
// (dx, dy) of one dz times 16 input channels, 5 K-steps per dz (the tenth
// tap slot is zero).  Block (bx, by, bz) owns x in [16*NS*bx, +16*NS),
// y in [4*by, +4) (wavefront w: row 4*by + w) and z in [bz*zchunk, +zchunk).
// out = act(sum + bias); bias may be null (zero).
template <int NS>
__global__ void __launch_bounds__(256) conv16_mfma_kernel(const float *in, const uint16_t *w,
                                                          const float *bias, int relu, float *out,
                                                          int nx, int ny, int nz, int zchunk) {
    constexpr int PX = 16 * NS + 2;  // halo plane extent in x and y
    constexpr int PY = kCY + 2;
    constexpr int NPT = PX * PY;
    constexpr int PTS = 48;          // bf16 per plane point: 3 terms x 16 channels
    constexpr int LOADS = (NPT + 255) / 256;
    __shared__ __attribute__((aligned(16))) __bf16 plane[2][NPT * PTS];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lrow = lane & 15, kg = lane >> 4;
    const int x0 = blockIdx.x * 16 * NS, y0 = blockIdx.y * kCY, z0 = blockIdx.z * zchunk;
    const int z1 = min(nz, z0 + zchunk);
    const long pl = (long)nx * ny, vol = pl * nz;

    // B operands: bw[dzi][j] = W[o = lane%16][c = 8(kg&1)+i][tau + 9 dzi], tau = 2j + kg/2
    bf8 bw[3][5];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            const int tau = 2 * j + (kg >> 1);
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int c = 8 * (kg & 1) + i;
                bw[d][j][i] = bf16_bits(tau < 9 ? w[(lrow * 16 + c) * 27 + tau + 9 * d] : 0);
            }
        }
    const float bo = bias ? bias[lrow] : 0.0f;

    f4 acc[3][NS];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[q][s] = f4{0.f, 0.f, 0.f, 0.f};

    // plane loads: thread point e = tid + 256 r, all 16 channels
    float raw[LOADS][16];
    auto load_plane = [&](int zi) {
#pragma unroll
        for (int r = 0; r < LOADS; ++r) {
            const int e = tid + 256 * r;
            const int ly = e / PX, lx = e - (e / PX) * PX;
            const int x = x0 - 1 + lx, y = y0 - 1 + ly;
            const bool ok = e < NPT && x >= 0 && x < nx && y >= 0 && y < ny && zi >= 0 && zi < nz;
            const long base = ok ? (long)zi * pl + (long)y * nx + x : 0;
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const float v = in[base + c * vol];
                raw[r][c] = ok ? v : 0.0f;
            }
        }
    };
    auto store_plane = [&](int buf) {
#pragma unroll
        for (int r = 0; r < LOADS; ++r) {
            const int e = tid + 256 * r;
            if (e >= NPT) continue;
            bf8 t[3][2];
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const float x = raw[r][c];
                const __bf16 h = (__bf16)x;
                const float r1 = x - (float)h;
                const __bf16 m = (__bf16)r1;
                const float r2 = r1 - (float)m;
                t[0][c >> 3][c & 7] = h;
                t[1][c >> 3][c & 7] = m;
                t[2][c >> 3][c & 7] = (__bf16)r2;
            }
            bf8 *dst = reinterpret_cast<bf8 *>(&plane[buf][e * PTS]);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                dst[2 * q] = t[q][0];
                dst[2 * q + 1] = t[q][1];
            }
        }
    };
    auto epilogue = [&](f4 (&a)[NS], int zo) {
        const int y = y0 + wave;
        if (y >= ny) return;
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int x = x0 + 16 * s + 4 * kg + r;
                if (x >= nx) continue;
                float v = a[s][r] + bo;
                if (relu && v < 0.0f) v = 0.0f;
                out[lrow * vol + (long)zo * pl + (long)y * nx + x] = v;
            }
    };

    const int nplanes = z1 - z0 + 2;  // input planes z0-1 .. z1
    load_plane(z0 - 1);
    store_plane(0);
    __syncthreads();

    // plane q = zi - (z0 - 1) feeds outputs zi+1, zi, zi-1 in slots (R+1)%3,
    // R, (R+2)%3 with R = q % 3; output zi-1 is complete afterwards.
    auto step = [&](int q, auto rc) {
        constexpr int R = decltype(rc)::value;
        constexpr int SN = (R + 1) % 3, SC = R, SP = (R + 2) % 3;
        if (q >= nplanes) return;
        const int zi = z0 - 1 + q;
        const int buf = q & 1;
        const bool more = q + 1 < nplanes;
        if (more) load_plane(zi + 1);
        const bool dn = zi + 1 >= z0 && zi + 1 < z1;  // output zi+1 (weights dz = -1)
        const bool dc = zi >= z0 && zi < z1;          // output zi    (dz = 0)
        const bool dp = zi - 1 >= z0 && zi - 1 < z1;  // output zi-1  (dz = +1)
        if (zi >= 0 && zi < nz) {
            const __bf16 *pb = plane[buf];
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int j = 0; j < 5; ++j) {
                    const int tau = min(2 * j + (kg >> 1), 8);  // slot 9 has zero weights
                    const int ly = wave + tau / 3, lx = 16 * s + lrow + tau % 3;
                    const bf8 *src =
                        reinterpret_cast<const bf8 *>(pb + (ly * PX + lx) * PTS + 8 * (kg & 1));
                    const bf8 a0 = src[0], a1 = src[2], a2 = src[4];
                    if (dn) {
                        acc[SN][s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, bw[0][j], acc[SN][s], 0, 0, 0);
                        acc[SN][s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, bw[0][j], acc[SN][s], 0, 0, 0);
                        acc[SN][s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, bw[0][j], acc[SN][s], 0, 0, 0);
                    }
                    if (dc) {
                        acc[SC][s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, bw[1][j], acc[SC][s], 0, 0, 0);
                        acc[SC][s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, bw[1][j], acc[SC][s], 0, 0, 0);
                        acc[SC][s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, bw[1][j], acc[SC][s], 0, 0, 0);
                    }
                    if (dp) {
                        acc[SP][s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, bw[2][j], acc[SP][s], 0, 0, 0);
                        acc[SP][s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, bw[2][j], acc[SP][s], 0, 0, 0);
                        acc[SP][s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, bw[2][j], acc[SP][s], 0, 0, 0);
                    }
                }
        }
        if (dp) epilogue(acc[SP], zi - 1);
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[SP][s] = f4{0.f, 0.f, 0.f, 0.f};
        if (more) store_plane(buf ^ 1);
        __syncthreads();
    };
    for (int q = 0; q < nplanes; q += 3) {
        step(q, std::integral_constant<int, 0>{});
        step(q + 1, std::integral_constant<int, 1>{});
        step(q + 2, std::integral_constant<int, 2>{});
    }
}

// W'[c][o][26 - t] = W[o][c][t] (bf16 bits): backward-data weights
__global__ void __launch_bounds__(256) transpose_mirror_kernel(const uint16_t *w, int cout,
                                                               int cin, uint16_t *wt) {
    const int n = cout * cin * 27;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const int t = e % 27, c = (e / 27) % cin, o = e / (27 * cin);
        wt[(c * cout + o) * 27 + 26 - t] = w[e];
    }
}

// x-subtiles per block for an x extent: least padded width, halo included
int conv_ns(int nx) {
    int best = 1;
    double best_cost = 1e30;
    for (int ns = 1; ns <= 4; ++ns) {
        const int bw = 16 * ns;
        const double cost = (double)((nx + bw - 1) / bw) * (bw + 2);
        if (cost < best_cost - 1e-9) { best_cost = cost; best = ns; }
    }
    return best;
}

int launch_conv16(dage_ctx ctx, int nx, int ny, int nz, const float *in, const uint16_t *w,
                  const float *bias, int relu, float *out) {
    const int ns = conv_ns(nx);
    const int bx = (nx + 16 * ns - 1) / (16 * ns), by = (ny + kCY - 1) / kCY;
    const int want = std::max(1, (2048 + bx * by - 1) / (bx * by));  // z-chunks for ~2048 blocks
    const int zchunk = std::max(2, (nz + want - 1) / want);
    const dim3 grid(bx, by, (nz + zchunk - 1) / zchunk);
    hipEvent_t ev;
    prof_begin(ctx, "conv16_mfma_kernel", &ev);
    switch (ns) {
    case 1: hipLaunchKernelGGL(conv16_mfma_kernel<1>, grid, dim3(256), 0, ctx->stream, in, w, bias, relu, out, nx, ny, nz, zchunk); break;
    case 2: hipLaunchKernelGGL(conv16_mfma_kernel<2>, grid, dim3(256), 0, ctx->stream, in, w, bias, relu, out, nx, ny, nz, zchunk); break;
    case 3: hipLaunchKernelGGL(conv16_mfma_kernel<3>, grid, dim3(256), 0, ctx->stream, in, w, bias, relu, out, nx, ny, nz, zchunk); break;
    default: hipLaunchKernelGGL(conv16_mfma_kernel<4>, grid, dim3(256), 0, ctx->stream, in, w, bias, relu, out, nx, ny, nz, zchunk); break;
    }
    DAGE_HIP_CHECK(hipGetLastError());
    prof_end(ctx, "conv16_mfma_kernel", ev);
    return DAGE_OK;
}

// Filter gradient, stage 1: block b takes x-rows (y, z) in [b*rows, ...);
// thread t = o*cin + c accumulates gW[o][c][0..26] and (c == 0) gb[o].  Per
// row and per dz the 3 input rows (y-1..y+1) of every channel are staged in
// LDS with a one-point halo in x.
__global__ void __launch_bounds__(256) conv_bwd_filter_kernel(const float *in, const float *gout,
                                                              int cin, int cout, int nx, int ny,
                                                              int nz, int rows_per_block,
                                                              float *part) {
    extern __shared__ float sm[];
    const int W = nx + 2;
    float *sin_ = sm;                  // [cin][3][W]
    float *sgo = sm + cin * 3 * W;     // [cout][nx]
    const int tid = threadIdx.x;
    const int o = tid / cin, c = tid % cin;
    const bool active = tid < cin * cout;
    float acc[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) acc[t] = 0.0f;
    float accb = 0.0f;
    const long vol = (long)nx * ny * nz;
    const long nrows = (long)ny * nz;
    const long r0 = (long)blockIdx.x * rows_per_block;
    const long r1 = r0 + rows_per_block < nrows ? r0 + rows_per_block : nrows;
    for (long row = r0; row < r1; ++row) {
        const int y = (int)(row % ny), z = (int)(row / ny);
#pragma unroll
        for (int kz = 0; kz < 3; ++kz) {
            __syncthreads();
            const int zz = z + kz - 1;
            for (int e = tid; e < cin * 3 * W; e += 256) {
                const int cc = e / (3 * W);
                const int rem = e - cc * 3 * W;
                const int ky = rem / W, xx = rem % W - 1;
                const int yy = y + ky - 1;
                const bool ok = xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz;
                sin_[e] = ok ? in[(long)cc * vol + xx + (long)nx * (yy + (long)ny * zz)] : 0.0f;
            }
            if (kz == 0)
                for (int e = tid; e < cout * nx; e += 256) {
                    const int oo = e / nx, xx = e % nx;
                    sgo[e] = gout[(long)oo * vol + xx + (long)nx * (y + (long)ny * z)];
                }
            __syncthreads();
            if (active) {
                const float *go = sgo + o * nx;
                const float *ic = sin_ + c * 3 * W;
                for (int x = 0; x < nx; ++x) {
                    const float g = go[x];
                    if (kz == 0 && c == 0) accb += g;
#pragma unroll
                    for (int ky = 0; ky < 3; ++ky)
#pragma unroll
                        for (int dx = 0; dx < 3; ++dx)
                            acc[9 * kz + 3 * ky + dx] += g * ic[ky * W + x + dx];
                }
            }
        }
    }
    if (active) {
        float *pp = part + (long)blockIdx.x * (cout * cin * 27 + cout);
#pragma unroll
        for (int t = 0; t < 27; ++t) pp[(o * cin + c) * 27 + t] = acc[t];
        if (c == 0) pp[cout * cin * 27 + o] = accb;
    }
}

__global__ void __launch_bounds__(256) sum_blocks_kernel(const float *part, int nblocks, int len,
                                                         float *out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int b = 0; b < nblocks; ++b) s += part[(long)b * len + i];
        out[i] = s;
    }
}

__global__ void __launch_bounds__(256) relu_bwd_kernel(long n, const float *out, float *g) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        if (out[i] <= 0.0f) g[i] = 0.0f;
}

__global__ void __launch_bounds__(256) adam_kernel(int n, float *w, const float *g, float *m,
                                                   float *v, float lr, float b1, float b2,
                                                   float eps, float c1, float c2,
                                                   uint16_t *wbf) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        wbf[i] = f32_to_bf16(wi);
    }
}

__global__ void __launch_bounds__(256) to_bf16_kernel(int n, const float *w, uint16_t *wbf) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        wbf[i] = f32_to_bf16(w[i]);
}

// grad = scale*(pred - target); part[block] = sum over the block's range of (pred - target)^2
__global__ void __launch_bounds__(256) mse_kernel(long n, const float *pred, const float *target,
                                                  float scale, float *grad, double *part) {
    __shared__ double red[256];
    double s = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const float d = pred[i] - target[i];
        grad[i] = scale * d;
        s += (double)d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) cvt_f64_f32_kernel(long n, const double *x, double scale,
                                                          float *y) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        y[i] = (float)(scale * x[i]);
}

__global__ void __launch_bounds__(256) axpy_f32_f64_kernel(long n, double a, const float *x,
                                                           double *y) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        y[i] = y[i] + a * (double)x[i];
}

int check_dims(int nx, int ny, int nz, int cin, int cout) {
    DAGE_REQUIRE(nx > 0 && ny > 0 && nz > 0, DAGE_ERR_ARG, "empty volume");
    DAGE_REQUIRE(cin >= 1 && cin <= kMaxC && cout >= 1 && cout <= kMaxC, DAGE_ERR_UNSUPPORTED,
                 "channels %d -> %d outside 1..%d", cin, cout, kMaxC);
    return DAGE_OK;
}

}  // namespace
}  // namespace dage

using namespace dage;

extern "C" {

int dage_conv3d_forward(dage_ctx ctx, int nx, int ny, int nz, int cin, int cout, const float *in,
                        const uint16_t *w, const float *bias, int relu, float *out) {
    DAGE_TRY(use_device(ctx));
    DAGE_TRY(check_dims(nx, ny, nz, cin, cout));
    DAGE_REQUIRE(in && w && bias && out && in != out, DAGE_ERR_ARG, "bad argument");
    DAGE_REQUIRE(cout == 1 || cout == kMaxC, DAGE_ERR_UNSUPPORTED,
                 "forward supports 1 or %d output channels", kMaxC);
    const long vol = (long)nx * ny * nz;
    if (cin == kMaxC && cout == kMaxC)
        return launch_conv16(ctx, nx, ny, nz, in, w, bias, relu, out);
    void *scr = nullptr;  // FP32 weight table
    const int nw = cout * cin * 27;
    DAGE_TRY(ctx_scratch(ctx, sizeof(float) * nw, &scr));
    float *wf = static_cast<float *>(scr);
    hipLaunchKernelGGL(weights_f32_kernel, dim3((nw + 255) / 256), dim3(256), 0, ctx->stream, w,
                       cout, cin, wf);
    DAGE_HIP_CHECK(hipGetLastError());
    if (cout == 1)
        hipLaunchKernelGGL(conv_fwd_kernel<1>, dim3(grid_of(vol)), dim3(256), 0, ctx->stream, in,
                           cin, wf, bias, relu, out, nx, ny, nz);
    else
        hipLaunchKernelGGL(conv_fwd_kernel<kMaxC>, dim3(grid_of(vol)), dim3(256), 0, ctx->stream,
                           in, cin, wf, bias, relu, out, nx, ny, nz);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int dage_conv3d_backward_data(dage_ctx ctx, int nx, int ny, int nz, int cin, int cout,
                              const float *gout, const uint16_t *w, float *gin) {
    DAGE_TRY(use_device(ctx));
    DAGE_TRY(check_dims(nx, ny, nz, cin, cout));
    DAGE_REQUIRE(gout && w && gin && gout != gin, DAGE_ERR_ARG, "bad argument");
    const long vol = (long)nx * ny * nz;
    if (cin == kMaxC && cout == kMaxC) {
        void *scr = nullptr;  // transposed, mirrored weights
        DAGE_TRY(ctx_scratch(ctx, sizeof(uint16_t) * kMaxC * kMaxC * 27, &scr));
        uint16_t *wt = static_cast<uint16_t *>(scr);
        hipLaunchKernelGGL(transpose_mirror_kernel, dim3(27), dim3(256), 0, ctx->stream, w, cout,
                           cin, wt);
        DAGE_HIP_CHECK(hipGetLastError());
        return launch_conv16(ctx, nx, ny, nz, gout, wt, nullptr, 0, gin);
    }
    hipLaunchKernelGGL(conv_bwd_data_kernel, dim3(grid_of(vol)), dim3(256), 0, ctx->stream, gout,
                       cout, cin, w, gin, nx, ny, nz);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int dage_conv3d_backward_filter(dage_ctx ctx, int nx, int ny, int nz, int cin, int cout,
                                const float *in, const float *gout, float *gw, float *gb) {
    DAGE_TRY(use_device(ctx));
    DAGE_TRY(check_dims(nx, ny, nz, cin, cout));
    DAGE_REQUIRE(in && gout && gw && gb, DAGE_ERR_ARG, "bad argument");
    const size_t smem = sizeof(float) * ((size_t)cin * 3 * (nx + 2) + (size_t)cout * nx);
    DAGE_REQUIRE(smem <= 64 * 1024, DAGE_ERR_UNSUPPORTED, "x extent %d too long for one block",
                 nx);
    const long nrows = (long)ny * nz;
    const int rows = (int)std::max<long>(1, (nrows + 1023) / 1024);
    const int nblocks = (int)((nrows + rows - 1) / rows);
    const int len = cout * cin * 27 + cout;
    const size_t bpart = sizeof(float) * (size_t)nblocks * len;
    void *scr = nullptr;
    DAGE_TRY(ctx_scratch(ctx, (bpart + 255) / 256 * 256 + sizeof(float) * len, &scr));
    float *part = static_cast<float *>(scr);
    float *sums = reinterpret_cast<float *>(static_cast<char *>(scr) + (bpart + 255) / 256 * 256);
    hipLaunchKernelGGL(conv_bwd_filter_kernel, dim3(nblocks), dim3(256), smem, ctx->stream, in,
                       gout, cin, cout, nx, ny, nz, rows, part);
    DAGE_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(sum_blocks_kernel, dim3((len + 255) / 256), dim3(256), 0, ctx->stream,
                       part, nblocks, len, sums);
    DAGE_HIP_CHECK(hipGetLastError());
    DAGE_HIP_CHECK(hipMemcpyAsync(gw, sums, sizeof(float) * (len - cout), hipMemcpyDeviceToDevice,
                                  ctx->stream));
    DAGE_HIP_CHECK(hipMemcpyAsync(gb, sums + (len - cout), sizeof(float) * cout,
                                  hipMemcpyDeviceToDevice, ctx->stream));
    return DAGE_OK;
}

int dage_relu_backward(dage_ctx ctx, long n, const float *out, float *grad) {
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(n > 0 && out && grad, DAGE_ERR_ARG, "bad argument");
    hipLaunchKernelGGL(relu_bwd_kernel, dim3(grid_of(n)), dim3(256), 0, ctx->stream, n, out, grad);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int dage_adam_step(dage_ctx ctx, int n, float *w, const float *g, float *m, float *v, float lr,
                   float beta1, float beta2, float eps, int step, uint16_t *w_bf16) {
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(n > 0 && w && g && m && v && w_bf16 && step >= 1, DAGE_ERR_ARG, "bad argument");
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid_of(n)), dim3(256), 0, ctx->stream, n, w, g, m, v,
                       lr, beta1, beta2, eps, c1, c2, w_bf16);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int dage_weights_to_bf16(dage_ctx ctx, int n, const float *w, uint16_t *w_bf16) {
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(n > 0 && w && w_bf16, DAGE_ERR_ARG, "bad argument");
    hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_of(n)), dim3(256), 0, ctx->stream, n, w, w_bf16);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int dage_mse_loss(dage_ctx ctx, long n, const float *pred, const float *target, float *grad,
                  double *loss) {
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(n > 0 && pred && target && grad && loss, DAGE_ERR_ARG, "bad argument");
    const int nb = std::min<long>(1024, (n + 255) / 256);
    void *scr = nullptr;
    DAGE_TRY(ctx_scratch(ctx, sizeof(double) * nb, &scr));
    double *part = static_cast<double *>(scr);
    hipLaunchKernelGGL(mse_kernel, dim3(nb), dim3(256), 0, ctx->stream, n, pred, target,
                       (float)(2.0 / (double)n), grad, part);
    DAGE_HIP_CHECK(hipGetLastError());
    std::vector<double> h(nb);
    DAGE_HIP_CHECK(hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                  ctx->stream));
    DAGE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (double v : h) s += v;
    *loss = s / (double)n;
    return DAGE_OK;
}

int dage_cvt_f64_f32(dage_ctx ctx, long n, const double *x, double scale, float *y) {
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(n > 0 && x && y, DAGE_ERR_ARG, "bad argument");
    hipLaunchKernelGGL(cvt_f64_f32_kernel, dim3(grid_of(n)), dim3(256), 0, ctx->stream, n, x,
                       scale, y);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int dage_axpy_f32_f64(dage_ctx ctx, long n, double a, const float *x, double *y) {
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(n > 0 && x && y, DAGE_ERR_ARG, "bad argument");
    hipLaunchKernelGGL(axpy_f32_f64_kernel, dim3(grid_of(n)), dim3(256), 0, ctx->stream, n, a, x,
                       y);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int dage_device_alloc(dage_ctx ctx, size_t bytes, void **ptr) {
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(ptr != nullptr && bytes > 0, DAGE_ERR_ARG, "bad argument");
    DAGE_HIP_CHECK(hipMalloc(ptr, bytes));
    return DAGE_OK;
}

int dage_device_free(dage_ctx ctx, void *ptr) {
    DAGE_TRY(use_device(ctx));
    if (ptr) DAGE_HIP_CHECK(hipFree(ptr));
    return DAGE_OK;
}

int dage_memcpy(dage_ctx ctx, void *dst, const void *src, size_t bytes, int kind) {
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(dst && src, DAGE_ERR_ARG, "bad argument");
    DAGE_REQUIRE(kind >= 0 && kind <= 2, DAGE_ERR_ARG, "kind 0 (h2d), 1 (d2h) or 2 (d2d)");
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                                      : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    DAGE_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, k, ctx->stream));
    DAGE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return DAGE_OK;
}

}  // extern "C"
