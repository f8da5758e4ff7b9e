// Banded FP64 passes for the t-points whose 1D operator matrices are narrow
// bands (x, y and z mode products of the tensor-quadrature apply, the same
// sums as the three passes of apply_f64.hip over a contiguous range of t).
//
// At large t the Gaussian exp(-t^2 (x - x')^2) confines every nonzero entry
// of F_t (after the matrices' tails are stored as zeros) to |o - i| <= H_t
// with H_t of a few points: at 385 points and T = 40 the last 11 t-points
// have H_t <= 8, the last 8 H_t <= 6 (13 taps per output where a 16x16 MFMA
// tile of the band executes over 100; wider bands, whose sums run at a small
// fraction of the VALU peak here, stay on the MFMA passes).  These passes run on the FP64 VALU
// (same 78.6 TFLOP/s peak as FP64 MFMA on gfx950) instead:
//
//   * one thread per (line, R consecutive outputs along the contracted axis);
//     the 64 lanes of a wavefront are 64 consecutive x positions, so every
//     load and store is one contiguous 512-byte segment;
//   * the thread loads the R + 2H input values its outputs need into
//     registers once (each value is used by up to 2H + 1 outputs);
//   * the coefficients depend only on the output index along the axis, which
//     is uniform across the wavefront, so they come through scalar loads;
//   * the z pass sums the t-points of the chunk in registers and writes each
//     output once (partial slab 0 of the split-K reduction, or the output
//     itself when the z pass writes straight to it).
//
// Coefficient tables (built once per operator) hold F_t(o, o - kBandH + j),
// zero where that input index is outside the axis, so a kernel with
// half-width HC <= kBandH reads taps j in [kBandH - HC, kBandH + HC]: the y
// and z passes sum one output at a time from a row-major table (an output's
// taps are one contiguous scalar read); the x pass sums tap by tap over its
// R outputs from a transposed table (a tap's R coefficients are contiguous),
// which gives its FMAs R independent sums.
#include "dage_internal.hpp"

#include <algorithm>
#include <cstdlib>
#include <vector>

namespace dage {
namespace {

// Tile index -> logical order, so that each XCD (hardware hands block L to
// XCD L % 8) works through one contiguous range of tiles: neighbouring
// output blocks of the same lines, whose input windows overlap, then share
// that XCD's L2.
__device__ inline unsigned xcd_order(unsigned L, unsigned total) {
    const unsigned q = total >> 3, r = total & 7u, x = L & 7u;
    return x * q + min(x, r) + (L >> 3);
}

// F_{t0+tl}(o, o - kBandH + j), F_t(o, i) = f[o*so + i*si + t*st] (0 for i
// outside [0, nin) and for the padding o >= nout), stored
//   row-major  (nop == 0): tab[(tl*nout + o)*kBandW + j]
//   transposed (nop > 0):  tab[(tl*kBandW + j)*nop + o], o < nop
__global__ void __launch_bounds__(256) band_table_kernel(const double *f, long so, long si,
                                                         long st, int nout, int nin, int nop,
                                                         int t0, int nt, double *tab) {
    const int no = nop > 0 ? nop : nout;
    const long total = (long)nt * kBandW * no;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        int o, j, tl;
        if (nop > 0) {
            o = (int)(e % nop);
            const long r = e / nop;
            j = (int)(r % kBandW);
            tl = (int)(r / kBandW);
        } else {
            j = (int)(e % kBandW);
            const long r = e / kBandW;
            o = (int)(r % nout);
            tl = (int)(r / nout);
        }
        const int i = o - kBandH + j;
        tab[e] = (o < nout && i >= 0 && i < nin)
                     ? f[(long)o * so + (long)i * si + (long)(t0 + tl) * st]
                     : 0.0;
    }
}

struct BandY {
    const double *V;  // V_{b,tl}(xo, yi, zi) at V + (b*tc + tl)*Vt + xo + nx*(yi + nyi*zi)
    double *U;        // U_{b,tl}(xo, yo, zi) at U + (b*tc + tl)*Ut + xo + nx*(yo + nyo*zi)
    const double *w;  // weight of (b, t): w[b*wS + t0 + tl]
    long wS;
    long Vt, Ut;
    int nx, nyi, nyo, nz;
    int t0, tc, B;    // chunk start and length (tc: stride of b in V and U)
    int tl0, tr;      // this launch: t-points tl0..tl0+tr-1 of the chunk
    long tab_ts;      // nyo * kBandW (row-major table)
    int ngroups, nblk;  // line groups of 256, output blocks along y
};

// U_{b,t}(xo, yo, zi) = w_{b,t} sum_{|yi - yo| <= HC} F_t(yo, yi) V_{b,t}(xo, yi, zi)
template <int HC, int R>
__global__ void __launch_bounds__(256) band_y_kernel(BandY a, const double *__restrict__ tab) {
    const unsigned L = xcd_order(blockIdx.x, gridDim.x);
    const int yb = (int)(L % (unsigned)a.nblk);
    const unsigned rest = L / (unsigned)a.nblk;
    const int lg = (int)(rest % (unsigned)a.ngroups);
    const int bt = (int)(rest / (unsigned)a.ngroups);
    const int tl = a.tl0 + bt % a.tr, b = bt / a.tr;
    const long lines = (long)a.nx * a.nz;
    const long line = (long)lg * 256 + threadIdx.x;
    const bool ok = line < lines;
    const long l = ok ? line : lines - 1;
    const long xo = l % a.nx, zi = l / a.nx;
    const long bt_off = (long)b * a.tc + tl;
    const double *src = a.V + bt_off * a.Vt + xo + (long)a.nx * a.nyi * zi;
    double *dst = a.U + bt_off * a.Ut + xo + (long)a.nx * a.nyo * zi;
    const double alpha = a.w[(long)b * a.wS + a.t0 + tl];
    const int y0 = yb * R;
    const double *tt = tab + (long)(tl - a.tl0) * a.tab_ts + (kBandH - HC);
    double win[R + 2 * HC];
#pragma unroll
    for (int i = 0; i < R + 2 * HC; ++i) {
        const int yi = min(max(y0 - HC + i, 0), a.nyi - 1);  // taps outside carry 0
        win[i] = src[(long)a.nx * yi];
    }
    // one output at a time: its 2HC+1 coefficients are one contiguous scalar
    // read, and the stores stream out as the sums finish
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int yo = y0 + r;
        const double *c = tt + (long)min(yo, a.nyo - 1) * kBandW;
        double s = 0.0;
#pragma unroll
        for (int j = 0; j <= 2 * HC; ++j) s = fma(c[j], win[r + j], s);
        if (ok && yo < a.nyo) dst[(long)a.nx * yo] = alpha * s;
    }
}

struct BandX {
    const double *in;  // in_b(xi, line) at in + b*nin + xi + nxi*line, line = yi + nyi*zi
    double *V;         // V_{b,tl}(xo, line) at V + (b*tc + tl)*Vt + xo + nxo*line
    long nin, Vt, lines;
    int nxi, nxo;
    int tc, tl0, tr, B;  // chunk length (stride of b in V); this launch's t-points
    int nop;             // padded output count of the table rows
    long tab_ts;         // kBandW * nop
    int nxt, ngroups;    // x tiles, line groups of 64
};

// V_{b,t}(xo, line) = sum_{|xi - xo| <= HC} F_t(xo, xi) in_b(xi, line) for the
// launch's t-points.  The contracted axis is the contiguous one, so the
// block stages the input of 64 lines x (4R + 2HC) points in LDS (coalesced),
// the lanes of a wavefront take one line each (coefficients uniform again:
// the wavefront's R outputs share xo) and hold their R + 2HC window in
// registers for every t of the launch, and the 64 x 4R outputs go back
// through LDS so the stores run along x.  Odd LDS row strides keep the
// lane-per-line accesses on distinct banks.
template <int HC, int R>
__global__ void __launch_bounds__(256) band_x_kernel(BandX a, const double *__restrict__ tab) {
    constexpr int XT = 4 * R;        // outputs per block along x
    constexpr int WI = XT + 2 * HC;  // input points per line
    constexpr int LSI = WI | 1, LSO = XT | 1;
    // the input tile is in registers (win) before the first output tile is
    // written, so both share one LDS buffer
    __shared__ double S[64 * LSI];
    double *Sin = S, *Sout = S;
    const unsigned L = xcd_order(blockIdx.x, gridDim.x);
    const int xt = (int)(L % (unsigned)a.nxt);
    const unsigned rest = L / (unsigned)a.nxt;
    const int lg = (int)(rest % (unsigned)a.ngroups);
    const int b = (int)(rest / (unsigned)a.ngroups);
    const int x0 = xt * XT;
    const long l0 = (long)lg * 64;
    const int tid = threadIdx.x, lane = tid & 63;
    // wavefront-uniform, made visible to the compiler so that the coefficient
    // addresses stay scalar
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const double *src = a.in + (long)b * a.nin;
    for (int e = tid; e < 64 * WI; e += 256) {
        const int line = e / WI, xl = e - line * WI;
        const long l = min(l0 + line, a.lines - 1);
        const int x = min(max(x0 - HC + xl, 0), a.nxi - 1);  // taps outside carry 0
        Sin[line * LSI + xl] = src[x + (long)a.nxi * l];
    }
    __syncthreads();
    double win[R + 2 * HC];
#pragma unroll
    for (int i = 0; i < R + 2 * HC; ++i) win[i] = Sin[lane * LSI + wave * R + i];
    __syncthreads();  // every wavefront holds its window: the buffer takes outputs now
    for (int k = 0; k < a.tr; ++k) {
        const double *tt =
            tab + (long)k * a.tab_ts + (long)(kBandH - HC) * a.nop + x0 + wave * R;
        double acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = 0.0;
#pragma unroll
        for (int j = 0; j <= 2 * HC; ++j) {
            const double *c = tt + (long)j * a.nop;
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] = fma(c[r], win[r + j], acc[r]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) Sout[lane * LSO + wave * R + r] = acc[r];
        __syncthreads();
        double *dst = a.V + ((long)b * a.tc + a.tl0 + k) * a.Vt;
        for (int e = tid; e < 64 * XT; e += 256) {
            const int line = e / XT, xl = e - line * XT;
            const int xo = x0 + xl;
            const long l = l0 + line;
            if (xo < a.nxo && l < a.lines) dst[xo + (long)a.nxo * l] = Sout[line * LSO + xl];
        }
        __syncthreads();
    }
}

struct BandXY {
    const double *in;  // in_b(xi, yi, zi) at in + b*nin + xi + nxi*(yi + nyi*zi)
    double *U;         // U_{b,tl}(xo, yo, zi) at U + (b*tc + tl)*Ut + xo + nxo*(yo + nyo*zi)
    const double *w;   // weight of (b, t): w[b*wS + t0 + tl]
    long wS, nin, Ut;
    int nxi, nyi, nzi, nxo, nyo;
    int t0, tc, tl0, tr, B;  // chunk start and length; this launch's t-points
    int nopx;                // padded row length of the transposed x table
    long tabx_ts, taby_ts;   // per-t strides of the x (transposed) and y (row-major) tables
    int nxt, nyt;            // tiles along x and y
};

// The banded x and y passes of one (right-hand side, z-plane) tile without
// V in memory: U(xo, yo) = w_t sum_yi F_y(yo, yi) sum_xi F_x(xo, xi) in(xi, yi)
// for xo in a tile of 4R, yo in a tile of 4RY, and the launch's t-points.
// The block stages in(x-window, y-window) in LDS (64 rows = 4RY + 2HC y
// points); the x sums run as in band_x_kernel (one y row per lane, window in
// registers for every t, tap by tap over R outputs); V (64 rows x 4R) goes
// through the same LDS buffer to the y sums, which run as in band_y_kernel
// (one x point per lane, coefficients uniform, an output's taps in order).
// The same sums in the same order as the two kernels: bit-identical U.
template <int HC, int R, int RY>
__global__ void __launch_bounds__(256) band_xy_kernel(BandXY a, const double *__restrict__ tabx,
                                                      const double *__restrict__ taby) {
    constexpr int XT = 4 * R, YT = 4 * RY;
    static_assert(YT + 2 * HC == 64, "one y row per lane");
    constexpr int WI = XT + 2 * HC;  // input points per row
    constexpr int LSI = WI | 1, LSV = XT | 1;
    __shared__ double S[64 * LSI];
    const unsigned L = xcd_order(blockIdx.x, gridDim.x);
    const int xt = (int)(L % (unsigned)a.nxt);
    unsigned rest = L / (unsigned)a.nxt;
    const int yt = (int)(rest % (unsigned)a.nyt);
    rest /= (unsigned)a.nyt;
    const int zi = (int)(rest % (unsigned)a.nzi);
    const int b = (int)(rest / (unsigned)a.nzi);
    const int x0 = xt * XT, y0 = yt * YT;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const double *src = a.in + (long)b * a.nin + (long)a.nxi * a.nyi * zi;
    for (int e = tid; e < 64 * WI; e += 256) {
        const int row = e / WI, xl = e - row * WI;
        const int yi = min(max(y0 - HC + row, 0), a.nyi - 1);  // rows outside carry 0 weight
        const int x = min(max(x0 - HC + xl, 0), a.nxi - 1);
        S[row * LSI + xl] = src[x + (long)a.nxi * yi];
    }
    __syncthreads();
    double win[R + 2 * HC];
#pragma unroll
    for (int i = 0; i < R + 2 * HC; ++i) win[i] = S[lane * LSI + wave * R + i];
    __syncthreads();  // the buffer holds V from here on
    for (int k = 0; k < a.tr; ++k) {
        const int tl = a.tl0 + k;
        // x sums: V(row = lane, xo = x0 + wave*R + r)
        const double *tt =
            tabx + (long)k * a.tabx_ts + (long)(kBandH - HC) * a.nopx + x0 + wave * R;
        double acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = 0.0;
#pragma unroll
        for (int j = 0; j <= 2 * HC; ++j) {
            const double *c = tt + (long)j * a.nopx;
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] = fma(c[r], win[r + j], acc[r]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) S[lane * LSV + wave * R + r] = acc[r];
        __syncthreads();
        // y sums: U(xo = x0 + lane, yo = y0 + wave*RY + ry)
        if (lane < XT) {
            const int xo = x0 + lane;
            double wy[RY + 2 * HC];
#pragma unroll
            for (int i = 0; i < RY + 2 * HC; ++i) wy[i] = S[(wave * RY + i) * LSV + lane];
            const double alpha = a.w[(long)b * a.wS + a.t0 + tl];
            const double *ty = taby + (long)k * a.taby_ts + (kBandH - HC);
            double *dst = a.U + ((long)b * a.tc + tl) * a.Ut + xo + (long)a.nxo * a.nyo * zi;
#pragma unroll
            for (int ry = 0; ry < RY; ++ry) {
                const int yo = y0 + wave * RY + ry;
                const double *c = ty + (long)min(yo, a.nyo - 1) * kBandW;
                double s = 0.0;
#pragma unroll
                for (int j = 0; j <= 2 * HC; ++j) s = fma(c[j], wy[ry + j], s);
                if (xo < a.nxo && yo < a.nyo) dst[(long)a.nxo * yo] = alpha * s;
            }
        }
        __syncthreads();  // V read before the next t overwrites it
    }
}

struct BandZ {
    const double *U;  // U_{b,tl}(l, zi) at U + (b*tc + tl)*Ut + l + M3*zi, l = xo + nxo*yo
    double *C;        // C_b(l, zo) at C + b*cS + l + M3*zo
    long Ut, cS, M3;
    int nzi, nzo, tc, B;
    int accumulate;
    long tab_ts;      // nzo * kBandW (row-major table)
    int ngroups, nblk;
};

// C_b(l, zo) (+)= sum_tl sum_{|zi - zo| <= HC} F_t(zo, zi) U_{b,tl}(l, zi)
template <int HC, int R>
__global__ void __launch_bounds__(256) band_z_kernel(BandZ a, const double *__restrict__ tab) {
    const unsigned L = xcd_order(blockIdx.x, gridDim.x);
    const int zb = (int)(L % (unsigned)a.nblk);
    const unsigned rest = L / (unsigned)a.nblk;
    const int lg = (int)(rest % (unsigned)a.ngroups);
    const int b = (int)(rest / (unsigned)a.ngroups);
    const long line = (long)lg * 256 + threadIdx.x;
    const bool ok = line < a.M3;
    const long l = ok ? line : a.M3 - 1;
    const int z0 = zb * R;
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    for (int tl = 0; tl < a.tc; ++tl) {
        const double *src = a.U + ((long)b * a.tc + tl) * a.Ut + l;
        const double *tt = tab + (long)tl * a.tab_ts + (kBandH - HC);
        double win[R + 2 * HC];
#pragma unroll
        for (int i = 0; i < R + 2 * HC; ++i) {
            const int zi = min(max(z0 - HC + i, 0), a.nzi - 1);
            win[i] = src[a.M3 * zi];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double *c = tt + (long)min(z0 + r, a.nzo - 1) * kBandW;
            double s = acc[r];
#pragma unroll
            for (int j = 0; j <= 2 * HC; ++j) s = fma(c[j], win[r + j], s);
            acc[r] = s;
        }
    }
    double *dst = a.C + (long)b * a.cS + l;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int zo = z0 + r;
        if (ok && zo < a.nzo) {
            double *p = dst + a.M3 * zo;
            *p = a.accumulate ? *p + acc[r] : acc[r];
        }
    }
}

// Outputs per thread along the contracted axis for each half-width class:
// R + 2HC input values stay in registers.
template <int HC> constexpr int band_rows() { return 16; }

// One half-width class: every banded t-point runs the kBandH kernels (a
// narrower band only carries zero taps).
int band_class(int) { return kBandH; }

// Table rows padded past the last output (zeros) by more than the widest
// output tile (4R = 56 points in the x pass), so the coefficients a tile
// reads for its outputs past the axis end stay inside the row.
int band_nop(long n) { return (int)n + 64; }

// x pass: R outputs per wavefront along x (4R per block, 37 KB of LDS: four
// blocks per CU); 4R = 56 leaves little of a 385-point axis to a last,
// mostly empty tile
template <int HC> constexpr int band_rows_x() { return 14; }

template <int HC>
int launch_x(dage_ctx ctx, BandX a, const double *tab, const char *name) {
    constexpr int R = band_rows_x<HC>();
    a.nxt = (a.nxo + 4 * R - 1) / (4 * R);
    a.ngroups = (int)((a.lines + 63) / 64);
    const long blocks = (long)a.nxt * a.ngroups * a.B;
    DAGE_REQUIRE(blocks < (1L << 31), DAGE_ERR_ARG, "banded x pass too large (%ld blocks)", blocks);
    hipEvent_t ev;
    prof_begin(ctx, name, &ev);
    hipLaunchKernelGGL((band_x_kernel<HC, R>), dim3((unsigned)blocks), dim3(256), 0, ctx->stream, a,
                       tab);
    DAGE_HIP_CHECK(hipGetLastError());
    prof_end(ctx, name, ev);
    return DAGE_OK;
}

template <int HC>
int launch_y(dage_ctx ctx, BandY a, const double *tab, const char *name) {
    constexpr int R = band_rows<HC>();
    a.ngroups = (int)(((long)a.nx * a.nz + 255) / 256);
    a.nblk = (a.nyo + R - 1) / R;
    const long blocks = (long)a.ngroups * a.nblk * a.tr * a.B;
    DAGE_REQUIRE(blocks < (1L << 31), DAGE_ERR_ARG, "banded y pass too large (%ld blocks)", blocks);
    hipEvent_t ev;
    prof_begin(ctx, name, &ev);
    hipLaunchKernelGGL((band_y_kernel<HC, R>), dim3((unsigned)blocks), dim3(256), 0, ctx->stream, a,
                       tab);
    DAGE_HIP_CHECK(hipGetLastError());
    prof_end(ctx, name, ev);
    return DAGE_OK;
}

template <int HC>
int launch_z(dage_ctx ctx, BandZ a, const double *tab, const char *name) {
    constexpr int R = band_rows<HC>();
    a.ngroups = (int)((a.M3 + 255) / 256);
    a.nblk = (a.nzo + R - 1) / R;
    const long blocks = (long)a.ngroups * a.nblk * a.B;
    DAGE_REQUIRE(blocks < (1L << 31), DAGE_ERR_ARG, "banded z pass too large (%ld blocks)", blocks);
    hipEvent_t ev;
    prof_begin(ctx, name, &ev);
    hipLaunchKernelGGL((band_z_kernel<HC, R>), dim3((unsigned)blocks), dim3(256), 0, ctx->stream, a,
                       tab);
    DAGE_HIP_CHECK(hipGetLastError());
    prof_end(ctx, name, ev);
    return DAGE_OK;
}

// DAGE_APPLY_BAND: 0 no banded passes, 2 banded passes at every size,
// otherwise (default 1) where the launch is large enough to fill the GPU
int band_mode() {
    const char *e = std::getenv("DAGE_APPLY_BAND");
    if (!e) return 1;
    return e[0] == '0' ? 0 : (e[0] == '2' ? 2 : 1);
}

}  // namespace

int build_bands(dage_operator op, const std::vector<int> &span) {
    const long nxi = op->din[0], nyi = op->din[1], nzi = op->din[2];
    const long nxo = op->dout[0], nyo = op->dout[1], nzo = op->dout[2];
    const int T = op->ntp;
    const long rx = (long)T * nxi, ry = (long)T * nyi;
    // half-width of t: the largest |o - i| of a nonzero entry of Fx_t, Fy_t or Fz_t
    // (spans: first and last nonzero output of each input row)
    op->band_h.assign(T, 0);
    auto widen = [&](long row, long i, int t) {
        const int lo = span[2 * row], hi = span[2 * row + 1];
        if (hi < 0) return;
        op->band_h[t] = std::max(op->band_h[t], (int)std::max(i - lo, (long)hi - i));
    };
    for (int t = 0; t < T; ++t) {
        for (long xi = 0; xi < nxi; ++xi) widen(t * nxi + xi, xi, t);
        for (long yi = 0; yi < nyi; ++yi) widen(rx + t * nyi + yi, yi, t);
        for (long zi = 0; zi < nzi; ++zi) widen(rx + ry + zi + nzi * t, zi, t);
    }
    int t0 = T;
    while (t0 > 0 && op->band_h[t0 - 1] <= kBandH) --t0;
    if (op->d_band) DAGE_HIP_CHECK(hipFree(op->d_band));
    op->d_band = nullptr;
    op->band_t0 = T;
    op->band_mode = band_mode();
    if (t0 == T || op->band_mode == 0) return DAGE_OK;
    const int nt = T - t0;
    const size_t ny = (size_t)nt * kBandW * nyo, nz = (size_t)nt * kBandW * nzo;
    const size_t nx = (size_t)nt * kBandW * band_nop(nxo);
    DAGE_HIP_CHECK(hipMalloc((void **)&op->d_band, 8 * (ny + nz + nx)));
    hipStream_t s = op->ctx->stream;
    auto grid = [](size_t n) { return (int)std::min<size_t>((n + 255) / 256, 65536); };
    // Fy (nyi, nyo, T): F_t(yo, yi) = fy[yi + nyi*yo + nyi*nyo*t]
    hipLaunchKernelGGL(band_table_kernel, dim3(grid(ny)), dim3(256), 0, s, op->d_fy, nyi, 1L,
                       nyi * nyo, (int)nyo, (int)nyi, 0, t0, nt, op->d_band);
    DAGE_HIP_CHECK(hipGetLastError());
    // Fz (nzi, nzo, T): F_t(zo, zi) = fz[zi + nzi*zo + nzi*nzo*t]
    hipLaunchKernelGGL(band_table_kernel, dim3(grid(nz)), dim3(256), 0, s, op->d_fz, nzi, 1L,
                       nzi * nzo, (int)nzo, (int)nzi, 0, t0, nt, op->d_band + ny);
    DAGE_HIP_CHECK(hipGetLastError());
    // Fx (nxo, nxi, T): F_t(xo, xi) = fx[xo + nxo*xi + nxo*nxi*t]
    hipLaunchKernelGGL(band_table_kernel, dim3(grid(nx)), dim3(256), 0, s, op->d_fx, 1L, nxo,
                       nxo * nxi, (int)nxo, (int)nxi, band_nop(nxo), t0, nt,
                       op->d_band + ny + nz);
    DAGE_HIP_CHECK(hipGetLastError());
    op->band_off[0] = ny;
    op->band_off[1] = ny + nz;
    op->band_t0 = t0;
    return DAGE_OK;
}

namespace {

// Runs of consecutive t-points of one half-width class within [t0, t0+tc):
// calls fn(first, count, class) for each.
template <typename Fn>
int for_runs(const ApplyMats &m, int t0, int tc, Fn fn) {
    DAGE_REQUIRE(t0 >= m.band_t0 && t0 + tc <= m.ntp, DAGE_ERR_STATE,
                 "t-points %d..%d are not banded", t0, t0 + tc - 1);
    for (int a = t0; a < t0 + tc;) {
        DAGE_REQUIRE(m.band_h[a] <= kBandH, DAGE_ERR_STATE, "t-point %d is not banded", a);
        const int c = band_class(m.band_h[a]);
        int e = a + 1;
        while (e < t0 + tc && band_class(m.band_h[e]) == c) ++e;
        DAGE_TRY(fn(a, e - a, c));
        a = e;
    }
    return DAGE_OK;
}

}  // namespace

int band_pass_xy(dage_ctx ctx, const ApplyMats &m, const double *in, double *U, const double *w,
                 long wS, int t0, int tc, int B, const char *name) {
    constexpr int R = 14, RY = 12;
    BandXY a{};
    a.in = in;
    a.U = U;
    a.w = w;
    a.wS = wS;
    a.nxi = m.din[0]; a.nyi = m.din[1]; a.nzi = m.din[2];
    a.nxo = m.dout[0]; a.nyo = m.dout[1];
    a.nin = (long)a.nxi * a.nyi * a.nzi;
    a.Ut = (long)a.nxo * a.nyo * a.nzi;
    a.t0 = t0; a.tc = tc; a.B = B;
    a.nopx = band_nop(a.nxo);
    a.tabx_ts = (long)kBandW * a.nopx;
    a.taby_ts = (long)a.nyo * kBandW;
    a.nxt = (a.nxo + 4 * R - 1) / (4 * R);
    a.nyt = (a.nyo + 4 * RY - 1) / (4 * RY);
    return for_runs(m, t0, tc, [&](int first, int count, int) {
        BandXY r = a;
        r.tl0 = first - t0;
        r.tr = count;
        const double *tx = m.band_x + (long)(first - m.band_t0) * a.tabx_ts;
        const double *ty = m.band_y + (long)(first - m.band_t0) * a.taby_ts;
        const long blocks = (long)r.nxt * r.nyt * r.nzi * r.B;
        DAGE_REQUIRE(blocks < (1L << 31), DAGE_ERR_ARG, "banded x/y pass too large (%ld blocks)",
                     blocks);
        hipEvent_t ev;
        prof_begin(ctx, name, &ev);
        hipLaunchKernelGGL((band_xy_kernel<kBandH, R, RY>), dim3((unsigned)blocks), dim3(256), 0,
                           ctx->stream, r, tx, ty);
        DAGE_HIP_CHECK(hipGetLastError());
        prof_end(ctx, name, ev);
        return DAGE_OK;
    });
}

bool band_xy_fused() {
    const char *e = std::getenv("DAGE_APPLY_BANDXY");
    return !(e && e[0] == '0');
}

bool band_pays(const ApplyMats &m, int B) {
    if (!m.band_y) return false;
    if (m.band_mode == 2) return true;
    // the z pass has the fewest blocks (no t in its grid): below ~4 per CU
    // the MFMA passes, whose z pass splits K, fill the GPU better
    const long m3 = (long)m.dout[0] * m.dout[1];
    const long zblocks = (m3 + 255) / 256 * ((m.dout[2] + band_rows<kBandH>() - 1) / band_rows<kBandH>()) * B;
    return zblocks >= 1024;
}

int band_pass_x(dage_ctx ctx, const ApplyMats &m, const double *in, double *V, int t0, int tc,
                int B, const char *name) {
    BandX a{};
    a.in = in;
    a.V = V;
    a.nxi = m.din[0];
    a.nxo = m.dout[0];
    a.lines = (long)m.din[1] * m.din[2];
    a.nin = (long)a.nxi * a.lines;
    a.Vt = (long)a.nxo * a.lines;
    a.tc = tc;
    a.B = B;
    a.nop = band_nop(a.nxo);
    a.tab_ts = (long)kBandW * a.nop;
    return for_runs(m, t0, tc, [&](int first, int count, int c) {
        BandX r = a;
        r.tl0 = first - t0;
        r.tr = count;
        const double *tab = m.band_x + (long)(first - m.band_t0) * a.tab_ts;
        (void)c;
        return launch_x<kBandH>(ctx, r, tab, name);
    });
}

int band_pass_y(dage_ctx ctx, const ApplyMats &m, const double *V, double *U, const double *w,
                long wS, int t0, int tc, int B, const char *name) {
    const int nx = m.dout[0], nyi = m.din[1], nyo = m.dout[1], nz = m.din[2];
    BandY a{};
    a.V = V;
    a.U = U;
    a.w = w;
    a.wS = wS;
    a.Vt = (long)nx * nyi * nz;
    a.Ut = (long)nx * nyo * nz;
    a.nx = nx; a.nyi = nyi; a.nyo = nyo; a.nz = nz;
    a.t0 = t0; a.tc = tc; a.B = B;
    a.tab_ts = (long)nyo * kBandW;
    return for_runs(m, t0, tc, [&](int first, int count, int c) {
        BandY r = a;
        r.tl0 = first - t0;
        r.tr = count;
        const double *tab = m.band_y + (long)(first - m.band_t0) * a.tab_ts;
        (void)c;
        return launch_y<kBandH>(ctx, r, tab, name);
    });
}

int band_pass_z(dage_ctx ctx, const ApplyMats &m, const double *U, double *C, long cS,
                bool accumulate, int t0, int tc, int B, const char *name) {
    const int nzi = m.din[2], nzo = m.dout[2];
    int h = 0;
    for (int t = t0; t < t0 + tc; ++t) h = std::max(h, m.band_h[t]);
    DAGE_REQUIRE(t0 >= m.band_t0 && h <= kBandH, DAGE_ERR_STATE, "t-points %d..%d are not banded",
                 t0, t0 + tc - 1);
    BandZ a{};
    a.M3 = (long)m.dout[0] * m.dout[1];
    a.U = U;
    a.C = C;
    a.Ut = a.M3 * nzi;
    a.cS = cS;
    a.nzi = nzi; a.nzo = nzo; a.tc = tc; a.B = B;
    a.accumulate = accumulate ? 1 : 0;
    a.tab_ts = (long)nzo * kBandW;
    const double *tab = m.band_z + (long)(t0 - m.band_t0) * a.tab_ts;
    return launch_z<kBandH>(ctx, a, tab, name);
}

}  // namespace dage
