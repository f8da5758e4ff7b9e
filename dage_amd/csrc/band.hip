// Banded FP64 passes for the t-points whose 1D operator matrices are narrow
// bands (y and z mode products of the tensor-quadrature apply, the same sums
// as passes 2 and 3 of apply_f64.hip over a contiguous range of t).
//
// At large t the Gaussian exp(-t^2 (x - x')^2) confines every nonzero entry
// of F_t (after the matrices' tails are stored as zeros) to |o - i| <= H_t
// with H_t of a few points: at 385 points and T = 40 the last 16 t-points
// have H_t <= 24, the last 8 H_t <= 6 (13 taps per output where a 16x16 MFMA
// tile of the band executes over 100).  These passes run on the FP64 VALU
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
// Coefficient table (built once per operator): tab[((t - t0)*n_out + o)*kBandW
// + j] = F_t(o, o - kBandH + j), zero where that input index is outside the
// axis, so a kernel with half-width HC <= kBandH reads taps
// j in [kBandH - HC, kBandH + HC].
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

// tab[((tl*nout) + o)*kBandW + j] = F_{t0+tl}(o, o - kBandH + j) with
// F_t(o, i) = f[o*so + i*si + t*st] (0 for i outside [0, nin)).
__global__ void __launch_bounds__(256) band_table_kernel(const double *f, long so, long si,
                                                         long st, int nout, int nin, int t0,
                                                         int nt, double *tab) {
    const long total = (long)nt * nout * kBandW;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int j = (int)(e % kBandW);
        const long r = e / kBandW;
        const int o = (int)(r % nout);
        const int tl = (int)(r / nout);
        const int i = o - kBandH + j;
        tab[e] = (i >= 0 && i < nin) ? f[(long)o * so + (long)i * si + (long)(t0 + tl) * st] : 0.0;
    }
}

struct BandY {
    const double *V;  // V_{b,tl}(xo, yi, zi) at V + (b*tc + tl)*Vt + xo + nx*(yi + nyi*zi)
    double *U;        // U_{b,tl}(xo, yo, zi) at U + (b*tc + tl)*Ut + xo + nx*(yo + nyo*zi)
    const double *w;  // weight of (b, t): w[b*wS + t0 + tl]
    long wS;
    long Vt, Ut;
    int nx, nyi, nyo, nz;
    int t0, tc, B;
    long tab_ts;      // nyo * kBandW
    int ngroups, nblk;  // line groups of 256, output blocks along y
};

// U_{b,t}(xo, yo, zi) = w_{b,t} sum_{|yi - yo| <= HC} F_t(yo, yi) V_{b,t}(xo, yi, zi)
template <int HC, int R>
__global__ void __launch_bounds__(256) band_y_kernel(BandY a, const double *__restrict__ tab) {
    const unsigned L = xcd_order(blockIdx.x, gridDim.x);
    const int yb = (int)(L % (unsigned)a.nblk);
    const unsigned rest = L / (unsigned)a.nblk;
    const int lg = (int)(rest % (unsigned)a.ngroups);
    const int bt = (int)(rest / (unsigned)a.ngroups);  // b*tc + tl
    const int tl = bt % a.tc, b = bt / a.tc;
    const long lines = (long)a.nx * a.nz;
    const long line = (long)lg * 256 + threadIdx.x;
    const bool ok = line < lines;
    const long l = ok ? line : lines - 1;
    const long xo = l % a.nx, zi = l / a.nx;
    const double *src = a.V + (long)bt * a.Vt + xo + (long)a.nx * a.nyi * zi;
    double *dst = a.U + (long)bt * a.Ut + xo + (long)a.nx * a.nyo * zi;
    const double alpha = a.w[(long)b * a.wS + a.t0 + tl];
    const double *tt = tab + (long)tl * a.tab_ts + (kBandH - HC);
    const int y0 = yb * R;
    double win[R + 2 * HC];
#pragma unroll
    for (int i = 0; i < R + 2 * HC; ++i) {
        const int yi = min(max(y0 - HC + i, 0), a.nyi - 1);  // taps outside carry 0
        win[i] = src[(long)a.nx * yi];
    }
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

struct BandZ {
    const double *U;  // U_{b,tl}(l, zi) at U + (b*tc + tl)*Ut + l + M3*zi, l = xo + nxo*yo
    double *C;        // C_b(l, zo) at C + b*cS + l + M3*zo
    long Ut, cS, M3;
    int nzi, nzo, tc, B;
    int accumulate;
    long tab_ts;      // nzo * kBandW
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
template <int HC> constexpr int band_rows() { return HC <= 8 ? 16 : (HC <= 16 ? 12 : 8); }

int band_class(int h) { return h <= 8 ? 8 : (h <= 16 ? 16 : 24); }

template <int HC>
int launch_y(dage_ctx ctx, BandY a, const double *tab, const char *name) {
    constexpr int R = band_rows<HC>();
    a.ngroups = (int)(((long)a.nx * a.nz + 255) / 256);
    a.nblk = (a.nyo + R - 1) / R;
    const long blocks = (long)a.ngroups * a.nblk * a.tc * a.B;
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

bool band_enabled() {
    const char *e = std::getenv("DAGE_APPLY_BAND");
    return !(e && e[0] == '0');
}

}  // namespace

int build_bands(dage_operator op, const std::vector<int> &span) {
    const long nxi = op->din[0], nyi = op->din[1], nzi = op->din[2];
    const long nyo = op->dout[1], nzo = op->dout[2];
    const int T = op->ntp;
    const long rx = (long)T * nxi, ry = (long)T * nyi;
    // half-width of t: the largest |o - i| of a nonzero entry of Fy_t or Fz_t
    // (spans: first and last nonzero output of each input row)
    op->band_h.assign(T, 0);
    auto widen = [&](long row, long i, int t) {
        const int lo = span[2 * row], hi = span[2 * row + 1];
        if (hi < 0) return;
        op->band_h[t] = std::max(op->band_h[t], (int)std::max(i - lo, (long)hi - i));
    };
    for (int t = 0; t < T; ++t) {
        for (long yi = 0; yi < nyi; ++yi) widen(rx + t * nyi + yi, yi, t);
        for (long zi = 0; zi < nzi; ++zi) widen(rx + ry + zi + nzi * t, zi, t);
    }
    int t0 = T;
    while (t0 > 0 && op->band_h[t0 - 1] <= kBandH) --t0;
    if (op->d_band) DAGE_HIP_CHECK(hipFree(op->d_band));
    op->d_band = nullptr;
    op->band_t0 = T;
    if (t0 == T || !band_enabled()) return DAGE_OK;
    const int nt = T - t0;
    const size_t ny = (size_t)nt * nyo * kBandW, nz = (size_t)nt * nzo * kBandW;
    DAGE_HIP_CHECK(hipMalloc((void **)&op->d_band, 8 * (ny + nz)));
    hipStream_t s = op->ctx->stream;
    auto grid = [](size_t n) { return (int)std::min<size_t>((n + 255) / 256, 65536); };
    // Fy (nyi, nyo, T): F_t(yo, yi) = fy[yi + nyi*yo + nyi*nyo*t]
    hipLaunchKernelGGL(band_table_kernel, dim3(grid(ny)), dim3(256), 0, s, op->d_fy, nyi, 1L,
                       nyi * nyo, (int)nyo, (int)nyi, t0, nt, op->d_band);
    DAGE_HIP_CHECK(hipGetLastError());
    // Fz (nzi, nzo, T): F_t(zo, zi) = fz[zi + nzi*zo + nzi*nzo*t]
    hipLaunchKernelGGL(band_table_kernel, dim3(grid(nz)), dim3(256), 0, s, op->d_fz, nzi, 1L,
                       nzi * nzo, (int)nzo, (int)nzi, t0, nt, op->d_band + ny);
    DAGE_HIP_CHECK(hipGetLastError());
    op->band_t0 = t0;
    return DAGE_OK;
}

int band_pass_y(dage_ctx ctx, const ApplyMats &m, const double *V, double *U, const double *w,
                long wS, int t0, int tc, int B, const char *name) {
    const int nx = m.dout[0], nyi = m.din[1], nyo = m.dout[1], nz = m.din[2];
    int h = 0;
    for (int t = t0; t < t0 + tc; ++t) h = std::max(h, m.band_h[t]);
    DAGE_REQUIRE(t0 >= m.band_t0 && h <= kBandH, DAGE_ERR_STATE, "t-points %d..%d are not banded",
                 t0, t0 + tc - 1);
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
    const double *tab = m.band_y + (long)(t0 - m.band_t0) * a.tab_ts;
    switch (band_class(h)) {
    case 8: return launch_y<8>(ctx, a, tab, name);
    case 16: return launch_y<16>(ctx, a, tab, name);
    default: return launch_y<24>(ctx, a, tab, name);
    }
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
    switch (band_class(h)) {
    case 8: return launch_z<8>(ctx, a, tab, name);
    case 16: return launch_z<16>(ctx, a, tab, name);
    default: return launch_z<24>(ctx, a, tab, name);
    }
}

}  // namespace dage
