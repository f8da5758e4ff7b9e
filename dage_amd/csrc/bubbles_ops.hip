// Bubbles arithmetic on the GPU: allocation and copies of bubble sets that
// share radial grids, and the radial parts of the Coulomb and Helmholtz
// operators acting on bubbles:
//   Coulomb3D_evaluate_bubbles_potential (src/bubbles/coulomb.F90:222-336)
//     V_lm(r) = 4pi/(2l+1) [ r^-(l+1) int_0^r s^(l+2+k) f_lm(s) ds
//                            + r^l int_r^rmax s^(1-l+k) f_lm(s) ds ]
//   Helmholtz3D_transform_bubbles_sub (src/bubbles/helmholtz3d.F90:519-657)
//     V_lm(r) = 8 kappa [ k_l(kappa r) int_0^r s^(2+k) i_l(kappa s) f_lm ds
//                         + i_l(kappa r) int_r^rmax s^(2+k) k_l(kappa s) f_lm ds ]
// with the cumulative integrals of the radial LIP interpolant
// (InOutIntegrator1D, src/bubbles/evaluators.F90:1263-1320; cell tables
// LIPBasis_inward/outward_integrals, src/bubbles/lip.F90:414-485) and the
// modified spherical Bessel functions i_l, k_l (k_0 = pi/2 e^-x / x) of
// helmholtz3d.F90:879-1210.
//
// One wavefront per (bubble, l): the lanes evaluate the integrands on the
// radial points (Bessel values by the reference's series and recursions, in
// FP64 where the reference uses REAL128: ~1 ulp), then one lane per cell
// forms the cell's cumulative sums, and the carries between cells are added
// in the reference's sequential order, so every output point is summed
// exactly as the reference sums it.
#include <cmath>
#include <cstring>
#include <vector>

#include "dage_internal.hpp"

namespace dage {
namespace {

constexpr int kMaxRadial = 1024;  // radial points per bubble in the transforms

// x^n for integer n by binary powering (the gfortran/libgcc order of x**n)
__device__ __host__ inline double powi(double x, int m) {
    unsigned n = m < 0 ? -m : m;
    double y = (n % 2) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n % 2) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

// Bubbles_rpow (bubbles.F90:3363-3378): r^n, and 0 at r = 0 for n < 0
__device__ inline double rpow(double r, int i, int n) {
    return (n < 0 && i == 0) ? 0.0 : powi(r, n);
}

// FirstModSphBesselCollection_evaluate_small (helmholtz3d.F90:935-980):
// z^n/(2n+1)!! sum_i (z^2/2)^i / (i! (2n+3)(2n+5)...(2n+2i+1)), the same
// terms to the same tolerance.  The reference accumulates the power and the
// two factorials separately in REAL128; in FP64 those overflow beyond
// z ~ 100, so here each term is formed from the previous one (the terms
// themselves stay below e^z).
__device__ double bessel_i_series(double z, int n) {
    double prefactor = powi(z, n);
    double divider = 1.0;
    for (int i = 1; i <= 2 * n + 1; i += 2) divider = divider * i;
    prefactor = prefactor / divider;
    const double a = 0.5 * (z * z);
    double addition = 1.0, reslt = 1.0;
    for (int i = 1; addition > 1e-15 * reslt && i < 300; ++i) {
        addition = addition * (a / (i * (double)(2 * n + 2 * i + 1)));
        reslt = reslt + addition;
    }
    return prefactor * reslt;
}

// i_l(x) for one l: series at lmax and lmax-1, downward recursion to l
// (FirstModSphBesselCollection_evaluate, helmholtz3d.F90:879-933)
__device__ double bessel_i(int l, int lmax, double x) {
    if (!(fabs(x) > 0.0)) return l == 0 ? 1.0 : 0.0;
    if (l >= lmax - 1) return bessel_i_series(x, l);
    double r2 = bessel_i_series(x, lmax), r1 = bessel_i_series(x, lmax - 1);
    for (int j = lmax - 2; j >= l; --j) {
        const double r0 = (2 * j + 3) * r1 / x + r2;
        r2 = r1;
        r1 = r0;
    }
    return r1;
}

// k_l(x): closed forms for l = 0, 1 and upward recursion
// (SecondModSphBesselCollection_evaluate, helmholtz3d.F90:1133-1174)
__device__ double bessel_k(int l, double x) {
    if (!(fabs(x) > 0.0)) return 0.0;
    const double pi = 3.141592653589793;
    const double e = exp(-x);
    double k0 = e / x * pi / 2;
    if (l == 0) return k0;
    double k1 = e * (x + 1) / (x * x) * pi / 2;
    for (int j = 2; j <= l; ++j) {
        const double k2 = (2 * j - 1) * k1 / x + k0;
        k0 = k1;
        k1 = k2;
    }
    return k1;
}

// e^x k_l(x): the same recursion from (pi/2)/x and (pi/2)(x+1)/x^2
__device__ double bessel_k_scaled(int l, double x) {
    const double pi = 3.141592653589793;
    double k0 = pi / 2 / x;
    if (l == 0) return k0;
    double k1 = (x + 1) / (x * x) * pi / 2;
    for (int j = 2; j <= l; ++j) {
        const double k2 = (2 * j - 1) * k1 / x + k0;
        k0 = k1;
        k1 = k2;
    }
    return k1;
}

// Above this kappa * (cell length) the LIP interpolant of k_l(kappa s) f(s)
// or i_l(kappa s) f(s) over one cell carries errors of order exp(kappa L)
// into the result (the reference's scheme); such bubbles use the
// exponentially fitted one.
constexpr double kFittedKappaL = 4.0;

constexpr int kRadialChunk = 96;  // bubbles per launch (pointer tables by value)

struct RadialArgs {
    const BubbleDev *bubs;   // input bubbles (grids shared with the output)
    const double *fin[kRadialChunk];   // raw input values per bubble
    double *fout[kRadialChunk];        // output values per bubble (k = 0)
    int nbub, lmax, k, nlip;
    int helmholtz;
    double kappa;
    double ii[kMaxNlip * (kMaxNlip - 1)];  // inward integrals, (nlip, nlip-1)
    double io[kMaxNlip * (kMaxNlip - 1)];  // outward integrals
    // exponentially fitted weights per bubble (nullptr: the reference's
    // scheme): per cell the (nlip, nlip-1) inward then outward tables of
    // lip_inout_integrals_exp at mu = kappa * step
    const double *wfit[kRadialChunk];
};

__global__ void __launch_bounds__(64) bubbles_radial_kernel(RadialArgs A) {
    const int l = blockIdx.x, b = blockIdx.y;
    const int lane = threadIdx.x;
    const BubbleDev B = A.bubs[b];
    const int npts = B.npts, ncell = B.ncell, n = A.nlip - 1;
    __shared__ double r[kMaxRadial], p1[kMaxRadial], p2[kMaxRadial], p3[kMaxRadial],
        p4[kMaxRadial], g[kMaxRadial], cum[kMaxRadial], outw[kMaxRadial];
    __shared__ double carry_out[kMaxRadial / 4 + 1], carry_in[kMaxRadial / 4 + 1];
    // radial points: the cell starts plus j steps (grid.F90:446-537)
    for (int i = lane; i < npts; i += 64) {
        const int c = min(i / n, ncell - 1), j = i - c * n;
        r[i] = B.celld[c] + j * B.steps[c];
    }
    __syncthreads();
    if (A.helmholtz && A.wfit[b] != nullptr) {
        // Exponentially fitted form.  With i_l(x) = e^x I(x), k_l(x) = e^-x K(x):
        //   k_l(kr) int_0^r s^(2+k) i_l(ks) f ds = K(kr) int_0^r e^-k(r-s) u ds,
        //   i_l(kr) int_r^R s^(2+k) k_l(ks) f ds = I(kr) int_r^R e^-k(s-r) v ds,
        // u = s^(2+k) I(ks) f, v = s^(2+k) K(ks) f smooth and interpolated
        // cell by cell, the exponentials integrated exactly (wfit), and the
        // carries between cells scaled by e^(-k L_c): no term exceeds the
        // result's own scale.
        const double kap = A.kappa;
        const double *W = A.wfit[b];
        const int tab = A.nlip * n;
        for (int i = lane; i < npts; i += 64) {
            const double ri = r[i], x = kap * ri;
            const double rp = rpow(ri, i, 2 + A.k);
            const double I = bessel_i(l, A.lmax, x) * exp(-x);
            const double K = i == 0 ? 0.0 : bessel_k_scaled(l, x);
            p1[i] = K;          // K(kr_i)
            p2[i] = rp * I;     // u / f
            p3[i] = I;          // I(kr_i)
            p4[i] = rp * K;     // v / f
        }
        __syncthreads();
        for (int m = -l; m <= l; ++m) {
            const int lm = l * l + l + m;
            const double *f = A.fin[b] + (long)lm * npts;
            double *out = A.fout[b] + (long)lm * npts;
            for (int i = lane; i < npts; i += 64) g[i] = p2[i] * f[i];  // u
            __syncthreads();
            // outward partials: cum[s0 + 1 + j] = h int_0^(j+1) e^-mu(j+1-t) u
            for (int c = lane; c < ncell; c += 64) {
                const int s0 = c * n;
                const double *wo = W + (long)c * 2 * tab + tab;
                for (int j = 0; j < n; ++j) {
                    double sum = 0.0;
                    for (int i = 0; i <= n; ++i) sum += g[s0 + i] * wo[i + A.nlip * j];
                    cum[s0 + 1 + j] = sum * B.steps[c];
                }
            }
            __syncthreads();
            if (lane == 0) {  // D_c = int_0^a_c e^-k(a_c - s) u ds
                double D = 0.0;
                for (int c = 0; c < ncell; ++c) {
                    carry_out[c] = D;
                    D = exp(-kap * n * B.steps[c]) * D + cum[c * n + n];
                }
            }
            __syncthreads();
            for (int i = lane; i < npts; i += 64) {
                if (i == 0) outw[i] = 0.0;
                else {
                    const int c = (i - 1) / n, j = i - c * n;  // j = 1..n in cell c
                    outw[i] = exp(-kap * j * B.steps[c]) * carry_out[c] + cum[i];
                }
            }
            __syncthreads();
            // inward partials with v; the first value extrapolated for l = 0
            for (int i = lane; i < npts; i += 64) g[i] = p4[i] * f[i];
            __syncthreads();
            if (lane == 0) g[0] = (l == 0) ? 3.0 * g[1] - 3.0 * g[2] + 1.0 * g[3] : 0.0;
            __syncthreads();
            for (int c = lane; c < ncell; c += 64) {
                const int s0 = c * n;
                const double *wi = W + (long)c * 2 * tab;
                for (int j = 0; j < n; ++j) {
                    double sum = 0.0;
                    for (int i = 0; i <= n; ++i) sum += g[s0 + i] * wi[i + A.nlip * j];
                    cum[s0 + j] = sum * B.steps[c];
                }
            }
            __syncthreads();
            if (lane == 0) {  // C_c = int_(b_c)^R e^-k(s - b_c) v ds
                double C = 0.0;
                for (int c = ncell - 1; c >= 0; --c) {
                    carry_in[c] = C;
                    C = exp(-kap * n * B.steps[c]) * C + cum[c * n];
                }
            }
            __syncthreads();
            for (int i = lane; i < npts; i += 64) {
                double inw;
                if (i == npts - 1) inw = 0.0;
                else {
                    const int c = i / n, j = i - c * n;  // j = 0..n-1 in cell c
                    inw = cum[i] + exp(-kap * (n - j) * B.steps[c]) * carry_in[c];
                }
                const double v = p1[i] * outw[i] + p3[i] * inw;
                out[i] = v * 8.0 * kap;
            }
            __syncthreads();
        }
        return;
    }
    // per-point factors: p2/p4 multiply f before the outward/inward
    // integrals, p1/p3 the integrals afterwards
    for (int i = lane; i < npts; i += 64) {
        const double ri = r[i];
        if (A.helmholtz) {
            const double x = A.kappa * ri;
            const double I = bessel_i(l, A.lmax, x), K = bessel_k(l, x);
            const double rp = rpow(ri, i, 2 + A.k);
            p1[i] = K;
            p2[i] = rp * I;
            p3[i] = I;
            p4[i] = rp * K;
        } else {
            p1[i] = rpow(ri, i, -(l + 1));
            p2[i] = rpow(ri, i, l + 2 + A.k);
            p3[i] = rpow(ri, i, l);
            p4[i] = rpow(ri, i, -(l + 1) + 2 + A.k);
        }
    }
    __syncthreads();
    const double pre = A.helmholtz ? 1.0 : 4.0 * 3.141592653589793 / (2 * l + 1);
    for (int m = -l; m <= l; ++m) {
        const int lm = l * l + l + m;
        const double *f = A.fin[b] + (long)lm * npts;
        double *out = A.fout[b] + (long)lm * npts;
        // outwards(f p2): cell c holds cum[start+1+j] = h_c sum_i g io(i,j)
        for (int i = lane; i < npts; i += 64) g[i] = A.helmholtz ? (p2[i] * f[i]) : (f[i] * p2[i]);
        __syncthreads();
        for (int c = lane; c < ncell; c += 64) {
            const int s0 = c * n;
            for (int j = 0; j < n; ++j) {
                double sum = 0.0;
                for (int i = 0; i <= n; ++i) sum += g[s0 + i] * A.io[i + A.nlip * j];
                cum[s0 + 1 + j] = sum * B.steps[c];
            }
        }
        __syncthreads();
        if (lane == 0) {  // res(start_c) in the reference's order
            double carry = 0.0;
            for (int c = 0; c < ncell; ++c) {
                carry_out[c] = carry;
                carry = cum[c * n + n] + carry;
            }
        }
        __syncthreads();
        for (int i = lane; i < npts; i += 64) {
            if (i == 0) outw[i] = 0.0;
            else {
                const int c = (i - 1) / n;
                outw[i] = cum[i] + carry_out[c];
            }
        }
        __syncthreads();
        // inwards(f p4); for Helmholtz the first value of the integrand is
        // extrapolated for l = 0 (extrapolate_first_nlip7(3, equidistant),
        // evaluators.F90:1588-1640) and zero otherwise
        for (int i = lane; i < npts; i += 64) g[i] = A.helmholtz ? (p4[i] * f[i]) : (f[i] * p4[i]);
        __syncthreads();
        if (A.helmholtz && lane == 0)
            g[0] = (l == 0) ? 3.0 * g[1] - 3.0 * g[2] + 1.0 * g[3] : 0.0;
        __syncthreads();
        for (int c = lane; c < ncell; c += 64) {
            const int s0 = c * n;
            for (int j = 0; j < n; ++j) {
                double sum = 0.0;
                for (int i = 0; i <= n; ++i) sum += g[s0 + i] * A.ii[i + A.nlip * j];
                cum[s0 + j] = sum * B.steps[c];
            }
        }
        __syncthreads();
        if (lane == 0) {  // res(start_c + n) from the last cell down
            double carry = 0.0;
            for (int c = ncell - 1; c >= 0; --c) {
                carry_in[c] = carry;
                carry = cum[c * n] + carry;
            }
        }
        __syncthreads();
        for (int i = lane; i < npts; i += 64) {
            double inw;
            if (i == npts - 1) inw = 0.0;
            else {
                const int c = i / n;
                inw = cum[i] + carry_in[c];
            }
            const double v = p1[i] * outw[i] + p3[i] * inw;
            out[i] = A.helmholtz ? v * 8.0 * A.kappa : pre * v;
        }
        __syncthreads();
    }
}

// dst(lm) = a * src(lm) r^dk + beta * dst(lm) for lm < nlm_common, dst(lm)
// = beta * dst(lm) above (src has no such l); dk = k_src - k_dst >= 0 puts
// src into dst's r^k representation (Bubbles_add, bubbles.F90:2499-2535)
__global__ void __launch_bounds__(256) bubble_axpby_kernel(const double *src, int nlm_src,
                                                           double a, double *dst, int nlm_dst,
                                                           double beta, int npts, int dk,
                                                           const double *celld,
                                                           const double *steps, int n) {
    const long total = (long)npts * nlm_dst;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int lm = (int)(e / npts);
        double s = 0.0;
        if (lm < nlm_src) {
            double x = src[e];
            if (dk != 0) {
                const int i = (int)(e % npts);
                const int c = min(i / n, (npts - 1) / n - 1);  // cell, last point in the last
                const double r = celld[c] + (i - c * n) * steps[c];
                x = x * powi(r, dk);
            }
            s = a * x;
        }
        dst[e] = beta == 0.0 ? s : s + beta * dst[e];
    }
}

// fk = r^k f with r(0)^k = 0 for k < 0 (Bubbles_rpow)
__global__ void __launch_bounds__(256) fold_rk_kernel(const double *celld, const double *steps,
                                                      int ncell, int nlip, int npts, int nlm,
                                                      int k, const double *f, double *fk) {
    const int n = nlip - 1;
    const long total = (long)npts * nlm;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e % npts);
        const int c = min(i / n, ncell - 1), j = i - c * n;
        const double r = celld[c] + j * steps[c];
        const double rp = (k < 0 && i == 0) ? 0.0 : pow(r, (double)k);
        fk[e] = f[e] * rp;
    }
}

int grid_blocks(long total) { return (int)std::min<long>((total + 255) / 256, 4096); }

}  // namespace

int bubble_alloc_host(dage_bubbles b, int ibub, const double center[3], double charge,
                      const Axis &ax) {
    BubbleHost &h = b->host[ibub];
    const int ncell = ax.ncell, npts = ax.n();
    // cell of r_k = k*qmax/kCellTab: 1 + the number of cell starts 1..ncell-1
    // at or below r_k (the cell the bisection of grid.F90:1212-1246 finds);
    // the inject kernel starts from the entry one bin below r and steps up
    const double qmax = ax.points[npts - 1];
    std::vector<int> tab(kCellTab);
    for (int k = 0, ic = 1; k < kCellTab; ++k) {
        const double rk = k * qmax / kCellTab;
        while (ic < ncell && ax.celld[ic] <= rk) ++ic;
        tab[k] = ic;
    }
    DAGE_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    if (h.d_buf) DAGE_HIP_CHECK(hipFree(h.d_buf));
    if (h.d_tab) DAGE_HIP_CHECK(hipFree(h.d_tab));
    h.d_buf = nullptr;
    h.d_tab = nullptr;
    h.set = false;
    const size_t nd = b->buf_doubles(ncell, npts);
    DAGE_HIP_CHECK(hipMalloc((void **)&h.d_buf, sizeof(double) * nd));
    DAGE_HIP_CHECK(hipMemset(h.d_buf, 0, sizeof(double) * nd));
    DAGE_HIP_CHECK(hipMemcpy(h.d_buf, ax.celld.data(), sizeof(double) * ncell,
                             hipMemcpyHostToDevice));
    DAGE_HIP_CHECK(hipMemcpy(h.d_buf + ncell, ax.steps.data(), sizeof(double) * ncell,
                             hipMemcpyHostToDevice));
    DAGE_HIP_CHECK(hipMalloc((void **)&h.d_tab, sizeof(int) * tab.size()));
    DAGE_HIP_CHECK(hipMemcpy(h.d_tab, tab.data(), sizeof(int) * tab.size(),
                             hipMemcpyHostToDevice));
    h.h_celld = ax.celld;
    h.h_steps = ax.steps;
    for (int a = 0; a < 3; ++a) h.center[a] = center[a];
    h.charge = charge;
    h.ncell = ncell;
    h.npts = npts;
    h.qmax = qmax;
    b->dirty = true;
    return DAGE_OK;
}

int bubble_alloc_like(dage_bubbles b, int ibub, dage_bubbles src, int sbub) {
    const BubbleHost &s = src->host[sbub];
    BubbleHost &h = b->host[ibub];
    DAGE_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    if (h.d_buf) DAGE_HIP_CHECK(hipFree(h.d_buf));
    if (h.d_tab) DAGE_HIP_CHECK(hipFree(h.d_tab));
    h.d_buf = nullptr;
    h.d_tab = nullptr;
    h.set = false;
    const size_t nd = b->buf_doubles(s.ncell, s.npts);
    DAGE_HIP_CHECK(hipMalloc((void **)&h.d_buf, sizeof(double) * nd));
    DAGE_HIP_CHECK(hipMemset(h.d_buf, 0, sizeof(double) * nd));
    DAGE_HIP_CHECK(hipMemcpy(h.d_buf, s.d_buf, sizeof(double) * 2 * s.ncell,
                             hipMemcpyDeviceToDevice));
    DAGE_HIP_CHECK(hipMalloc((void **)&h.d_tab, sizeof(int) * kCellTab));
    DAGE_HIP_CHECK(hipMemcpy(h.d_tab, s.d_tab, sizeof(int) * kCellTab, hipMemcpyDeviceToDevice));
    h.h_celld = s.h_celld;
    h.h_steps = s.h_steps;
    for (int a = 0; a < 3; ++a) h.center[a] = s.center[a];
    h.charge = s.charge;
    h.ncell = s.ncell;
    h.npts = s.npts;
    h.qmax = s.qmax;
    h.set = true;
    b->dirty = true;
    return DAGE_OK;
}

int bubbles_fold_rk(dage_bubbles b) {
    if (b->k == 0) return DAGE_OK;
    for (int i = 0; i < b->nbub; ++i) {
        const BubbleHost &h = b->host[i];
        const long total = (long)h.npts * b->nlm();
        hipLaunchKernelGGL(fold_rk_kernel, dim3(grid_blocks(total)), dim3(256), 0, b->ctx->stream,
                           h.celld(), h.steps(), h.ncell, b->nlip, h.npts, b->nlm(), b->k, b->f(i),
                           h.fk());
        DAGE_HIP_CHECK(hipGetLastError());
    }
    return DAGE_OK;
}

namespace {

// a and b hold bubbles on the same radial grids (same count, centres within
// rounding, same cell steps)
int require_same_grids(dage_bubbles a, dage_bubbles b) {
    DAGE_REQUIRE(a->ctx == b->ctx, DAGE_ERR_ARG, "bubbles belong to different contexts");
    DAGE_REQUIRE(a->nbub == b->nbub && a->nlip == b->nlip, DAGE_ERR_ARG,
                 "bubble sets differ in count (%d, %d) or nlip (%d, %d)", a->nbub, b->nbub,
                 a->nlip, b->nlip);
    for (int i = 0; i < a->nbub; ++i) {
        const BubbleHost &x = a->host[i], &y = b->host[i];
        DAGE_REQUIRE(x.set && y.set, DAGE_ERR_STATE, "bubble %d has not been set", i);
        DAGE_REQUIRE(x.ncell == y.ncell && x.qmax == y.qmax, DAGE_ERR_ARG,
                     "bubble %d: radial grids differ", i);
        for (int c = 0; c < 3; ++c)
            DAGE_REQUIRE(x.center[c] == y.center[c], DAGE_ERR_ARG, "bubble %d: centres differ",
                         i);
    }
    return DAGE_OK;
}

int use_ctx(dage_ctx ctx) {
    DAGE_REQUIRE(ctx != nullptr, DAGE_ERR_ARG, "null context handle");
    DAGE_HIP_CHECK(hipSetDevice(ctx->device));
    return DAGE_OK;
}

int radial_transform(dage_bubbles in, dage_bubbles out, bool helmholtz, double energy) {
    DAGE_TRY(require_same_grids(in, out));
    DAGE_REQUIRE(out->k == 0, DAGE_ERR_ARG, "result bubbles must have k = 0 (got %d)", out->k);
    DAGE_REQUIRE(out->lmax == in->lmax, DAGE_ERR_ARG, "result lmax %d != input lmax %d",
                 out->lmax, in->lmax);
    DAGE_REQUIRE(in->nlip == 7 || !helmholtz, DAGE_ERR_UNSUPPORTED,
                 "the Helmholtz bubble transform needs 7-point radial cells (got %d)", in->nlip);
    DAGE_REQUIRE(!helmholtz || (std::isfinite(energy) && energy < 0.0), DAGE_ERR_ARG,
                 "the Helmholtz bubble transform needs a negative energy (got %g)", energy);
    // the i_l series (300 terms, as the reference) converges to 1e-15 up to
    // kappa r ~ 400; i_l itself (~ e^(kappa r) / (2 kappa r)) fits FP64 to ~700
    if (helmholtz)
        for (int i = 0; i < in->nbub; ++i)
            DAGE_REQUIRE(std::sqrt(-2.0 * energy) * in->host[i].qmax <= 400.0, DAGE_ERR_UNSUPPORTED,
                         "Helmholtz bubble transform at energy %g: kappa * r_max = %g exceeds "
                         "400 (bubble %d), where the i_l series stops converging", energy,
                         std::sqrt(-2.0 * energy) * in->host[i].qmax, i);
    for (int i = 0; i < in->nbub; ++i)
        DAGE_REQUIRE(in->host[i].npts <= kMaxRadial && in->host[i].ncell * 4 <= kMaxRadial,
                     DAGE_ERR_UNSUPPORTED, "bubble %d has more than %d radial points", i,
                     kMaxRadial);
    if (in->nbub == 0) return DAGE_OK;
    dage_ctx ctx = in->ctx;
    DAGE_TRY(bubbles_sync_device(in));
    RadialArgs A{};
    A.nbub = in->nbub;
    A.lmax = in->lmax;
    A.k = in->k;
    A.nlip = in->nlip;
    A.helmholtz = helmholtz ? 1 : 0;
    A.kappa = helmholtz ? std::sqrt(-2.0 * energy) : 0.0;
    DAGE_REQUIRE(lip_inout_integrals(in->nlip, A.ii, A.io) == 0, DAGE_ERR_ARG,
                 "unsupported radial nlip %d", in->nlip);
    // bubbles with a cell of kappa L > kFittedKappaL take the exponentially
    // fitted scheme: weight tables per cell at mu = kappa * step
    const int n = in->nlip - 1, tab = in->nlip * n;
    std::vector<long> woff(in->nbub, -1);
    std::vector<double> wts;
    if (helmholtz) {
        for (int i = 0; i < in->nbub; ++i) {
            const BubbleHost &h = in->host[i];
            const std::vector<double> &st = h.h_steps;
            double kl = 0.0;
            for (double x : st) kl = std::max(kl, A.kappa * n * x);
            if (kl <= kFittedKappaL) continue;
            woff[i] = (long)wts.size();
            wts.resize(wts.size() + (size_t)h.ncell * 2 * tab);
            for (int c = 0; c < h.ncell; ++c) {
                double *w = wts.data() + woff[i] + (size_t)c * 2 * tab;
                DAGE_REQUIRE(lip_inout_integrals_exp(in->nlip, A.kappa * st[c], w, w + tab) == 0,
                             DAGE_ERR_ARG, "unsupported radial nlip %d", in->nlip);
            }
        }
    }
    double *d_w = nullptr;
    if (!wts.empty()) {
        void *scr = nullptr;
        DAGE_TRY(ctx_scratch(ctx, 8 * wts.size(), &scr));
        d_w = (double *)scr;
        DAGE_HIP_CHECK(hipMemcpyAsync(d_w, wts.data(), 8 * wts.size(), hipMemcpyHostToDevice,
                                      ctx->stream));
    }
    hipEvent_t ev;
    prof_begin(ctx, "bubbles_radial_kernel", &ev);
    for (int b0 = 0; b0 < in->nbub; b0 += kRadialChunk) {
        const int nb = std::min(kRadialChunk, in->nbub - b0);
        A.bubs = in->d_desc + b0;
        for (int i = 0; i < nb; ++i) {
            A.fin[i] = in->f(b0 + i);
            A.fout[i] = out->f(b0 + i);
            A.wfit[i] = woff[b0 + i] >= 0 ? d_w + woff[b0 + i] : nullptr;
        }
        hipLaunchKernelGGL(bubbles_radial_kernel, dim3(in->lmax + 1, nb), dim3(64), 0,
                           ctx->stream, A);
        DAGE_HIP_CHECK(hipGetLastError());
    }
    prof_end(ctx, "bubbles_radial_kernel", ev);
    return DAGE_OK;
}

}  // namespace
}  // namespace dage

using namespace dage;

extern "C" {

int dage_bubbles_create_like(dage_bubbles src, int lmax, int k, dage_bubbles *out) {
    DAGE_REQUIRE(src && out, DAGE_ERR_ARG, "null argument");
    *out = nullptr;
    for (int i = 0; i < src->nbub; ++i)
        DAGE_REQUIRE(src->host[i].set, DAGE_ERR_STATE, "bubble %d has not been set", i);
    dage_bubbles b = nullptr;
    DAGE_TRY(dage_bubbles_create(src->ctx, src->nbub, lmax, k, src->nlip, &b));
    for (int i = 0; i < src->nbub; ++i) {
        const int rc = bubble_alloc_like(b, i, src, i);
        if (rc != DAGE_OK) {
            dage_bubbles_destroy(b);
            return rc;
        }
    }
    *out = b;
    return DAGE_OK;
}

int dage_bubbles_info(dage_bubbles b, int *nbub, int *lmax, int *k, int *nlip) {
    DAGE_REQUIRE(b != nullptr, DAGE_ERR_ARG, "null bubbles");
    if (nbub) *nbub = b->nbub;
    if (lmax) *lmax = b->lmax;
    if (k) *k = b->k;
    if (nlip) *nlip = b->nlip;
    return DAGE_OK;
}

int dage_bubbles_npoints(dage_bubbles b, int ibub, int *npts) {
    DAGE_REQUIRE(b && npts && ibub >= 0 && ibub < b->nbub, DAGE_ERR_ARG, "bad argument");
    DAGE_REQUIRE(b->host[ibub].set, DAGE_ERR_STATE, "bubble %d has not been set", ibub);
    *npts = b->host[ibub].npts;
    return DAGE_OK;
}

int dage_bubbles_get(dage_bubbles b, int ibub, double *center, double *charge, double *steps,
                     double *bf) {
    DAGE_REQUIRE(b && ibub >= 0 && ibub < b->nbub, DAGE_ERR_ARG, "bad argument");
    const BubbleHost &h = b->host[ibub];
    DAGE_REQUIRE(h.set, DAGE_ERR_STATE, "bubble %d has not been set", ibub);
    DAGE_TRY(use_ctx(b->ctx));
    DAGE_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    if (center)
        for (int a = 0; a < 3; ++a) center[a] = h.center[a];
    if (charge) *charge = h.charge;
    if (steps)
        DAGE_HIP_CHECK(hipMemcpy(steps, h.steps(), sizeof(double) * h.ncell,
                                 hipMemcpyDeviceToHost));
    if (bf)
        DAGE_HIP_CHECK(hipMemcpy(bf, b->f(ibub), sizeof(double) * h.npts * b->nlm(),
                                 hipMemcpyDeviceToHost));
    return DAGE_OK;
}

int dage_bubbles_axpby(double a, dage_bubbles x, double beta, dage_bubbles y) {
    DAGE_REQUIRE(x && y, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(require_same_grids(x, y));
    DAGE_REQUIRE(x->k >= y->k, DAGE_ERR_ARG,
                 "x has k %d below y's %d (the sum would need y's k lowered)", x->k, y->k);
    DAGE_TRY(use_ctx(y->ctx));
    for (int i = 0; i < y->nbub; ++i) {
        const BubbleHost &h = y->host[i];
        const long total = (long)h.npts * y->nlm();
        hipLaunchKernelGGL(bubble_axpby_kernel, dim3(grid_blocks(total)), dim3(256), 0,
                           y->ctx->stream, x->f(i), std::min(x->nlm(), y->nlm()), a, y->f(i),
                           y->nlm(), beta, h.npts, x->k - y->k, h.celld(), h.steps(),
                           y->nlip - 1);
        DAGE_HIP_CHECK(hipGetLastError());
    }
    return bubbles_fold_rk(y);
}

int dage_bubbles_coulomb_transform(dage_bubbles in, dage_bubbles out) {
    DAGE_REQUIRE(in && out, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(use_ctx(in->ctx));
    return radial_transform(in, out, false, 0.0);
}

int dage_bubbles_helmholtz_transform(dage_bubbles in, double energy, dage_bubbles out) {
    DAGE_REQUIRE(in && out, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(use_ctx(in->ctx));
    return radial_transform(in, out, true, energy);
}

}  // extern "C"

extern "C" {

int dage_bubbles_copy_lmax(dage_bubbles src, int lmax, dage_bubbles *out) {
    DAGE_REQUIRE(src && out, DAGE_ERR_ARG, "null argument");
    *out = nullptr;
    DAGE_REQUIRE(lmax >= 0, DAGE_ERR_ARG, "negative lmax");
    DAGE_TRY(use_ctx(src->ctx));
    dage_bubbles b = nullptr;
    DAGE_TRY(dage_bubbles_create_like(src, lmax, src->k, &b));
    const int rc = dage_bubbles_axpby(1.0, src, 0.0, b);
    if (rc != DAGE_OK) {
        dage_bubbles_destroy(b);
        return rc;
    }
    *out = b;
    return DAGE_OK;
}

int dage_function3d_inject_extra_bubbles(dage_grid3d grid, double *cube_dev, dage_bubbles in,
                                         int lmax, dage_bubbles *kept) {
    DAGE_REQUIRE(grid && cube_dev && in && kept, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE(lmax >= 0, DAGE_ERR_ARG, "negative lmax");
    *kept = nullptr;
    DAGE_TRY(dage_bubbles_inject_device(in, grid, lmax + 1, 1.0, cube_dev));
    return dage_bubbles_copy_lmax(in, std::min(lmax, in->lmax), kept);
}

int dage_operator_operate_on(dage_operator op, const double *cube_in, dage_bubbles bub_in,
                             double *cube_out, dage_bubbles bub_out, int only_cube) {
    DAGE_REQUIRE(op && cube_in && cube_out, DAGE_ERR_ARG, "null argument");
    const bool bubbles = !only_cube && bub_in != nullptr;
    DAGE_REQUIRE(!bubbles || bub_out != nullptr, DAGE_ERR_ARG,
                 "operating on bubbles needs result bubbles");
    DAGE_TRY(use_ctx(op->ctx));
    if (bubbles) {
        if (op->helmholtz) {
            DAGE_REQUIRE(op->energy < 0.0, DAGE_ERR_STATE,
                         "the Helmholtz bubble transform needs a negative energy (got %g)",
                         op->energy);
            DAGE_TRY(radial_transform(bub_in, bub_out, true, op->energy));
        } else {
            DAGE_TRY(radial_transform(bub_in, bub_out, false, 0.0));
        }
    }
    DAGE_TRY(dage_operator_apply_device(op, cube_in, cube_out, DAGE_PREC_FP64));
    // Helmholtz3D_operate resets the operator to energy 0 afterwards
    // (helmholtz3d.F90:479)
    if (op->helmholtz) DAGE_TRY(dage_operator_set_energy(op, 0.0));
    return DAGE_OK;
}

}  // extern "C"

// ------------------------------------------------------------ bubble products
// BubblesMultiplier_multiply (src/bubbles/bubbles.F90:3553-3760, the
// simple product without Taylor-series terms) and the cube part of
// Function3D_multiply_sub (src/bubbles/function3d.F90:1097-1308, non-CUDA
// branch :1272-1295):
//   bubbles: g_LM(r) += c(l1m1, l2m2 -> LM) f1_l1m1(r) f2_l2m2(r), k = k1 + k2
//   cube:    (cube2 + B2) (cube1 + B1) - B_result, B = bubbles on the grid.
// c are the expansion coefficients of products of the Racah real spherical
// harmonics Y_lm = S_lm / r^l (YProduct, src/bubbles/harmonic.F90:884-1000):
// Y1 Y2 = sum_LM c Y_LM with c = (2L+1)/(4 pi) int Y1 Y2 Y_LM dOmega,
// computed here by a product quadrature on the sphere that is exact for these
// degrees (Gauss-Legendre in cos(theta), uniform in phi) in long double.
#include <map>
#include <mutex>

namespace dage {
namespace {

typedef long double ld;

struct YProd {
    int l1max = 0, l2max = 0;
    std::vector<int> first, count;   // per pair (lm1 * nlm2 + lm2): term range
    std::vector<int> out_lm;
    std::vector<double> coef;
};

void solid_ld(int lmax, ld x, ld y, ld z, ld *S) {
    S[0] = 1.0L;
    if (lmax < 1) return;
    S[1] = y;
    S[2] = z;
    S[3] = x;
    const ld r2 = x * x + y * y + z * z;
    for (int l = 2; l <= lmax; ++l) {
        const int c = l * l + l, p = (l - 1) * (l - 1) + (l - 1), q = (l - 2) * (l - 2) + (l - 2);
        const ld f = sqrtl((2 * l - 1.0L) / (2.0L * l));
        S[c - l] = f * (y * S[p + (l - 1)] + x * S[p - (l - 1)]);
        for (int m = -(l - 1); m <= l - 1; ++m) {
            ld v = (2 * l - 1.0L) / sqrtl(1.0L * (l + m) * (l - m)) * (z * S[p + m]);
            if (l - std::abs(m) > 1)
                v -= sqrtl((l + m - 1.0L) * (l - m - 1.0L) / ((ld)(l + m) * (l - m))) * (r2 * S[q + m]);
            S[c + m] = v;
        }
        S[c + l] = f * (x * S[p + (l - 1)] - y * S[p - (l - 1)]);
    }
}

// Gauss-Legendre nodes and weights on [-1, 1] (Newton on the recurrence)
void gauss_legendre_ld(int n, std::vector<ld> &x, std::vector<ld> &w) {
    x.assign(n, 0.0L);
    w.assign(n, 0.0L);
    const ld pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        ld z = cosl(pi * (i + 0.75L) / (n + 0.5L)), pp = 0.0L;
        for (int it = 0; it < 100; ++it) {
            ld p1 = 1.0L, p2 = 0.0L;
            for (int j = 0; j < n; ++j) {
                const ld p3 = p2;
                p2 = p1;
                p1 = ((2 * j + 1) * z * p2 - j * p3) / (j + 1);
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0L);
            const ld z1 = z;
            z = z1 - p1 / pp;
            if (fabsl(z - z1) < 1e-19L) break;
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2.0L / ((1.0L - z * z) * pp * pp);
    }
}

const YProd &yproduct(int l1max, int l2max) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, YProd> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({l1max, l2max});
    if (it != cache.end()) return it->second;
    YProd &P = cache[{l1max, l2max}];
    P.l1max = l1max;
    P.l2max = l2max;
    const int L = l1max + l2max;
    const int n1 = (l1max + 1) * (l1max + 1), n2 = (l2max + 1) * (l2max + 1),
              nL = (L + 1) * (L + 1);
    const int nt = L + 1, np = 2 * L + 2;
    std::vector<ld> xt, wt;
    gauss_legendre_ld(nt, xt, wt);
    const ld pi = 3.141592653589793238462643383279502884L;
    std::vector<ld> Y((size_t)nt * np * nL), wq((size_t)nt * np);
    for (int a = 0; a < nt; ++a)
        for (int b = 0; b < np; ++b) {
            const ld ct = xt[a], st = sqrtl(1.0L - ct * ct), ph = 2.0L * pi * b / np;
            solid_ld(L, st * cosl(ph), st * sinl(ph), ct, &Y[((size_t)a * np + b) * nL]);
            wq[(size_t)a * np + b] = wt[a] * (2.0L * pi / np);
        }
    P.first.assign((size_t)n1 * n2, 0);
    P.count.assign((size_t)n1 * n2, 0);
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j) {
            const int li = (int)std::sqrt((double)i), lj = (int)std::sqrt((double)j);
            const size_t pair = (size_t)i * n2 + j;
            P.first[pair] = (int)P.coef.size();
            for (int lo = std::abs(li - lj); lo <= li + lj; lo += 2)
                for (int mo = -lo; mo <= lo; ++mo) {
                    const int o = lo * lo + lo + mo;
                    ld s = 0.0L;
                    for (size_t q = 0; q < wq.size(); ++q) {
                        const ld *y = &Y[q * nL];
                        s += wq[q] * y[i] * y[j] * y[o];
                    }
                    const ld c = s * (2 * lo + 1) / (4.0L * pi);
                    if (fabsl(c) > 1e-10L) {  // zero up to quadrature noise
                        P.out_lm.push_back(o);
                        P.coef.push_back((double)c);
                    }
                }
            P.count[pair] = (int)P.coef.size() - P.first[pair];
        }
    return P;
}

constexpr int kProdChunk = 64;  // bubbles per product launch

struct ProdArgs {
    const double *f1[kProdChunk], *f2[kProdChunk];
    double *out[kProdChunk];
    int npts[kProdChunk];
    int nb, n1, n2;
    const int *first, *count, *out_lm;
    const double *coef;
};

// one thread per (radial point, bubble); the pairs in the reference's loop
// order (l1, m1 outer, l2, m2 inner), so each output sums in that order
__global__ void __launch_bounds__(256) bubble_product_kernel(ProdArgs A) {
    const int b = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.npts[b]) return;
    const int np = A.npts[b];
    const double *f1 = A.f1[b], *f2 = A.f2[b];
    double *out = A.out[b];
    for (int p = 0; p < A.n1; ++p) {
        const double v1 = f1[(long)p * np + i];
        for (int q = 0; q < A.n2; ++q) {
            const int pair = p * A.n2 + q;
            const int t0 = A.first[pair], nt = A.count[pair];
            if (nt == 0) continue;
            const double v2 = f2[(long)q * np + i];
            for (int t = t0; t < t0 + nt; ++t) {
                double *o = out + (long)A.out_lm[t] * np + i;
                *o = *o + A.coef[t] * v1 * v2;
            }
        }
    }
}

__global__ void __launch_bounds__(256) cube_mul_kernel(double *a, const double *b, long n) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x)
        a[e] = a[e] * b[e];
}

}  // namespace

int bubble_product_acc(dage_ctx ctx, int nb, const double *const *f1, int l1max,
                       const double *const *f2, int l2max, double *const *out, const int *npts) {
    if (nb == 0) return DAGE_OK;
    const YProd &P = yproduct(l1max, l2max);
    const size_t pairs = P.first.size(), terms = P.coef.size();
    const size_t bytes = sizeof(int) * (2 * pairs + terms) + sizeof(double) * terms + 64;
    void *scr = nullptr;
    DAGE_TRY(ctx_scratch(ctx, bytes, &scr));
    double *d_coef = (double *)scr;
    int *d_first = (int *)(d_coef + terms), *d_count = d_first + pairs, *d_out = d_count + pairs;
    hipStream_t s = ctx->stream;
    DAGE_HIP_CHECK(hipMemcpyAsync(d_coef, P.coef.data(), sizeof(double) * terms,
                                  hipMemcpyHostToDevice, s));
    DAGE_HIP_CHECK(hipMemcpyAsync(d_first, P.first.data(), sizeof(int) * pairs,
                                  hipMemcpyHostToDevice, s));
    DAGE_HIP_CHECK(hipMemcpyAsync(d_count, P.count.data(), sizeof(int) * pairs,
                                  hipMemcpyHostToDevice, s));
    DAGE_HIP_CHECK(hipMemcpyAsync(d_out, P.out_lm.data(), sizeof(int) * terms,
                                  hipMemcpyHostToDevice, s));
    hipEvent_t ev;
    prof_begin(ctx, "bubble_product_kernel", &ev);
    for (int b0 = 0; b0 < nb; b0 += kProdChunk) {
        ProdArgs A{};
        A.nb = std::min(kProdChunk, nb - b0);
        A.n1 = (l1max + 1) * (l1max + 1);
        A.n2 = (l2max + 1) * (l2max + 1);
        A.first = d_first;
        A.count = d_count;
        A.out_lm = d_out;
        A.coef = d_coef;
        int maxp = 1;
        for (int i = 0; i < A.nb; ++i) {
            A.f1[i] = f1[b0 + i];
            A.f2[i] = f2[b0 + i];
            A.out[i] = out[b0 + i];
            A.npts[i] = npts[b0 + i];
            maxp = std::max(maxp, A.npts[i]);
        }
        hipLaunchKernelGGL(bubble_product_kernel, dim3((maxp + 255) / 256, A.nb), dim3(256), 0, s,
                           A);
        DAGE_HIP_CHECK(hipGetLastError());
    }
    prof_end(ctx, "bubble_product_kernel", ev);
    return DAGE_OK;
}

}  // namespace dage

extern "C" {

int dage_bubbles_multiply(dage_bubbles a, dage_bubbles b, dage_bubbles out) {
    DAGE_REQUIRE(a && b && out, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(require_same_grids(a, b));
    DAGE_TRY(require_same_grids(a, out));
    DAGE_REQUIRE(out->lmax == a->lmax + b->lmax && out->k == a->k + b->k, DAGE_ERR_ARG,
                 "result bubbles need lmax %d and k %d (got %d, %d)", a->lmax + b->lmax,
                 a->k + b->k, out->lmax, out->k);
    DAGE_REQUIRE(out != a && out != b, DAGE_ERR_ARG, "the product cannot overwrite a factor");
    DAGE_TRY(use_ctx(a->ctx));
    dage_ctx ctx = a->ctx;
    // result = 0, then the products
    DAGE_TRY(dage_bubbles_axpby(0.0, out, 0.0, out));
    std::vector<const double *> f1(a->nbub), f2(a->nbub);
    std::vector<double *> fo(a->nbub);
    std::vector<int> np(a->nbub);
    for (int i = 0; i < a->nbub; ++i) {
        f1[i] = a->f(i);
        f2[i] = b->f(i);
        fo[i] = out->f(i);
        np[i] = a->host[i].npts;
    }
    DAGE_TRY(bubble_product_acc(ctx, a->nbub, f1.data(), a->lmax, f2.data(), b->lmax, fo.data(),
                                np.data()));
    return bubbles_fold_rk(out);
}

int dage_function3d_multiply(dage_grid3d grid, const double *cube1, dage_bubbles b1,
                             const double *cube2, dage_bubbles b2, double *cube_out,
                             dage_bubbles bub_out) {
    DAGE_REQUIRE(grid && cube1 && cube2 && cube_out, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE((b1 == nullptr) == (b2 == nullptr) || bub_out == nullptr, DAGE_ERR_ARG,
                 "result bubbles need bubbles in both factors");
    DAGE_REQUIRE(cube_out != cube1, DAGE_ERR_ARG, "the result cube cannot be the first factor");
    dage_ctx ctx = grid->ctx;
    DAGE_TRY(use_ctx(ctx));
    const long n = (long)grid->shape(0) * grid->shape(1) * grid->shape(2);
    hipStream_t s = ctx->stream;
    if (b1 && b2 && bub_out) DAGE_TRY(dage_bubbles_multiply(b1, b2, bub_out));
    // cube_out = cube2 + B2
    if (cube_out != cube2)
        DAGE_HIP_CHECK(hipMemcpyAsync(cube_out, cube2, 8 * n, hipMemcpyDeviceToDevice, s));
    if (b2) DAGE_TRY(dage_bubbles_inject_device(b2, grid, 0, 1.0, cube_out));
    // tmp = cube1 + B1 (context scratch), then cube_out *= tmp
    const double *fac = cube1;
    if (b1) {
        void *scr = nullptr;
        DAGE_TRY(ctx_scratch(ctx, 8 * (size_t)n, &scr));
        double *tmp = (double *)scr;
        DAGE_HIP_CHECK(hipMemcpyAsync(tmp, cube1, 8 * n, hipMemcpyDeviceToDevice, s));
        DAGE_TRY(dage_bubbles_inject_device(b1, grid, 0, 1.0, tmp));
        fac = tmp;
    }
    hipLaunchKernelGGL(cube_mul_kernel, dim3(grid_blocks(n)), dim3(256), 0, s, cube_out, fac, n);
    DAGE_HIP_CHECK(hipGetLastError());
    // minus the product bubbles
    if (bub_out && b1 && b2) DAGE_TRY(dage_bubbles_inject_device(bub_out, grid, 0, -1.0, cube_out));
    return DAGE_OK;
}

}  // extern "C"

extern "C" {

int dage_bubbles_integrate(dage_bubbles b, double *result) {
    DAGE_REQUIRE(b && result, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(use_ctx(b->ctx));
    DAGE_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    double ints[kMaxNlip];
    DAGE_REQUIRE(lip_integrals(b->nlip, 1, ints) == 0, DAGE_ERR_ARG, "unsupported nlip");
    const double fourpi = 4.0 * 3.141592653589793;
    const int n = b->nlip - 1;
    double res = 0.0;
    for (int i = 0; i < b->nbub; ++i) {
        const BubbleHost &h = b->host[i];
        DAGE_REQUIRE(h.set, DAGE_ERR_STATE, "bubble %d has not been set", i);
        std::vector<double> st(h.ncell), f(h.npts), cd(h.ncell);
        DAGE_HIP_CHECK(hipMemcpy(st.data(), h.steps(), 8 * h.ncell, hipMemcpyDeviceToHost));
        DAGE_HIP_CHECK(hipMemcpy(cd.data(), h.celld(), 8 * h.ncell, hipMemcpyDeviceToHost));
        DAGE_HIP_CHECK(hipMemcpy(f.data(), b->f(i), 8 * h.npts, hipMemcpyDeviceToHost));
        // Bubbles_integrate_pick (bubbles.F90:3146-3180): only the s part
        // survives the angular integral; 4 pi int r^(2+k) f_00 dr with the
        // radial LIP weights (Integrator1D_eval, evaluators.F90:1130-1145)
        double sum = 0.0;
        for (int c = 0; c < h.ncell; ++c) {
            double s = 0.0;
            for (int j = 0; j <= n; ++j) {
                const int p = c * n + j;
                const double r = cd[c] + j * st[c];
                const double rp = (2 + b->k < 0 && p == 0) ? 0.0 : powi(r, 2 + b->k);
                s += f[p] * rp * ints[j];
            }
            sum += s * st[c];
        }
        res += fourpi * sum;
    }
    *result = res;
    return DAGE_OK;
}

}  // extern "C"
