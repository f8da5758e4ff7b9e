// FP64 generation of the 1D Coulomb operator matrices on the GPU.
//
// Restates make_Coulomb3D_matrices / exponentials (src/bubbles/
// exponentials.F90:178-346, 602-639): entry (out point x2, in point i) is
//   h_c * sum_k LIP(j,k) * int_first^last q^(n-1-k) exp(-(t h_c)^2 (q-p)^2) dq
// summed over the cells c that contain in-point i (local index j), with p the
// cell coordinate of x2 (grid.F90:1768-1776).  The monomial integrals use the
// reference's closed forms expo_0..expo_6 for t h >= 0.1 and its 4x12-point
// Gauss-Legendre rule below that (exponentials.F90:647-666, 695-1297,
// 2287-2429), term by term, so the matrices carry the same rounding as the
// reference's.  FMA contraction is off in this file for that reason.
#include <cmath>

#include "dage_internal.hpp"

#pragma clang fp contract(off)

namespace dage {
namespace {

__device__ inline double sq(double x) { return x * x; }
__device__ inline double series4(double q) {
    return q + q * q / 2.0 + q * q * q / 6.0 + q * q * q * q / 24.0;
}

// int_a^b x^n exp(-t^2 (x-p)^2) dx, closed forms (exponentials.F90:695-1297).
// SP is sqrt(acos(-1)) as the host computes it (globals.F90:57).
__device__ double expo_closed(int n, double a, double b, double t, double p, double SP) {
    switch (n) {
    case 0: {
        if (t == 0.0 || t < 1.0e-6) return b - a;
        return SP * (erf(b * t - t * p) - erf(a * t - t * p)) / t / 2.0;
    }
    case 1: {
        if (t == 0.0 || t < 1.0e-6) return (b * b - a * a) / 2.0;
        double v1 = p * SP * erf(t * (b - p));
        double v2;
        if (t > 1.0e-3) {
            v2 = (exp(-(t * t) * sq(a - p)) - exp(-(t * t) * sq(b - p))) / t;
        } else {
            double q = -(t * t) * sq(a - p), r = -(t * t) * sq(b - p);
            v2 = (q + q * q / 2.0 + q * q * q / 6.0 + q * q * q * q / 24.0 - r - r * r / 2.0 -
                  r * r * r / 6.0 - r * r * r * r / 24.0) / t;
        }
        double v3 = -(p * SP * erf(a * t - t * p));
        return (v1 + v2 + v3) / t / 2.0;
    }
    case 2: {
        if (t == 0.0 || t < 0.5e-4) return (b * b * b - a * a * a) / 3.0;
        double v1 = p * p * SP * erf(b * t - t * p) / 2.0;
        double v2, v3;
        if (t > 1.0e-3) {
            double Ea = exp(-(t * t) * sq(a - p)), Eb = exp(-(t * t) * sq(b - p));
            v2 = (Ea * a - Eb * b);
            v3 = (Ea - Eb) * p / t / 2.0;
        } else {
            double q = -(t * t) * sq(a - p), r = -(t * t) * sq(b - p);
            v2 = series4(q) * a - series4(r) * b;
            v2 = (v2 + a - b);
            v3 = series4(q) - series4(r);
            v3 = v3 * p / t / 2.0;
        }
        double v5 = SP * (erf(t * (b - p)) - erf(t * (a - p))) / t / 2.0;
        v2 = (v2 + v5) / t / 2.0;
        double v4 = -(p * p * SP * erf(a * t - t * p) / 2.0);
        return (v1 + v2 + v3 + v4 + 0.0) / t;
    }
    case 3: {
        if (t == 0.0 || t < 1.0e-4) return (b * b * b * b - a * a * a * a) / 4.0;
        double v1, v2, v3, v5;
        if (t > 1.0e-3) {
            double Ea = exp(-(t * t) * sq(a - p)), Eb = exp(-(t * t) * sq(b - p));
            v1 = 2.0 * (Ea * (a * a) - Eb * (b * b)) / (t * t);
            v2 = 2.0 * p * (a * Ea - b * Eb) * t;
            v3 = 2.0 * (Ea * (p * p) - Eb * (p * p)) / (t * t);
            v5 = 2.0 * (Ea - Eb) / t;
        } else {
            double q = -(t * t) * sq(a - p), r = -(t * t) * sq(b - p);
            v1 = (1.0 + series4(q)) * a * a - (1.0 + series4(r)) * b * b;
            v1 = 2.0 * v1 / t / t;
            v2 = series4(q) * a - series4(r) * b;
            v2 = v2 + a - b;
            v2 = 2.0 * v2 * p * t;
            v3 = series4(q) - series4(r);
            v3 = 2.0 * v3 * p * p / t / t;
            v5 = series4(q) - series4(r);
            v5 = 2.0 * v5 / t;
        }
        double dE = erf(b * t - t * p) - erf(a * t - t * p);
        double v4 = 2.0 * (p * p * p) * SP * dE / t;
        double v6 = 3.0 * p * SP * dE;
        v6 = (v6 + v2 + v5) / t / t / t;
        return 0.25 * (v1 + 0.0 + v3 + v4 + 0.0 + v6);
    }
    case 4: {
        if (t == 0.0 || t < 1.0e-3) return (pow(b, 5.0) - pow(a, 5.0)) / 5.0;
        double v1, v2, v3, v4, v7, v8;
        double p2 = p * p, p3 = p * p * p;
        if (t > 1.0e-3) {
            double Ea = exp(-(t * t) * sq(a - p)), Eb = exp(-(t * t) * sq(b - p));
            v1 = 4.0 * (Ea * (a * a * a) - Eb * (b * b * b));
            v2 = 4.0 * (Ea * p * (a * a) - Eb * p * (b * b));
            v3 = 4.0 * (Ea * p2 * a - Eb * p2 * b);
            v4 = 4.0 * (Ea * p3 - Eb * p3);
            v7 = 10.0 * p * (Ea - Eb) / t / t;
            v8 = 6.0 * (Ea * a - Eb * b);
        } else {
            double q = -(t * t) * sq(a - p), r = -(t * t) * sq(b - p);
            double Sq = 1.0 + series4(q), Sr = 1.0 + series4(r);
            v1 = 4.0 * (Sq * (a * a * a) - Sr * (b * b * b));
            v2 = 4.0 * (Sq * p * a * a - Sr * p * b * b);
            v3 = 4.0 * (Sq * p * p * a - Sr * p * p * b);
            v4 = 4.0 * (series4(q) - series4(r)) * p3;
            v7 = 10.0 * (series4(q) - series4(r)) * p / t / t;
            v8 = 6.0 * (Sq * a - Sr * b);
        }
        double dE = erf(a * t - t * p) - erf(b * t - t * p);
        double v5 = -(4.0 * t * (p2 * p2) * SP * dE);
        double v6 = -(12.0 / t * p2 * SP * dE);
        double v9 = -(3.0 * SP * dE / t);
        v2 = v2 / (t * t);
        v5 = (v4 + v5) / (t * t);
        v7 = (v3 + v6 + v7) / (t * t);
        v8 = ((v8 + v9) / t / t + v1) / (t * t);
        return 0.125 * (0.0 + v2 + 0.0 + 0.0 + v5 + 0.0 + v7 + v8 + 0.0);
    }
    case 5: {
        if (t == 0.0 || t < 2.5e-3) return (pow(b, 6.0) - pow(a, 6.0)) / 6.0;
        double v1, v2, v3, v4, v5, v8, v9, v11, v12;
        double p2 = p * p, p3 = p * p * p, p4 = p2 * p2, p5 = p4 * p;
        double a2 = a * a, a3 = a * a * a, a4 = a2 * a2;
        double b2 = b * b, b3 = b * b * b, b4 = b2 * b2;
        double t2 = t * t, t4 = t2 * t2;
        if (t > 1.0e-2) {
            double Ea = exp(-(t * t) * sq(a - p)), Eb = exp(-(t * t) * sq(b - p));
            v1 = 4.0 * (Ea * a4 - Eb * b4);
            v2 = 4.0 * (Ea * p * a3 - Eb * p * b3);
            v3 = 4.0 * (Ea * p2 * a2 - Eb * p2 * b2);
            v4 = 4.0 * (Ea * p3 * a - Eb * p3 * b);
            v5 = 4.0 * (Ea * p4 - Eb * p4);
            v8 = 18.0 * (Ea - Eb) * p2 / t2;
            v9 = 14.0 * (Ea * p * a - Eb * p * b) / t2;
            v11 = 8.0 * (Ea * a2 - Eb * b2) / t2;
            v12 = 8.0 * (Ea - Eb) / t4;
        } else {
            double q = -(t * t) * sq(a - p), r = -(t * t) * sq(b - p);
            double Sq = 1.0 + series4(q), Sr = 1.0 + series4(r);
            v1 = 4.0 * (Sq * a4 - Sr * b4);
            v2 = 4.0 * (Sq * p * a3 - Sr * p * b3);
            v3 = 4.0 * (Sq * p2 * a2 - Sr * p2 * b2);
            v4 = 4.0 * (Sq * p3 * a - Sr * p3 * b);
            v5 = 4.0 * (series4(q) * p4 - series4(r) * p4);
            v8 = 18.0 * ((series4(q) - series4(r)) * p2 / t2);
            v9 = 14.0 * ((Sq * a * p - Sr * b * p) / t2);
            v11 = 8.0 * ((Sq * a * a - Sr * b * b) / t2);
            v12 = 8.0 * ((series4(q) - series4(r)) / t4);
        }
        double dE = erf(a * t - t * p) - erf(b * t - t * p);
        double v6 = -(4.0 * t * p5 * SP * dE);
        double v7 = -(20.0 / t * p3 * SP * dE);
        double v10 = -(15.0 * p * SP * dE / (t2 * t));
        double value = v1 + v2 + v3 + v4 + v5 + v6;
        value = value + v7 + v8 + v9 + v10 + v11 + v12;
        value = value / t2;
        return 0.125 * value;
    }
    case 6: {
        if (t == 0.0 || t < 5.0e-3) return (pow(b, 7.0) - pow(a, 7.0)) / 7.0;
        double p2 = p * p, p3 = p * p * p, p4 = p2 * p2, p5 = p4 * p, p6 = p3 * p3;
        double a2 = a * a, a3 = a * a * a, a4 = a2 * a2, a5 = a4 * a;
        double b2 = b * b, b3 = b * b * b, b4 = b2 * b2, b5 = b4 * b;
        double Ea = exp(-(t * t) * sq(a - p)), Eb = exp(-(t * t) * sq(b - p));
        double dE = erf(b * t - t * p) - erf(a * t - t * p);
        double v1 = 30.0 * (Ea * a - Eb * b);
        double v2 = 15.0 * SP * dE;
        v1 = (v1 + v2 / t) / t / t / t;
        double v3 = 90.0 * p2 * SP * dE / t / t;
        double v4 = 56.0 * (Ea - Eb) * p3 / t;
        double v5 = 66.0 * (Ea - Eb) * p / t / t / t;
        double v6 = 8.0 * (Ea - Eb) * p5 * t;
        double v7 = 20.0 * (Ea * a3 - Eb * b3) / t;
        double v8 = 36.0 * (Ea * a2 - Eb * b2) * p / t;
        double v9 = 48.0 * (Ea * a - Eb * b) * p2 / t;
        double v10 = 8.0 * p6 * SP * dE * (t * t);
        double v11 = 8.0 * (Ea * a4 - Eb * b4) * p * t;
        double v12 = 8.0 * (Ea * a3 - Eb * b3) * p2 * t;
        double v13 = 8.0 * (Ea * a5 - Eb * b5) * t;
        double v14 = 8.0 * (Ea * a - Eb * b) * p4 * t;
        double v15 = 8.0 * (Ea * a2 - Eb * b2) * p3 * t;
        double v16 = 60.0 * p4 * SP * dE;
        double value = v1 + 0.0 + v3 + v4 + v5 + v6 + v7 + v8;
        value = value + v9 + v10 + v11 + v12 + v13 + v14 + v15 + v16;
        value = value / (t * t * t);
        return 0.0625 * value;
    }
    default:
        return 1.23456789e10;  // expo_7 placeholder of the reference
    }
}

// 4 sub-intervals x 12-point Gauss-Legendre with the reference's literals
// (exponentials.F90:2287-2429).
__device__ double expo_gauss(int n, double a, double b, double t, double p) {
    const double wk[12] = {0.047175336386512, 0.106939325995318, 0.160078328543346,
                           0.203167426723066, 0.233492536538355, 0.249147045813403,
                           0.249147045813403, 0.233492536538355, 0.203167426723066,
                           0.160078328543346, 0.106939325995318, 0.047175336386512};
    const double tk[12] = {-0.981560634246719, -0.904117256370475, -0.769902674194305,
                           -0.587317954286617, -0.367831498998180, -0.125233408511469,
                           0.125233408511469,  0.367831498998180,  0.587317954286617,
                           0.769902674194305,  0.904117256370475,  0.981560634246719};
    if (t == 0.0) {
        if (n == 0) return b - a;
        return (pow(b, double(n + 1)) - pow(a, double(n + 1))) / (n + 1);
    }
    double limit[5];
    for (int i = 1; i <= 5; ++i) limit[i - 1] = ((5 - i) * a + (i - 1) * b) / 4;
    double summa = 0.0;
    for (int i = 0; i < 4; ++i) {
        const double a0 = limit[i], a1 = limit[i + 1];
        const double pr = 1.0 - a1 * 2.0 / (a1 - a0);
        const double qr = 2.0 / (a1 - a0);
        for (int j = 0; j < 12; ++j) {
            const double x = (tk[j] - pr) / qr;
            const double wgt = wk[j] / qr;
            const double e = exp(-(t * t) * (x - p) * (x - p));
            double xn = 1.0;
            for (int k = 0; k < n; ++k) xn *= x;
            summa = (n > 0) ? summa + xn * e * wgt : summa + e * wgt;
        }
    }
    return summa;
}

struct MatArgs {
    const double *celld_in;   // ncell_in
    const double *steps_in;   // ncell_in
    const double *points_out; // nout
    const double *tp;         // T
    double *out;
    int nin, nout, ncell, nlip, T;
    int transpose;
    double first, last, sqrtpi;
    double lip[kMaxNlip * kMaxNlip];  // lip[j + nlip*k]
};

// Contribution of cell c (local basis index j) to entry (x2, in-point).
__device__ double cell_entry(const MatArgs &A, int c, int j, double x2, double t) {
    const double h = A.steps_in[c];
    const double shft = t * h;
    const double p = (x2 - A.celld_in[c]) / h + A.first;
    double vals[kMaxNlip];
    const int n = A.nlip;
    for (int i = 0; i < n; ++i)
        vals[n - 1 - i] = (shft < 0.1) ? expo_gauss(i, A.first, A.last, shft, p)
                                       : expo_closed(i, A.first, A.last, shft, p, A.sqrtpi);
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += A.lip[j + n * k] * vals[k];
    return s * h;
}

// One thread per matrix entry (out point, in point, t).
__global__ void __launch_bounds__(256) coulomb_matrix_kernel(MatArgs A) {
    const long total = (long)A.nin * A.nout * A.T;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        int xi, xo;
        long rest;
        if (A.transpose) {  // (nin, nout, T)
            xi = (int)(e % A.nin);
            rest = e / A.nin;
            xo = (int)(rest % A.nout);
        } else {            // (nout, nin, T)
            xo = (int)(e % A.nout);
            rest = e / A.nout;
            xi = (int)(rest % A.nin);
        }
        const int it = (int)(rest / (A.transpose ? A.nout : A.nin));
        const double t = A.tp[it];
        const double x2 = A.points_out[xo];
        const int step = A.nlip - 1;
        int c = xi / step;
        int j = xi - c * step;
        if (c == A.ncell) {  // last point belongs to the last cell only
            c = A.ncell - 1;
            j = step;
        }
        double v = 0.0;
        if (j == 0 && c > 0) v = v + cell_entry(A, c - 1, step, x2, t);
        v = v + cell_entry(A, c, j, x2, t);
        A.out[e] = v;
    }
}

int gen_axis(dage_operator op, int axis, double *dst, int transpose) {
    const Axis &ai = op->gin->axis[axis];
    const Axis &ao = op->gout->axis[axis];
    MatArgs A{};
    A.nin = ai.n();
    A.nout = ao.n();
    A.ncell = ai.ncell;
    A.nlip = ai.nlip;
    A.T = op->ntp;
    A.transpose = transpose;
    lip_first_last(ai.nlip, &A.first, &A.last);
    A.sqrtpi = std::sqrt(pi());
    DAGE_REQUIRE(lip_coeffs(ai.nlip, ai.grid_type, A.lip) == 0, DAGE_ERR_ARG,
                 "unsupported LIP basis (nlip %d, grid_type %d)", ai.nlip, ai.grid_type);
    hipStream_t s = op->ctx->stream;
    // small staging buffer for the per-axis tables
    const size_t nb = sizeof(double) * (2 * ai.ncell + ao.n() + op->ntp);
    double *tab = nullptr;
    DAGE_HIP_CHECK(hipMallocAsync((void **)&tab, nb, s));
    std::vector<double> host;
    host.reserve(nb / sizeof(double));
    host.insert(host.end(), ai.celld.begin(), ai.celld.end());
    host.insert(host.end(), ai.steps.begin(), ai.steps.end());
    host.insert(host.end(), ao.points.begin(), ao.points.end());
    host.insert(host.end(), op->tp.begin(), op->tp.end());
    DAGE_HIP_CHECK(hipMemcpyAsync(tab, host.data(), nb, hipMemcpyHostToDevice, s));
    A.celld_in = tab;
    A.steps_in = tab + ai.ncell;
    A.points_out = tab + 2 * ai.ncell;
    A.tp = tab + 2 * ai.ncell + ao.n();
    A.out = dst;
    const long total = (long)A.nin * A.nout * A.T;
    const int block = 256;
    const long want = (total + block - 1) / block;
    const int grid = (int)(want < 65536 ? want : 65536);
    hipEvent_t ev;
    prof_begin(op->ctx, "coulomb_matrix_kernel", &ev);
    hipLaunchKernelGGL(coulomb_matrix_kernel, dim3(grid), dim3(block), 0, s, A);
    DAGE_HIP_CHECK(hipGetLastError());
    prof_end(op->ctx, "coulomb_matrix_kernel", ev);
    // host vector must outlive the async copy
    DAGE_HIP_CHECK(hipStreamSynchronize(s));
    DAGE_HIP_CHECK(hipFreeAsync(tab, s));
    return DAGE_OK;
}

}  // namespace

int generate_matrices(dage_operator op) {
    DAGE_TRY(gen_axis(op, 0, op->d_fx, 0));
    DAGE_TRY(gen_axis(op, 1, op->d_fy, 1));
    DAGE_TRY(gen_axis(op, 2, op->d_fz, 1));
    return DAGE_OK;
}

}  // namespace dage
