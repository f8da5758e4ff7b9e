// Cube algebra on device cubes for the SCF cycle: the cube part of the
// Function3D arithmetic the Helmholtz SCF update uses
// (HelmholtzSCFCycle_update_orbitals_worker, scf_cycle.F90:2927-3054:
// product_in_place_REAL64, subtract_in_place, add_in_place, the density
// build of calculate_electron_density_worker :2333-2388, the overlap matrix
// of calculate_overlap_matrix :2399-2424 and the orbital rotations of
// get_linear_combination_of_orbitals :936-968).  Plain elementwise HIP
// kernels (HBM-bound, grid-stride, x fastest) and one deterministic Gram-
// matrix reduction with the grid's LIP weights (FP64: overlaps and energies
// stay FP64).
#include <initializer_list>
#include <vector>

#include "dage_internal.hpp"

namespace dage {
namespace {

int use_device(dage_ctx ctx) {
    DAGE_REQUIRE(ctx != nullptr, DAGE_ERR_ARG, "null context handle");
    DAGE_HIP_CHECK(hipSetDevice(ctx->device));
    return DAGE_OK;
}

int grid_of(long n) { return (int)std::min<long>((n + 255) / 256, 65536); }

__global__ void __launch_bounds__(256) axpby_kernel(long n, double a, const double *x, double b,
                                                    double *y) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        y[i] = a * x[i] + b * y[i];
}

__global__ void __launch_bounds__(256) mul_kernel(long n, const double *x, const double *y,
                                                  double *out) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        out[i] = x[i] * y[i];
}

__global__ void __launch_bounds__(256) square_acc_kernel(long n, double occ, const double *x,
                                                         double *rho) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        rho[i] = rho[i] + occ * (x[i] * x[i]);
}

// Slater (LDA) exchange: v = -(3 rho/pi)^(1/3), e = -(3/4)(3/pi)^(1/3) rho^(4/3)
__global__ void __launch_bounds__(256) lda_x_kernel(long n, const double *rho, double cv,
                                                    double ce, double *v, double *e) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const double r = rho[i] > 0.0 ? rho[i] : 0.0;
        const double c = cbrt(r);
        v[i] = -cv * c;
        if (e) e[i] = -ce * (r * c);
    }
}

// out_j = sum_i coef[i + nin*j] in_i  for j < nout
__global__ void __launch_bounds__(256) lincomb_kernel(long n, int nin, const double *const *in,
                                                     const double *coef, int nout,
                                                     double *const *out) {
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < n;
         p += (long)gridDim.x * blockDim.x) {
        for (int j = 0; j < nout; ++j) {
            double s = 0.0;
            for (int i = 0; i < nin; ++i) s += coef[i + (long)nin * j] * in[i][p];
            out[j][p] = s;
        }
    }
}

// part[(pair * nz + z)] = sum over the z-plane of wx wy f_i g_j (times wz)
// Gram tiles: block (z, tile) sums w(x,y) f_i g_j over plane z for the 8 x 8
// pairs (i0 + a, j0 + b) of its tile, reading each function's plane once;
// every thread keeps the 64 sums of its points in registers, then a fixed
// shuffle tree per wavefront and a fixed sum over the 4 wavefronts
// (deterministic).  part[(i + nf*j)*nz + z] = wz[z] * plane sum.
constexpr int kGT = 8;

__global__ void __launch_bounds__(256) gram_kernel(const double *const *f, const double *const *g,
                                                   int nf, int ng, const double *wx,
                                                   const double *wy, const double *wz, int nx,
                                                   int ny, int nz, double *part) {
    __shared__ double red[4][kGT * kGT];
    const int z = blockIdx.x;
    const int tilesf = (nf + kGT - 1) / kGT;
    const int i0 = (blockIdx.y % tilesf) * kGT, j0 = (blockIdx.y / tilesf) * kGT;
    const long off = (long)nx * ny * z;
    const double *fp[kGT], *gp[kGT];
#pragma unroll
    for (int a = 0; a < kGT; ++a) {
        fp[a] = f[min(i0 + a, nf - 1)] + off;  // out-of-range rows read a valid cube, unused
        gp[a] = g[min(j0 + a, ng - 1)] + off;
    }
    double acc[kGT][kGT];
#pragma unroll
    for (int a = 0; a < kGT; ++a)
#pragma unroll
        for (int b = 0; b < kGT; ++b) acc[a][b] = 0.0;
    for (int e = threadIdx.x; e < nx * ny; e += 256) {
        const int x = e % nx, y = e / nx;
        const double w = wx[x] * wy[y];
        double fv[kGT], gv[kGT];
#pragma unroll
        for (int a = 0; a < kGT; ++a) {
            fv[a] = w * fp[a][e];
            gv[a] = gp[a][e];
        }
#pragma unroll
        for (int a = 0; a < kGT; ++a)
#pragma unroll
            for (int b = 0; b < kGT; ++b) acc[a][b] += fv[a] * gv[b];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < kGT; ++a)
#pragma unroll
        for (int b = 0; b < kGT; ++b) {
            double v = acc[a][b];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
            if (lane == 0) red[wave][a * kGT + b] = v;
        }
    __syncthreads();
    if (threadIdx.x < kGT * kGT) {
        const int a = threadIdx.x / kGT, b = threadIdx.x % kGT;
        const double v = (red[0][threadIdx.x] + red[1][threadIdx.x]) +
                         (red[2][threadIdx.x] + red[3][threadIdx.x]);
        if (i0 + a < nf && j0 + b < ng)
            part[((long)(i0 + a) + (long)nf * (j0 + b)) * nz + z] = v * wz[z];
    }
}

// df = d f / d(axis) from the cells' Lagrange interpolants: at a point
// inside a cell the derivative of that cell's interpolant, at a point shared
// by two cells the mean of both.  D[j*nlip + k] = L_k'(node j) for unit node
// spacing; the cell's step is recovered from its end points.
__global__ void __launch_bounds__(256) gradient_kernel(const double *f, double *df, int nx, int ny,
                                                       int nz, int axis, const double *pts,
                                                       int nlip, const double *D) {
    const long total = (long)nx * ny * nz;
    const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const long stride = axis == 0 ? 1 : (axis == 1 ? (long)nx : (long)nx * ny);
    const int np1 = nlip - 1;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int x = (int)(e % nx);
        const long r = e / nx;
        const int y = (int)(r % ny);
        const int z = (int)(r / ny);
        const int i = axis == 0 ? x : (axis == 1 ? y : z);
        const long base = e - (long)i * stride;  // the line's first point
        double sum = 0.0;
        int cnt = 0;
        for (int side = 0; side < 2; ++side) {
            // side 0: the cell on the left of a shared point / the cell
            // containing an interior point; side 1: the cell on the right
            int c;
            if (side == 0) {
                if (i % np1 == 0) {
                    if (i == 0) continue;
                    c = i / np1 - 1;
                } else {
                    c = i / np1;
                }
            } else {
                if (i % np1 != 0 || i == n - 1) continue;
                c = i / np1;
            }
            const int j = i - c * np1;
            const double h = (pts[(c + 1) * np1] - pts[c * np1]) / np1;
            double d = 0.0;
            for (int k = 0; k < nlip; ++k) d += D[j * nlip + k] * f[base + (long)(c * np1 + k) * stride];
            sum += d / h;
            ++cnt;
        }
        df[e] = sum / cnt;
    }
}

// out = (K f) / w along `axis`: K the assembled stiffness matrix of the
// cells' Lagrange interpolants, (K f)_i = sum over the cells c holding i of
// sum_k Kc[j*nlip + k] f(c, k) / h_c with Kc[a*nlip + b] = int L_a' L_b' over
// a cell of unit node spacing; w the point's LIP weight along the axis.  Then
// sum_p wx wy wz g (out) = sum_lines w_perp int g' f' (the interpolants'
// exact kinetic integral along that axis).
__global__ void __launch_bounds__(256) stiffness_kernel(const double *f, double *out, int nx,
                                                        int ny, int nz, int axis,
                                                        const double *pts, const double *wts,
                                                        int nlip, const double *Kc) {
    const long total = (long)nx * ny * nz;
    const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const long stride = axis == 0 ? 1 : (axis == 1 ? (long)nx : (long)nx * ny);
    const int np1 = nlip - 1;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int x = (int)(e % nx);
        const long r = e / nx;
        const int y = (int)(r % ny);
        const int z = (int)(r / ny);
        const int i = axis == 0 ? x : (axis == 1 ? y : z);
        const long base = e - (long)i * stride;
        double sum = 0.0;
        for (int side = 0; side < 2; ++side) {
            int c;
            if (side == 0) {
                if (i % np1 == 0) {
                    if (i == 0) continue;
                    c = i / np1 - 1;
                } else {
                    c = i / np1;
                }
            } else {
                if (i % np1 != 0 || i == n - 1) continue;
                c = i / np1;
            }
            const int j = i - c * np1;
            const double h = (pts[(c + 1) * np1] - pts[c * np1]) / np1;
            double d = 0.0;
            for (int k = 0; k < nlip; ++k) d += Kc[j * nlip + k] * f[base + (long)(c * np1 + k) * stride];
            sum += d / h;
        }
        out[e] = sum / wts[i];
    }
}

// Carves 256-byte-aligned pieces out of the context scratch.
struct Carver {
    char *base = nullptr;
    size_t off = 0;
    template <typename T>
    T *take(size_t n) {
        T *p = reinterpret_cast<T *>(base + off);
        off += (sizeof(T) * (n ? n : 1) + 255) / 256 * 256;
        return p;
    }
};

size_t carve_size(std::initializer_list<size_t> bytes) {
    size_t s = 0;
    for (size_t b : bytes) s += (b + 255) / 256 * 256;
    return s;
}

// D[j*nlip + k] = L_k'(node j) for unit node spacing (barycentric form).
std::vector<double> deriv_table(const Axis &ax) {
    const int nlip = ax.nlip;
    std::vector<double> node(nlip), wb(nlip), D((size_t)nlip * nlip, 0.0);
    for (int k = 0; k < nlip; ++k) node[k] = (ax.points[k] - ax.points[0]) / ax.steps[0];
    for (int k = 0; k < nlip; ++k) {
        double p = 1.0;
        for (int m = 0; m < nlip; ++m)
            if (m != k) p *= node[k] - node[m];
        wb[k] = 1.0 / p;
    }
    for (int j = 0; j < nlip; ++j) {
        double diag = 0.0;
        for (int k = 0; k < nlip; ++k)
            if (k != j) {
                D[j * nlip + k] = (wb[k] / wb[j]) / (node[j] - node[k]);
                diag -= D[j * nlip + k];
            }
        D[j * nlip + j] = diag;
    }
    return D;
}

// Kc[a*nlip + b] = int_0^{nlip-1} L_a' L_b' (nodes 0..nlip-1), by the
// nlip-point Gauss-Legendre rule (exact: the integrand has degree 2 nlip - 4).
std::vector<double> stiffness_table(int nlip) {
    std::vector<double> wb(nlip), Kc((size_t)nlip * nlip, 0.0);
    for (int k = 0; k < nlip; ++k) {
        double p = 1.0;
        for (int m = 0; m < nlip; ++m)
            if (m != k) p *= (double)(k - m);
        wb[k] = 1.0 / p;
    }
    std::vector<double> gx(nlip), gw(nlip);
    gauleg(nlip, gx.data(), gw.data());  // on [-1, 1]
    const double half = 0.5 * (nlip - 1);
    for (int q = 0; q < nlip; ++q) {
        const double xi = half * (gx[q] + 1.0), wq = half * gw[q];
        // L_a'(xi) = sum over m != a of wb[a] * prod_{mm != a, m} (xi - mm)
        std::vector<double> dL(nlip, 0.0);
        for (int a = 0; a < nlip; ++a) {
            double sum = 0.0;
            for (int m = 0; m < nlip; ++m) {
                if (m == a) continue;
                double p = wb[a];
                for (int mm = 0; mm < nlip; ++mm)
                    if (mm != a && mm != m) p *= xi - mm;
                sum += p;
            }
            dL[a] = sum;
        }
        for (int a = 0; a < nlip; ++a)
            for (int b = 0; b < nlip; ++b) Kc[a * nlip + b] += wq * dL[a] * dL[b];
    }
    return Kc;
}

// The grid's device copy of an axis table, built on first use.
int grid_table(dage_grid3d grid, int axis, bool stiff, const double **out) {
    double *&slot = stiff ? grid->d_stiff[axis] : grid->d_deriv[axis];
    if (!slot) {
        const Axis &ax = grid->axis[axis];
        const std::vector<double> t = stiff ? stiffness_table(ax.nlip) : deriv_table(ax);
        DAGE_HIP_CHECK(hipMalloc((void **)&slot, sizeof(double) * t.size()));
        DAGE_HIP_CHECK(hipMemcpy(slot, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
    }
    *out = slot;
    return DAGE_OK;
}

}  // namespace
}  // namespace dage

using namespace dage;

extern "C" {

int dage_vec_axpby(dage_ctx ctx, long n, double alpha, const double *x, double beta, double *y) {
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(n >= 0 && (n == 0 || (x && y)), DAGE_ERR_ARG, "bad argument");
    if (n == 0) return DAGE_OK;
    hipLaunchKernelGGL(axpby_kernel, dim3(grid_of(n)), dim3(256), 0, ctx->stream, n, alpha, x,
                       beta, y);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int dage_vec_mul(dage_ctx ctx, long n, const double *x, const double *y, double *out) {
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(n >= 0 && (n == 0 || (x && y && out)), DAGE_ERR_ARG, "bad argument");
    if (n == 0) return DAGE_OK;
    hipLaunchKernelGGL(mul_kernel, dim3(grid_of(n)), dim3(256), 0, ctx->stream, n, x, y, out);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int dage_vec_square_acc(dage_ctx ctx, long n, double occ, const double *x, double *rho) {
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(n >= 0 && (n == 0 || (x && rho)), DAGE_ERR_ARG, "bad argument");
    if (n == 0) return DAGE_OK;
    hipLaunchKernelGGL(square_acc_kernel, dim3(grid_of(n)), dim3(256), 0, ctx->stream, n, occ, x,
                       rho);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int dage_vec_lda_x(dage_ctx ctx, long n, const double *rho, double *vx, double *ex) {
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(n >= 0 && (n == 0 || (rho && vx)), DAGE_ERR_ARG, "bad argument");
    if (n == 0) return DAGE_OK;
    const double pi = acos(-1.0);
    const double cv = cbrt(3.0 / pi), ce = 0.75 * cbrt(3.0 / pi);
    hipLaunchKernelGGL(lda_x_kernel, dim3(grid_of(n)), dim3(256), 0, ctx->stream, n, rho, cv, ce,
                       vx, ex);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int dage_vec_lincomb(dage_ctx ctx, long n, int nin, const double *const *in, const double *coef,
                     int nout, double *const *out) {
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(n > 0 && nin > 0 && nout > 0 && in && coef && out, DAGE_ERR_ARG, "bad argument");
    for (int j = 0; j < nout; ++j)
        for (int i = 0; i < nin; ++i)
            DAGE_REQUIRE(out[j] != in[i], DAGE_ERR_ARG, "output %d aliases input %d", j, i);
    const size_t bi = sizeof(void *) * nin, bo = sizeof(void *) * nout,
                 bc = sizeof(double) * nin * nout;
    void *scr = nullptr;
    DAGE_TRY(ctx_scratch(ctx, carve_size({bi, bo, bc}), &scr));
    Carver cv{static_cast<char *>(scr)};
    const double **d_in = cv.take<const double *>(nin);
    double **d_out = cv.take<double *>(nout);
    double *d_coef = cv.take<double>((size_t)nin * nout);
    DAGE_HIP_CHECK(hipMemcpyAsync(d_in, in, bi, hipMemcpyHostToDevice, ctx->stream));
    DAGE_HIP_CHECK(hipMemcpyAsync(d_out, out, bo, hipMemcpyHostToDevice, ctx->stream));
    DAGE_HIP_CHECK(hipMemcpyAsync(d_coef, coef, bc, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(lincomb_kernel, dim3(grid_of(n)), dim3(256), 0, ctx->stream, n, nin, d_in,
                       d_coef, nout, d_out);
    DAGE_HIP_CHECK(hipGetLastError());
    // the host arrays are the caller's: finish before returning
    DAGE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return DAGE_OK;
}

int dage_cube_gradient(dage_grid3d grid, int axis, const double *f, double *df) {
    DAGE_REQUIRE(grid && f && df && axis >= 0 && axis < 3, DAGE_ERR_ARG, "bad argument");
    DAGE_REQUIRE(f != df, DAGE_ERR_ARG, "gradient output aliases its input");
    dage_ctx ctx = grid->ctx;
    DAGE_TRY(use_device(ctx));
    // the node template is shared by every cell: equidistant nodes only
    DAGE_REQUIRE(grid->axis[axis].grid_type == 1, DAGE_ERR_UNSUPPORTED,
                 "gradient needs equidistant cell nodes");
    const double *d_D = nullptr;
    DAGE_TRY(grid_table(grid, axis, false, &d_D));
    const long total = (long)grid->shape(0) * grid->shape(1) * grid->shape(2);
    hipLaunchKernelGGL(gradient_kernel, dim3(grid_of(total)), dim3(256), 0, ctx->stream, f, df,
                       grid->shape(0), grid->shape(1), grid->shape(2), axis,
                       grid->d_points[axis], grid->axis[axis].nlip, d_D);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int dage_cube_stiffness(dage_grid3d grid, int axis, const double *f, double *out) {
    DAGE_REQUIRE(grid && f && out && axis >= 0 && axis < 3, DAGE_ERR_ARG, "bad argument");
    DAGE_REQUIRE(f != out, DAGE_ERR_ARG, "stiffness output aliases its input");
    dage_ctx ctx = grid->ctx;
    DAGE_TRY(use_device(ctx));
    const Axis &ax = grid->axis[axis];
    DAGE_REQUIRE(ax.grid_type == 1, DAGE_ERR_UNSUPPORTED, "stiffness needs equidistant cell nodes");
    const double *d_K = nullptr;
    DAGE_TRY(grid_table(grid, axis, true, &d_K));
    const long total = (long)grid->shape(0) * grid->shape(1) * grid->shape(2);
    hipLaunchKernelGGL(stiffness_kernel, dim3(grid_of(total)), dim3(256), 0, ctx->stream, f, out,
                       grid->shape(0), grid->shape(1), grid->shape(2), axis,
                       grid->d_points[axis], grid->d_weights[axis], ax.nlip, d_K);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int dage_cube_gram(dage_grid3d grid, int nf, const double *const *f, int ng,
                   const double *const *g, double *S) {
    DAGE_REQUIRE(grid && f && g && S && nf > 0 && ng > 0, DAGE_ERR_ARG, "bad argument");
    dage_ctx ctx = grid->ctx;
    DAGE_TRY(use_device(ctx));
    const int nx = grid->shape(0), ny = grid->shape(1), nz = grid->shape(2);
    const int npair = nf * ng;
    const int ntile = ((nf + kGT - 1) / kGT) * ((ng + kGT - 1) / kGT);
    DAGE_REQUIRE(ntile <= 65535, DAGE_ERR_ARG, "too many pairs (%d)", npair);
    const size_t bf = sizeof(void *) * nf, bg = sizeof(void *) * ng,
                 bp = sizeof(double) * (size_t)npair * nz;
    void *scr = nullptr;
    DAGE_TRY(ctx_scratch(ctx, carve_size({bf, bg, bp}), &scr));
    Carver cv{static_cast<char *>(scr)};
    const double **d_f = cv.take<const double *>(nf);
    const double **d_g = cv.take<const double *>(ng);
    double *d_part = cv.take<double>((size_t)npair * nz);
    std::vector<double> part((size_t)npair * nz);
    DAGE_HIP_CHECK(hipMemcpyAsync(d_f, f, bf, hipMemcpyHostToDevice, ctx->stream));
    DAGE_HIP_CHECK(hipMemcpyAsync(d_g, g, bg, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(gram_kernel, dim3(nz, ntile), dim3(256), 0, ctx->stream, d_f, d_g, nf, ng,
                       grid->d_weights[0], grid->d_weights[1], grid->d_weights[2], nx, ny, nz,
                       d_part);
    DAGE_HIP_CHECK(hipGetLastError());
    DAGE_HIP_CHECK(hipMemcpyAsync(part.data(), d_part, bp, hipMemcpyDeviceToHost, ctx->stream));
    DAGE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (int p = 0; p < npair; ++p) {
        double sum = 0.0;
        for (int z = 0; z < nz; ++z) sum += part[(size_t)p * nz + z];
        S[p] = sum;
    }
    return DAGE_OK;
}

}  // extern "C"
