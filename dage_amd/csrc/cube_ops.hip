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
__global__ void __launch_bounds__(256) gram_kernel(const double *const *f, const double *const *g,
                                                   int nf, const double *wx, const double *wy,
                                                   const double *wz, int nx, int ny, int nz,
                                                   double *part) {
    __shared__ double red[256];
    const int pair = blockIdx.y, z = blockIdx.x;
    const int i = pair % nf, j = pair / nf;
    const double *fi = f[i] + (long)nx * ny * z;
    const double *gj = g[j] + (long)nx * ny * z;
    double s = 0.0;
    for (int e = threadIdx.x; e < nx * ny; e += 256) {
        const int x = e % nx, y = e / nx;
        s += (wx[x] * wy[y]) * (fi[e] * gj[e]);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(long)pair * nz + z] = red[0] * wz[z];
}

template <typename T>
int to_device(const T *host, size_t n, T **dev) {
    DAGE_HIP_CHECK(hipMalloc((void **)dev, sizeof(T) * (n ? n : 1)));
    DAGE_HIP_CHECK(hipMemcpy(*dev, host, sizeof(T) * n, hipMemcpyHostToDevice));
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
    const double **d_in = nullptr;
    double **d_out = nullptr;
    double *d_coef = nullptr;
    int rc = to_device(in, nin, &d_in);
    if (rc == DAGE_OK) rc = to_device(out, nout, &d_out);
    if (rc == DAGE_OK) rc = to_device(coef, (size_t)nin * nout, &d_coef);
    if (rc == DAGE_OK) {
        hipLaunchKernelGGL(lincomb_kernel, dim3(grid_of(n)), dim3(256), 0, ctx->stream, n, nin,
                           d_in, d_coef, nout, d_out);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
            set_error("linear combination failed");
            rc = DAGE_ERR_HIP;
        }
    }
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    (void)hipFree(d_coef);
    return rc;
}

int dage_cube_gram(dage_grid3d grid, int nf, const double *const *f, int ng,
                   const double *const *g, double *S) {
    DAGE_REQUIRE(grid && f && g && S && nf > 0 && ng > 0, DAGE_ERR_ARG, "bad argument");
    dage_ctx ctx = grid->ctx;
    DAGE_TRY(use_device(ctx));
    const int nx = grid->shape(0), ny = grid->shape(1), nz = grid->shape(2);
    const int npair = nf * ng;
    DAGE_REQUIRE(npair <= 65535, DAGE_ERR_ARG, "too many pairs (%d)", npair);
    const double **d_f = nullptr, **d_g = nullptr;
    double *d_part = nullptr;
    int rc = to_device(f, nf, &d_f);
    if (rc == DAGE_OK) rc = to_device(g, ng, &d_g);
    if (rc == DAGE_OK && hipMalloc((void **)&d_part, 8 * (size_t)npair * nz) != hipSuccess) {
        set_error("device allocation failed");
        rc = DAGE_ERR_HIP;
    }
    std::vector<double> part((size_t)npair * nz);
    if (rc == DAGE_OK) {
        hipLaunchKernelGGL(gram_kernel, dim3(nz, npair), dim3(256), 0, ctx->stream, d_f, d_g, nf,
                           grid->d_weights[0], grid->d_weights[1], grid->d_weights[2], nx, ny, nz,
                           d_part);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(part.data(), d_part, 8 * part.size(), hipMemcpyDeviceToHost,
                           ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            set_error("Gram matrix failed");
            rc = DAGE_ERR_HIP;
        }
    }
    if (rc == DAGE_OK)
        for (int p = 0; p < npair; ++p) {
            double s = 0.0;
            for (int z = 0; z < nz; ++z) s += part[(size_t)p * nz + z];
            S[p] = s;
        }
    (void)hipFree(d_f);
    (void)hipFree(d_g);
    (void)hipFree(d_part);
    return rc;
}

}  // extern "C"
