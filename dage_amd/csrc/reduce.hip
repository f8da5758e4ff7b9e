// FP64 LIP-weighted cube integral  sum_ijk wx_i wy_j wz_k f_ijk [g_ijk]
// (Integrator3D reduce3d, src/bubbles/integrator.cu:320-373; the cube part of
// Function3D_dot_product, src/bubbles/function3d.F90:1497-1530).
// Deterministic: one block per z-slice reduces with a fixed tree, then one
// block sums the slices in order.
#include "dage_internal.hpp"

namespace dage {
namespace {

__device__ inline double block_sum(double v, double *sh) {
    // wavefront reduction with DPP-friendly shuffles, then across waves
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
    return s;
}

__global__ void __launch_bounds__(256) slice_kernel(const double *f, const double *g,
                                                    const double *wx, const double *wy, int nx,
                                                    int ny, double *partial) {
    __shared__ double sh[4];
    const int k = blockIdx.x;
    const long base = (long)nx * ny * k;
    double acc = 0.0;
    for (int e = threadIdx.x; e < nx * ny; e += blockDim.x) {
        const int i = e % nx, j = e / nx;
        const double v = g ? f[base + e] * g[base + e] : f[base + e];
        acc += wx[i] * wy[j] * v;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[k] = s;
}

__global__ void __launch_bounds__(256) final_kernel(const double *partial, const double *wz,
                                                    int nz, double *result) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int k = threadIdx.x; k < nz; k += blockDim.x) acc += wz[k] * partial[k];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) *result = s;
}

__global__ void __launch_bounds__(256) fill_zero_kernel(double *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

__global__ void __launch_bounds__(256) cube_copy_kernel(double2 *dst, const double2 *src,
                                                        size_t n2) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n2;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void cube_copy_tail_kernel(double *dst, const double *src, size_t i) { dst[i] = src[i]; }

}  // namespace

int copy_device(dage_ctx ctx, double *dst, const double *src, size_t n) {
    if (n == 0) return DAGE_OK;
    const size_t n2 = n / 2;
    hipEvent_t ev;
    prof_begin(ctx, "cube_copy_kernel", &ev);
    if (n2) {
        const size_t blocks = (n2 + 255) / 256;
        const int grid = (int)(blocks < 8192 ? blocks : 8192);
        hipLaunchKernelGGL(cube_copy_kernel, dim3(grid), dim3(256), 0, ctx->stream,
                           (double2 *)dst, (const double2 *)src, n2);
    }
    if (n & 1)
        hipLaunchKernelGGL(cube_copy_tail_kernel, dim3(1), dim3(1), 0, ctx->stream, dst, src,
                           n - 1);
    DAGE_HIP_CHECK(hipGetLastError());
    prof_end(ctx, "cube_copy_kernel", ev);
    return DAGE_OK;
}

int fill_zero(dage_ctx ctx, double *p, size_t n) {
    if (n == 0) return DAGE_OK;
    const size_t blocks = (n + 255) / 256;
    const int grid = (int)(blocks < 65536 ? blocks : 65536);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(grid), dim3(256), 0, ctx->stream, p, n);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}

int cube_integrate(dage_grid3d grid, const double *f, const double *g, double *result) {
    dage_ctx ctx = grid->ctx;
    const int nx = grid->shape(0), ny = grid->shape(1), nz = grid->shape(2);
    double *buf = nullptr;
    DAGE_HIP_CHECK(hipMallocAsync((void **)&buf, sizeof(double) * (nz + 1), ctx->stream));
    hipEvent_t ev;
    prof_begin(ctx, "cube_integrate", &ev);
    hipLaunchKernelGGL(slice_kernel, dim3(nz), dim3(256), 0, ctx->stream, f, g,
                       grid->d_weights[0], grid->d_weights[1], nx, ny, buf);
    hipLaunchKernelGGL(final_kernel, dim3(1), dim3(256), 0, ctx->stream, buf,
                       grid->d_weights[2], nz, buf + nz);
    DAGE_HIP_CHECK(hipGetLastError());
    prof_end(ctx, "cube_integrate", ev);
    DAGE_HIP_CHECK(hipMemcpyAsync(result, buf + nz, sizeof(double), hipMemcpyDeviceToHost,
                                  ctx->stream));
    DAGE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    DAGE_HIP_CHECK(hipFreeAsync(buf, ctx->stream));
    return DAGE_OK;
}

}  // namespace dage
