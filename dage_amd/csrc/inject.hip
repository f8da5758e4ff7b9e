// Bubbles inject: cube(x,y,z) += multiplier * sum_bubbles sum_{l>=lmin,m}
//                 interp(r^k f_lm)(r) * S_lm(x-c)/r^l
// (Bubbles_eval_3dgrid, src/bubbles/bubbles.F90:2148-2211, and the CUDA
// Bubbles::inject it dispatches to, src/bubbles/bubbles_cuda.cu:3090-3340).
//
// One fused launch for all bubbles and all (l,m): one thread per cube point,
// a wavefront per 4 x 4 x 4 patch of points.  Per bubble the thread finds
// the radial cell (grid.F90:1212-1246, from a lookup table), evaluates the
// nlip cell polynomials once by Horner (lip.F90:551-566), reuses them for every
// (l,m); the Racah solid harmonics come from the recursion of
// harmonics.F90:586-659, fully unrolled for the compile-time lmax.
#include <cmath>
#include <cstring>

#include "dage_internal.hpp"

namespace dage {
namespace {

// Global-address-space views of the bubble tables: the pointers come out of
// device memory (the descriptors), where the compiler may otherwise treat
// them as generic and emit flat loads, or vector loads for uniform addresses.
typedef const double __attribute__((address_space(1))) *gdptr;
typedef const int __attribute__((address_space(1))) *giptr;

struct InjectArgs {
    const double *gx, *gy, *gz;  // cube grid points
    int nx, ny, nz;
    const BubbleDev *bubs;
    int nbub;
    int lmin;
    int nlip;
    double first;
    double multiplier;
    double *cube;
    double lip[kMaxNlip * kMaxNlip];  // lip[j + nlip*k], decreasing powers
};

template <int LMAX>
__device__ inline void solid_harmonics(double x, double y, double z, double *S) {
    S[0] = 1.0;
    if constexpr (LMAX >= 1) {
        S[1] = y;
        S[2] = z;
        S[3] = x;
    }
    const double r2 = x * x + y * y + z * z;
#pragma unroll
    for (int l = 2; l <= LMAX; ++l) {
        const int c = l * l + l, p = (l - 1) * (l - 1) + (l - 1), q = (l - 2) * (l - 2) + (l - 2);
        const double f = sqrt((2 * l - 1.0) / (2.0 * l));
        S[c - l] = f * (y * S[p + (l - 1)] + x * S[p - (l - 1)]);
#pragma unroll
        for (int m = -(l - 1); m <= l - 1; ++m) {
            double v = (2 * l - 1.0) / sqrt(1.0 * (l + m) * (l - m)) * (z * S[p + m]);
            if (l - (m < 0 ? -m : m) > 1)
                v = v - sqrt((l + m - 1.0) * (l - m - 1.0) / ((l + m) * (l - m))) * (r2 * S[q + m]);
            S[c + m] = v;
        }
        S[c + l] = f * (x * S[p + (l - 1)] - y * S[p - (l - 1)]);
    }
}

// Grid1D_get_icell (grid.F90:1212-1246): the 1-based cell c with
// celld[c-1] <= r < celld[c] (the last cell from its start up to qmax), 0
// outside [0, qmax].  The reference bisects; here the lookup-table entry one
// bin below r (a margin for the rounding of r*tscale) is at or below r's
// cell, and upward steps stop at the first cell start above r: the same
// cell, in one or two dependent loads instead of about log2(ncell).
__device__ inline int radial_icell_tab(const BubbleDev &B, double r) {
    if (!(r >= 0.0 && r <= B.qmax)) return 0;  // also rejects NaN
    const gdptr celld = (gdptr)B.celld;
    if (celld[B.ncell - 1] <= r) return B.ncell;
    const int k = min(max((int)(r * B.tscale) - 1, 0), kCellTab - 1);
    int ic = ((giptr)B.ctab)[k];
    while (celld[ic] <= r) ++ic;
    return ic;
}

// NLIP > 0: compile-time cell size (unrolled Horner and dot products);
// NLIP == 0: A.nlip at run time.  The reference divides every S_lm by r^l
// (RealSphericalHarmonics_eval, harmonics.F90:1360-1370); here each l's sum
// over m is scaled once by (1/r)^l, which moves the result by rounding only
// (1e-16 relative per term).
//
// A wavefront covers a 4 x 4 x 4 patch of points (lane l: x = l & 3,
// y = (l >> 2) & 3, z = l >> 4), a block four patches along x.  The points
// of one patch almost always fall in one or two radial cells of a bubble,
// so the wavefront loops over the distinct cells present (ballot, lowest
// lane first): the cell index is wavefront-uniform inside the loop and the
// 25 x nlip radial coefficients of the cell are read once per wavefront
// through scalar loads instead of once per lane.  Each point still sees
// exactly the arithmetic of its own cell, in the same order.
template <int LMAX, int NLIP>
__global__ void __launch_bounds__(256) bubbles_inject_kernel(InjectArgs A) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ix = blockIdx.x * 16 + wave * 4 + (lane & 3);
    const int iy = blockIdx.y * 4 + ((lane >> 2) & 3);
    const int iz = blockIdx.z * 4 + (lane >> 4);
    // no early exit: every lane takes part in the wavefront-uniform cell loop
    const bool inside = ix < A.nx && iy < A.ny && iz < A.nz;
    constexpr int NMAX = NLIP > 0 ? NLIP : kMaxNlip;
    const int nlip = NLIP > 0 ? NLIP : A.nlip;
    const double px = A.gx[inside ? ix : 0], py = A.gy[inside ? iy : 0],
                 pz = A.gz[inside ? iz : 0];
    constexpr double kEps = 2.220446049250313e-16;  // epsilon(0.d0)
    double total = 0.0;
    for (int b = 0; b < A.nbub; ++b) {
        const BubbleDev B = A.bubs[b];
        const double dx = px - B.cx, dy = py - B.cy, dz = pz - B.cz;
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        const int ic = inside ? radial_icell_tab(B, r) : 0;
        unsigned long long todo = __ballot(ic != 0);
        if (todo == 0ull) continue;
        const int icl = ic > 0 ? ic - 1 : 0;
        const double crd = (r - ((gdptr)B.celld)[icl]) / ((gdptr)B.steps)[icl] + A.first;
        double polys[NMAX];
#pragma unroll
        for (int j = 0; j < NMAX; ++j) {
            if (j >= nlip) break;
            double v = A.lip[j];
#pragma unroll
            for (int q = 1; q < NMAX; ++q) {
                if (q >= nlip) break;
                v = v * crd + A.lip[j + nlip * q];
            }
            polys[j] = v;
        }
        double S[(LMAX + 1) * (LMAX + 1)];
        solid_harmonics<LMAX>(dx, dy, dz, S);
        const double inv = (r > kEps) ? 1.0 / r : 1.0;  // r <= eps: r^l taken as 1
        while (todo != 0ull) {
            const int leader = __builtin_ctzll(todo);
            const int c = __builtin_amdgcn_readlane(ic, leader);  // uniform cell
            const bool mine = ic == c;
            todo &= ~__ballot(mine);
            if (!mine) continue;
            // re-read from the first active lane so the address stays scalar
            // (inside the branch the compiler may otherwise substitute the
            // per-lane ic for c)
            const int cs = __builtin_amdgcn_readfirstlane(c);
            const gdptr f = (gdptr)(B.f + (long)(cs - 1) * (nlip - 1));
            double res = 0.0;
            double ipow = 1.0;
#pragma unroll
            for (int l = 0; l <= LMAX; ++l) {
                if (l >= A.lmin) {
                    double sl = 0.0;
#pragma unroll
                    for (int m = -l; m <= l; ++m) {
                        const int lm = l * l + l + m;
                        const gdptr fl = f + (long)lm * B.npts;
                        double fv = 0.0;
#pragma unroll
                        for (int j = 0; j < NMAX; ++j) {
                            if (j >= nlip) break;
                            fv += fl[j] * polys[j];
                        }
                        sl += fv * S[lm];
                    }
                    res += sl * ipow;
                }
                ipow *= inv;
            }
            total += res;
        }
    }
    if (inside) {
        const long id = ix + (long)A.nx * (iy + (long)A.ny * iz);
        A.cube[id] = A.cube[id] + A.multiplier * total;
    }
}

// Bubbles with 6 < lmax <= 12 (products of two bubble sets, l > lmax parts
// injected after a transform): the (lmax+1)^2 solid harmonics would not fit
// in registers, so each l's 2l+1 values are generated from the previous two
// l's by the same recursion (harmonics.F90:586-659) right before they are
// consumed, per radial cell present in the wavefront.
template <int LMAX, int NLIP>
__global__ void __launch_bounds__(256) bubbles_inject_hi_kernel(InjectArgs A) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ix = blockIdx.x * 16 + wave * 4 + (lane & 3);
    const int iy = blockIdx.y * 4 + ((lane >> 2) & 3);
    const int iz = blockIdx.z * 4 + (lane >> 4);
    const bool inside = ix < A.nx && iy < A.ny && iz < A.nz;
    constexpr int NMAX = NLIP > 0 ? NLIP : kMaxNlip;
    constexpr int W = 2 * LMAX + 1;
    const int nlip = NLIP > 0 ? NLIP : A.nlip;
    const double px = A.gx[inside ? ix : 0], py = A.gy[inside ? iy : 0],
                 pz = A.gz[inside ? iz : 0];
    constexpr double kEps = 2.220446049250313e-16;  // epsilon(0.d0)
    double total = 0.0;
    for (int b = 0; b < A.nbub; ++b) {
        const BubbleDev B = A.bubs[b];
        const double dx = px - B.cx, dy = py - B.cy, dz = pz - B.cz;
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        const int ic = inside ? radial_icell_tab(B, r) : 0;
        unsigned long long todo = __ballot(ic != 0);
        if (todo == 0ull) continue;
        const int icl = ic > 0 ? ic - 1 : 0;
        const double crd = (r - ((gdptr)B.celld)[icl]) / ((gdptr)B.steps)[icl] + A.first;
        double polys[NMAX];
#pragma unroll
        for (int j = 0; j < NMAX; ++j) {
            if (j >= nlip) break;
            double v = A.lip[j];
#pragma unroll
            for (int q = 1; q < NMAX; ++q) {
                if (q >= nlip) break;
                v = v * crd + A.lip[j + nlip * q];
            }
            polys[j] = v;
        }
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double inv = (r > kEps) ? 1.0 / r : 1.0;
        while (todo != 0ull) {
            const int leader = __builtin_ctzll(todo);
            const int c = __builtin_amdgcn_readlane(ic, leader);
            const bool mine = ic == c;
            todo &= ~__ballot(mine);
            if (!mine) continue;
            const int cs = __builtin_amdgcn_readfirstlane(c);
            const gdptr f = (gdptr)(B.f + (long)(cs - 1) * (nlip - 1));
            // rolling solid harmonics: pp = S_(l-2), p = S_(l-1), q = S_l,
            // entry m + l of each
            double pp[W], p[W], q[W];
            double res = 0.0, ipow = 1.0;
#pragma unroll
            for (int l = 0; l <= LMAX; ++l) {
                if (l == 0) {
                    q[0] = 1.0;
                } else if (l == 1) {
                    q[0] = dy;
                    q[1] = dz;
                    q[2] = dx;
                } else {
                    const double fl = sqrt((2 * l - 1.0) / (2.0 * l));
                    q[0] = fl * (dy * p[2 * (l - 1)] + dx * p[0]);
#pragma unroll
                    for (int m = -(l - 1); m <= l - 1; ++m) {
                        double v = (2 * l - 1.0) / sqrt(1.0 * (l + m) * (l - m)) *
                                   (dz * p[m + l - 1]);
                        if (l - (m < 0 ? -m : m) > 1)
                            v = v - sqrt((l + m - 1.0) * (l - m - 1.0) / ((l + m) * (l - m))) *
                                        (r2 * pp[m + l - 2]);
                        q[m + l] = v;
                    }
                    q[2 * l] = fl * (dx * p[2 * (l - 1)] - dy * p[0]);
                }
                if (l >= A.lmin) {
                    double sl = 0.0;
#pragma unroll
                    for (int m = -l; m <= l; ++m) {
                        const gdptr fl_ = f + (long)(l * l + l + m) * B.npts;
                        double fv = 0.0;
#pragma unroll
                        for (int j = 0; j < NMAX; ++j) {
                            if (j >= nlip) break;
                            fv += fl_[j] * polys[j];
                        }
                        sl += fv * q[m + l];
                    }
                    res += sl * ipow;
                }
                ipow *= inv;
#pragma unroll
                for (int m = 0; m < W; ++m) {
                    pp[m] = p[m];
                    p[m] = q[m];
                }
            }
            total += res;
        }
    }
    if (inside) {
        const long id = ix + (long)A.nx * (iy + (long)A.ny * iz);
        A.cube[id] = A.cube[id] + A.multiplier * total;
    }
}

template <int LMAX>
void launch_inject_hi(const InjectArgs &A, dim3 g, dim3 block, hipStream_t s) {
    if (A.nlip == 7)
        hipLaunchKernelGGL((bubbles_inject_hi_kernel<LMAX, 7>), g, block, 0, s, A);
    else if (A.nlip == 5)
        hipLaunchKernelGGL((bubbles_inject_hi_kernel<LMAX, 5>), g, block, 0, s, A);
    else
        hipLaunchKernelGGL((bubbles_inject_hi_kernel<LMAX, 0>), g, block, 0, s, A);
}

template <int LMAX>
void launch_inject(const InjectArgs &A, dim3 g, dim3 block, hipStream_t s) {
    if (A.nlip == 7)
        hipLaunchKernelGGL((bubbles_inject_kernel<LMAX, 7>), g, block, 0, s, A);
    else if (A.nlip == 5)
        hipLaunchKernelGGL((bubbles_inject_kernel<LMAX, 5>), g, block, 0, s, A);
    else
        hipLaunchKernelGGL((bubbles_inject_kernel<LMAX, 0>), g, block, 0, s, A);
}

}  // namespace

int bubbles_sync_device(dage_bubbles bub) {
    if (!bub->dirty) return DAGE_OK;
    std::vector<BubbleDev> desc(bub->nbub);
    for (int b = 0; b < bub->nbub; ++b) {
        const BubbleHost &h = bub->host[b];
        BubbleDev &d = desc[b];
        d.cx = h.center[0];
        d.cy = h.center[1];
        d.cz = h.center[2];
        d.qmax = h.qmax;
        d.ncell = h.ncell;
        d.npts = h.ncell * (bub->nlip - 1) + 1;
        d.celld = h.d_buf;
        d.steps = h.d_buf + h.ncell;
        d.f = h.d_buf + 2 * h.ncell;
        d.ctab = h.d_tab;
        d.tscale = kCellTab / h.qmax;
    }
    DAGE_HIP_CHECK(hipMemcpy(bub->d_desc, desc.data(), sizeof(BubbleDev) * bub->nbub,
                             hipMemcpyHostToDevice));
    bub->dirty = false;
    return DAGE_OK;
}

int bubbles_inject(dage_bubbles bub, dage_grid3d grid, int lmin, double multiplier,
                   double *cube) {
    dage_ctx ctx = bub->ctx;
    DAGE_TRY(bubbles_sync_device(bub));
    InjectArgs A{};
    A.gx = grid->d_points[0];
    A.gy = grid->d_points[1];
    A.gz = grid->d_points[2];
    A.nx = grid->shape(0);
    A.ny = grid->shape(1);
    A.nz = grid->shape(2);
    A.bubs = bub->d_desc;
    A.nbub = bub->nbub;
    A.lmin = lmin;
    A.nlip = bub->nlip;
    double last;
    lip_first_last(bub->nlip, &A.first, &last);
    DAGE_REQUIRE(lip_coeffs(bub->nlip, 1, A.lip) == 0, DAGE_ERR_ARG, "unsupported bubble nlip");
    A.multiplier = multiplier;
    A.cube = cube;
    const dim3 block(256, 1, 1);  // four 4x4x4 patches along x
    const dim3 g((A.nx + 15) / 16, (A.ny + 3) / 4, (A.nz + 3) / 4);
    hipEvent_t ev;
    prof_begin(ctx, "bubbles_inject_kernel", &ev);
    switch (bub->lmax) {
    case 0: launch_inject<0>(A, g, block, ctx->stream); break;
    case 1: launch_inject<1>(A, g, block, ctx->stream); break;
    case 2: launch_inject<2>(A, g, block, ctx->stream); break;
    case 3: launch_inject<3>(A, g, block, ctx->stream); break;
    case 4: launch_inject<4>(A, g, block, ctx->stream); break;
    case 5: launch_inject<5>(A, g, block, ctx->stream); break;
    case 6: launch_inject<6>(A, g, block, ctx->stream); break;
    case 7: launch_inject_hi<7>(A, g, block, ctx->stream); break;
    case 8: launch_inject_hi<8>(A, g, block, ctx->stream); break;
    case 9: launch_inject_hi<9>(A, g, block, ctx->stream); break;
    case 10: launch_inject_hi<10>(A, g, block, ctx->stream); break;
    case 11: launch_inject_hi<11>(A, g, block, ctx->stream); break;
    case 12: launch_inject_hi<12>(A, g, block, ctx->stream); break;
    default:
        set_error("bubbles lmax %d outside 0..%d", bub->lmax, kMaxBubbleLmax);
        return DAGE_ERR_UNSUPPORTED;
    }
    DAGE_HIP_CHECK(hipGetLastError());
    prof_end(ctx, "bubbles_inject_kernel", ev);
    return DAGE_OK;
}

}  // namespace dage
