// C-ABI entry points (include/dage_ops.h).  Argument checking, handle
// lifetime, host<->device staging and profiling live here; the numerics are in
// host_math.cpp, matrices.hip, apply_f64.hip and reduce.hip.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>

#include "dage_internal.hpp"

namespace dage {

namespace {
thread_local char g_err[1024] = "";
}

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

const char *last_error() { return g_err; }

namespace {
std::mutex g_ref_mutex;
}

int ctx_scratch(dage_ctx ctx, size_t bytes, void **p) {
    if (ctx->scratch_stream && ctx->scratch_stream != ctx->stream)
        DAGE_HIP_CHECK(hipStreamSynchronize(ctx->scratch_stream));
    ctx->scratch_stream = ctx->stream;
    if (bytes > ctx->scratch_bytes) {
        if (ctx->scratch) {
            DAGE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            (void)hipFree(ctx->scratch);
        }
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
        const size_t n = std::max(bytes, (size_t)1 << 20);
        DAGE_HIP_CHECK(hipMalloc(&ctx->scratch, n));
        ctx->scratch_bytes = n;
    }
    *p = ctx->scratch;
    return DAGE_OK;
}

int ctx_aux(dage_ctx ctx, size_t bytes, void **p) {
    if (bytes > ctx->aux_bytes) {
        if (ctx->aux) {
            DAGE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            (void)hipFree(ctx->aux);
        }
        ctx->aux = nullptr;
        ctx->aux_bytes = 0;
        const size_t n = std::max(bytes, (size_t)1 << 20);
        DAGE_HIP_CHECK(hipMalloc(&ctx->aux, n));
        ctx->aux_bytes = n;
    }
    *p = ctx->aux;
    return DAGE_OK;
}

void ctx_retain(dage_ctx ctx) {
    std::lock_guard<std::mutex> lock(g_ref_mutex);
    ++ctx->refs;
}

void ctx_release(dage_ctx ctx) {
    bool last;
    {
        std::lock_guard<std::mutex> lock(g_ref_mutex);
        last = (--ctx->refs == 0);
    }
    if (!last) return;
    (void)hipSetDevice(ctx->device);
    // an external stream set by the caller may already be gone; only the
    // library's own stream is known to be alive here
    (void)hipStreamSynchronize(ctx->own_stream);
    for (auto &p : ctx->pending) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    if (ctx->timer_start) (void)hipEventDestroy(ctx->timer_start);
    if (ctx->timer_stop) (void)hipEventDestroy(ctx->timer_stop);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->aux) (void)hipFree(ctx->aux);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

void prof_begin(dage_ctx ctx, const char *name, hipEvent_t *start) {
    (void)name;
    *start = nullptr;
    if (!ctx->profiling) return;
    if (hipEventCreate(start) != hipSuccess) {
        *start = nullptr;
        return;
    }
    (void)hipEventRecord(*start, ctx->stream);
}

void prof_end(dage_ctx ctx, const char *name, hipEvent_t start) {
    if (!ctx->profiling || !start) return;
    hipEvent_t stop;
    if (hipEventCreate(&stop) != hipSuccess) {
        (void)hipEventDestroy(start);
        return;
    }
    (void)hipEventRecord(stop, ctx->stream);
    ctx->pending.push_back({start, stop, name});
}

int prof_resolve(dage_ctx ctx) {
    if (ctx->pending.empty()) return DAGE_OK;
    DAGE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (auto &p : ctx->pending) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, p.start, p.stop);
        auto &e = ctx->profile[p.name];
        e.total_ms += ms;
        e.launches += 1;
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    ctx->pending.clear();
    return DAGE_OK;
}

}  // namespace dage

using namespace dage;

// ------------------------------------------------ allocation-free apply
// dage_operator_apply_device_noalloc keeps its intermediates in this fixed
// array: one copy per GPU, held for the life of the process.  Host code never
// passes the array's name to a kernel (on the host it is only a stand-in);
// it uses the device address hipGetSymbolAddress returns, looked up once per
// device at context creation with that device current.
constexpr size_t kNoAllocScratchBytes = 128ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mutex;

// This device's address of `scratch` (the device must be current).
int scratch_address(int device, void **p) {
    DAGE_REQUIRE(device >= 0 && device < kMaxDevices, DAGE_ERR_UNSUPPORTED,
                 "device %d beyond the %d the library tracks", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_dev[device]) {
        void *ptr = nullptr;
        DAGE_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_dev[device] = ptr;
    }
    *p = g_scratch_dev[device];
    return DAGE_OK;
}
}  // namespace

namespace {

int use_device(dage_ctx ctx) {
    DAGE_REQUIRE(ctx != nullptr, DAGE_ERR_ARG, "null context handle");
    DAGE_HIP_CHECK(hipSetDevice(ctx->device));
    return DAGE_OK;
}

void free_operator_buffers(dage_operator op) {
    double **bufs[] = {&op->d_fx, &op->d_fy, &op->d_fz, &op->d_gz,
                       &op->d_w,  &op->d_v,  &op->d_in, &op->d_out};
    for (double **b : bufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    if (op->d_planes) (void)hipFree(op->d_planes);
    op->d_planes = nullptr;
    if (op->d_mask) (void)hipFree(op->d_mask);
    op->d_mask = nullptr;
    if (op->d_tmask) (void)hipFree(op->d_tmask);
    op->d_tmask = nullptr;
    if (op->d_xband) (void)hipFree(op->d_xband);
    op->d_xband = nullptr;
    if (op->d_yband) (void)hipFree(op->d_yband);
    op->d_yband = nullptr;
    if (op->d_band) (void)hipFree(op->d_band);
    op->d_band = nullptr;
    if (op->d_lr) (void)hipFree(op->d_lr);
    op->d_lr = nullptr;
    op->lr = LowRank{};
    op->planes_valid = false;
    op->ws_bytes = 0;
}

int alloc_dev(double **p, size_t elems) {
    DAGE_HIP_CHECK(hipMalloc((void **)p, sizeof(double) * (elems ? elems : 1)));
    return DAGE_OK;
}

bool same_axis(const Axis &a, const Axis &b) {
    return a.ncell == b.ncell && a.nlip == b.nlip && a.grid_type == b.grid_type &&
           a.points == b.points;
}

int upload_weights(dage_operator op) {
    DAGE_HIP_CHECK(hipMemcpyAsync(op->d_w, op->w.data(), sizeof(double) * op->ntp,
                                  hipMemcpyHostToDevice, op->ctx->stream));
    // op->w is a member vector that outlives the copy; later host writes to
    // it go through set_weights, which synchronises first.
    return DAGE_OK;
}

}  // namespace

extern "C" {

int dage_abi_version(void) { return DAGE_ABI_VERSION; }

const char *dage_last_error(void) { return last_error(); }

// ------------------------------------------------------------------ context
int dage_ctx_create(int device, dage_ctx *out) {
    DAGE_REQUIRE(out != nullptr, DAGE_ERR_ARG, "null output pointer");
    *out = nullptr;
    int count = 0;
    DAGE_HIP_CHECK(hipGetDeviceCount(&count));
    DAGE_REQUIRE(device >= 0 && device < count, DAGE_ERR_ARG,
                 "device %d out of range (%d visible)", device, count);
    DAGE_HIP_CHECK(hipSetDevice(device));
    auto *ctx = new (std::nothrow) dage_ctx_s();
    DAGE_REQUIRE(ctx != nullptr, DAGE_ERR_ALLOC, "out of host memory");
    ctx->device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        set_error("hipStreamCreateWithFlags failed: %s", hipGetErrorString(e));
        return DAGE_ERR_HIP;
    }
    ctx->stream = ctx->own_stream;
    int rc = prepare_fused_kernels();
    if (rc == DAGE_OK) rc = scratch_address(device, &ctx->noalloc_scratch);
    if (rc != DAGE_OK) {
        (void)hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return DAGE_OK;
}

int dage_ctx_destroy(dage_ctx ctx) {
    if (!ctx) return DAGE_OK;
    ctx_release(ctx);
    return DAGE_OK;
}

int dage_ctx_set_stream(dage_ctx ctx, void *stream) {
    DAGE_TRY(use_device(ctx));
    DAGE_TRY(prof_resolve(ctx));
    // work on the outgoing stream may still use the scratch buffer, and that
    // stream may be destroyed by its owner once switched away: wait for it
    // here and forget it, so ctx_scratch never touches a stale handle
    DAGE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->scratch_stream = nullptr;
    ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
    return DAGE_OK;
}

int dage_ctx_get_stream(dage_ctx ctx, void **stream) {
    DAGE_REQUIRE(ctx && stream, DAGE_ERR_ARG, "null argument");
    *stream = ctx->stream;
    return DAGE_OK;
}

int dage_ctx_synchronize(dage_ctx ctx) {
    DAGE_TRY(use_device(ctx));
    DAGE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return DAGE_OK;
}

int dage_ctx_profile(dage_ctx ctx, int enable) {
    DAGE_TRY(use_device(ctx));
    DAGE_TRY(prof_resolve(ctx));
    ctx->profile.clear();
    ctx->profiling = enable != 0;
    return DAGE_OK;
}

int dage_ctx_profile_query(dage_ctx ctx, const char *name, double *total_ms, int *launches) {
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(name && total_ms && launches, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(prof_resolve(ctx));
    auto it = ctx->profile.find(name);
    *total_ms = it == ctx->profile.end() ? 0.0 : it->second.total_ms;
    *launches = it == ctx->profile.end() ? 0 : it->second.launches;
    return DAGE_OK;
}

int dage_ctx_time_begin(dage_ctx ctx) {
    DAGE_TRY(use_device(ctx));
    if (!ctx->timer_start) DAGE_HIP_CHECK(hipEventCreate(&ctx->timer_start));
    if (!ctx->timer_stop) DAGE_HIP_CHECK(hipEventCreate(&ctx->timer_stop));
    DAGE_HIP_CHECK(hipEventRecord(ctx->timer_start, ctx->stream));
    return DAGE_OK;
}

int dage_ctx_time_end(dage_ctx ctx, double *ms) {
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(ms != nullptr, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE(ctx->timer_start && ctx->timer_stop, DAGE_ERR_STATE,
                 "dage_ctx_time_end without dage_ctx_time_begin");
    DAGE_HIP_CHECK(hipEventRecord(ctx->timer_stop, ctx->stream));
    DAGE_HIP_CHECK(hipEventSynchronize(ctx->timer_stop));
    float f = 0.f;
    DAGE_HIP_CHECK(hipEventElapsedTime(&f, ctx->timer_start, ctx->timer_stop));
    *ms = f;
    return DAGE_OK;
}

// -------------------------------------------------------------- device cube
int dage_cube_create(dage_ctx ctx, const int shape[3], dage_cube *out) {
    DAGE_REQUIRE(out != nullptr, DAGE_ERR_ARG, "null output pointer");
    *out = nullptr;
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(shape && shape[0] > 0 && shape[1] > 0 && shape[2] > 0, DAGE_ERR_ARG,
                 "cube shape must be positive");
    auto *c = new (std::nothrow) dage_cube_s();
    DAGE_REQUIRE(c != nullptr, DAGE_ERR_ALLOC, "out of host memory");
    c->ctx = ctx;
    for (int a = 0; a < 3; ++a) c->shape[a] = shape[a];
    hipError_t e = hipMalloc((void **)&c->d, sizeof(double) * c->count());
    if (e != hipSuccess) {
        delete c;
        set_error("hipMalloc of a %d x %d x %d cube failed: %s", shape[0], shape[1], shape[2],
                  hipGetErrorString(e));
        return DAGE_ERR_ALLOC;
    }
    // a new cube holds zeros (cube += bubbles and the other accumulating
    // calls may be its first writers)
    const int rc = fill_zero(ctx, c->d, c->count());
    if (rc != DAGE_OK) {
        (void)hipFree(c->d);
        delete c;
        return rc;
    }
    ctx_retain(ctx);
    *out = c;
    return DAGE_OK;
}

int dage_cube_destroy(dage_cube c) {
    if (!c) return DAGE_OK;
    dage_ctx ctx = c->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(c->d);
    delete c;
    ctx_release(ctx);
    return DAGE_OK;
}

int dage_cube_upload(dage_cube c, const double *host) {
    DAGE_REQUIRE(c && host, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(use_device(c->ctx));
    DAGE_HIP_CHECK(hipMemcpyAsync(c->d, host, sizeof(double) * c->count(),
                                  hipMemcpyHostToDevice, c->ctx->stream));
    DAGE_HIP_CHECK(hipStreamSynchronize(c->ctx->stream));
    return DAGE_OK;
}

int dage_cube_download(dage_cube c, double *host) {
    DAGE_REQUIRE(c && host, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(use_device(c->ctx));
    DAGE_HIP_CHECK(hipMemcpyAsync(host, c->d, sizeof(double) * c->count(),
                                  hipMemcpyDeviceToHost, c->ctx->stream));
    DAGE_HIP_CHECK(hipStreamSynchronize(c->ctx->stream));
    return DAGE_OK;
}

int dage_cube_set_to_zero(dage_cube c) {
    DAGE_REQUIRE(c != nullptr, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(use_device(c->ctx));
    DAGE_TRY(fill_zero(c->ctx, c->d, c->count()));
    return DAGE_OK;
}

int dage_cube_copy(dage_cube dst, dage_cube src) {
    DAGE_REQUIRE(dst && src, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE(dst->ctx == src->ctx, DAGE_ERR_ARG, "cubes belong to different contexts");
    DAGE_REQUIRE(dst->shape[0] == src->shape[0] && dst->shape[1] == src->shape[1] &&
                     dst->shape[2] == src->shape[2],
                 DAGE_ERR_ARG, "cube shapes differ");
    DAGE_TRY(use_device(dst->ctx));
    if (dst == src) return DAGE_OK;
    return copy_device(dst->ctx, dst->d, src->d, dst->count());
}

int dage_cube_device_pointer(dage_cube c, double **dev) {
    DAGE_REQUIRE(c && dev, DAGE_ERR_ARG, "null argument");
    *dev = c->d;
    return DAGE_OK;
}

// --------------------------------------------------------------- quadrature
int dage_gauss_quad(int nlin, int nlog, double tstart, double tlin, double tlog, double *tp,
                    double *tw) {
    DAGE_REQUIRE(nlin >= 1 && nlog >= 0 && tp && tw, DAGE_ERR_ARG, "invalid quadrature sizes");
    DAGE_REQUIRE(tlin > tstart && (nlog == 0 || (tlog > tlin && tlin > 0.0)), DAGE_ERR_ARG,
                 "quadrature needs tstart < tlin < tlog");
    gauss_quad(nlin, nlog, tstart, tlin, tlog, tp, tw);
    return DAGE_OK;
}

// --------------------------------------------------------------------- grid
int dage_grid3d_create(dage_ctx ctx, const double qmin[3], const int ncell[3], int nlip,
                       const double *sx, const double *sy, const double *sz, int grid_type,
                       dage_grid3d *out) {
    DAGE_REQUIRE(out != nullptr, DAGE_ERR_ARG, "null output pointer");
    *out = nullptr;
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(qmin && ncell && sx && sy && sz, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE(nlip >= 2 && nlip <= 7, DAGE_ERR_UNSUPPORTED,
                 "nlip %d outside 2..7 (the reference's closed forms stop at expo_6)", nlip);
    DAGE_REQUIRE(grid_type == 1 || (grid_type == 2 && (nlip == 6 || nlip == 7)),
                 DAGE_ERR_UNSUPPORTED, "grid_type %d with nlip %d is not supported", grid_type,
                 nlip);
    const double *steps[3] = {sx, sy, sz};
    for (int a = 0; a < 3; ++a) {
        DAGE_REQUIRE(ncell[a] >= 1, DAGE_ERR_ARG, "ncell[%d] = %d", a, ncell[a]);
        for (int c = 0; c < ncell[a]; ++c)
            DAGE_REQUIRE(steps[a][c] > 0.0 && std::isfinite(steps[a][c]), DAGE_ERR_ARG,
                         "step %d of axis %d is not positive", c, a);
    }
    auto *g = new (std::nothrow) dage_grid3d_s();
    DAGE_REQUIRE(g != nullptr, DAGE_ERR_ALLOC, "out of host memory");
    g->ctx = ctx;
    ctx_retain(ctx);
    for (int a = 0; a < 3; ++a) {
        if (make_axis(qmin[a], ncell[a], nlip, steps[a], grid_type, &g->axis[a])) {
            ctx_release(ctx);
            delete g;
            set_error("invalid grid axis %d", a);
            return DAGE_ERR_ARG;
        }
    }
    for (int a = 0; a < 3; ++a) {
        const Axis &ax = g->axis[a];
        double base[kMaxNlip];
        lip_integrals(nlip, grid_type, base);
        std::vector<double> w(ax.n(), 0.0);
        for (int c = 0; c < ax.ncell; ++c)
            for (int j = 0; j < nlip; ++j) w[c * (nlip - 1) + j] += base[j] * ax.steps[c];
        if (hipMalloc((void **)&g->d_weights[a], sizeof(double) * ax.n()) != hipSuccess ||
            hipMemcpy(g->d_weights[a], w.data(), sizeof(double) * ax.n(),
                      hipMemcpyHostToDevice) != hipSuccess ||
            hipMalloc((void **)&g->d_points[a], sizeof(double) * ax.n()) != hipSuccess ||
            hipMemcpy(g->d_points[a], ax.points.data(), sizeof(double) * ax.n(),
                      hipMemcpyHostToDevice) != hipSuccess) {
            for (auto &p : g->d_weights)
                if (p) (void)hipFree(p);
            for (auto &p : g->d_points)
                if (p) (void)hipFree(p);
            ctx_release(ctx);
            delete g;
            set_error("device allocation for grid weights failed");
            return DAGE_ERR_ALLOC;
        }
    }
    *out = g;
    return DAGE_OK;
}

int dage_grid3d_create_equidistant(dage_ctx ctx, const double ranges[6], int nlip,
                                   double stepmax, int grid_type, dage_grid3d *out) {
    DAGE_REQUIRE(ranges && stepmax > 0.0 && nlip >= 2, DAGE_ERR_ARG, "invalid arguments");
    double qmin[3];
    int ncell[3];
    std::vector<double> steps[3];
    for (int a = 0; a < 3; ++a) {
        const double r0 = ranges[2 * a], r1 = ranges[2 * a + 1];
        DAGE_REQUIRE(r1 > r0, DAGE_ERR_ARG, "empty range on axis %d", a);
        ncell[a] = (int)std::ceil((r1 - r0) / (stepmax * (nlip - 1)));
        const double s = (r1 - r0) / (ncell[a] * (nlip - 1));
        steps[a].assign(ncell[a], s);
        qmin[a] = r0;
    }
    return dage_grid3d_create(ctx, qmin, ncell, nlip, steps[0].data(), steps[1].data(),
                              steps[2].data(), grid_type, out);
}

int dage_grid3d_destroy(dage_grid3d g) {
    if (!g) return DAGE_OK;
    dage_ctx ctx = g->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &p : g->d_weights)
        if (p) (void)hipFree(p);
    for (auto &p : g->d_points)
        if (p) (void)hipFree(p);
    for (int a = 0; a < 3; ++a) {
        if (g->d_deriv[a]) (void)hipFree(g->d_deriv[a]);
        if (g->d_stiff[a]) (void)hipFree(g->d_stiff[a]);
    }
    delete g;
    ctx_release(ctx);
    return DAGE_OK;
}

int dage_grid3d_shape(dage_grid3d g, int shape[3]) {
    DAGE_REQUIRE(g && shape, DAGE_ERR_ARG, "null argument");
    for (int a = 0; a < 3; ++a) shape[a] = g->shape(a);
    return DAGE_OK;
}

int dage_grid3d_coord(dage_grid3d g, int axis, double *points) {
    DAGE_REQUIRE(g && points && axis >= 0 && axis < 3, DAGE_ERR_ARG, "invalid argument");
    std::memcpy(points, g->axis[axis].points.data(), sizeof(double) * g->shape(axis));
    return DAGE_OK;
}

int dage_grid3d_weights(dage_grid3d g, int axis, double *weights) {
    DAGE_REQUIRE(g && weights && axis >= 0 && axis < 3, DAGE_ERR_ARG, "invalid argument");
    DAGE_TRY(use_device(g->ctx));
    DAGE_HIP_CHECK(hipMemcpy(weights, g->d_weights[axis], sizeof(double) * g->shape(axis),
                             hipMemcpyDeviceToHost));
    return DAGE_OK;
}

// ----------------------------------------------------------------- operator
int dage_coulomb3d_create(dage_ctx ctx, dage_grid3d gin, dage_grid3d gout, const double *tp,
                          const double *tw, int ntp, double tlog, int flags,
                          dage_operator *out) {
    DAGE_REQUIRE(out != nullptr, DAGE_ERR_ARG, "null output pointer");
    *out = nullptr;
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(gin && gout && tp && tw && ntp >= 1, DAGE_ERR_ARG, "invalid arguments");
    DAGE_REQUIRE(gin->ctx == ctx && gout->ctx == ctx, DAGE_ERR_ARG,
                 "grids belong to another context");
    DAGE_REQUIRE(tlog > 0.0, DAGE_ERR_ARG, "tlog must be positive");
    for (int a = 0; a < 3; ++a)
        DAGE_REQUIRE(gin->axis[a].nlip == gout->axis[a].nlip &&
                         gin->axis[a].grid_type == gout->axis[a].grid_type,
                     DAGE_ERR_ARG, "input and output grids must share nlip and grid_type");
    for (int i = 0; i < ntp; ++i)
        DAGE_REQUIRE(std::isfinite(tp[i]) && tp[i] >= 0.0 && std::isfinite(tw[i]), DAGE_ERR_ARG,
                     "t-point %d is not a finite non-negative value", i);
    auto *op = new (std::nothrow) dage_operator_s();
    DAGE_REQUIRE(op != nullptr, DAGE_ERR_ALLOC, "out of host memory");
    op->ctx = ctx;
    op->gin = gin;
    op->gout = gout;
    op->ntp = ntp;
    op->tp.assign(tp, tp + ntp);
    op->tw.assign(tw, tw + ntp);
    op->w.resize(ntp);
    const double c = two_over_sqrtpi();
    for (int i = 0; i < ntp; ++i) op->w[i] = c * tw[i];
    op->coda = pi() / (tlog * tlog);
    op->flags = flags;
    bool same = true;
    for (int a = 0; a < 3; ++a) {
        op->din[a] = gin->shape(a);
        op->dout[a] = gout->shape(a);
        same = same && same_axis(gin->axis[a], gout->axis[a]);
    }
    if (!same && !(flags & DAGE_OP_NO_CODA)) {
        delete op;
        set_error("the coda term needs identical input and output grids (cube_project onto "
                  "another grid is not implemented); pass DAGE_OP_NO_CODA");
        return DAGE_ERR_UNSUPPORTED;
    }
    const size_t nfx = (size_t)op->dout[0] * op->din[0] * ntp;
    const size_t nfy = (size_t)op->din[1] * op->dout[1] * ntp;
    const size_t nfz = (size_t)op->din[2] * op->dout[2] * ntp;
    int rc = alloc_dev(&op->d_fx, nfx);
    if (rc == DAGE_OK) rc = alloc_dev(&op->d_fy, nfy);
    if (rc == DAGE_OK) rc = alloc_dev(&op->d_fz, nfz);
    if (rc == DAGE_OK) rc = alloc_dev(&op->d_gz, nfz);
    if (rc == DAGE_OK) rc = alloc_dev(&op->d_w, ntp);
    if (rc == DAGE_OK) rc = upload_weights(op);
    if (rc == DAGE_OK) rc = generate_matrices(op);
    if (rc == DAGE_OK) rc = flush_matrices(op);
    if (rc == DAGE_OK) rc = relayout_fz(op);
    if (rc == DAGE_OK) rc = build_masks(op);
    if (rc == DAGE_OK) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            set_error("operator setup failed: %s", hipGetErrorString(e));
            rc = DAGE_ERR_HIP;
        }
    }
    if (rc != DAGE_OK) {
        free_operator_buffers(op);
        delete op;
        return rc;
    }
    // the grids are only read while the matrices are generated
    op->gin = op->gout = nullptr;
    ctx_retain(ctx);
    *out = op;
    return DAGE_OK;
}

int dage_operator_set_weights(dage_operator op, const double *w) {
    DAGE_REQUIRE(op && w, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(use_device(op->ctx));
    DAGE_HIP_CHECK(hipStreamSynchronize(op->ctx->stream));
    op->w.assign(w, w + op->ntp);
    return upload_weights(op);
}

int dage_operator_get_weights(dage_operator op, double *w) {
    DAGE_REQUIRE(op && w, DAGE_ERR_ARG, "null argument");
    std::memcpy(w, op->w.data(), sizeof(double) * op->ntp);
    return DAGE_OK;
}

int dage_operator_set_energy(dage_operator op, double energy) {
    DAGE_REQUIRE(op != nullptr, DAGE_ERR_ARG, "null operator");
    DAGE_REQUIRE(std::isfinite(energy), DAGE_ERR_ARG, "energy is not finite");
    DAGE_TRY(use_device(op->ctx));
    DAGE_HIP_CHECK(hipStreamSynchronize(op->ctx->stream));
    const double c = two_over_sqrtpi();
    for (int i = 0; i < op->ntp; ++i) {
        double wi = op->tw[i];
        if (energy < -0.005) wi = wi * std::exp(2.0 * energy / (4.0 * (op->tp[i] * op->tp[i])));
        op->w[i] = c * wi;
    }
    op->helmholtz = true;
    op->energy = energy;
    return upload_weights(op);
}

int dage_operator_set_matrices(dage_operator op, const double *fx, const double *fy,
                               const double *fz) {
    DAGE_REQUIRE(op && fx && fy && fz, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(use_device(op->ctx));
    const size_t T = op->ntp;
    hipStream_t s = op->ctx->stream;
    DAGE_HIP_CHECK(hipMemcpyAsync(op->d_fx, fx, 8 * (size_t)op->dout[0] * op->din[0] * T,
                                  hipMemcpyHostToDevice, s));
    DAGE_HIP_CHECK(hipMemcpyAsync(op->d_fy, fy, 8 * (size_t)op->din[1] * op->dout[1] * T,
                                  hipMemcpyHostToDevice, s));
    DAGE_HIP_CHECK(hipMemcpyAsync(op->d_fz, fz, 8 * (size_t)op->din[2] * op->dout[2] * T,
                                  hipMemcpyHostToDevice, s));
    DAGE_TRY(flush_matrices(op));
    DAGE_TRY(relayout_fz(op));
    DAGE_TRY(build_masks(op));
    op->planes_valid = false;
    DAGE_HIP_CHECK(hipStreamSynchronize(s));
    return DAGE_OK;
}

int dage_operator_get_matrices(dage_operator op, double *fx, double *fy, double *fz) {
    DAGE_REQUIRE(op && fx && fy && fz, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(use_device(op->ctx));
    const size_t T = op->ntp;
    hipStream_t s = op->ctx->stream;
    DAGE_HIP_CHECK(hipMemcpyAsync(fx, op->d_fx, 8 * (size_t)op->dout[0] * op->din[0] * T,
                                  hipMemcpyDeviceToHost, s));
    DAGE_HIP_CHECK(hipMemcpyAsync(fy, op->d_fy, 8 * (size_t)op->din[1] * op->dout[1] * T,
                                  hipMemcpyDeviceToHost, s));
    DAGE_HIP_CHECK(hipMemcpyAsync(fz, op->d_fz, 8 * (size_t)op->din[2] * op->dout[2] * T,
                                  hipMemcpyDeviceToHost, s));
    DAGE_HIP_CHECK(hipStreamSynchronize(s));
    return DAGE_OK;
}

int dage_operator_apply_device(dage_operator op, const double *in, double *out, int precision) {
    DAGE_REQUIRE(op && in && out, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE(precision == DAGE_PREC_FP64 || precision == DAGE_PREC_BF16X6 ||
                     precision == DAGE_PREC_BF16X3,
                 DAGE_ERR_UNSUPPORTED, "precision mode %d is not implemented", precision);
    DAGE_REQUIRE((const void *)in != (const void *)out, DAGE_ERR_ARG,
                 "in-place apply is not supported");
    DAGE_TRY(use_device(op->ctx));
    if (precision == DAGE_PREC_BF16X6) return apply_mixed(op, in, out, 6);
    if (precision == DAGE_PREC_BF16X3) return apply_mixed(op, in, out, 3);
    return apply_f64(op, in, out);
}

namespace {
// element (i, j, k) of an nx x ny x nz box: src[i + sx (j + sy k)] ->
// dst[i + dx (j + dy k)] (column-major sub-cubes with leading dimensions)
__global__ void __launch_bounds__(256) copy_strided_kernel(const double *__restrict__ src, long sx,
                                                           long sy, double *__restrict__ dst,
                                                           long dx, long dy, int nx, int ny,
                                                           long total) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const long i = e % nx, j = (e / nx) % ny, k = e / ((long)nx * ny);
        dst[i + dx * (j + dy * k)] = src[i + sx * (j + sy * k)];
    }
}

int copy_strided(dage_ctx ctx, const double *src, long sx, long sy, double *dst, long dx, long dy,
                 const int *dims) {
    const long total = (long)dims[0] * dims[1] * dims[2];
    if (total == 0) return DAGE_OK;
    const int blocks = (int)std::min<long>((total + 255) / 256, 16384);
    hipLaunchKernelGGL(copy_strided_kernel, dim3(blocks), dim3(256), 0, ctx->stream, src, sx, sy,
                       dst, dx, dy, dims[0], dims[1], total);
    DAGE_HIP_CHECK(hipGetLastError());
    return DAGE_OK;
}
}  // namespace

int dage_operator_apply_device_strided(dage_operator op, const double *in, int ldx_in, int ldy_in,
                                       double *out, int ldx_out, int ldy_out, int precision) {
    DAGE_REQUIRE(op && in && out, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE(ldx_in >= op->din[0] && ldy_in >= op->din[1], DAGE_ERR_ARG,
                 "input leading dimensions (%d, %d) below the input grid's (%d, %d)", ldx_in,
                 ldy_in, op->din[0], op->din[1]);
    DAGE_REQUIRE(ldx_out >= op->dout[0] && ldy_out >= op->dout[1], DAGE_ERR_ARG,
                 "output leading dimensions (%d, %d) below the output grid's (%d, %d)", ldx_out,
                 ldy_out, op->dout[0], op->dout[1]);
    const bool pin = ldx_in != op->din[0] || ldy_in != op->din[1];
    const bool pout = ldx_out != op->dout[0] || ldy_out != op->dout[1];
    if (!pin && !pout) return dage_operator_apply_device(op, in, out, precision);
    DAGE_TRY(use_device(op->ctx));
    dage_ctx ctx = op->ctx;
    const size_t nin = (size_t)op->din[0] * op->din[1] * op->din[2];
    const size_t nout = (size_t)op->dout[0] * op->dout[1] * op->dout[2];
    void *ap = nullptr;
    DAGE_TRY(ctx_aux(ctx, 8 * ((pin ? nin : 0) + (pout ? nout : 0)), &ap));
    double *d_in = pin ? (double *)ap : const_cast<double *>(in);
    double *d_out = pout ? (double *)ap + (pin ? nin : 0) : out;
    if (pin) DAGE_TRY(copy_strided(ctx, in, ldx_in, ldy_in, d_in, op->din[0], op->din[1], op->din));
    DAGE_TRY(dage_operator_apply_device(op, d_in, d_out, precision));
    if (pout)
        DAGE_TRY(copy_strided(ctx, d_out, op->dout[0], op->dout[1], out, ldx_out, ldy_out,
                              op->dout));
    return DAGE_OK;
}

int dage_bubbles_inject_device_strided(dage_bubbles bubbles, dage_grid3d grid, int lmin,
                                       double multiplier, double *cube, int ldx, int ldy) {
    DAGE_REQUIRE(bubbles && grid && cube, DAGE_ERR_ARG, "null argument");
    const int dims[3] = {grid->shape(0), grid->shape(1), grid->shape(2)};
    DAGE_REQUIRE(ldx >= dims[0] && ldy >= dims[1], DAGE_ERR_ARG,
                 "leading dimensions (%d, %d) below the grid's (%d, %d)", ldx, ldy, dims[0],
                 dims[1]);
    if (ldx == dims[0] && ldy == dims[1])
        return dage_bubbles_inject_device(bubbles, grid, lmin, multiplier, cube);
    DAGE_TRY(use_device(grid->ctx));
    dage_ctx ctx = grid->ctx;
    void *ap = nullptr;
    DAGE_TRY(ctx_aux(ctx, 8 * (size_t)dims[0] * dims[1] * dims[2], &ap));
    double *box = (double *)ap;
    DAGE_TRY(copy_strided(ctx, cube, ldx, ldy, box, dims[0], dims[1], dims));
    DAGE_TRY(dage_bubbles_inject_device(bubbles, grid, lmin, multiplier, box));
    return copy_strided(ctx, box, dims[0], dims[1], cube, ldx, ldy, dims);
}

int dage_operator_apply_device_noalloc(dage_operator op, const double *in, double *out) {
    DAGE_REQUIRE(op && in && out, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE((const void *)in != (const void *)out, DAGE_ERR_ARG,
                 "in-place apply is not supported");
    DAGE_TRY(use_device(op->ctx));
    DAGE_REQUIRE(op->ctx->noalloc_scratch != nullptr, DAGE_ERR_STATE,
                 "no fixed scratch on device %d", op->ctx->device);
    return apply_f64_fixed(op, in, out, op->ctx->noalloc_scratch, kNoAllocScratchBytes);
}

int dage_operator_noalloc_limit(size_t *bytes) {
    DAGE_REQUIRE(bytes != nullptr, DAGE_ERR_ARG, "null argument");
    *bytes = kNoAllocScratchBytes;
    return DAGE_OK;
}

int dage_operator_apply(dage_operator op, const double *in, double *out, int precision) {
    DAGE_REQUIRE(op && in && out, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(use_device(op->ctx));
    const size_t nin = (size_t)op->din[0] * op->din[1] * op->din[2];
    const size_t nout = (size_t)op->dout[0] * op->dout[1] * op->dout[2];
    if (!op->d_in) DAGE_TRY(alloc_dev(&op->d_in, nin));
    if (!op->d_out) DAGE_TRY(alloc_dev(&op->d_out, nout));
    hipStream_t s = op->ctx->stream;
    DAGE_HIP_CHECK(hipMemcpyAsync(op->d_in, in, 8 * nin, hipMemcpyHostToDevice, s));
    DAGE_TRY(dage_operator_apply_device(op, op->d_in, op->d_out, precision));
    DAGE_HIP_CHECK(hipMemcpyAsync(out, op->d_out, 8 * nout, hipMemcpyDeviceToHost, s));
    DAGE_HIP_CHECK(hipStreamSynchronize(s));
    return DAGE_OK;
}

int dage_operator_workspace_bytes(dage_operator op, size_t *bytes) {
    DAGE_REQUIRE(op && bytes, DAGE_ERR_ARG, "null argument");
    return apply_f64_workspace(op, bytes);
}

int dage_operator_get_tiling(dage_operator op, int nb[3], int pitch[3], int skip[3]) {
    DAGE_REQUIRE(op && nb && pitch && skip, DAGE_ERR_ARG, "null argument");
    for (int p = 0; p < 3; ++p) {
        pitch[p] = op->pitch[p];
        nb[p] = op->pitch[p] > 0 ? op->nbp[p] : op->nb_pick[p];
        skip[p] = op->mask_used[p] ? 1 : 0;
    }
    return DAGE_OK;
}

int dage_operator_get_band(dage_operator op, int *t0) {
    DAGE_REQUIRE(op && t0, DAGE_ERR_ARG, "null argument");
    *t0 = op->d_band ? op->band_t0 : op->ntp;
    return DAGE_OK;
}

int dage_operator_get_lowrank(dage_operator op, int *t1, int *kx_sum) {
    DAGE_REQUIRE(op && t1 && kx_sum, DAGE_ERR_ARG, "null argument");
    *t1 = op->d_lr ? op->lr.t1 : 0;
    *kx_sum = op->d_lr ? op->lr.Kx : 0;
    return DAGE_OK;
}

int dage_operator_destroy(dage_operator op) {
    if (!op) return DAGE_OK;
    dage_ctx ctx = op->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_operator_buffers(op);
    delete op;
    ctx_release(ctx);
    return DAGE_OK;
}

// ------------------------------------------------------------------ bubbles
int dage_radial_grid_steps(double z, int ncell, int nlip, double r_max, double *steps) {
    DAGE_REQUIRE(steps && z > 0.0 && ncell >= 1 && nlip >= 2 && r_max > 0.0, DAGE_ERR_ARG,
                 "invalid radial grid parameters");
    double x0 = 0.0;
    const double cst = 8.0 * std::pow(z, -1.5);
    const double dx = r_max / ncell;
    for (int i = 1; i <= ncell; ++i) {
        const double x1 = cst / ((r_max + cst) / (i * dx) - 1.0);
        steps[i - 1] = (x1 - x0) / (nlip - 1);
        x0 = x1;
    }
    return DAGE_OK;
}

int dage_bubbles_create(dage_ctx ctx, int nbub, int lmax, int k, int nlip, dage_bubbles *out) {
    DAGE_REQUIRE(out != nullptr, DAGE_ERR_ARG, "null output pointer");
    *out = nullptr;
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(nbub >= 0, DAGE_ERR_ARG, "negative bubble count");
    DAGE_REQUIRE(lmax >= 0 && lmax <= kMaxBubbleLmax, DAGE_ERR_UNSUPPORTED,
                 "bubbles lmax %d outside 0..%d", lmax, kMaxBubbleLmax);
    DAGE_REQUIRE(nlip >= 2 && nlip <= 7, DAGE_ERR_UNSUPPORTED, "bubble nlip %d outside 2..7",
                 nlip);
    auto *b = new (std::nothrow) dage_bubbles_s();
    DAGE_REQUIRE(b != nullptr, DAGE_ERR_ALLOC, "out of host memory");
    b->ctx = ctx;
    b->nbub = nbub;
    b->lmax = lmax;
    b->k = k;
    b->nlip = nlip;
    b->host.resize(nbub);
    if (nbub > 0 &&
        hipMalloc((void **)&b->d_desc, sizeof(BubbleDev) * nbub) != hipSuccess) {
        delete b;
        set_error("device allocation of bubble descriptors failed");
        return DAGE_ERR_ALLOC;
    }
    ctx_retain(ctx);
    *out = b;
    return DAGE_OK;
}

int dage_bubbles_set(dage_bubbles b, int ibub, const double center[3], double charge, int ncell,
                     const double *steps, const double *bf) {
    DAGE_REQUIRE(b && center && steps && bf, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE(ibub >= 0 && ibub < b->nbub, DAGE_ERR_ARG, "bubble index %d outside 0..%d",
                 ibub, b->nbub - 1);
    DAGE_REQUIRE(ncell >= 1, DAGE_ERR_ARG, "radial grid needs at least one cell");
    DAGE_TRY(use_device(b->ctx));
    Axis ax;
    DAGE_REQUIRE(make_axis(0.0, ncell, b->nlip, steps, 1, &ax) == 0, DAGE_ERR_ARG,
                 "invalid radial grid");
    for (int c = 0; c < ncell; ++c)
        DAGE_REQUIRE(steps[c] > 0.0 && std::isfinite(steps[c]), DAGE_ERR_ARG,
                     "radial step %d is not positive", c);
    DAGE_TRY(bubble_alloc_host(b, ibub, center, charge, ax));
    const int npts = ax.n();
    const size_t nv = (size_t)npts * b->nlm();
    // r^k folded into the radial values on the grid (Bubbles_rpow,
    // bubbles.F90:3363-3378: r(1)^k = 0 for k < 0); the raw values are kept
    // beside them when k != 0 (the radial transforms and products use them)
    std::vector<double> fk(nv);
    for (int lm = 0; lm < b->nlm(); ++lm)
        for (int i = 0; i < npts; ++i) {
            const double v = bf[i + (size_t)npts * lm];
            double rp = 1.0;
            if (b->k > 0) rp = std::pow(ax.points[i], b->k);
            else if (b->k < 0) rp = (i == 0) ? 0.0 : std::pow(ax.points[i], b->k);
            fk[i + (size_t)npts * lm] = (b->k == 0) ? v : v * rp;
        }
    BubbleHost &h = b->host[ibub];
    DAGE_HIP_CHECK(hipMemcpy(h.fk(), fk.data(), sizeof(double) * nv, hipMemcpyHostToDevice));
    if (b->k != 0)
        DAGE_HIP_CHECK(hipMemcpy(b->f(ibub), bf, sizeof(double) * nv, hipMemcpyHostToDevice));
    h.set = true;
    return DAGE_OK;
}

int dage_bubbles_inject_device(dage_bubbles b, dage_grid3d grid, int lmin, double multiplier,
                               double *cube) {
    DAGE_REQUIRE(b && grid && cube, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE(grid->ctx == b->ctx, DAGE_ERR_ARG, "grid belongs to another context");
    DAGE_REQUIRE(lmin >= 0, DAGE_ERR_ARG, "negative lmin");
    DAGE_TRY(use_device(b->ctx));
    for (int i = 0; i < b->nbub; ++i)
        DAGE_REQUIRE(b->host[i].set, DAGE_ERR_STATE, "bubble %d has not been set", i);
    if (b->nbub == 0 || lmin > b->lmax) return DAGE_OK;
    return bubbles_inject(b, grid, lmin, multiplier, cube);
}

int dage_bubbles_inject(dage_bubbles b, dage_grid3d grid, int lmin, double multiplier,
                        double *cube) {
    DAGE_REQUIRE(b && grid && cube, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE(grid->ctx == b->ctx, DAGE_ERR_ARG, "grid belongs to another context");
    DAGE_TRY(use_device(b->ctx));
    const size_t n = (size_t)grid->shape(0) * grid->shape(1) * grid->shape(2);
    double *d = nullptr;
    DAGE_TRY(ctx_scratch(b->ctx, 8 * n, (void **)&d));
    hipStream_t s = b->ctx->stream;
    DAGE_HIP_CHECK(hipMemcpyAsync(d, cube, 8 * n, hipMemcpyHostToDevice, s));
    DAGE_TRY(dage_bubbles_inject_device(b, grid, lmin, multiplier, d));
    DAGE_HIP_CHECK(hipMemcpyAsync(cube, d, 8 * n, hipMemcpyDeviceToHost, s));
    DAGE_HIP_CHECK(hipStreamSynchronize(s));
    return DAGE_OK;
}

int dage_bubbles_destroy(dage_bubbles b) {
    if (!b) return DAGE_OK;
    dage_ctx ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &h : b->host) {
        if (h.d_buf) (void)hipFree(h.d_buf);
        if (h.d_tab) (void)hipFree(h.d_tab);
    }
    if (b->d_desc) (void)hipFree(b->d_desc);
    delete b;
    ctx_release(ctx);
    return DAGE_OK;
}

// ---------------------------------------------------------------- cube sums
int dage_cube_integrate_device(dage_grid3d g, const double *f, const double *h, double *result) {
    DAGE_REQUIRE(g && f && result, DAGE_ERR_ARG, "null argument");
    DAGE_TRY(use_device(g->ctx));
    return cube_integrate(g, f, h, result);
}

}  // extern "C"
