// Batched Helmholtz applies: many right-hand sides (orbitals), each with its
// own energy E_i (kernel exp(-k_i r)/r, k_i^2 = -2 E_i) and its own cube grid,
// applied together (the per-orbital loop of HelmholtzSCFCycle_update_orbitals_worker,
// scf_cycle.F90:2927-3054, which calls Helmholtz3D_set_energy and one apply
// per orbital, helmholtz3d.F90:411-481).
//
// Right-hand sides on identical grids form a group that shares one operator's
// matrices; per group every mode pass is ONE launch over (t, rhs) with the
// per-(rhs, t) weights w_{i,t} = 2/sqrt(pi) tw_t exp(2 E_i / (4 t^2)).  The
// whole batch (all groups, all passes, the reductions) is captured once into
// a hipGraph and replayed with a single hipGraphLaunch; energies live in
// device memory, so changing them needs no re-capture.  The batch owns its
// input and output cubes (contiguous per group) so the graph's pointers are
// fixed.  dage_batch_tune times every GEMM column-block count (and the
// operators' own tiling) per pass and group and re-captures the graph with
// the fastest.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "dage_internal.hpp"

struct dage_batch_s {
    struct Group {
        dage_operator op = nullptr;
        std::vector<int> members;  // rhs indices
        double *d_in = nullptr, *d_out = nullptr, *d_w = nullptr;
        long vol = 0;              // points per cube
        int nb[3] = {0, 0, 0};     // tuned column blocks (0: heuristic)
        int slice = 0;             // right-hand sides per apply call (workspace cap)
    };
    dage_ctx ctx = nullptr;
    int nrhs = 0;
    int precision = DAGE_PREC_FP64;
    int flags = 0;
    std::vector<Group> groups;
    std::vector<int> group_of, slot_of;  // per rhs
    std::vector<double> energies, tp, tw;
    dage::Workspace ws;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace dage {
namespace {

int use_device(dage_ctx ctx) {
    DAGE_REQUIRE(ctx != nullptr, DAGE_ERR_ARG, "null context handle");
    DAGE_HIP_CHECK(hipSetDevice(ctx->device));
    return DAGE_OK;
}

bool same_grid(dage_grid3d a, dage_grid3d b) {
    if (a == b) return true;
    for (int i = 0; i < 3; ++i) {
        const Axis &x = a->axis[i], &y = b->axis[i];
        if (x.ncell != y.ncell || x.nlip != y.nlip || x.grid_type != y.grid_type ||
            x.points != y.points)
            return false;
    }
    return true;
}

int upload_weights(dage_batch b) {
    const int T = (int)b->tp.size();
    const double c = two_over_sqrtpi();
    for (auto &g : b->groups) {
        std::vector<double> w(g.members.size() * (size_t)T);
        for (size_t j = 0; j < g.members.size(); ++j) {
            const double e = b->energies[g.members[j]];
            for (int t = 0; t < T; ++t) {
                double wi = b->tw[t];
                if (e < -0.005) wi = wi * std::exp(2.0 * e / (4.0 * (b->tp[t] * b->tp[t])));
                w[j * T + t] = c * wi;
            }
        }
        DAGE_HIP_CHECK(hipMemcpy(g.d_w, w.data(), 8 * w.size(), hipMemcpyHostToDevice));
    }
    return DAGE_OK;
}

int group_workspace(dage_batch b, dage_batch_s::Group &g, int B, size_t *n) {
    if (b->precision == DAGE_PREC_FP64) return apply_f64_batch_workspace(operator_mats(g.op), B, n);
    return apply_mixed_batch_workspace(g.op, B, n);
}

// Right-hand sides per apply call: all of the group's, halved until the
// workspace fits the cap (the t-chunking has a floor: one t-point in FP64,
// up to 8 in split-BF16, so very large groups run in slices).
int pick_slice(dage_batch b, dage_batch_s::Group &g) {
    size_t cap = 64ull << 30;
    if (const char *e = std::getenv("DAGE_BATCH_WORKSPACE_MB")) cap = (size_t)atol(e) << 20;
    int s = (int)g.members.size();
    while (s > 1) {
        size_t n = 0;
        DAGE_TRY(group_workspace(b, g, s, &n));
        if (n <= cap) break;
        s = (s + 1) / 2;
    }
    g.slice = s;
    return DAGE_OK;
}

// One group's applies, slice by slice; nb: column-block counts (0: heuristic).
int run_group(dage_batch b, dage_batch_s::Group &g, const int nb[3], bool no_alloc) {
    const int B = (int)g.members.size(), T = (int)b->tp.size();
    const int S = g.slice > 0 ? g.slice : B;
    for (int s0 = 0; s0 < B; s0 += S) {
        BatchSpec bs;
        bs.B = std::min(S, B - s0);
        bs.w = g.d_w + (size_t)s0 * T;
        const bool coda = (g.op->coda != 0.0) && !(b->flags & DAGE_OP_NO_CODA);
        bs.coda = coda ? g.op->coda : 0.0;
        for (int i = 0; i < 3; ++i) bs.nb[i] = nb[i];
        bs.no_alloc = no_alloc;
        const double *in = g.d_in + (size_t)s0 * g.vol;
        double *out = g.d_out + (size_t)s0 * g.vol;
        if (b->precision == DAGE_PREC_FP64)
            DAGE_TRY(apply_f64_batch(b->ctx, operator_mats(g.op), in, out, bs, &b->ws));
        else
            DAGE_TRY(apply_mixed_batch(g.op, in, out, b->precision == DAGE_PREC_BF16X6 ? 6 : 3,
                                       bs, &b->ws));
    }
    return DAGE_OK;
}

// All groups, one after the other, on the context stream.
int run_batch(dage_batch b, bool no_alloc) {
    for (auto &g : b->groups) DAGE_TRY(run_group(b, g, g.nb, no_alloc));
    return DAGE_OK;
}

int reserve_workspace(dage_batch b) {
    size_t need = 0;
    for (auto &g : b->groups) {
        size_t n = 0;
        DAGE_TRY(pick_slice(b, g));
        DAGE_TRY(group_workspace(b, g, g.slice, &n));
        need = std::max(need, n);
    }
    if (need > b->ws.bytes) {
        if (b->ws.p) DAGE_HIP_CHECK(hipFree(b->ws.p));
        b->ws.p = nullptr;
        b->ws.bytes = 0;
        DAGE_HIP_CHECK(hipMalloc((void **)&b->ws.p, need));
        b->ws.bytes = need;
    }
    return DAGE_OK;
}

void drop_graph(dage_batch b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Capture the whole batch into a graph.  Profiling events are not recorded
// inside the capture; nothing in the captured work allocates or memsets
// (zero fills are kernels).
int capture(dage_batch b) {
    drop_graph(b);
    DAGE_TRY(reserve_workspace(b));
    for (auto &g : b->groups)
        if (b->precision != DAGE_PREC_FP64) DAGE_TRY(prepare_mixed(g.op));
    DAGE_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    const bool prof = b->ctx->profiling;
    b->ctx->profiling = false;
    DAGE_HIP_CHECK(hipStreamBeginCapture(b->ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = run_batch(b, true);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(b->ctx->stream, &graph);
    b->ctx->profiling = prof;
    if (rc != DAGE_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    DAGE_HIP_CHECK(ec);
    b->graph = graph;
    DAGE_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return DAGE_OK;
}

void free_batch(dage_batch b) {
    drop_graph(b);
    for (auto &g : b->groups) {
        if (g.op) dage_operator_destroy(g.op);
        for (double *p : {g.d_in, g.d_out, g.d_w})
            if (p) (void)hipFree(p);
    }
    if (b->ws.p) (void)hipFree(b->ws.p);
    if (b->ctx) ctx_release(b->ctx);
    delete b;
}

const char *kPassName[2][3] = {{"gemm_f64_xmode", "gemm_f64_ymode", "gemm_f64_zmode"},
                               {"gemm_mixed_xmode", "gemm_mixed_ymode", "gemm_mixed_zmode"}};

}  // namespace
}  // namespace dage

using namespace dage;

extern "C" {

int dage_helmholtz_batch_create(dage_ctx ctx, int nrhs, const dage_grid3d *grids,
                                const double *energies, const double *tp, const double *tw,
                                int ntp, double tlog, int flags, int precision, dage_batch *out) {
    DAGE_REQUIRE(out != nullptr, DAGE_ERR_ARG, "null output pointer");
    *out = nullptr;
    DAGE_TRY(use_device(ctx));
    DAGE_REQUIRE(nrhs > 0 && grids && energies && tp && tw && ntp > 0, DAGE_ERR_ARG,
                 "null or empty argument");
    DAGE_REQUIRE(precision == DAGE_PREC_FP64 || precision == DAGE_PREC_BF16X6 ||
                     precision == DAGE_PREC_BF16X3,
                 DAGE_ERR_ARG, "unknown precision mode %d", precision);
    for (int i = 0; i < nrhs; ++i) {
        DAGE_REQUIRE(grids[i] != nullptr && grids[i]->ctx == ctx, DAGE_ERR_ARG,
                     "grid %d is null or belongs to another context", i);
        DAGE_REQUIRE(std::isfinite(energies[i]), DAGE_ERR_ARG, "energy %d is not finite", i);
    }
    auto *b = new (std::nothrow) dage_batch_s();
    DAGE_REQUIRE(b != nullptr, DAGE_ERR_ALLOC, "out of host memory");
    b->ctx = ctx;
    ctx_retain(ctx);
    b->nrhs = nrhs;
    b->precision = precision;
    b->flags = flags;
    b->energies.assign(energies, energies + nrhs);
    b->tp.assign(tp, tp + ntp);
    b->tw.assign(tw, tw + ntp);
    b->group_of.assign(nrhs, -1);
    b->slot_of.assign(nrhs, -1);
    std::vector<dage_grid3d> ggrid;
    for (int i = 0; i < nrhs; ++i) {
        int gi = -1;
        for (size_t g = 0; g < ggrid.size(); ++g)
            if (same_grid(ggrid[g], grids[i])) gi = (int)g;
        if (gi < 0) {
            gi = (int)ggrid.size();
            ggrid.push_back(grids[i]);
            b->groups.emplace_back();
        }
        b->group_of[i] = gi;
        b->slot_of[i] = (int)b->groups[gi].members.size();
        b->groups[gi].members.push_back(i);
    }
    int rc = DAGE_OK;
    for (size_t g = 0; g < b->groups.size() && rc == DAGE_OK; ++g) {
        auto &G = b->groups[g];
        rc = dage_coulomb3d_create(ctx, ggrid[g], ggrid[g], tp, tw, ntp, tlog, flags, &G.op);
        if (rc != DAGE_OK) break;
        G.vol = (long)G.op->din[0] * G.op->din[1] * G.op->din[2];
        const size_t bytes = 8 * (size_t)G.vol * G.members.size();
        if (hipMalloc((void **)&G.d_in, bytes) != hipSuccess ||
            hipMalloc((void **)&G.d_out, bytes) != hipSuccess ||
            hipMalloc((void **)&G.d_w, 8 * G.members.size() * (size_t)ntp) != hipSuccess) {
            set_error("device allocation for group %zu failed", g);
            rc = DAGE_ERR_HIP;
        }
    }
    if (rc == DAGE_OK) rc = upload_weights(b);
    if (rc == DAGE_OK) rc = capture(b);
    if (rc != DAGE_OK) {
        free_batch(b);
        return rc;
    }
    *out = b;
    return DAGE_OK;
}

int dage_batch_input(dage_batch b, int i, double **dev) {
    DAGE_REQUIRE(b && dev && i >= 0 && i < b->nrhs, DAGE_ERR_ARG, "bad argument");
    const auto &G = b->groups[b->group_of[i]];
    *dev = G.d_in + (size_t)b->slot_of[i] * G.vol;
    return DAGE_OK;
}

int dage_batch_output(dage_batch b, int i, double **dev) {
    DAGE_REQUIRE(b && dev && i >= 0 && i < b->nrhs, DAGE_ERR_ARG, "bad argument");
    const auto &G = b->groups[b->group_of[i]];
    *dev = G.d_out + (size_t)b->slot_of[i] * G.vol;
    return DAGE_OK;
}

int dage_batch_upload(dage_batch b, int i, const double *host) {
    double *dev = nullptr;
    DAGE_TRY(dage_batch_input(b, i, &dev));
    DAGE_REQUIRE(host != nullptr, DAGE_ERR_ARG, "null host pointer");
    DAGE_TRY(use_device(b->ctx));
    const auto &G = b->groups[b->group_of[i]];
    DAGE_HIP_CHECK(hipMemcpyAsync(dev, host, 8 * (size_t)G.vol, hipMemcpyHostToDevice,
                                  b->ctx->stream));
    DAGE_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    return DAGE_OK;
}

int dage_batch_download(dage_batch b, int i, double *host) {
    double *dev = nullptr;
    DAGE_TRY(dage_batch_output(b, i, &dev));
    DAGE_REQUIRE(host != nullptr, DAGE_ERR_ARG, "null host pointer");
    DAGE_TRY(use_device(b->ctx));
    const auto &G = b->groups[b->group_of[i]];
    DAGE_HIP_CHECK(hipMemcpyAsync(host, dev, 8 * (size_t)G.vol, hipMemcpyDeviceToHost,
                                  b->ctx->stream));
    DAGE_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    return DAGE_OK;
}

int dage_batch_set_energies(dage_batch b, const double *energies) {
    DAGE_REQUIRE(b && energies, DAGE_ERR_ARG, "null argument");
    for (int i = 0; i < b->nrhs; ++i)
        DAGE_REQUIRE(std::isfinite(energies[i]), DAGE_ERR_ARG, "energy %d is not finite", i);
    DAGE_TRY(use_device(b->ctx));
    DAGE_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    b->energies.assign(energies, energies + b->nrhs);
    return upload_weights(b);
}

int dage_batch_apply(dage_batch b) {
    DAGE_REQUIRE(b && b->exec, DAGE_ERR_STATE, "batch has no captured graph");
    DAGE_TRY(use_device(b->ctx));
    DAGE_HIP_CHECK(hipGraphLaunch(b->exec, b->ctx->stream));
    return DAGE_OK;
}

int dage_batch_apply_eager(dage_batch b) {
    DAGE_REQUIRE(b != nullptr, DAGE_ERR_ARG, "null batch");
    DAGE_TRY(use_device(b->ctx));
    return run_batch(b, true);
}

int dage_batch_tune(dage_batch b, int reps, int *tiles) {
    DAGE_REQUIRE(b != nullptr && reps > 0, DAGE_ERR_ARG, "bad argument");
    DAGE_TRY(use_device(b->ctx));
    dage_ctx ctx = b->ctx;
    const int kind = b->precision == DAGE_PREC_FP64 ? 0 : 1;
    const bool prof = ctx->profiling;
    DAGE_TRY(prof_resolve(ctx));
    const auto saved = ctx->profile;  // the caller's profile survives tuning
    for (auto &g : b->groups) {
        double best[3] = {1e30, 1e30, 1e30};
        int pick[3] = {0, 0, 0};
        for (int nb = 0; nb <= 8; ++nb) {  // 0: the operator's own tiling
            for (int i = 0; i < 3; ++i) g.nb[i] = nb;
            ctx->profiling = true;
            ctx->profile.clear();
            int rc = DAGE_OK;
            for (int r = 0; r < reps + 1 && rc == DAGE_OK; ++r) {
                if (r == 1) {  // first run warms up
                    DAGE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
                    (void)prof_resolve(ctx);
                    ctx->profile.clear();
                }
                rc = run_group(b, g, g.nb, true);
            }
            DAGE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            (void)prof_resolve(ctx);
            ctx->profiling = prof;
            DAGE_TRY(rc);
            for (int i = 0; i < 3; ++i) {
                auto it = ctx->profile.find(kPassName[kind][i]);
                const double ms = (it == ctx->profile.end()) ? 1e30 : it->second.total_ms;
                if (ms < best[i]) {
                    best[i] = ms;
                    pick[i] = nb;
                }
            }
        }
        for (int i = 0; i < 3; ++i) g.nb[i] = pick[i];
    }
    ctx->profile = saved;
    if (tiles)
        for (size_t g = 0; g < b->groups.size(); ++g)
            for (int i = 0; i < 3; ++i) tiles[3 * g + i] = b->groups[g].nb[i];
    return capture(b);
}

int dage_batch_group_count(dage_batch b, int *ngroups) {
    DAGE_REQUIRE(b && ngroups, DAGE_ERR_ARG, "null argument");
    *ngroups = (int)b->groups.size();
    return DAGE_OK;
}

int dage_batch_destroy(dage_batch b) {
    if (!b) return DAGE_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    free_batch(b);
    return DAGE_OK;
}

}  // extern "C"
