// Low-rank t-points of the tensor-quadrature apply.
//
// At small t the Gaussian exp(-t^2 (x - x')^2) varies little across the
// box, and the 1D matrices F_t are numerically of low rank: at 385 points
// over 38.4 bohr the first ten t-points of a T = 40 quadrature have ranks
// 6-137 at the truncation used here.  For those t the operator is applied
// through factors F_t ~= Q_t R_t (Q_t orthonormal, n_out x k; R_t = Q_t^T F_t,
// k x n_in) on each axis:
//
//   1. A1 = [R_x,t]_t . in          reduce x, all low-rank t stacked (one GEMM)
//   2. A2_t = A1_t . R_y,t^T        reduce y (batched over z-planes)
//   3. A3_t = A2_t . R_z,t^T        reduce z: the k_x x k_y x k_z core
//   4. E1_t = w_t A3_t . Q_z,t^T    expand z, times the weight of (b, t)
//   5. E2_t = E1_t . Q_y,t^T        expand y (batched over z-planes)
//   6. out  = [Q_x,t]_t . E2        expand x and sum over t (one GEMM)
//
// Steps 1 and 6 cost 2 K n^3 each (K = sum of the x ranks) where the three
// dense passes cost 6 n^4 per t-point.  Every step is a column-major FP64
// MFMA GEMM (gemm_f64_kernel, PASS 3).
//
// Truncation: column-pivoted Gram-Schmidt on F_t (with re-orthogonalization)
// stops when the largest residual column norm is at most kLowRankTol times
// the largest column norm of F_t.  The spectral truncation error is then
// 2-7e-15 of ||F_t|| (measured on the 385-point axis), under the rounding
// bound of the dense n-term FP64 sums it replaces (n * 2^-53 = 4.3e-14).  A
// t-point is low-rank only while every axis's rank stays at most half the
// axis length (beyond that the dense passes are cheaper); the low-rank set is
// t = 0..lr_t1-1.
#include "dage_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

namespace dage {
namespace {

constexpr double kLowRankTol = 1e-14;
constexpr int kLowRankMaxN = 800;  // longest axis factored at set-up

// Factor M (no x ni, column-major, M(o, i) = m[o + no*i]) as Q R with Q
// orthonormal (no x k) and R = Q^T M (k x ni).  Returns k, or -1 when the
// rank would exceed kmax.
int factor(const std::vector<double> &m, int no, int ni, int kmax, std::vector<double> &Q,
           std::vector<double> &R) {
    std::vector<double> A(m);
    std::vector<double> nrm(ni);
    double big = 0.0;
    for (int i = 0; i < ni; ++i) {
        double s = 0.0;
        for (int o = 0; o < no; ++o) s += A[o + (size_t)no * i] * A[o + (size_t)no * i];
        nrm[i] = s;
        big = std::max(big, s);
    }
    const double stop = kLowRankTol * std::sqrt(big);
    Q.clear();
    int k = 0;
    std::vector<double> q(no);
    while (true) {
        int p = 0;
        for (int i = 1; i < ni; ++i)
            if (nrm[i] > nrm[p]) p = i;
        if (std::sqrt(nrm[p]) <= stop || k == std::min(no, ni)) break;
        if (k >= kmax) return -1;
        for (int o = 0; o < no; ++o) q[o] = A[o + (size_t)no * p];
        // two rounds of classical Gram-Schmidt against the columns so far
        for (int pass = 0; pass < 2; ++pass)
            for (int l = 0; l < k; ++l) {
                const double *ql = &Q[(size_t)no * l];
                double d = 0.0;
                for (int o = 0; o < no; ++o) d += ql[o] * q[o];
                for (int o = 0; o < no; ++o) q[o] -= d * ql[o];
            }
        double s = 0.0;
        for (int o = 0; o < no; ++o) s += q[o] * q[o];
        s = std::sqrt(s);
        if (!(s > 0.0)) break;
        for (int o = 0; o < no; ++o) q[o] /= s;
        Q.insert(Q.end(), q.begin(), q.end());
        ++k;
        // remove the new direction from every column, refresh the norms
        for (int i = 0; i < ni; ++i) {
            double *a = &A[(size_t)no * i];
            double d = 0.0;
            for (int o = 0; o < no; ++o) d += q[o] * a[o];
            double t = 0.0;
            for (int o = 0; o < no; ++o) {
                a[o] -= d * q[o];
                t += a[o] * a[o];
            }
            nrm[i] = t;
        }
    }
    R.assign((size_t)k * ni, 0.0);
    for (int i = 0; i < ni; ++i)
        for (int a = 0; a < k; ++a) {
            const double *qa = &Q[(size_t)no * a];
            double d = 0.0;
            for (int o = 0; o < no; ++o) d += qa[o] * m[o + (size_t)no * i];
            R[a + (size_t)k * i] = d;
        }
    return k;
}

// DAGE_APPLY_LOWRANK: 0 none, 2 at every size, otherwise (default 1) where
// the apply is large enough to pay for its ~4 small GEMM launches per t-point
int lowrank_mode() {
    const char *e = std::getenv("DAGE_APPLY_LOWRANK");
    if (!e) return 1;
    return e[0] == '0' ? 0 : (e[0] == '2' ? 2 : 1);
}

ColGemm plain_gemm(int M, int N, int K) {
    ColGemm g;
    g.M = M; g.N = N; g.K = K;
    return g;
}

}  // namespace

int build_lowrank(dage_operator op) {
    LowRank &L = op->lr;
    if (op->d_lr) DAGE_HIP_CHECK(hipFree(op->d_lr));
    op->d_lr = nullptr;
    L = LowRank{};
    const int T = op->ntp;
    const int tmax = std::min(T, op->d_band ? op->band_t0 : T);
    L.mode = lowrank_mode();
    if (L.mode == 0 || tmax == 0) return DAGE_OK;
    const int n_in[3] = {op->din[0], op->din[1], op->din[2]};
    const int n_out[3] = {op->dout[0], op->dout[1], op->dout[2]};
    // the host factorization costs O(n^2 k) per t and axis: fine at a few
    // hundred points (385: ~0.1 s per t), too slow for set-up beyond ~800
    for (int a = 0; a < 3; ++a)
        if (n_in[a] > kLowRankMaxN || n_out[a] > kLowRankMaxN) return DAGE_OK;
    hipStream_t s = op->ctx->stream;
    // host copies of the (flushed) matrices
    std::vector<double> h[3];
    const double *src[3] = {op->d_fx, op->d_fy, op->d_fz};
    for (int a = 0; a < 3; ++a) {
        h[a].resize((size_t)n_in[a] * n_out[a] * T);
        DAGE_HIP_CHECK(hipMemcpyAsync(h[a].data(), src[a], 8 * h[a].size(), hipMemcpyDeviceToHost, s));
    }
    DAGE_HIP_CHECK(hipStreamSynchronize(s));
    // per t and axis: Q (n_out x k), R (k x n_in)
    std::vector<std::vector<double>> Qs[3], Rs[3];
    std::vector<double> m;
    int t1 = 0;
    for (int t = 0; t < tmax; ++t) {
        std::vector<double> Q[3], R[3];
        bool ok = true;
        int k[3];
        for (int a = 0; a < 3 && ok; ++a) {
            const int no = n_out[a], ni = n_in[a];
            m.resize((size_t)no * ni);
            for (int i = 0; i < ni; ++i)
                for (int o = 0; o < no; ++o)
                    m[o + (size_t)no * i] = a == 0 ? h[0][o + (size_t)no * i + (size_t)no * ni * t]
                                                   : h[a][i + (size_t)ni * o + (size_t)no * ni * t];
            k[a] = factor(m, no, ni, std::min(no, ni) / 2, Q[a], R[a]);
            ok = k[a] > 0;
        }
        if (!ok) break;
        for (int a = 0; a < 3; ++a) {
            L.k[a].push_back(k[a]);
            Qs[a].push_back(std::move(Q[a]));
            Rs[a].push_back(std::move(R[a]));
        }
        t1 = t + 1;
    }
    if (t1 == 0) return DAGE_OK;
    // device layouts (column-major):
    //   RxAll (Kx x nxi), rows koff[t]..: R_x,t;   QxAll (nxo x Kx), columns koff[t]..: Q_x,t
    //   per t: RyT (nyi x ky), QyT (ky x nyo), RzT (nzi x kz), QzT (kz x nzo)
    int Kx = 0;
    for (int t = 0; t < t1; ++t) {
        L.koff.push_back(Kx);
        Kx += L.k[0][t];
    }
    L.Kx = Kx;
    std::vector<double> buf;
    const size_t nx_in = n_in[0], nx_out = n_out[0];
    L.off_rx = buf.size();
    buf.resize(buf.size() + (size_t)Kx * nx_in, 0.0);
    for (int t = 0; t < t1; ++t) {
        const int kx = L.k[0][t];
        for (size_t i = 0; i < nx_in; ++i)
            for (int a = 0; a < kx; ++a)
                buf[L.off_rx + (L.koff[t] + a) + (size_t)Kx * i] = Rs[0][t][a + (size_t)kx * i];
    }
    L.off_qx = buf.size();
    buf.resize(buf.size() + nx_out * Kx, 0.0);
    for (int t = 0; t < t1; ++t) {
        const int kx = L.k[0][t];
        for (int a = 0; a < kx; ++a)
            for (size_t o = 0; o < nx_out; ++o)
                buf[L.off_qx + o + nx_out * (L.koff[t] + a)] = Qs[0][t][o + nx_out * a];
    }
    for (int t = 0; t < t1; ++t) {
        for (int a = 1; a < 3; ++a) {
            const int kk = L.k[a][t], ni = n_in[a], no = n_out[a];
            // R^T (ni x k)
            L.off_rt[a - 1].push_back(buf.size());
            buf.resize(buf.size() + (size_t)ni * kk);
            double *rt = &buf[L.off_rt[a - 1].back()];
            for (int c = 0; c < kk; ++c)
                for (int i = 0; i < ni; ++i) rt[i + (size_t)ni * c] = Rs[a][t][c + (size_t)kk * i];
            // Q^T (k x no)
            L.off_qt[a - 1].push_back(buf.size());
            buf.resize(buf.size() + (size_t)kk * no);
            double *qt = &buf[L.off_qt[a - 1].back()];
            for (int o = 0; o < no; ++o)
                for (int c = 0; c < kk; ++c) qt[c + (size_t)kk * o] = Qs[a][t][o + (size_t)no * c];
        }
    }
    DAGE_HIP_CHECK(hipMalloc((void **)&op->d_lr, 8 * buf.size()));
    DAGE_HIP_CHECK(hipMemcpyAsync(op->d_lr, buf.data(), 8 * buf.size(), hipMemcpyHostToDevice, s));
    DAGE_HIP_CHECK(hipStreamSynchronize(s));
    L.t1 = t1;
    L.base = op->d_lr;
    return DAGE_OK;
}

bool lowrank_pays(const ApplyMats &m, int B) {
    if (!m.lr) return false;
    if (m.lr->mode == 2) return true;
    // below ~8 M output points per call the per-t launches outweigh the
    // saved work (97^3, one right-hand side: 1.21 against 0.83 ms)
    return (long)B * m.dout[0] * m.dout[1] * m.dout[2] >= (1L << 23);
}

// Workspace (doubles) of the low-rank apply for B right-hand sides: the first
// region (A1, A2, A3, E1) and the second (E2).
void lowrank_workspace(const ApplyMats &m, int B, size_t *first, size_t *second) {
    const LowRank &L = *m.lr;
    const size_t nyi = m.din[1], nzi = m.din[2], nyo = m.dout[1], nzo = m.dout[2];
    size_t a2 = 0, a3 = 0, e1 = 0;
    for (int t = 0; t < L.t1; ++t) {
        const size_t kx = L.k[0][t], ky = L.k[1][t], kz = L.k[2][t];
        a2 += kx * ky * nzi;
        a3 += kx * ky * kz;
        e1 += kx * ky * nzo;
    }
    *first = ((size_t)L.Kx * nyi * nzi + a2 + a3 + e1) * B;
    *second = (size_t)L.Kx * nyo * nzo * B;
}

int lowrank_apply(dage_ctx ctx, const ApplyMats &m, const double *in, double *C, long cS, int B,
                  const double *W, long wS, double *ws1, double *ws2, const char *name) {
    const LowRank &L = *m.lr;
    const int nxi = m.din[0], nyi = m.din[1], nzi = m.din[2];
    const int nxo = m.dout[0], nyo = m.dout[1], nzo = m.dout[2];
    const int Kx = L.Kx;
    const double *base = L.base;
    double *A1 = ws1;
    double *A2 = A1 + (size_t)Kx * nyi * nzi * B;
    size_t a2 = 0, a3 = 0, e1 = 0;
    for (int t = 0; t < L.t1; ++t) {
        const size_t kx = L.k[0][t], ky = L.k[1][t];
        a2 += kx * ky * nzi * B;
    }
    double *A3 = A2 + a2;
    for (int t = 0; t < L.t1; ++t) a3 += (size_t)L.k[0][t] * L.k[1][t] * L.k[2][t] * B;
    double *E1 = A3 + a3;
    double *E2 = ws2;
    // 1. A1 (Kx x nyi*nzi*B) = RxAll (Kx x nxi) . in (nxi x nyi*nzi*B)
    {
        ColGemm g = plain_gemm(Kx, nyi * nzi * B, nxi);
        g.A = base + L.off_rx; g.sAk = Kx;
        g.B = in; g.sBn = nxi;
        g.C = A1; g.sCn = Kx;
        DAGE_TRY(gemm_colmajor_general(ctx, name, g, 1));
    }
    size_t p2 = 0, p3 = 0, p1 = 0;
    for (int t = 0; t < L.t1; ++t) {
        const int kx = L.k[0][t], ky = L.k[1][t], kz = L.k[2][t];
        double *a2t = A2 + p2, *a3t = A3 + p3, *e1t = E1 + p1;
        // 2. per (zi, b): A2_t (kx x ky) = A1[koff..koff+kx, :, zi, b] (kx x nyi) . RyT (nyi x ky)
        {
            ColGemm g = plain_gemm(kx, ky, nyi);
            g.A = A1 + L.koff[t]; g.sAk = Kx; g.bA = (long)Kx * nyi;
            g.B = base + L.off_rt[0][t]; g.sBn = nyi;
            g.C = a2t; g.sCn = kx; g.bC = (long)kx * ky;
            DAGE_TRY(gemm_colmajor_general(ctx, name, g, nzi * B));
        }
        // 3. per b: A3_t (kx*ky x kz) = A2_t (kx*ky x nzi) . RzT (nzi x kz)
        {
            ColGemm g = plain_gemm(kx * ky, kz, nzi);
            g.A = a2t; g.sAk = (long)kx * ky; g.bA = (long)kx * ky * nzi;
            g.B = base + L.off_rt[1][t]; g.sBn = nzi;
            g.C = a3t; g.sCn = (long)kx * ky; g.bC = (long)kx * ky * kz;
            DAGE_TRY(gemm_colmajor_general(ctx, name, g, B));
        }
        // 4. per b: E1_t (kx*ky x nzo) = w_{b,t} A3_t (kx*ky x kz) . QzT (kz x nzo)
        {
            ColGemm g = plain_gemm(kx * ky, nzo, kz);
            g.A = a3t; g.sAk = (long)kx * ky;
            g.B = base + L.off_qt[1][t]; g.sBn = kz;
            g.C = e1t; g.sCn = (long)kx * ky;
            g.Z0 = 1;  // the batch index is the right-hand side (z_hi)
            g.bA1 = (long)kx * ky * kz; g.bC1 = (long)kx * ky * nzo;
            g.alpha = W + t; g.aS1 = wS;
            DAGE_TRY(gemm_colmajor_general(ctx, name, g, B));
        }
        // 5. per (zo, b): E2[koff.., :, zo, b] (kx x nyo) = E1_t (kx x ky) . QyT (ky x nyo)
        {
            ColGemm g = plain_gemm(kx, nyo, ky);
            g.A = e1t; g.sAk = kx; g.bA = (long)kx * ky;
            g.B = base + L.off_qt[0][t]; g.sBn = ky;
            g.C = E2 + L.koff[t]; g.sCn = Kx; g.bC = (long)Kx * nyo;
            DAGE_TRY(gemm_colmajor_general(ctx, name, g, nzo * B));
        }
        p2 += (size_t)kx * ky * nzi * B;
        p3 += (size_t)kx * ky * kz * B;
        p1 += (size_t)kx * ky * nzo * B;
    }
    // 6. per b: C_b (nxo x nyo*nzo) = QxAll (nxo x Kx) . E2_b (Kx x nyo*nzo)
    {
        ColGemm g = plain_gemm(nxo, nyo * nzo, Kx);
        g.A = base + L.off_qx; g.sAk = nxo;
        g.B = E2; g.sBn = Kx; g.bB = (long)Kx * nyo * nzo;
        g.C = C; g.sCn = nxo; g.bC = cS;
        DAGE_TRY(gemm_colmajor_general(ctx, name, g, B));
    }
    return DAGE_OK;
}

}  // namespace dage
