// Taylor-series bubbles of a Function3D and the product that uses them
// (Function3D_get_taylor_series_bubbles, src/bubbles/function3d.F90:1338-1404;
// Function3D_get_cube_contaminants, function3d.F90:1570-1622;
// Bubbles_get_contaminants, src/bubbles/bubbles.F90:2844-2903;
// Bubbles_make_taylor, bubbles.F90:1435-1533; the product with them,
// BubblesMultiplier_multiply, bubbles.F90:3567-3740 and
// bubbles_multiplier.cu:233-295).
//
// At every centre B the "contaminants" of a function f = cube + sum_A B_A are
// the Cartesian derivatives d_kappa = d^kappa (cube + sum_{A != B} B_A)(R_B),
// |kappa| <= t (t = 2, the reference's default order), of the cube's LIP
// interpolant and of the other centres' bubbles.  The Taylor bubble at B is
//   T_B(x) = w(r) sum_kappa d_kappa / |kappa|! (x - R_B)^kappa,
//   w(r) = erfc(r - 2/r) / 2 (1 at r = 0),
// expanded in the Racah harmonics C_lm (l <= t) and stored, as any bubble of
// the same k, as f_lm(r) with r^k f_lm = the expansion.  1/|kappa|! (not
// 1/kappa!) is the reference's weighting of the mixed terms; the product's
// value is the same for any weighting, only its split between cube and
// bubbles changes.
//
// Product of two cusp-type functions (both carry Taylor bubbles T1, T2):
//   result bubbles  = (B1 + T1)(B2 + T2) - T1 T2      (centre by centre,
//                     formed as B1 B2 + B1 T2 + T1 B2, the Taylor values in
//                     the context's auxiliary buffer: no allocation)
//   result cube     = (c1 + B1)(c2 + B2) - result bubbles (on the grid)
// so near a nucleus A the cube holds B1_A (f2 - T2_A) + ..., in which the
// Taylor series cancels the low orders of f2 that a singular B1_A (e.g. the
// nuclear -Z/r) would otherwise leave in the cube.
//
// The derivatives of a bubble term g(x) = P_lm(r) C_lm(x/r), P_lm the radial
// LIP interpolant of r^k f_lm (the values the evaluation interpolates,
// Bubbles_get_radial_derivatives, bubbles.F90:2272-2305), are taken exactly:
// g = G(r) S_lm(x) with G = r^-l P_lm and S_lm = r^l C_lm, S_lm carried with
// its gradient and Hessian through the harmonic recursion.  Host work (per
// centre pair and lm, a few hundred flops); the cube's LIP cells around the
// centres are gathered on the GPU.
#include <array>
#include <cmath>
#include <vector>

#include "dage_internal.hpp"

namespace dage {
namespace {

constexpr int kMaxTaylorOrder = 2;

// CartIter order (src/bubbles/cartiter.F90:196-222): x exponent outermost,
// z innermost, |kappa| <= t.
std::vector<std::array<int, 3>> cart_indices(int t) {
    std::vector<std::array<int, 3>> k;
    for (int a = 0; a <= t; ++a)
        for (int b = 0; b <= t - a; ++b)
            for (int c = 0; c <= t - a - b; ++c) k.push_back({a, b, c});
    return k;
}

// value, gradient and Hessian of a function of x
struct Jet {
    double v = 0.0;
    double g[3] = {0.0, 0.0, 0.0};
    double h[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
};

Jet jet_mul_var(const Jet &p, int a, const double x[3]) {  // p * x_a
    Jet r;
    r.v = p.v * x[a];
    for (int i = 0; i < 3; ++i) r.g[i] = p.g[i] * x[a] + (i == a ? p.v : 0.0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.h[i][j] = p.h[i][j] * x[a] + (j == a ? p.g[i] : 0.0) + (i == a ? p.g[j] : 0.0);
    return r;
}

Jet jet_mul_r2(const Jet &p, const double x[3]) {  // p * |x|^2
    const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    Jet r;
    r.v = p.v * r2;
    for (int i = 0; i < 3; ++i) r.g[i] = p.g[i] * r2 + 2.0 * p.v * x[i];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.h[i][j] = p.h[i][j] * r2 + 2.0 * p.g[i] * x[j] + 2.0 * p.g[j] * x[i] +
                        (i == j ? 2.0 * p.v : 0.0);
    return r;
}

Jet jet_axpy(double a, const Jet &p, double b, const Jet &q) {  // a p + b q
    Jet r;
    r.v = a * p.v + b * q.v;
    for (int i = 0; i < 3; ++i) r.g[i] = a * p.g[i] + b * q.g[i];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.h[i][j] = a * p.h[i][j] + b * q.h[i][j];
    return r;
}

// Racah solid harmonics S_lm(x) = r^l C_lm with first and second derivatives:
// the recursion of solid_harmonics_host on jets (harmonics.F90:586-659).
void solid_harmonics_jet(int lmax, const double x[3], std::vector<Jet> &S) {
    S.assign((size_t)(lmax + 1) * (lmax + 1), Jet());
    S[0].v = 1.0;
    if (lmax < 1) return;
    Jet one;
    one.v = 1.0;
    S[1] = jet_mul_var(one, 1, x);  // y
    S[2] = jet_mul_var(one, 2, x);  // z
    S[3] = jet_mul_var(one, 0, x);  // x
    for (int l = 2; l <= lmax; ++l) {
        const int c = l * l + l, p = (l - 1) * (l - 1) + (l - 1), q = (l - 2) * (l - 2) + (l - 2);
        const double f = std::sqrt((2 * l - 1.0) / (2.0 * l));
        S[c - l] = jet_axpy(f, jet_mul_var(S[p + (l - 1)], 1, x), f,
                            jet_mul_var(S[p - (l - 1)], 0, x));
        for (int m = -(l - 1); m <= l - 1; ++m) {
            Jet v = jet_mul_var(S[p + m], 2, x);
            const double a = (2 * l - 1.0) / std::sqrt(1.0 * (l + m) * (l - m));
            if (l - std::abs(m) > 1) {
                const double b = std::sqrt((l + m - 1.0) * (l - m - 1.0) / ((l + m) * (l - m)));
                v = jet_axpy(a, v, -b, jet_mul_r2(S[q + m], x));
            } else {
                v = jet_axpy(a, v, 0.0, v);
            }
            S[c + m] = v;
        }
        S[c + l] = jet_axpy(f, jet_mul_var(S[p + (l - 1)], 0, x), -f,
                            jet_mul_var(S[p - (l - 1)], 1, x));
    }
}

double jet_derivative(const Jet &j, const std::array<int, 3> &kap) {
    const int n = kap[0] + kap[1] + kap[2];
    if (n == 0) return j.v;
    int ax[2] = {-1, -1}, na = 0;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < kap[c]; ++r) ax[na++] = c;
    return n == 1 ? j.g[ax[0]] : j.h[ax[0]][ax[1]];
}

// LIP basis at cell coordinate t: L[d][j] = d^d/dt^d of basis j, d = 0..2
// (lip_coeffs: c[j + n*k], decreasing powers).
void lip_basis(int nlip, const double *coef, double t, double L[3][kMaxNlip]) {
    for (int j = 0; j < nlip; ++j) {
        // Horner: p0 = p(t), p1 = p'(t), p2 = p''(t) / 2
        double p0 = 0.0, p1 = 0.0, p2 = 0.0;
        for (int k = 0; k < nlip; ++k) {
            p2 = p2 * t + p1;
            p1 = p1 * t + p0;
            p0 = p0 * t + coef[j + nlip * k];
        }
        L[0][j] = p0;
        L[1][j] = p1;
        L[2][j] = 2.0 * p2;
    }
}

// Grid1D_get_icell (src/bubbles/grid.F90:1212-1246): the cell c (0-based)
// with celld[c] <= x < celld[c+1], the last cell for x >= its start; -1
// outside [qmin, qmax].
int find_cell(const double *celld, int ncell, double qmin, double qmax, double x) {
    if (x > qmax || x < qmin) return -1;
    if (celld[ncell - 1] <= x) return ncell - 1;
    int lo = 0, hi = ncell - 1;  // celld[lo] <= x < celld[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (celld[mid] <= x)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__global__ void gather_cells_kernel(const double *__restrict__ cube, int nx, int ny,
                                    const int *__restrict__ starts, int nlip,
                                    double *__restrict__ out) {
    const int c = blockIdx.x, n3 = nlip * nlip * nlip;
    const int sx = starts[3 * c], sy = starts[3 * c + 1], sz = starts[3 * c + 2];
    for (int e = threadIdx.x; e < n3; e += blockDim.x) {
        const int i = e % nlip, j = (e / nlip) % nlip, k = e / (nlip * nlip);
        out[(long)c * n3 + e] = cube[(long)(sx + i) + (long)nx * ((sy + j) + (long)ny * (sz + k))];
    }
}

// d^kappa of the cube's LIP interpolant at the centres (nder x nc, column
// per centre); centres outside the grid get zeros (Interpolator_eval,
// src/bubbles/evaluators.F90:908-990).
int cube_contaminants(dage_grid3d grid, const double *cube, const std::vector<double> &centres,
                      int t, std::vector<double> &out) {
    const int nc = (int)centres.size() / 3, nlip = grid->axis[0].nlip;
    const auto kap = cart_indices(t);
    const int nd = (int)kap.size();
    out.assign((size_t)nd * nc, 0.0);
    double coef[kMaxNlip * kMaxNlip];
    DAGE_REQUIRE(lip_coeffs(nlip, grid->axis[0].grid_type, coef) == 0, DAGE_ERR_ARG,
                 "unsupported nlip %d", nlip);
    double first, last;
    lip_first_last(nlip, &first, &last);
    std::vector<int> starts;
    std::vector<int> inside;
    std::vector<std::array<std::array<double, kMaxNlip>, 9>> basis;  // [axis*3 + order][node]
    for (int c = 0; c < nc; ++c) {
        std::array<std::array<double, kMaxNlip>, 9> B{};
        int st[3];
        bool ok = true;
        for (int a = 0; a < 3 && ok; ++a) {
            const Axis &ax = grid->axis[a];
            const double x = centres[3 * c + a];
            const int cell = find_cell(ax.celld.data(), ax.ncell, ax.points.front(),
                                       ax.points.back(), x);
            if (cell < 0) {
                ok = false;
                break;
            }
            const double h = ax.steps[cell];
            const double tl = (x - ax.celld[cell]) / h + first;
            double L[3][kMaxNlip];
            lip_basis(nlip, coef, tl, L);
            for (int d = 0; d < 3; ++d)
                for (int j = 0; j < nlip; ++j) B[3 * a + d][j] = L[d][j] / std::pow(h, d);
            st[a] = cell * (nlip - 1);
        }
        if (!ok) continue;
        inside.push_back(c);
        starts.insert(starts.end(), st, st + 3);
        basis.push_back(B);
    }
    const int ni = (int)inside.size();
    if (ni == 0) return DAGE_OK;
    dage_ctx ctx = grid->ctx;
    const int n3 = nlip * nlip * nlip;
    const size_t bytes = sizeof(double) * (size_t)ni * n3 + sizeof(int) * 3 * (size_t)ni + 64;
    void *scr = nullptr;
    DAGE_TRY(ctx_scratch(ctx, bytes, &scr));
    double *d_blk = (double *)scr;
    int *d_starts = (int *)(d_blk + (size_t)ni * n3);
    hipStream_t s = ctx->stream;
    DAGE_HIP_CHECK(hipMemcpyAsync(d_starts, starts.data(), sizeof(int) * 3 * ni,
                                  hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(gather_cells_kernel, dim3(ni), dim3(256), 0, s, cube, grid->shape(0),
                       grid->shape(1), d_starts, nlip, d_blk);
    DAGE_HIP_CHECK(hipGetLastError());
    std::vector<double> blk((size_t)ni * n3);
    DAGE_HIP_CHECK(hipMemcpyAsync(blk.data(), d_blk, sizeof(double) * blk.size(),
                                  hipMemcpyDeviceToHost, s));
    DAGE_HIP_CHECK(hipStreamSynchronize(s));
    for (int q = 0; q < ni; ++q) {
        const auto &B = basis[q];
        const double *v = blk.data() + (size_t)q * n3;
        for (int d = 0; d < nd; ++d) {
            const auto &kp = kap[d];
            double sum = 0.0;
            for (int k = 0; k < nlip; ++k) {
                double sj = 0.0;
                for (int j = 0; j < nlip; ++j) {
                    double si = 0.0;
                    for (int i = 0; i < nlip; ++i)
                        si += v[i + nlip * (j + nlip * k)] * B[kp[0]][i];
                    sj += si * B[3 + kp[1]][j];
                }
                sum += sj * B[6 + kp[2]][k];
            }
            out[(size_t)nd * inside[q] + d] = sum;
        }
    }
    return DAGE_OK;
}

struct RadialHost {
    std::vector<double> celld, steps, fk;  // fk: npts x nlm, r^k f
    int ncell = 0, npts = 0;
    double qmax = 0.0;
};

int download_bubble(dage_bubbles b, int i, RadialHost &R) {
    const BubbleHost &h = b->host[i];
    R.ncell = h.ncell;
    R.npts = h.npts;
    R.qmax = h.qmax;
    R.celld = h.h_celld;
    R.steps = h.h_steps;
    R.fk.resize((size_t)h.npts * b->nlm());
    hipStream_t s = b->ctx->stream;
    DAGE_HIP_CHECK(hipMemcpyAsync(R.fk.data(), h.fk(), 8 * R.fk.size(), hipMemcpyDeviceToHost, s));
    return DAGE_OK;
}

// sum over the other centres' bubbles of d^kappa [r^k P_lm(r) C_lm] at each
// centre (nder x nbub); points beyond a bubble's radial grid contribute 0.
int bubble_contaminants(dage_bubbles b, int t, std::vector<double> &out) {
    const auto kap = cart_indices(t);
    const int nd = (int)kap.size(), nb = b->nbub, lmax = b->lmax, nlm = b->nlm();
    out.assign((size_t)nd * nb, 0.0);
    if (nb < 2) return DAGE_OK;
    std::vector<RadialHost> R(nb);
    for (int i = 0; i < nb; ++i) DAGE_TRY(download_bubble(b, i, R[i]));
    DAGE_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    double coef[kMaxNlip * kMaxNlip];
    DAGE_REQUIRE(lip_coeffs(b->nlip, 1, coef) == 0, DAGE_ERR_ARG, "unsupported nlip %d", b->nlip);
    double first, last;
    lip_first_last(b->nlip, &first, &last);
    const int nl = b->nlip;
    std::vector<Jet> S;
    std::vector<double> P0(nlm), P1(nlm), P2(nlm);
    for (int ib = 0; ib < nb; ++ib) {      // centre where the derivatives are taken
        for (int ia = 0; ia < nb; ++ia) {  // the bubble
            if (ia == ib) continue;
            const double *ca = b->host[ia].center, *cb = b->host[ib].center;
            const double x[3] = {cb[0] - ca[0], cb[1] - ca[1], cb[2] - ca[2]};
            const double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
            const RadialHost &A = R[ia];
            const int cell = find_cell(A.celld.data(), A.ncell, 0.0, A.qmax, r);
            if (cell < 0 || r == 0.0) continue;
            const double h = A.steps[cell];
            double L[3][kMaxNlip];
            lip_basis(nl, coef, (r - A.celld[cell]) / h + first, L);
            const int p0 = cell * (nl - 1);
            for (int lm = 0; lm < nlm; ++lm) {
                double a0 = 0.0, a1 = 0.0, a2 = 0.0;
                for (int j = 0; j < nl; ++j) {
                    const double v = A.fk[(size_t)lm * A.npts + p0 + j];
                    a0 += v * L[0][j];
                    a1 += v * L[1][j];
                    a2 += v * L[2][j];
                }
                P0[lm] = a0;
                P1[lm] = a1 / h;
                P2[lm] = a2 / (h * h);
            }
            solid_harmonics_jet(lmax, x, S);
            Jet tot;
            for (int l = 0; l <= lmax; ++l) {
                // G = r^a P, a = -l, and its r-derivatives
                const double a = -l;
                const double ra = std::pow(r, a), ra1 = ra / r, ra2 = ra1 / r;
                for (int m = -l; m <= l; ++m) {
                    const int lm = l * l + l + m;
                    const double G = ra * P0[lm];
                    const double G1 = a * ra1 * P0[lm] + ra * P1[lm];
                    const double G2 = a * (a - 1.0) * ra2 * P0[lm] + 2.0 * a * ra1 * P1[lm] +
                                      ra * P2[lm];
                    Jet g;  // G(|x|)
                    g.v = G;
                    for (int i = 0; i < 3; ++i) g.g[i] = G1 * x[i] / r;
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j)
                            g.h[i][j] = G2 * x[i] * x[j] / (r * r) +
                                        G1 * ((i == j ? 1.0 / r : 0.0) - x[i] * x[j] / (r * r * r));
                    const Jet &s = S[lm];
                    tot.v += g.v * s.v;
                    for (int i = 0; i < 3; ++i) tot.g[i] += g.g[i] * s.v + g.v * s.g[i];
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j)
                            tot.h[i][j] += g.h[i][j] * s.v + g.g[i] * s.g[j] + g.g[j] * s.g[i] +
                                           g.v * s.h[i][j];
                }
            }
            for (int d = 0; d < nd; ++d) out[(size_t)nd * ib + d] += jet_derivative(tot, kap[d]);
        }
    }
    return DAGE_OK;
}

// c2s(kappa, lm): u^kappa = sum_lm c2s C_lm(u) on the unit sphere, l <= |kappa|,
// by projection (2l+1)/(4 pi) int u^kappa C_lm dOmega with a product rule
// exact to degree 7.
const std::vector<double> &cart_to_sph(int t) {
    static std::vector<double> tab[kMaxTaylorOrder + 1];
    static bool done[kMaxTaylorOrder + 1] = {false, false, false};
    if (done[t]) return tab[t];
    const auto kap = cart_indices(t);
    const int nd = (int)kap.size(), N = (t + 1) * (t + 1);
    std::vector<double> c((size_t)nd * N, 0.0);
    const int nt = 6, np = 12;
    double ct[nt], wt[nt];
    gauleg(nt, ct, wt);
    std::vector<double> S(N);
    for (int i = 0; i < nt; ++i)
        for (int j = 0; j < np; ++j) {
            const double st = std::sqrt(1.0 - ct[i] * ct[i]), ph = 2.0 * pi() * j / np;
            const double u[3] = {st * std::cos(ph), st * std::sin(ph), ct[i]};
            const double w = wt[i] * 2.0 * pi() / np;
            solid_harmonics_host(t, u[0], u[1], u[2], S.data());
            for (int d = 0; d < nd; ++d) {
                const double mono = std::pow(u[0], kap[d][0]) * std::pow(u[1], kap[d][1]) *
                                    std::pow(u[2], kap[d][2]);
                for (int l = 0; l <= t; ++l)
                    for (int m = -l; m <= l; ++m) {
                        const int lm = l * l + l + m;
                        c[(size_t)d * N + lm] += (2.0 * l + 1.0) / (4.0 * pi()) * w * mono * S[lm];
                    }
            }
        }
    for (auto &v : c)
        if (std::fabs(v) < 1e-14) v = 0.0;
    tab[t] = c;
    done[t] = true;
    return tab[t];
}

// r^p with 0^p = 0 for p < 0 (Bubbles_rpow)
double rpow(double r, int p) {
    if (r == 0.0) return p == 0 ? 1.0 : 0.0;
    return std::pow(r, p);
}

// Bubbles_make_taylor: radial values (raw, k of b) of the Taylor bubbles,
// lmax max(lmax, t), bubble i at vals[off[i]], npts x (lmax_t+1)^2.
void taylor_values(dage_bubbles b, const std::vector<double> &contam, int t,
                   std::vector<double> &vals, std::vector<size_t> &off) {
    const int lt = std::max(b->lmax, t), nlt = (lt + 1) * (lt + 1), k = b->k;
    const auto kap = cart_indices(t);
    const int nd = (int)kap.size(), Nt = (t + 1) * (t + 1);
    const std::vector<double> &c2s = cart_to_sph(t);
    const double fact[kMaxTaylorOrder + 1] = {1.0, 1.0, 0.5};
    off.resize(b->nbub);
    size_t total = 0;
    for (int i = 0; i < b->nbub; ++i) {
        off[i] = total;
        total += (size_t)b->host[i].npts * nlt;
    }
    vals.assign(total, 0.0);
    const int nl = b->nlip;
    for (int i = 0; i < b->nbub; ++i) {
        const BubbleHost &h = b->host[i];
        double *f = vals.data() + off[i];
        for (int p = 0; p < h.npts; ++p) {
            const int cell = std::min(p / (nl - 1), h.ncell - 1);
            const double r = h.h_celld[cell] + (p - cell * (nl - 1)) * h.h_steps[cell];
            const double w = r > 2.220446049250313e-16 ? 0.5 * std::erfc(r - 2.0 / r) : 1.0;
            for (int d = 0; d < nd; ++d) {
                const int n = kap[d][0] + kap[d][1] + kap[d][2];
                const double v = contam[(size_t)nd * i + d] * fact[n] * w * rpow(r, n - k);
                if (v == 0.0) continue;
                for (int lm = 0; lm < Nt; ++lm) {
                    const double cc = c2s[(size_t)d * Nt + lm];
                    if (cc != 0.0) f[(size_t)lm * h.npts + p] += cc * v;
                }
            }
        }
    }
}

int make_taylor(dage_bubbles b, const std::vector<double> &contam, int t, dage_bubbles *out) {
    const int lt = std::max(b->lmax, t);
    std::vector<double> vals;
    std::vector<size_t> off;
    taylor_values(b, contam, t, vals, off);
    DAGE_TRY(dage_bubbles_create_like(b, lt, b->k, out));
    dage_bubbles T = *out;
    const int nlt = (lt + 1) * (lt + 1);
    for (int i = 0; i < b->nbub; ++i)
        DAGE_HIP_CHECK(hipMemcpy(T->f(i), vals.data() + off[i],
                                 8 * (size_t)b->host[i].npts * nlt, hipMemcpyHostToDevice));
    T->dirty = true;
    return bubbles_fold_rk(T);
}

int centres_of(dage_bubbles b, std::vector<double> &c) {
    c.resize(3 * (size_t)b->nbub);
    for (int i = 0; i < b->nbub; ++i)
        for (int a = 0; a < 3; ++a) c[3 * (size_t)i + a] = b->host[i].center[a];
    return DAGE_OK;
}

int use_ctx(dage_ctx ctx) {
    DAGE_REQUIRE(ctx != nullptr, DAGE_ERR_ARG, "null context handle");
    DAGE_HIP_CHECK(hipSetDevice(ctx->device));
    return DAGE_OK;
}

int contaminants(dage_grid3d grid, const double *cube, dage_bubbles b, int t,
                 std::vector<double> &out) {
    std::vector<double> cen, cc;
    centres_of(b, cen);
    DAGE_TRY(bubble_contaminants(b, t, out));
    if (cube) {
        DAGE_TRY(cube_contaminants(grid, cube, cen, t, cc));
        for (size_t i = 0; i < out.size(); ++i) out[i] += cc[i];
    }
    return DAGE_OK;
}

}  // namespace
}  // namespace dage

using namespace dage;

extern "C" {

int dage_function3d_contaminants(dage_grid3d grid, const double *cube, dage_bubbles b,
                                 int taylor_order, double *out) {
    DAGE_REQUIRE(b && out, DAGE_ERR_ARG, "null argument");
    DAGE_REQUIRE(taylor_order >= 0 && taylor_order <= kMaxTaylorOrder, DAGE_ERR_ARG,
                 "taylor_order %d outside 0..%d", taylor_order, kMaxTaylorOrder);
    DAGE_REQUIRE(cube == nullptr || grid != nullptr, DAGE_ERR_ARG, "a cube needs its grid");
    for (int i = 0; i < b->nbub; ++i)
        DAGE_REQUIRE(b->host[i].set, DAGE_ERR_STATE, "bubble %d has not been set", i);
    if (grid) DAGE_REQUIRE(grid->ctx == b->ctx, DAGE_ERR_ARG, "grid and bubbles in different contexts");
    DAGE_TRY(use_ctx(b->ctx));
    std::vector<double> c;
    DAGE_TRY(contaminants(grid, cube, b, taylor_order, c));
    std::copy(c.begin(), c.end(), out);
    return DAGE_OK;
}

int dage_function3d_taylor_bubbles(dage_grid3d grid, const double *cube, dage_bubbles b,
                                   int taylor_order, dage_bubbles *out) {
    DAGE_REQUIRE(b && out, DAGE_ERR_ARG, "null argument");
    *out = nullptr;
    DAGE_REQUIRE(taylor_order >= 0 && taylor_order <= kMaxTaylorOrder, DAGE_ERR_ARG,
                 "taylor_order %d outside 0..%d", taylor_order, kMaxTaylorOrder);
    DAGE_REQUIRE(cube == nullptr || grid != nullptr, DAGE_ERR_ARG, "a cube needs its grid");
    for (int i = 0; i < b->nbub; ++i)
        DAGE_REQUIRE(b->host[i].set, DAGE_ERR_STATE, "bubble %d has not been set", i);
    if (grid) DAGE_REQUIRE(grid->ctx == b->ctx, DAGE_ERR_ARG, "grid and bubbles in different contexts");
    DAGE_TRY(use_ctx(b->ctx));
    std::vector<double> c;
    DAGE_TRY(contaminants(grid, cube, b, taylor_order, c));
    return make_taylor(b, c, taylor_order, out);
}

int dage_function3d_multiply_taylor(dage_grid3d grid, const double *cube1, dage_bubbles b1,
                                    const double *cube2, dage_bubbles b2, int taylor_order,
                                    double *cube_out, dage_bubbles bub_out) {
    DAGE_REQUIRE(grid && cube1 && cube2 && cube_out && b1 && b2 && bub_out, DAGE_ERR_ARG,
                 "null argument");
    DAGE_REQUIRE(taylor_order >= 0 && taylor_order <= kMaxTaylorOrder, DAGE_ERR_ARG,
                 "taylor_order %d outside 0..%d", taylor_order, kMaxTaylorOrder);
    DAGE_REQUIRE(cube_out != cube1, DAGE_ERR_ARG, "the result cube cannot be the first factor");
    const int t = taylor_order;
    const int l1 = std::max(b1->lmax, t), l2 = std::max(b2->lmax, t);
    DAGE_REQUIRE(bub_out->lmax == l1 + l2 && bub_out->k == b1->k + b2->k, DAGE_ERR_ARG,
                 "result bubbles need lmax %d and k %d (got %d, %d)", l1 + l2, b1->k + b2->k,
                 bub_out->lmax, bub_out->k);
    DAGE_REQUIRE(bub_out != b1 && bub_out != b2, DAGE_ERR_ARG,
                 "the product cannot overwrite a factor");
    DAGE_REQUIRE(b1->nbub == b2->nbub && b1->nbub == bub_out->nbub, DAGE_ERR_ARG,
                 "bubble counts differ (%d, %d, %d)", b1->nbub, b2->nbub, bub_out->nbub);
    for (int i = 0; i < b1->nbub; ++i)
        DAGE_REQUIRE(b1->host[i].set && b2->host[i].set && bub_out->host[i].set &&
                         b1->host[i].npts == b2->host[i].npts &&
                         b1->host[i].npts == bub_out->host[i].npts,
                     DAGE_ERR_ARG, "bubble %d: radial grids differ or are not set", i);
    DAGE_TRY(use_ctx(grid->ctx));
    dage_ctx ctx = grid->ctx;
    // the Taylor bubbles of both factors (contaminants read before the
    // result overwrites anything, cube_out may be cube2), kept in the
    // context's auxiliary buffer
    const bool same = (cube1 == cube2 && b1 == b2);
    std::vector<double> c1, c2, v1, v2;
    std::vector<size_t> o1, o2;
    DAGE_TRY(contaminants(grid, cube1, b1, t, c1));
    taylor_values(b1, c1, t, v1, o1);
    if (!same) {
        DAGE_TRY(contaminants(grid, cube2, b2, t, c2));
        taylor_values(b2, c2, t, v2, o2);
    }
    void *ap = nullptr;
    DAGE_TRY(ctx_aux(ctx, 8 * (v1.size() + v2.size()), &ap));
    double *d_t1 = (double *)ap, *d_t2 = same ? d_t1 : d_t1 + v1.size();
    DAGE_HIP_CHECK(hipMemcpyAsync(d_t1, v1.data(), 8 * v1.size(), hipMemcpyHostToDevice,
                                  ctx->stream));
    if (!same)
        DAGE_HIP_CHECK(hipMemcpyAsync(d_t2, v2.data(), 8 * v2.size(), hipMemcpyHostToDevice,
                                      ctx->stream));
    const int nb = b1->nbub;
    std::vector<const double *> f1(nb), f2(nb), t1(nb), t2(nb);
    std::vector<double *> fo(nb);
    std::vector<int> np(nb);
    for (int i = 0; i < nb; ++i) {
        f1[i] = b1->f(i);
        f2[i] = b2->f(i);
        t1[i] = d_t1 + o1[i];
        t2[i] = d_t2 + (same ? o1[i] : o2[i]);
        fo[i] = bub_out->f(i);
        np[i] = b1->host[i].npts;
    }
    // result bubbles: B1 B2 + B1 T2 + T1 B2  (= (B1 + T1)(B2 + T2) - T1 T2)
    DAGE_TRY(dage_bubbles_axpby(0.0, bub_out, 0.0, bub_out));
    DAGE_TRY(bubble_product_acc(ctx, nb, f1.data(), b1->lmax, f2.data(), b2->lmax, fo.data(),
                                np.data()));
    DAGE_TRY(bubble_product_acc(ctx, nb, f1.data(), b1->lmax, t2.data(), l2, fo.data(),
                                np.data()));
    DAGE_TRY(bubble_product_acc(ctx, nb, t1.data(), l1, f2.data(), b2->lmax, fo.data(),
                                np.data()));
    DAGE_TRY(bubbles_fold_rk(bub_out));
    // cube: (c1 + B1)(c2 + B2) - result bubbles
    DAGE_TRY(dage_function3d_multiply(grid, cube1, b1, cube2, b2, cube_out, nullptr));
    DAGE_TRY(dage_bubbles_inject_device(bub_out, grid, 0, -1.0, cube_out));
    // the host vectors go out of scope: the uploads must have completed
    DAGE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return DAGE_OK;
}

}  // extern "C"
