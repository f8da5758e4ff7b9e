// FMM translation (M2M/L2L) and interaction (M2L) matrices for Racah-
// normalised real regular solid harmonics S_lm, built from polynomial algebra.
//
// Conventions (those of the reference's multipole.F90:33-560, which this
// replaces):
//   multipoles  q_lm(c)  = integral rho(r) S_lm(r - c)
//   potential   V(r)     = sum_lm q_lm I_lm(r - c),  I_lm(v) = S_lm(v)/|v|^(2l+1)
//   local exp.  V(c + y) = sum_jk L_jk S_jk(y)
//   M2L: L_target = T(D) q_source,  D = c_source - c_target
//   M2M: q(new) = W(old - new) q(old);  L2L: L(child) = W(child - parent)^T L(parent)
//
// Derivation used here (not the reference's coefficient tables):
//  * Fischer inner product <P,Q> = sum_abc a! b! c! p_abc q_abc on homogeneous
//    polynomials; the S_lm of one degree are orthogonal under it, and it is
//    P(grad)Q for equal degrees.
//  * Translation: S_lm(y + d) = sum_{j<=l,k} S_jk(y) [S_jk(grad) S_lm](d) / <S_jk,S_jk>,
//    and S_jk(grad) S_lm is harmonic of degree l-j, expanded in S_{l-j,n}.
//  * Interaction (Hobson: H(grad) 1/r = (-1)^l (2l-1)!! H(r)/r^(2l+1) for
//    harmonic H of degree l):
//      I_lm(y - D) = sum_jk S_jk(y) (-1)^l (2l+2j-1)!!/((2l-1)!!(2j-1)!!)
//                    sum_n g_n I_{l+j,n}(D),
//    g_n the coefficients of the harmonic part of S_lm S_jk in S_{l+j,n}.
// Tables are computed once per lmax in long double and cached.
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "dage_internal.hpp"

namespace dage {
namespace {

typedef long double ld;

// Homogeneous polynomial of degree n: c[a*(n+1)+b] multiplies x^a y^b z^(n-a-b).
struct Poly {
    int n = 0;
    std::vector<ld> c;
    explicit Poly(int deg = 0) : n(deg), c((size_t)(deg + 1) * (deg + 1), 0.0L) {}
    ld &at(int a, int b) { return c[(size_t)a * (n + 1) + b]; }
    ld get(int a, int b) const { return c[(size_t)a * (n + 1) + b]; }
};

Poly mul_var(const Poly &p, int var) {  // var 0: x, 1: y, 2: z
    Poly r(p.n + 1);
    for (int a = 0; a <= p.n; ++a)
        for (int b = 0; a + b <= p.n; ++b) {
            const ld v = p.get(a, b);
            if (v == 0.0L) continue;
            if (var == 0) r.at(a + 1, b) += v;
            else if (var == 1) r.at(a, b + 1) += v;
            else r.at(a, b) += v;
        }
    return r;
}

Poly mul_r2(const Poly &p) {
    Poly r(p.n + 2);
    for (int a = 0; a <= p.n; ++a)
        for (int b = 0; a + b <= p.n; ++b) {
            const ld v = p.get(a, b);
            if (v == 0.0L) continue;
            r.at(a + 2, b) += v;
            r.at(a, b + 2) += v;
            r.at(a, b) += v;
        }
    return r;
}

Poly axpy(ld alpha, const Poly &x, ld beta, const Poly &y) {
    Poly r(x.n);
    for (size_t i = 0; i < r.c.size(); ++i) r.c[i] = alpha * x.c[i] + beta * y.c[i];
    return r;
}

Poly product(const Poly &p, const Poly &q) {
    Poly r(p.n + q.n);
    for (int a = 0; a <= p.n; ++a)
        for (int b = 0; a + b <= p.n; ++b) {
            const ld v = p.get(a, b);
            if (v == 0.0L) continue;
            for (int a2 = 0; a2 <= q.n; ++a2)
                for (int b2 = 0; a2 + b2 <= q.n; ++b2) {
                    const ld w = q.get(a2, b2);
                    if (w != 0.0L) r.at(a + a2, b + b2) += v * w;
                }
        }
    return r;
}

ld factl(int n) {
    ld f = 1.0L;
    for (int i = 2; i <= n; ++i) f *= (ld)i;
    return f;
}

ld falling(int n, int k) {  // n!/(n-k)!
    ld f = 1.0L;
    for (int i = 0; i < k; ++i) f *= (ld)(n - i);
    return f;
}

ld fischer(const Poly &p, const Poly &q) {
    ld s = 0.0L;
    for (int a = 0; a <= p.n; ++a)
        for (int b = 0; a + b <= p.n; ++b) {
            const ld v = p.get(a, b) * q.get(a, b);
            if (v != 0.0L) s += v * factl(a) * factl(b) * factl(p.n - a - b);
        }
    return s;
}

// P(grad) Q for deg P <= deg Q: homogeneous of degree Q.n - P.n.
Poly apply_grad(const Poly &p, const Poly &q) {
    Poly r(q.n - p.n);
    for (int a = 0; a <= p.n; ++a)
        for (int b = 0; a + b <= p.n; ++b) {
            const ld v = p.get(a, b);
            if (v == 0.0L) continue;
            const int c = p.n - a - b;
            for (int A = a; A <= q.n; ++A)
                for (int B = b; A + B <= q.n; ++B) {
                    const int C = q.n - A - B;
                    if (C < c) continue;
                    const ld w = q.get(A, B);
                    if (w == 0.0L) continue;
                    r.at(A - a, B - b) += v * w * falling(A, a) * falling(B, b) * falling(C, c);
                }
        }
    return r;
}

// Racah real regular solid harmonics as polynomials, index l*l + l + m.
std::vector<Poly> solid_polys(int lmax) {
    std::vector<Poly> S;
    S.reserve((size_t)(lmax + 1) * (lmax + 1));
    Poly one(0);
    one.at(0, 0) = 1.0L;
    S.push_back(one);
    if (lmax >= 1) {
        Poly y(1), z(1), x(1);
        y.at(0, 1) = 1.0L;
        z.at(0, 0) = 1.0L;
        x.at(1, 0) = 1.0L;
        S.push_back(y);
        S.push_back(z);
        S.push_back(x);
    }
    for (int l = 2; l <= lmax; ++l) {
        const int p = (l - 1) * (l - 1) + (l - 1), q = (l - 2) * (l - 2) + (l - 2);
        const ld f = std::sqrt((ld)(2 * l - 1) / (ld)(2 * l));
        std::vector<Poly> shell(2 * l + 1);
        shell[0] = axpy(f, mul_var(S[p + (l - 1)], 1), f, mul_var(S[p - (l - 1)], 0));
        for (int m = -(l - 1); m <= l - 1; ++m) {
            const ld a = (ld)(2 * l - 1) / std::sqrt((ld)(l + m) * (l - m));
            Poly v = mul_var(S[p + m], 2);
            for (auto &e : v.c) e *= a;
            if (l - std::abs(m) > 1) {
                const ld b = std::sqrt((ld)(l + m - 1) * (l - m - 1) / ((ld)(l + m) * (l - m)));
                v = axpy(1.0L, v, -b, mul_r2(S[q + m]));
            }
            shell[m + l] = v;
        }
        shell[2 * l] = axpy(f, mul_var(S[p + (l - 1)], 0), -f, mul_var(S[p - (l - 1)], 1));
        for (auto &e : shell) S.push_back(e);
    }
    return S;
}

// Coefficients of a harmonic degree-L polynomial in S_{L,n}, trying the
// orders n that a product or derivative of S_lm and S_jk can produce.
void project(const Poly &h, int L, int m, int k, const std::vector<Poly> &S,
             const std::vector<ld> &norm, std::vector<std::pair<int, ld>> *out) {
    int cand[4] = {std::abs(m) + std::abs(k), std::abs(std::abs(m) - std::abs(k)), 0, 0};
    cand[2] = -cand[0];
    cand[3] = -cand[1];
    for (int i = 0; i < 4; ++i) {
        const int n = cand[i];
        if (std::abs(n) > L) continue;
        bool dup = false;
        for (int j = 0; j < i; ++j) dup |= (cand[j] == n && std::abs(cand[j]) <= L);
        if (dup) continue;
        const int idx = L * L + L + n;
        const ld g = fischer(h, S[idx]) / norm[idx];
        if (std::fabs(g) > 1e-30L) out->push_back({idx, g});
    }
}

ld dfact(int n) {  // n!!, 1 for n <= 0
    ld f = 1.0L;
    for (int i = n; i > 1; i -= 2) f *= (ld)i;
    return f;
}

}  // namespace

void solid_harmonics_host(int lmax, double x, double y, double z, double *S) {
    S[0] = 1.0;
    if (lmax < 1) return;
    S[1] = y;
    S[2] = z;
    S[3] = x;
    const double r2 = x * x + y * y + z * z;
    for (int l = 2; l <= lmax; ++l) {
        const int c = l * l + l, p = (l - 1) * (l - 1) + (l - 1), q = (l - 2) * (l - 2) + (l - 2);
        const double f = std::sqrt((2 * l - 1.0) / (2.0 * l));
        S[c - l] = f * (y * S[p + (l - 1)] + x * S[p - (l - 1)]);
        for (int m = -(l - 1); m <= l - 1; ++m) {
            double v = (2 * l - 1.0) / std::sqrt(1.0 * (l + m) * (l - m)) * (z * S[p + m]);
            if (l - std::abs(m) > 1)
                v = v - std::sqrt((l + m - 1.0) * (l - m - 1.0) / ((l + m) * (l - m))) *
                            (r2 * S[q + m]);
            S[c + m] = v;
        }
        S[c + l] = f * (x * S[p + (l - 1)] - y * S[p - (l - 1)]);
    }
}

const FmmTables &fmm_tables(int lmax) {
    static std::mutex mu;
    static std::map<int, std::unique_ptr<FmmTables>> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(lmax);
    if (it != cache.end()) return *it->second;

    auto tab = std::make_unique<FmmTables>();
    tab->lmax = lmax;
    const int N = (lmax + 1) * (lmax + 1);
    const std::vector<Poly> S = solid_polys(2 * lmax);
    std::vector<ld> norm(S.size());
    for (size_t i = 0; i < S.size(); ++i) norm[i] = fischer(S[i], S[i]);

    std::vector<std::pair<int, ld>> terms;
    // interaction: entry (row jk, column lm)
    tab->t_start.assign((size_t)N * N + 1, 0);
    for (int col = 0; col < N; ++col) {
        const int l = (int)std::sqrt((double)col), m = col - l * l - l;
        for (int row = 0; row < N; ++row) {
            const int j = (int)std::sqrt((double)row), k = row - j * j - j;
            terms.clear();
            project(product(S[col], S[row]), l + j, m, k, S, norm, &terms);
            const ld coef = ((l % 2) ? -1.0L : 1.0L) * dfact(2 * l + 2 * j - 1) /
                            (dfact(2 * l - 1) * dfact(2 * j - 1));
            for (auto &t : terms) {
                tab->t_idx.push_back(t.first);
                tab->t_val.push_back((double)(coef * t.second));
            }
            tab->t_start[(size_t)row + (size_t)N * col + 1] = (int)tab->t_idx.size();
        }
    }
    // translation: entry (row lm, column jk), j <= l
    tab->w_start.assign((size_t)N * N + 1, 0);
    for (int col = 0; col < N; ++col) {
        const int j = (int)std::sqrt((double)col), k = col - j * j - j;
        for (int row = 0; row < N; ++row) {
            const int l = (int)std::sqrt((double)row), m = row - l * l - l;
            terms.clear();
            if (j <= l && row > col) {
                project(apply_grad(S[col], S[row]), l - j, m, k, S, norm, &terms);
                for (auto &t : terms) {
                    tab->w_idx.push_back(t.first);
                    tab->w_val.push_back((double)(t.second / norm[col]));
                }
            }
            tab->w_start[(size_t)row + (size_t)N * col + 1] = (int)tab->w_idx.size();
        }
    }
    const FmmTables &ref = *tab;
    cache[lmax] = std::move(tab);
    return ref;
}

void fmm_interaction(const FmmTables &tab, const double D[3], double *T) {
    const int lmax = tab.lmax, N = (lmax + 1) * (lmax + 1), L2 = 2 * lmax;
    std::vector<double> I((size_t)(L2 + 1) * (L2 + 1));
    solid_harmonics_host(L2, D[0], D[1], D[2], I.data());
    const double r = std::sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
    double rp = r;  // r^(2L+1)
    for (int L = 0; L <= L2; ++L) {
        for (int n = -L; n <= L; ++n) I[L * L + L + n] /= rp;
        rp *= r * r;
    }
    for (size_t e = 0; e < (size_t)N * N; ++e) {
        double v = 0.0;
        for (int t = tab.t_start[e]; t < tab.t_start[e + 1]; ++t) v += tab.t_val[t] * I[tab.t_idx[t]];
        T[e] = v;
    }
}

void fmm_translation(const FmmTables &tab, const double d[3], double *W) {
    const int lmax = tab.lmax, N = (lmax + 1) * (lmax + 1);
    std::vector<double> S(N);
    solid_harmonics_host(lmax, d[0], d[1], d[2], S.data());
    for (int col = 0; col < N; ++col)
        for (int row = 0; row < N; ++row) {
            const size_t e = (size_t)row + (size_t)N * col;
            double v = (row == col) ? 1.0 : 0.0;
            for (int t = tab.w_start[e]; t < tab.w_start[e + 1]; ++t) v += tab.w_val[t] * S[tab.w_idx[t]];
            W[e] = v;
        }
}

}  // namespace dage
