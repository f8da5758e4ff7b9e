// Host math of the screened-Coulomb (Yukawa) far field, the kernel
// G(r) = exp(-kappa r)/r of Helmholtz3D at energy E = -kappa^2/2
// (GBFMMHelmholtz3D, src/gbfmm/gbfmm_helmholtz3d.F90:241-515, whose
// translations helmholtz_multipole_tools.F90:127-678 rotates and shifts with
// Bessel recursions; here the same expansions are built by projection).
//
// Bases (C_lm = S_lm(v)/|v|^l, the Racah-normalised real harmonics of
// solid_harmonics_host; i_l, k_l modified spherical Bessel functions with
// k_0(x) = (pi/2) exp(-x)/x):
//   regular   R_lm(v) = i_l(kappa |v|) C_lm(v)
//   singular  K_lm(v) = k_l(kappa |v|) C_lm(v)
// Addition theorem, |y - c| < |x - c|:
//   G(|x - y|) = (2 kappa / pi) sum_lm (2l + 1) R_lm(y - c) K_lm(x - c).
// Multipoles M_lm(c) = sum_y rho(y) w(y) R_lm(y - c); local expansions
// phi(x) = sum_lm L_lm R_lm(x - d).
//
// A function f that is regular inside the sphere |x - d| <= rho has
// coefficients L_lm = (2l + 1) / (4 pi i_l(kappa rho)) int f(d + rho W) C_lm(W) dW;
// the integral is taken with a product rule (Gauss-Legendre in cos(theta),
// uniform in phi) exact for harmonics of degree 2 nt - 1, nt = lmax + 16, so
// the neglected components of f (falling as (rho/|D|)^l for a source at
// distance |D|, |D| >= 2.3 rho between well-separated boxes) alias below
// rounding.  M2L:  T = P K  with P the projector above and K the sampled
// singular fields of the source multipoles; M2M and L2L: the regular field
// R_lm(z + d) re-expanded about the child centre, A(lm, l'm').
#include <algorithm>
#include <cmath>
#include <vector>

#include "dage_internal.hpp"

namespace dage {

// g_l(x) = i_l(x) / x^l, l = 0..lmax, 0 <= x <= 600.  x <= 20: the
// all-positive series g_l(x) = sum_k (x^2/2)^k / (k! (2l + 2k + 1)!!) for the
// two highest l, then the downward recurrence g_{l-1} = (2l + 1) g_l +
// x^2 g_{l+1} (from i_{l-1} = i_{l+1} + (2l+1)/x i_l), positive terms only.
// x > 20: Miller's downward recurrence from l = lmax + x + 40 with arbitrary
// start values, normalised by i_0(x) = sinh(x)/x.
void bessel_i_scaled(int lmax, double x, double *g) {
    if (x <= 20.0) {
        const double h = 0.5 * x * x;
        auto series = [&](int l) {
            double df = 1.0;  // (2l+1)!!
            for (int i = 3; i <= 2 * l + 1; i += 2) df *= i;
            double term = 1.0 / df, sum = term;
            for (int k = 1; k < 4000; ++k) {
                term *= h / (k * (2.0 * l + 2.0 * k + 1.0));
                sum += term;
                if (term < 1e-18 * sum && (double)k > h) break;
            }
            return sum;
        };
        g[lmax] = series(lmax);
        if (lmax == 0) return;
        g[lmax - 1] = series(lmax - 1);
        for (int l = lmax - 1; l >= 1; --l) g[l - 1] = (2.0 * l + 1.0) * g[l] + x * x * g[l + 1];
        return;
    }
    const int L = lmax + (int)x + 40;
    double ip1 = 0.0, il = 1e-200;  // i_{l+1}, i_l up to a common factor
    for (int l = L; l >= 1; --l) {
        const double im1 = ip1 + (2.0 * l + 1.0) / x * il;
        ip1 = il;
        il = im1;  // i_{l-1}
        if (l - 1 <= lmax) g[l - 1] = il;
        if (il > 1e200) {
            il *= 1e-200;
            ip1 *= 1e-200;
            for (int j = l - 1; j <= lmax; ++j) g[j] *= 1e-200;
        }
    }
    const double scale = std::sinh(x) / x / g[0];
    double xl = 1.0;
    for (int l = 0; l <= lmax; ++l) {
        g[l] *= scale / xl;
        xl *= x;
    }
}

// k_l(x), l = 0..lmax, x > 0: k_0 = (pi/2) e^-x / x, k_1 = (pi/2) e^-x (1+x)/x^2,
// k_{l+1} = k_{l-1} + (2l+1)/x k_l (upward; all terms positive).
void bessel_k(int lmax, double x, double *k) {
    const double e = 0.5 * pi() * std::exp(-x) / x;
    k[0] = e;
    if (lmax == 0) return;
    k[1] = e * (1.0 + 1.0 / x);
    for (int l = 1; l < lmax; ++l) k[l + 1] = k[l - 1] + (2.0 * l + 1.0) / x * k[l];
}

// R_lm(v) for all lm (N = (lmax+1)^2): S_lm(v) kappa^l g_l(kappa |v|).
void yukawa_regular(int lmax, double kappa, const double v[3], double *R) {
    double g[kMaxFmmBasisL + 1];
    const double r = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    bessel_i_scaled(lmax, kappa * r, g);
    solid_harmonics_host(lmax, v[0], v[1], v[2], R);
    double kl = 1.0;
    for (int l = 0; l <= lmax; ++l) {
        const double f = kl * g[l];
        for (int m = -l; m <= l; ++m) R[l * l + l + m] *= f;
        kl *= kappa;
    }
}

// K_lm(v) for all lm, v != 0: S_lm(v) k_l(kappa |v|) / |v|^l.
void yukawa_singular(int lmax, double kappa, const double v[3], double *K) {
    double k[kMaxFmmBasisL + 1];
    const double r = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    bessel_k(lmax, kappa * r, k);
    solid_harmonics_host(lmax, v[0], v[1], v[2], K);
    double rl = 1.0;
    for (int l = 0; l <= lmax; ++l) {
        const double f = k[l] / rl;
        for (int m = -l; m <= l; ++m) K[l * l + l + m] *= f;
        rl *= r;
    }
}

YukawaSphere yukawa_sphere(int lmax) {
    YukawaSphere s;
    s.lmax = lmax;
    const int nt = lmax + 16, np = 2 * nt, N = (lmax + 1) * (lmax + 1);
    std::vector<double> ct(nt), wt(nt);
    gauleg(nt, ct.data(), wt.data());
    s.Q = nt * np;
    s.nodes.resize(3 * (size_t)s.Q);
    s.w.resize(s.Q);
    s.C.resize((size_t)N * s.Q);
    const double dphi = 2.0 * pi() / np;
    std::vector<double> S(N);
    for (int i = 0; i < nt; ++i) {
        const double st = std::sqrt(std::max(0.0, 1.0 - ct[i] * ct[i]));
        for (int j = 0; j < np; ++j) {
            const int q = i * np + j;
            const double n[3] = {st * std::cos(j * dphi), st * std::sin(j * dphi), ct[i]};
            for (int a = 0; a < 3; ++a) s.nodes[3 * (size_t)q + a] = n[a];
            s.w[q] = wt[i] * dphi;
            solid_harmonics_host(lmax, n[0], n[1], n[2], S.data());  // |n| = 1: S = C
            for (int lm = 0; lm < N; ++lm) s.C[(size_t)lm + (size_t)N * q] = S[lm];
        }
    }
    return s;
}

void yukawa_projector(const YukawaSphere &s, double kappa, double rho, double *P) {
    const int lmax = s.lmax, N = (lmax + 1) * (lmax + 1);
    double g[kMaxFmmBasisL + 1];
    bessel_i_scaled(lmax, kappa * rho, g);
    std::vector<double> f(lmax + 1);
    double xl = 1.0;
    for (int l = 0; l <= lmax; ++l) {
        f[l] = (2.0 * l + 1.0) / (4.0 * pi() * xl * g[l]);  // i_l(x) = x^l g_l(x)
        xl *= kappa * rho;
    }
    for (int q = 0; q < s.Q; ++q)
        for (int l = 0; l <= lmax; ++l)
            for (int m = -l; m <= l; ++m) {
                const int lm = l * l + l + m;
                P[(size_t)lm + (size_t)N * q] = f[l] * s.w[q] * s.C[(size_t)lm + (size_t)N * q];
            }
}

void yukawa_m2l_samples(const YukawaSphere &s, double kappa, const double D[3], double rho,
                        double *Kq) {
    const int lmax = s.lmax, N = (lmax + 1) * (lmax + 1);
    std::vector<double> K(N);
    for (int q = 0; q < s.Q; ++q) {
        const double v[3] = {rho * s.nodes[3 * (size_t)q] - D[0],
                             rho * s.nodes[3 * (size_t)q + 1] - D[1],
                             rho * s.nodes[3 * (size_t)q + 2] - D[2]};
        yukawa_singular(lmax, kappa, v, K.data());
        for (int l = 0; l <= lmax; ++l) {
            const double c = 2.0 * kappa / pi() * (2.0 * l + 1.0);
            for (int m = -l; m <= l; ++m) {
                const int lm = l * l + l + m;
                Kq[(size_t)q + (size_t)s.Q * lm] = c * K[lm];
            }
        }
    }
}

void yukawa_shift(const YukawaSphere &s, double kappa, const double d[3], double rho, double *A) {
    const int lmax = s.lmax, N = (lmax + 1) * (lmax + 1);
    std::vector<double> P((size_t)N * s.Q), Rq((size_t)N * s.Q), R(N);
    yukawa_projector(s, kappa, rho, P.data());
    for (int q = 0; q < s.Q; ++q) {
        const double v[3] = {rho * s.nodes[3 * (size_t)q] + d[0],
                             rho * s.nodes[3 * (size_t)q + 1] + d[1],
                             rho * s.nodes[3 * (size_t)q + 2] + d[2]};
        yukawa_regular(lmax, kappa, v, R.data());
        std::copy(R.begin(), R.end(), Rq.begin() + (size_t)N * q);
    }
    // A(lm, l'm') = sum_q R_lm(rho W_q + d) P(l'm', q)
    for (int c = 0; c < N; ++c)
        for (int r = 0; r < N; ++r) {
            double v = 0.0;
            for (int q = 0; q < s.Q; ++q)
                v += Rq[(size_t)r + (size_t)N * q] * P[(size_t)c + (size_t)N * q];
            A[(size_t)r + (size_t)N * c] = v;
        }
}

void yukawa_m2l(const YukawaSphere &s, double kappa, const double D[3], double rho, double *T) {
    const int lmax = s.lmax, N = (lmax + 1) * (lmax + 1);
    std::vector<double> P((size_t)N * s.Q), Kq((size_t)s.Q * N);
    yukawa_projector(s, kappa, rho, P.data());
    yukawa_m2l_samples(s, kappa, D, rho, Kq.data());
    for (int c = 0; c < N; ++c)
        for (int r = 0; r < N; ++r) {
            double v = 0.0;
            for (int q = 0; q < s.Q; ++q)
                v += P[(size_t)r + (size_t)N * q] * Kq[(size_t)q + (size_t)s.Q * c];
            T[(size_t)r + (size_t)N * c] = v;
        }
}

}  // namespace dage
