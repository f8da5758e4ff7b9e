// Host-side numerics of the library: t-quadrature, LIP basis and cell grids.
// These run once per operator/grid on the CPU and feed the GPU kernels.
// Built with -ffp-contract=off so results round exactly as written.
#include <cmath>
#include <cstring>

#include "dage_internal.hpp"

namespace dage {

// globals.F90:54-61: PI = dacos(-1), TWOOVERSQRTPI = 2/dsqrt(PI)
double pi() { return std::acos(-1.0); }
double two_over_sqrtpi() { return 2.0 / std::sqrt(pi()); }

// gauss.F90:337-384: Legendre roots by Newton iteration from the cosine
// guess, weights from the derivative of the last-but-one iterate.
void gauleg(int ngp, double *x, double *w) {
    const double eps = 3.0e-15;
    const double p = pi();
    const int half = (ngp + 1) / 2;
    for (int i = 1; i <= half; ++i) {
        double z = std::cos(p * (i - 0.25) / (ngp + 0.5));
        double dz, deriv;
        do {
            double pn = 1.0, pm = 0.0;
            for (int j = 1; j <= ngp; ++j) {
                double pmm = pm;
                pm = pn;
                pn = ((2.0 * j - 1.0) * z * pm - (j - 1.0) * pmm) / j;
            }
            deriv = ngp * (z * pn - pm) / (z * z - 1.0);
            double znew = z - pn / deriv;
            dz = znew - z;
            z = znew;
        } while (std::fabs(dz) > eps);
        x[i - 1] = -z;
        x[ngp - i] = z;
        w[i - 1] = 2.0 / ((1.0 - z * z) * deriv * deriv);
        w[ngp - i] = w[i - 1];
    }
}

// gauss.F90:81-122 with gauleg_int (:307-329): linear Gauss-Legendre part on
// [tstart, tlin], logarithmic part on [tlin, tlog].
void gauss_quad(int nlin, int nlog, double tstart, double tlin, double tlog,
                double *tp, double *tw) {
    gauleg(nlin, tp, tw);
    for (int i = 0; i < nlin; ++i) {
        tp[i] = 0.5 * ((tlin + tstart) + (tlin - tstart) * tp[i]);
        tw[i] = 0.5 * (tlin - tstart) * tw[i];
    }
    double *lp = tp + nlin, *lw = tw + nlin;
    gauleg(nlog, lp, lw);
    const double ratio = tlog / tlin;
    for (int i = 0; i < nlog; ++i) {
        lp[i] = tlin * std::pow(ratio, 0.5 * (lp[i] + 1.0));
        lw[i] = 0.5 * std::log(ratio) * lw[i] * lp[i];
    }
}

// gauss.F90:409-446 (orders 6 and 7 are the ones the reference tabulates).
int gauss_lobatto_grid(int n, double left, double right, double *tp) {
    if (n == 6) {
        const double a = std::sqrt((7.0 + 2 * std::sqrt(7.0)) / 21.0);
        const double b = std::sqrt((7.0 - 2 * std::sqrt(7.0)) / 21.0);
        const double v[6] = {-1.0, -a, -b, b, a, 1.0};
        std::memcpy(tp, v, sizeof v);
    } else if (n == 7) {
        const double a = std::sqrt((15.0 + 2 * std::sqrt(15.0)) / 33.0);
        const double b = std::sqrt((15.0 - 2 * std::sqrt(15.0)) / 33.0);
        const double v[7] = {-1.0, -a, -b, 0.0, b, a, 1.0};
        std::memcpy(tp, v, sizeof v);
    } else {
        return -1;
    }
    for (int i = 0; i < n; ++i) tp[i] = 0.5 * ((right + left) + (right - left) * tp[i]);
    return 0;
}

// lip.F90:118-142 (integer divisions as in the reference).
void lip_first_last(int nlip, double *first, double *last) {
    *first = static_cast<double>(-((nlip - 1) / 2));
    *last = static_cast<double>(nlip / 2);
}

// lip.F90:285-344: Lagrange basis coefficients, decreasing powers, built by
// multiplying in one factor (x - x_i)/(x_j - x_i) at a time.
int lip_coeffs(int nlip, int grid_type, double *c) {
    const int n = nlip;
    double node[kMaxNlip];
    if (n < 2 || n > kMaxNlip) return -1;
    if (grid_type == 1) {
        for (int i = 0; i < n; ++i) node[i] = static_cast<double>(i + 1 - (n + 1) / 2);
    } else if (grid_type == 2) {
        double first, last;
        lip_first_last(n, &first, &last);
        if (gauss_lobatto_grid(n, first, last, node)) return -1;
    } else {
        return -1;
    }
    for (int i = 0; i < n * n; ++i) c[i] = 0.0;
    for (int j = 0; j < n; ++j) c[j + n * (n - 1)] = 1.0;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            if (i == j) continue;
            const double dx = (grid_type == 1) ? static_cast<double>(j - i) : node[j] - node[i];
            const double xi = node[i];
            for (int k = 0; k + 1 < n; ++k)
                c[j + n * k] = (c[j + n * (k + 1)] - c[j + n * k] * xi) / dx;
            c[j + n * (n - 1)] = -c[j + n * (n - 1)] * xi / dx;
        }
    }
    return 0;
}

// int_0^T exp(-mu u) u^p du for mu >= 0 (the lower incomplete gamma
// function gamma(p+1, mu T) / mu^(p+1)), by its all-positive series up to
// mu T = 40 and by the finite upper-gamma sum beyond.
static long double exp_moment(int p, long double mu, long double T) {
    const long double x = mu * T;
    long double Tp = 1.0L;
    for (int i = 0; i <= p; ++i) Tp *= T;
    if (x == 0.0L) return Tp / (p + 1);
    if (x <= 40.0L) {
        long double term = 1.0L / (p + 1), sum = term;
        for (int m = 1; m < 400; ++m) {
            term *= x / (p + 1 + m);
            sum += term;
            if (term < 1e-22L * sum) break;
        }
        return Tp * std::exp(-x) * sum;
    }
    long double fact = 1.0L, mp = 1.0L, part = 0.0L, xm = 1.0L;
    for (int m = 0; m <= p; ++m) {
        if (m > 0) {
            xm *= x / m;
            fact *= m;
        }
        part += xm;
        mp *= mu;
    }
    return fact / mp * (1.0L - std::exp(-x) * part);
}

// Exponentially weighted inward / outward LIP integrals of one cell in cell
// units (nodes t = 0..n, n = nlip - 1, basis L_r on those nodes):
//   ii[r + nlip*j] = int_j^n exp(-mu (t - j)) L_r(t) dt,  j = 0..n-1
//   io[r + nlip*j] = int_0^(j+1) exp(-mu (j+1 - t)) L_r(t) dt,  j = 0..n-1
// (mu = 0: lip_inout_integrals).  Each basis polynomial is re-expanded about
// the node and integrated term by term against exp_moment, in long double.
int lip_inout_integrals_exp(int nlip, double mu, double *ii, double *io) {
    double c[kMaxNlip * kMaxNlip];
    if (lip_coeffs(nlip, 1, c) || mu < 0.0) return -1;
    double first, last;
    lip_first_last(nlip, &first, &last);
    const int n = nlip - 1;
    for (int r = 0; r < nlip; ++r) {
        for (int j = 0; j <= n; ++j) {
            // Taylor coefficients of L_r about node j (LIP coordinate first + j)
            long double a[kMaxNlip];
            for (int k = 0; k < nlip; ++k) a[k] = c[r + nlip * k];  // decreasing powers
            const long double x0 = first + j;
            long double d[kMaxNlip];
            for (int q = 0; q < nlip; ++q) {  // synthetic division: d[q] = P^(q)(x0)/q!
                long double v = 0.0L;
                for (int k = 0; k < nlip - q; ++k) {
                    v = v * x0 + a[k];
                    a[k] = v;
                }
                d[q] = v;
            }
            if (j < n) {
                long double s = 0.0L;
                for (int q = 0; q < nlip; ++q) s += d[q] * exp_moment(q, mu, (long double)(n - j));
                ii[r + nlip * j] = (double)s;
            }
            if (j > 0) {
                long double s = 0.0L;
                for (int q = 0; q < nlip; ++q)
                    s += ((q % 2) ? -d[q] : d[q]) * exp_moment(q, mu, (long double)j);
                io[r + nlip * (j - 1)] = (double)s;
            }
        }
    }
    return 0;
}

// lip.F90:367-415: antiderivative coefficients, evaluated at last and first.
int lip_integrals(int nlip, int grid_type, double *v) {
    double c[kMaxNlip * kMaxNlip];
    if (lip_coeffs(nlip, grid_type, c)) return -1;
    double first, last;
    lip_first_last(nlip, &first, &last);
    for (int j = 0; j < nlip; ++j) {
        double lo = 0.0, hi = 0.0;
        for (int k = 0; k < nlip; ++k) {
            const double ck = c[j + nlip * k] / (nlip - k);
            lo = (lo + ck) * first;
            hi = (hi + ck) * last;
        }
        v[j] = hi - lo;
    }
    return 0;
}

// lip.F90:414-485: integrals of each basis polynomial from node x_j to the
// cell end (ii, LIPBasis_inward_integrals) and from the cell start to node
// x_(j+1) (io, LIPBasis_outward_integrals), both (nlip, nlip-1)
// column-major, nodes at first + j (equidistant cells).
int lip_inout_integrals(int nlip, double *ii, double *io) {
    double c[kMaxNlip * kMaxNlip];
    if (lip_coeffs(nlip, 1, c)) return -1;
    double first, last;
    lip_first_last(nlip, &first, &last);
    auto antider = [&](int j, double x) {  // Horner over the antiderivative
        double v = 0.0;
        for (int k = 0; k < nlip; ++k) v = (v + c[j + nlip * k] / (nlip - k)) * x;
        return v;
    };
    for (int r = 0; r < nlip; ++r) {
        const double at_last = antider(r, last), at_first = antider(r, first);
        double x = first;
        for (int j = 0; j < nlip - 1; ++j) {
            ii[r + nlip * j] = at_last - antider(r, x);
            x = x + 1;
        }
        x = first;
        for (int j = 0; j < nlip - 1; ++j) {
            x = x + 1.0;
            io[r + nlip * j] = antider(r, x) - at_first;
        }
    }
    return 0;
}

// grid.F90:446-537: points of each cell measured from the cell start.
int make_axis(double qmin, int ncell, int nlip, const double *steps,
              int grid_type, Axis *ax) {
    if (ncell < 1 || nlip < 2 || nlip > kMaxNlip) return -1;
    double templ[kMaxNlip];
    if (grid_type == 1) {
        for (int k = 0; k < nlip; ++k) templ[k] = static_cast<double>(k);
    } else if (grid_type == 2) {
        if (gauss_lobatto_grid(nlip, 0.0, static_cast<double>(nlip - 1), templ)) return -1;
    } else {
        return -1;
    }
    ax->qmin = qmin;
    ax->ncell = ncell;
    ax->nlip = nlip;
    ax->grid_type = grid_type;
    ax->steps.assign(steps, steps + ncell);
    ax->points.assign(ax->n(), 0.0);
    ax->celld.assign(ncell, 0.0);
    int ip = 0;
    ax->points[0] = qmin;
    for (int c = 0; c < ncell; ++c) {
        ax->celld[c] = ax->points[ip];
        for (int k = 1; k < nlip; ++k) ax->points[++ip] = ax->celld[c] + templ[k] * steps[c];
    }
    return 0;
}

}  // namespace dage
