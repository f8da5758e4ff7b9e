/*
 * dage_oracle.c -- CPU FP64 restatement of DAGE's tensor-quadrature
 * Coulomb3D/Helmholtz3D operator (the hot path's oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP library; the
 * product never links or calls it.  See dage_oracle.h for the reference
 * routine each function restates.  The formulas are kept term for term in the
 * reference's order (including its cancellation-prone closed forms) because
 * the GPU library must reproduce the reference's numbers, not better ones.
 *
 * Built with -ffp-contract=off so that every product and sum rounds as
 * written.
 */
#include "dage_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* globals.F90:54-61: PI=dacos(-1), SQRTPI=dsqrt(PI), TWOOVERSQRTPI=2/SQRTPI */
static double o_pi(void) { return acos(-1.0); }
static double o_sqrtpi(void) { return sqrt(o_pi()); }

double dage_oracle_pi(void) { return o_pi(); }

/* ---------------------------------------------------------------- quadrature
 * gauss.F90:337-384.  Roots of P_n by Newton from the cos() guess; the weight
 * uses the derivative pp evaluated before the last Newton update, exactly as
 * the reference loop leaves it. */
void dage_oracle_gauleg(int ngp, double *xabsc, double *weig) {
    const double EPS = 3.0e-15;
    const double pi = o_pi();
    int m = (ngp + 1) / 2;
    for (int i = 1; i <= m; ++i) {
        double z = cos(pi * (i - 0.25) / (ngp + 0.5));
        double z1, pp, p1, p2, p3;
        for (;;) {
            p1 = 1.0;
            p2 = 0.0;
            for (int j = 1; j <= ngp; ++j) {
                p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = ngp * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
            if (!(fabs(z - z1) > EPS)) break;
        }
        xabsc[i - 1] = -z;
        xabsc[ngp - i] = z;
        weig[i - 1] = 2.0 / ((1.0 - z * z) * pp * pp);
        weig[ngp - i] = weig[i - 1];
    }
}

/* gauss.F90:307-329 (gauleg_int), linear or logarithmic map of [-1,1]. */
static void gauleg_int(int n, double *tp, double *tw, double a, double b,
                       int log_p) {
    dage_oracle_gauleg(n, tp, tw);
    for (int i = 0; i < n; ++i) {
        if (log_p) {
            tp[i] = a * pow(b / a, 0.5 * (tp[i] + 1.0));
            tw[i] = 0.5 * log(b / a) * tw[i] * tp[i];
        } else {
            tp[i] = 0.5 * ((b + a) + (b - a) * tp[i]);
            tw[i] = 0.5 * (b - a) * tw[i];
        }
    }
}

/* gauss.F90:81-122 */
void dage_oracle_gauss_quad(int nlin, int nlog, double tstart, double tlin,
                            double tlog, double *tp, double *tw) {
    gauleg_int(nlin, tp, tw, tstart, tlin, 0);
    gauleg_int(nlog, tp + nlin, tw + nlin, tlin, tlog, 1);
}

/* gauss.F90:409-446 */
int dage_oracle_gauss_lobatto_grid(int n, double left, double right,
                                   double *tp) {
    if (n == 6) {
        tp[0] = -1.0;
        tp[1] = -sqrt((7.0 + 2 * sqrt(7.0)) / 21.0);
        tp[2] = -sqrt((7.0 - 2 * sqrt(7.0)) / 21.0);
        tp[3] = sqrt((7.0 - 2 * sqrt(7.0)) / 21.0);
        tp[4] = sqrt((7.0 + 2 * sqrt(7.0)) / 21.0);
        tp[5] = 1.0;
    } else if (n == 7) {
        tp[0] = -1.0;
        tp[1] = -sqrt((15.0 + 2 * sqrt(15.0)) / 33.0);
        tp[2] = -sqrt((15.0 - 2 * sqrt(15.0)) / 33.0);
        tp[3] = 0.0;
        tp[4] = sqrt((15.0 - 2 * sqrt(15.0)) / 33.0);
        tp[5] = sqrt((15.0 + 2 * sqrt(15.0)) / 33.0);
        tp[6] = 1.0;
    } else {
        return -1;
    }
    for (int i = 0; i < n; ++i)
        tp[i] = 0.5 * ((right + left) + (right - left) * tp[i]);
    return 0;
}

/* ------------------------------------------------------------------ LIP basis
 * lip.F90:118-142: first = -(nlip-1)/2, last = nlip/2 in integer arithmetic
 * for the equidistant (1) and Gauss-Lobatto (2) cells. */
void dage_oracle_lip_first_last(int nlip, int grid_type, double *first,
                                double *last) {
    (void)grid_type;
    *first = (double)(-((nlip - 1) / 2));
    *last = (double)(nlip / 2);
}

/* lip.F90:285-344: product form of the Lagrange polynomials, built up one
 * node at a time in decreasing-power coefficient order. */
int dage_oracle_lip_coeffs(int nlip, int grid_type, double *c) {
    int n = nlip;
    double nodes[16];
    if (n < 2 || n > 15) return -1;
    if (grid_type == 1) {
        for (int i = 1; i <= n; ++i) nodes[i - 1] = (double)(i - (n + 1) / 2);
    } else if (grid_type == 2) {
        double first, last;
        dage_oracle_lip_first_last(n, grid_type, &first, &last);
        if (dage_oracle_gauss_lobatto_grid(n, first, last, nodes)) return -1;
    } else {
        return -1;
    }
    memset(c, 0, sizeof(double) * n * n);
    for (int j = 0; j < n; ++j) c[j + n * (n - 1)] = 1.0;
    for (int i = 1; i <= n; ++i) {
        double x_i = nodes[i - 1];
        for (int j = 1; j <= n; ++j) {
            if (i == j) continue;
            double dx = (grid_type == 1) ? (double)(j - i)
                                         : nodes[j - 1] - nodes[i - 1];
            double *row = c + (j - 1);
            /* p(j,1:n-1) = (p(j,2:n) - p(j,1:n-1)*x_i)/dx, old values on the
             * right; ascending k reads p(k+1) before it is overwritten. */
            for (int k = 0; k < n - 1; ++k)
                row[n * k] = (row[n * (k + 1)] - row[n * k] * x_i) / dx;
            row[n * (n - 1)] = -row[n * (n - 1)] * x_i / dx;
        }
    }
    return 0;
}

/* lip.F90:367-415: integral coefficients c(:,k)/(n-k+1), then Horner at
 * last minus Horner at first. */
int dage_oracle_lip_integrals(int nlip, int grid_type, double *integrals) {
    int n = nlip;
    double c[256], ic[256];
    double first, last;
    if (dage_oracle_lip_coeffs(n, grid_type, c)) return -1;
    dage_oracle_lip_first_last(n, grid_type, &first, &last);
    for (int k = 0; k < n; ++k)          /* 0-based k <-> Fortran column k+1 */
        for (int j = 0; j < n; ++j) ic[j + n * k] = c[j + n * k] / (n - k);
    for (int j = 0; j < n; ++j) {
        double t1 = 0.0, t2 = 0.0;
        for (int k = 0; k < n; ++k) t1 = (t1 + ic[j + n * k]) * first;
        for (int k = 0; k < n; ++k) t2 = (t2 + ic[j + n * k]) * last;
        integrals[j] = t2 - t1;
    }
    return 0;
}

/* ----------------------------------------------------------------------- grid
 * grid.F90:446-537: points accumulate from the start of each cell. */
int dage_oracle_grid1d(double qmin, int ncell, int nlip, const double *steps,
                       int grid_type, double *points, double *celld) {
    double templ[16];
    if (nlip < 2 || nlip > 15) return -1;
    if (grid_type == 1) {
        for (int k = 1; k <= nlip - 1; ++k) templ[k] = (double)k;
    } else if (grid_type == 2) {
        double gl[16];
        if (dage_oracle_gauss_lobatto_grid(nlip, 0.0, (double)(nlip - 1), gl))
            return -1;
        for (int k = 1; k <= nlip - 1; ++k) templ[k] = gl[k];
    } else {
        return -1;
    }
    int ip = 0;
    points[0] = qmin;
    for (int ic = 0; ic < ncell; ++ic) {
        celld[ic] = points[ip];
        for (int k = 1; k <= nlip - 1; ++k) {
            ++ip;
            points[ip] = celld[ic] + templ[k] * steps[ic];
        }
    }
    return 0;
}

/* grid.F90:539-567 */
void dage_oracle_equidistant_cells(double r0, double r1, int nlip,
                                   double stepmax, int *ncell, double *step) {
    int nc = (int)ceil((r1 - r0) / (stepmax * (nlip - 1)));
    *ncell = nc;
    *step = (r1 - r0) / (nc * (nlip - 1));
}

/* grid.F90:1398-1424 */
int dage_oracle_grid1d_ints(int ncell, int nlip, const double *steps,
                            int grid_type, double *ints) {
    double base[16];
    if (dage_oracle_lip_integrals(nlip, grid_type, base)) return -1;
    int n = ncell * (nlip - 1) + 1;
    for (int i = 0; i < n; ++i) ints[i] = 0.0;
    for (int ic = 0; ic < ncell; ++ic)
        for (int j = 0; j < nlip; ++j)
            ints[ic * (nlip - 1) + j] += base[j] * steps[ic];
    return 0;
}

/* ------------------------------------------------------- exponential integrals
 * exponentials.F90:695-1297.  Closed forms of int_a^b x^n e^{-t^2(x-p)^2} dx,
 * restated term by term.  Ea/Eb stand for dexp(-t**2*(a-p)**2) and
 * dexp(-t**2*(b-p)**2); ErA/ErB for derf(a*t-t*p) and derf(b*t-t*p). */
static double series4(double q) {
    return q + q * q / 2.0 + q * q * q / 6.0 + q * q * q * q / 24.0;
}

static double expo_0(double a, double b, double t, double p) {
    const double thr = 1.0e-6;
    if (t == 0.0) return b - a;
    if (t < thr) return b - a;
    return o_sqrtpi() * (erf(b * t - t * p) - erf(a * t - t * p)) / t / 2.0;
}

static double expo_1(double a, double b, double t, double p) {
    const double thr = 1.0e-6;
    const double SP = o_sqrtpi();
    if (t == 0.0 || t < thr) return (b * b - a * a) / 2.0;
    double v1 = p * SP * erf(t * (b - p));
    double v2;
    if (t > 1.0e-3) {
        v2 = (exp(-(t * t) * ((a - p) * (a - p))) -
              exp(-(t * t) * ((b - p) * (b - p)))) / t;
    } else {
        double q = -(t * t) * ((a - p) * (a - p));
        double r = -(t * t) * ((b - p) * (b - p));
        v2 = (q + q * q / 2.0 + q * q * q / 6.0 + q * q * q * q / 24.0 - r -
              r * r / 2.0 - r * r * r / 6.0 - r * r * r * r / 24.0) / t;
    }
    double v3 = -(p * SP * erf(a * t - t * p));
    return (v1 + v2 + v3) / t / 2.0;
}

static double expo_2(double a, double b, double t, double p) {
    const double thr = 0.5e-4;
    const double SP = o_sqrtpi();
    if (t == 0.0 || t < thr) return (b * b * b - a * a * a) / 3.0;
    double v1 = p * p * SP * erf(b * t - t * p) / 2.0;
    double v2, v3;
    if (t > 1.0e-3) {
        double Ea = exp(-(t * t) * ((a - p) * (a - p)));
        double Eb = exp(-(t * t) * ((b - p) * (b - p)));
        v2 = (Ea * a - Eb * b);
        v3 = (Ea - Eb) * p / t / 2.0;
    } else {
        double q = -(t * t) * ((a - p) * (a - p));
        double r = -(t * t) * ((b - p) * (b - p));
        v2 = series4(q) * a - series4(r) * b;
        v2 = (v2 + a - b);
        v3 = series4(q) - series4(r);
        v3 = v3 * p / t / 2.0;
    }
    double v5 = SP * (erf(t * (b - p)) - erf(t * (a - p))) / t / 2.0;
    v2 = (v2 + v5) / t / 2.0;
    v5 = 0.0;
    double v4 = -(p * p * SP * erf(a * t - t * p) / 2.0);
    return (v1 + v2 + v3 + v4 + v5) / t;
}

static double expo_3(double a, double b, double t, double p) {
    const double thr = 1.0e-4;
    const double SP = o_sqrtpi();
    if (t == 0.0 || t < thr) return (b * b * b * b - a * a * a * a) / 4.0;
    double v1, v2, v3, v5;
    if (t > 1.0e-3) {
        double Ea = exp(-(t * t) * ((a - p) * (a - p)));
        double Eb = exp(-(t * t) * ((b - p) * (b - p)));
        v1 = 2.0 * (Ea * (a * a) - Eb * (b * b)) / (t * t);
        v2 = 2.0 * p * (a * Ea - b * Eb) * t;
        v3 = 2.0 * (Ea * (p * p) - Eb * (p * p)) / (t * t);
        v5 = 2.0 * (Ea - Eb) / t;
    } else {
        double q = -(t * t) * ((a - p) * (a - p));
        double r = -(t * t) * ((b - p) * (b - p));
        v1 = (1.0 + series4(q)) * a * a - (1.0 + series4(r)) * b * b;
        v1 = 2.0 * v1 / t / t;
        v2 = series4(q) * a - series4(r) * b;
        v2 = v2 + a - b;
        v2 = 2.0 * v2 * p * t;
        v3 = series4(q) - series4(r);
        v3 = 2.0 * v3 * p * p / t / t;
        v5 = series4(q) - series4(r);
        v5 = 2.0 * v5 / t;
    }
    double dE = erf(b * t - t * p) - erf(a * t - t * p);
    double v4 = 2.0 * (p * p * p) * SP * dE / t;
    double v6 = 3.0 * p * SP * dE;
    v6 = (v6 + v2 + v5) / t / t / t;
    v2 = 0.0;
    v5 = 0.0;
    return 0.25 * (v1 + v2 + v3 + v4 + v5 + v6);
}

static double expo_4(double a, double b, double t, double p) {
    const double thr = 1.0e-3;
    const double SP = o_sqrtpi();
    if (t == 0.0 || t < thr) return (pow(b, 5) - pow(a, 5)) / 5.0;
    double v1, v2, v3, v4, v7, v8;
    double p2 = p * p, p3 = p * p * p;
    if (t > 1.0e-3) {
        double Ea = exp(-(t * t) * ((a - p) * (a - p)));
        double Eb = exp(-(t * t) * ((b - p) * (b - p)));
        v1 = 4.0 * (Ea * (a * a * a) - Eb * (b * b * b));
        v2 = 4.0 * (Ea * p * (a * a) - Eb * p * (b * b));
        v3 = 4.0 * (Ea * p2 * a - Eb * p2 * b);
        v4 = 4.0 * (Ea * p3 - Eb * p3);
        v7 = 10.0 * p * (Ea - Eb) / t / t;
        v8 = 6.0 * (Ea * a - Eb * b);
    } else {
        double q = -(t * t) * ((a - p) * (a - p));
        double r = -(t * t) * ((b - p) * (b - p));
        double Sq = 1.0 + series4(q), Sr = 1.0 + series4(r);
        v1 = 4.0 * (Sq * (a * a * a) - Sr * (b * b * b));
        v2 = 4.0 * (Sq * p * a * a - Sr * p * b * b);
        v3 = 4.0 * (Sq * p * p * a - Sr * p * p * b);
        v4 = 4.0 * (series4(q) - series4(r)) * p3;
        v7 = 10.0 * (series4(q) - series4(r)) * p / t / t;
        v8 = 6.0 * (Sq * a - Sr * b);
    }
    double dE = erf(a * t - t * p) - erf(b * t - t * p);
    double v5 = -(4.0 * t * (p2 * p2) * SP * dE);
    double v6 = -(12.0 / t * p2 * SP * dE);
    double v9 = -(3.0 * SP * dE / t);
    v2 = v2 / (t * t);
    v5 = (v4 + v5) / (t * t);
    v7 = (v3 + v6 + v7) / (t * t);
    v8 = ((v8 + v9) / t / t + v1) / (t * t);
    v1 = v3 = v4 = v6 = v9 = 0.0;
    return 0.125 * (v1 + v2 + v3 + v4 + v5 + v6 + v7 + v8 + v9);
}

static double expo_5(double a, double b, double t, double p) {
    const double thr = 2.5e-3;
    const double SP = o_sqrtpi();
    if (t == 0.0 || t < thr) return (pow(b, 6) - pow(a, 6)) / 6.0;
    double v1, v2, v3, v4, v5, v8, v9, v11, v12;
    double p2 = p * p, p3 = p * p * p, p4 = p2 * p2, p5 = p4 * p;
    double a2 = a * a, a3 = a * a * a, a4 = a2 * a2;
    double b2 = b * b, b3 = b * b * b, b4 = b2 * b2;
    double t2 = t * t, t4 = t2 * t2;
    if (t > 1.0e-2) {
        double Ea = exp(-(t * t) * ((a - p) * (a - p)));
        double Eb = exp(-(t * t) * ((b - p) * (b - p)));
        v1 = 4.0 * (Ea * a4 - Eb * b4);
        v2 = 4.0 * (Ea * p * a3 - Eb * p * b3);
        v3 = 4.0 * (Ea * p2 * a2 - Eb * p2 * b2);
        v4 = 4.0 * (Ea * p3 * a - Eb * p3 * b);
        v5 = 4.0 * (Ea * p4 - Eb * p4);
        v8 = 18.0 * (Ea - Eb) * p2 / t2;
        v9 = 14.0 * (Ea * p * a - Eb * p * b) / t2;
        v11 = 8.0 * (Ea * a2 - Eb * b2) / t2;
        v12 = 8.0 * (Ea - Eb) / t4;
    } else {
        double q = -(t * t) * ((a - p) * (a - p));
        double r = -(t * t) * ((b - p) * (b - p));
        double Sq = 1.0 + series4(q), Sr = 1.0 + series4(r);
        v1 = 4.0 * (Sq * a4 - Sr * b4);
        v2 = 4.0 * (Sq * p * a3 - Sr * p * b3);
        v3 = 4.0 * (Sq * p2 * a2 - Sr * p2 * b2);
        v4 = 4.0 * (Sq * p3 * a - Sr * p3 * b);
        v5 = 4.0 * ((series4(q)) * p4 - (series4(r)) * p4);
        v8 = 18.0 * (((series4(q)) - (series4(r))) * p2 / t2);
        v9 = 14.0 * ((Sq * a * p - Sr * b * p) / t2);
        v11 = 8.0 * ((Sq * a * a - Sr * b * b) / t2);
        v12 = 8.0 * (((series4(q)) - (series4(r))) / t4);
    }
    double dE = erf(a * t - t * p) - erf(b * t - t * p);
    double v6 = -(4.0 * t * p5 * SP * dE);
    double v7 = -(20.0 / t * p3 * SP * dE);
    double v10 = -(15.0 * p * SP * dE / (t2 * t));
    double value = v1 + v2 + v3 + v4 + v5 + v6;
    value = value + v7 + v8 + v9 + v10 + v11 + v12;
    value = value / t2;
    return 0.125 * value;
}

static double expo_6(double a, double b, double t, double p) {
    const double thr = 5.0e-3;
    const double SP = o_sqrtpi();
    if (t == 0.0 || t < thr) return (pow(b, 7) - pow(a, 7)) / 7.0;
    double p2 = p * p, p3 = p * p * p, p4 = p2 * p2, p5 = p4 * p,
           p6 = p3 * p3;
    double a2 = a * a, a3 = a * a * a, a4 = a2 * a2, a5 = a4 * a;
    double b2 = b * b, b3 = b * b * b, b4 = b2 * b2, b5 = b4 * b;
    double Ea = exp(-(t * t) * ((a - p) * (a - p)));
    double Eb = exp(-(t * t) * ((b - p) * (b - p)));
    double dE = erf(b * t - t * p) - erf(a * t - t * p);
    double v1 = 30.0 * (Ea * a - Eb * b);
    double v2 = 15.0 * SP * dE;
    v1 = (v1 + v2 / t) / t / t / t;
    v2 = 0.0;
    double v3 = 90.0 * p2 * SP * dE / t / t;
    double v4 = 56.0 * (Ea - Eb) * p3 / t;
    double v5 = 66.0 * (Ea - Eb) * p / t / t / t;
    double v6 = 8.0 * (Ea - Eb) * p5 * t;
    double v7 = 20.0 * (Ea * a3 - Eb * b3) / t;
    double v8 = 36.0 * (Ea * a2 - Eb * b2) * p / t;
    double v9 = 48.0 * (Ea * a - Eb * b) * p2 / t;
    double v10 = 8.0 * p6 * SP * dE * (t * t);
    double v11 = 8.0 * (Ea * a4 - Eb * b4) * p * t;
    double v12 = 8.0 * (Ea * a3 - Eb * b3) * p2 * t;
    double v13 = 8.0 * (Ea * a5 - Eb * b5) * t;
    double v14 = 8.0 * (Ea * a - Eb * b) * p4 * t;
    double v15 = 8.0 * (Ea * a2 - Eb * b2) * p3 * t;
    double v16 = 60.0 * p4 * SP * dE;
    double value = v1 + v2 + v3 + v4 + v5 + v6 + v7 + v8;
    value = value + v9 + v10 + v11 + v12 + v13 + v14 + v15 + v16;
    value = value / (t * t * t);
    return 0.0625 * value;
}

double dage_oracle_expo(int n, double a, double b, double t, double p) {
    switch (n) {
    case 0: return expo_0(a, b, t, p);
    case 1: return expo_1(a, b, t, p);
    case 2: return expo_2(a, b, t, p);
    case 3: return expo_3(a, b, t, p);
    case 4: return expo_4(a, b, t, p);
    case 5: return expo_5(a, b, t, p);
    case 6: return expo_6(a, b, t, p);
    default: return 1.23456789e10; /* expo_7, exponentials.F90:1299-1318 */
    }
}

/* exponentials.F90:2287-2429: 4 equal sub-intervals x 12-point
 * Gauss-Legendre with the reference's 15-digit node/weight literals. */
double dage_oracle_gauss_expo(int n, double a, double b, double t, double p) {
    static const double wk[12] = {
        0.047175336386512, 0.106939325995318, 0.160078328543346,
        0.203167426723066, 0.233492536538355, 0.249147045813403,
        0.249147045813403, 0.233492536538355, 0.203167426723066,
        0.160078328543346, 0.106939325995318, 0.047175336386512};
    static const double tk[12] = {
        -0.981560634246719, -0.904117256370475, -0.769902674194305,
        -0.587317954286617, -0.367831498998180, -0.125233408511469,
        0.125233408511469,  0.367831498998180,  0.587317954286617,
        0.769902674194305,  0.904117256370475,  0.981560634246719};
    if (t == 0.0) {
        if (n == 0) return b - a;
        return (pow(b, n + 1) - pow(a, n + 1)) / (n + 1);
    }
    const int interval = 4;
    double limit[5];
    for (int i = 1; i <= interval + 1; ++i)
        limit[i - 1] = ((interval + 1 - i) * a + (i - 1) * b) / interval;
    double summa = 0.0;
    for (int i = 1; i <= interval; ++i) {
        double a0 = limit[i - 1], a1 = limit[i];
        double pr = 1.0 - a1 * 2.0 / (a1 - a0);
        double qr = 2.0 / (a1 - a0);
        for (int j = 0; j < 12; ++j) {
            double tp = (tk[j] - pr) / qr;
            double tw = wk[j] / qr;
            double e = exp(-(t * t) * (tp - p) * (tp - p));
            if (n > 0) {
                double tpn = 1.0;
                for (int k = 0; k < n; ++k) tpn *= tp;
                summa = summa + tpn * e * tw;
            } else {
                summa = summa + e * tw;
            }
        }
    }
    return summa;
}

/* exponentials.F90:647-666 */
void dage_oracle_integrals(double a, double b, double t, double p, int n,
                           double *values) {
    for (int i = 0; i < n; ++i)
        values[n - 1 - i] = (t < 0.1) ? dage_oracle_gauss_expo(i, a, b, t, p)
                                      : dage_oracle_expo(i, a, b, t, p);
}

/* exponentials.F90:178-346 for an empty PBC string: one image (idisp = 0),
 * pst0 = x2cell(x2, icell) (grid.F90:1768-1776), LIP combination of the
 * monomial integrals times the cell scale, accumulated cell by cell. */
int dage_oracle_exponentials(double qmin_in, int ncell_in,
                             const double *steps_in, int nlip, int grid_type,
                             const double *points_out, int nout,
                             const double *tp, int ntp, int transpose,
                             double *out) {
    int nin = ncell_in * (nlip - 1) + 1;
    double *pin = (double *)malloc(sizeof(double) * nin);
    double *celld = (double *)malloc(sizeof(double) * ncell_in);
    double lipc[256], vals[16], temp[16];
    double first, last;
    if (!pin || !celld) { free(pin); free(celld); return -1; }
    if (dage_oracle_grid1d(qmin_in, ncell_in, nlip, steps_in, grid_type, pin,
                           celld) ||
        dage_oracle_lip_coeffs(nlip, grid_type, lipc)) {
        free(pin); free(celld);
        return -1;
    }
    dage_oracle_lip_first_last(nlip, grid_type, &first, &last);
    memset(out, 0, sizeof(double) * (size_t)nin * nout * ntp);
    for (int it = 0; it < ntp; ++it) {
        double t = tp[it];
        for (int ic = 0; ic < ncell_in; ++ic) {
            double h = steps_in[ic];
            double shft = t * h;
            int fg = ic * (nlip - 1);
            for (int ix = 0; ix < nout; ++ix) {
                double pst0 = (points_out[ix] - celld[ic]) / h + first;
                dage_oracle_integrals(first, last, shft, pst0, nlip, vals);
                for (int j = 0; j < nlip; ++j) {
                    double s = 0.0;
                    for (int k = 0; k < nlip; ++k) s += lipc[j + nlip * k] * vals[k];
                    temp[j] = s * h;
                }
                for (int j = 0; j < nlip; ++j) {
                    size_t idx = transpose
                        ? (size_t)(fg + j) + (size_t)nin * (ix + (size_t)nout * it)
                        : (size_t)ix + (size_t)nout * ((fg + j) + (size_t)nin * it);
                    out[idx] += temp[j];
                }
            }
        }
    }
    free(pin);
    free(celld);
    return 0;
}

/* -------------------------------------------------------------- tensor apply
 * function3d.F90:4301-4314: for each t, per input z-slice
 *   tmp1 = Fx_t . in(:,:,iz);  tmp_u(:,:,iz) = tmp1 . Fy_t
 * then per output y-row  out(:,iy,:) += w_t * tmp_u(:,iy,:) . Fz_t,
 * followed by out += coda*in (:4327-4332). */
int dage_oracle_transform_cube(const int di[3], const int dout[3], int ntp,
                               const double *fx, const double *fy,
                               const double *fz, const double *w, double coda,
                               const double *in, double *out, int t0, int t1,
                               int add_coda) {
    const size_t nxi = di[0], nyi = di[1], nzi = di[2];
    const size_t nxo = dout[0], nyo = dout[1], nzo = dout[2];
    if (add_coda && coda != 0.0 && (nxi != nxo || nyi != nyo || nzi != nzo))
        return -2; /* cube_project onto a different grid: not restated */
    double *tmpu = (double *)malloc(sizeof(double) * nxo * nyo * nzi);
    if (!tmpu) return -1;
    int fail = 0;
    memset(out, 0, sizeof(double) * nxo * nyo * nzo);
    for (int ip = t0; ip < t1 && ip < ntp; ++ip) {
        const double *Fx = fx + nxo * nxi * ip;
        const double *Fy = fy + nyi * nyo * ip;
        const double *Fz = fz + nzi * nzo * ip;
        /* The z-slices (and below the output y-lines) are independent: with
         * OpenMP each thread takes whole slices with its own temporaries, and
         * every element is summed in the same order as the serial loop. */
#pragma omp parallel reduction(| : fail)
        {
            double *tmp1 = (double *)malloc(sizeof(double) * nxo * nyi);
            if (!tmp1) fail = 1;
#pragma omp for schedule(static)
            for (size_t iz = 0; iz < nzi; ++iz) {
                if (!tmp1) continue;
                const double *slice = in + nxi * nyi * iz;
                memset(tmp1, 0, sizeof(double) * nxo * nyi);
                for (size_t yi = 0; yi < nyi; ++yi)
                    for (size_t xi = 0; xi < nxi; ++xi) {
                        double s = slice[xi + nxi * yi];
                        const double *col = Fx + nxo * xi;
                        double *dst = tmp1 + nxo * yi;
                        for (size_t xo = 0; xo < nxo; ++xo) dst[xo] += col[xo] * s;
                    }
                double *tu = tmpu + nxo * nyo * iz;
                memset(tu, 0, sizeof(double) * nxo * nyo);
                for (size_t yo = 0; yo < nyo; ++yo)
                    for (size_t yi = 0; yi < nyi; ++yi) {
                        double f = Fy[yi + nyi * yo];
                        const double *src = tmp1 + nxo * yi;
                        double *dst = tu + nxo * yo;
                        for (size_t xo = 0; xo < nxo; ++xo) dst[xo] += src[xo] * f;
                    }
            }
            free(tmp1);
        }
        if (fail) break;
#pragma omp parallel reduction(| : fail)
        {
            double *m = (double *)malloc(sizeof(double) * nxo * nzo);
            if (!m) fail = 1;
#pragma omp for schedule(static)
            for (size_t iy = 0; iy < nyo; ++iy) {
                if (!m) continue;
                memset(m, 0, sizeof(double) * nxo * nzo);
                for (size_t zo = 0; zo < nzo; ++zo)
                    for (size_t zi = 0; zi < nzi; ++zi) {
                        double f = Fz[zi + nzi * zo];
                        const double *src = tmpu + nxo * (iy + nyo * zi);
                        double *dst = m + nxo * zo;
                        for (size_t xo = 0; xo < nxo; ++xo) dst[xo] += src[xo] * f;
                    }
                for (size_t zo = 0; zo < nzo; ++zo) {
                    double *o = out + nxo * (iy + nyo * zo);
                    const double *src = m + nxo * zo;
                    for (size_t xo = 0; xo < nxo; ++xo) o[xo] = o[xo] + w[ip] * src[xo];
                }
            }
            free(m);
        }
        if (fail) break;
    }
    free(tmpu);
    if (fail) return -1;
    if (add_coda && coda != 0.0) {
        size_t n = nxo * nyo * nzo;
        for (size_t i = 0; i < n; ++i) out[i] = out[i] + coda * in[i];
    }
    return 0;
}

/* helmholtz3d.F90:411-444, coulomb.F90:150 */
void dage_oracle_helmholtz_weights(const double *tp, const double *tw, int ntp,
                                   double energy, double *w) {
    const double two_over_sqrtpi = 2.0 / o_sqrtpi();
    for (int i = 0; i < ntp; ++i) {
        double wi = tw[i];
        if (energy < -0.005) wi = wi * exp(2.0 * energy / (4.0 * (tp[i] * tp[i])));
        w[i] = two_over_sqrtpi * wi;
    }
}

/* gauss_generator.F90:291-318 */
void dage_oracle_gauss_density(int ngau, const double *q, const double *alpha,
                               const double *c, const double *x, int nx,
                               const double *y, int ny, const double *z,
                               int nz, double *cube) {
    const double pi = o_pi();
    double *ex = (double *)malloc(sizeof(double) * (nx + ny + nz));
    double *ey = ex + nx, *ez = ey + ny;
    memset(cube, 0, sizeof(double) * (size_t)nx * ny * nz);
    for (int p = 0; p < ngau; ++p) {
        double norm = pow(alpha[p] / pi, 1.5);
        for (int i = 0; i < nx; ++i) ex[i] = exp(-alpha[p] * ((x[i] - c[3 * p]) * (x[i] - c[3 * p])));
        for (int j = 0; j < ny; ++j) ey[j] = exp(-alpha[p] * ((y[j] - c[3 * p + 1]) * (y[j] - c[3 * p + 1])));
        for (int k = 0; k < nz; ++k) ez[k] = exp(-alpha[p] * ((z[k] - c[3 * p + 2]) * (z[k] - c[3 * p + 2])));
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    size_t id = i + (size_t)nx * (j + (size_t)ny * k);
                    cube[id] = cube[id] + q[p] * norm * ex[i] * ey[j] * ez[k];
                }
    }
    free(ex);
}

/* gauss_generator.F90:384-429 (gausspot: erf(sqrt(a) r)/r with a series
 * near r = 0 and 1/r far away). */
static double gausspot(double alpha, double r) {
    double sa = sqrt(alpha);
    double temp = r * sa;
    if (temp < 0.01)
        return 2.0 / o_sqrtpi() * sa *
               (1.0 - 0.33333333333333333 * (temp * temp) + 0.1 * (temp * temp * temp * temp));
    if (temp < 10.0) return erf(r * sa) / r;
    return 1.0 / r;
}

void dage_oracle_gauss_potential(int ngau, const double *q,
                                 const double *alpha, const double *c,
                                 const double *x, int nx, const double *y,
                                 int ny, const double *z, int nz,
                                 double *cube) {
    memset(cube, 0, sizeof(double) * (size_t)nx * ny * nz);
    for (int p = 0; p < ngau; ++p)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    double dx = c[3 * p] - x[i], dy = c[3 * p + 1] - y[j],
                           dz = c[3 * p + 2] - z[k];
                    double r = sqrt(dx * dx + dy * dy + dz * dz);
                    size_t id = i + (size_t)nx * (j + (size_t)ny * k);
                    cube[id] = cube[id] + q[p] * gausspot(alpha[p], r);
                }
}

double dage_oracle_cube_integral(const double *wx, int nx, const double *wy,
                                 int ny, const double *wz, int nz,
                                 const double *f, const double *g) {
    double total = 0.0;
    for (int k = 0; k < nz; ++k) {
        double sk = 0.0;
        for (int j = 0; j < ny; ++j) {
            double sj = 0.0;
            for (int i = 0; i < nx; ++i) {
                size_t id = i + (size_t)nx * (j + (size_t)ny * k);
                double v = g ? f[id] * g[id] : f[id];
                sj += wx[i] * v;
            }
            sk += wy[j] * sj;
        }
        total += wz[k] * sk;
    }
    return total;
}

/* ------------------------------------------------------------------- bubbles
 * grid.F90:569-603 (Grid1D_init_radial) */
void dage_oracle_radial_steps(double z, int ncell, int nlip, double r_max,
                              double *steps) {
    double x0 = 0.0;
    const double cst = 8.0 * pow(z, -1.5);
    const double dx = r_max / ncell;
    for (int i = 1; i <= ncell; ++i) {
        double x1 = cst / ((r_max + cst) / (i * dx) - 1.0);
        steps[i - 1] = (x1 - x0) / (nlip - 1);
        x0 = x1;
    }
}

/* grid.F90:1212-1246 */
int dage_oracle_get_icell(const double *celld, int ncell, double qmin,
                          double qmax, double x) {
    if (x > qmax || x < qmin) return 0;
    if (celld[ncell - 1] <= x) return ncell;
    int icell = ncell / 2, maxi = ncell - 1, mini = 1;
    for (;;) {
        if (x >= celld[icell]) {          /* celld(icell+1), 1-based */
            if (x < celld[icell + 1]) return icell + 1;
            mini = icell + 2;
            icell = (mini + maxi) / 2;
        } else {
            if (x >= celld[icell - 1]) return icell;
            maxi = icell - 1;
            icell = (mini + maxi) / 2;
        }
    }
}

/* harmonics.F90:586-659: S_00 = 1, S_1 = (y, z, x), then for l >= 2
 *   S_l,-l = sqrt((2l-1)/2l) (y S_l-1,l-1 + x S_l-1,-(l-1))
 *   S_l,m  = (2l-1)/sqrt((l+m)(l-m)) z S_l-1,m
 *            - sqrt((l+m-1)(l-m-1)/((l+m)(l-m))) r^2 S_l-2,m   (|m| < l-1)
 *   S_l,l  = sqrt((2l-1)/2l) (x S_l-1,l-1 - y S_l-1,-(l-1))
 * evaluated on values (the reference builds the same polynomials). */
void dage_oracle_solid_harmonics(int lmax, double x, double y, double z,
                                 double *S) {
    S[0] = 1.0;
    if (lmax < 1) return;
    S[1] = y;
    S[2] = z;
    S[3] = x;
    const double r2 = x * x + y * y + z * z;
    for (int l = 2; l <= lmax; ++l) {
        const int c = l * l + l, p = (l - 1) * (l - 1) + (l - 1), q = (l - 2) * (l - 2) + (l - 2);
        const double f = sqrt((2 * l - 1.0) / (2.0 * l));
        S[c - l] = f * (y * S[p + (l - 1)] + x * S[p - (l - 1)]);
        for (int m = -(l - 1); m <= l - 1; ++m) {
            double v = (2 * l - 1.0) / sqrt(1.0 * (l + m) * (l - m)) * (z * S[p + m]);
            if (l - abs(m) > 1)
                v = v - sqrt((l + m - 1.0) * (l - m - 1.0) / ((l + m) * (l - m))) * (r2 * S[q + m]);
            S[c + m] = v;
        }
        S[c + l] = f * (x * S[p + (l - 1)] - y * S[p - (l - 1)]);
    }
}

int dage_oracle_bubbles_eval_grid(int nbub, const double *centers, int lmax,
                                  int k, int lmin, const int *ncell, int nlip,
                                  const double *const *steps,
                                  const double *const *bf, const double *x,
                                  int nx, const double *y, int ny,
                                  const double *z, int nz, double *cube) {
    const int nlm = (lmax + 1) * (lmax + 1);
    double lipc[256], first, last, S[256], polys[16];
    if (lmax > 14 || dage_oracle_lip_coeffs(nlip, 1, lipc)) return -1;
    dage_oracle_lip_first_last(nlip, 1, &first, &last);
    const double eps = 2.220446049250313e-16; /* epsilon(0.d0), harmonics.F90:1345 */
    for (int b = 0; b < nbub; ++b) {
        const int nc = ncell[b], npts = nc * (nlip - 1) + 1;
        double *pts = (double *)malloc(sizeof(double) * npts);
        double *celld = (double *)malloc(sizeof(double) * nc);
        double *fk = (double *)malloc(sizeof(double) * (size_t)npts * nlm);
        if (!pts || !celld || !fk) { free(pts); free(celld); free(fk); return -1; }
        dage_oracle_grid1d(0.0, nc, nlip, steps[b], 1, pts, celld);
        /* bubble_values = bf * rpow (bubbles.F90:2118-2136, rpow :3363-3378) */
        for (int lm = 0; lm < nlm; ++lm)
            for (int i = 0; i < npts; ++i) {
                double rp = 1.0;
                if (k > 0) rp = pow(pts[i], k);
                else if (k < 0) rp = (i == 0) ? 0.0 : pow(pts[i], k);
                fk[i + (size_t)npts * lm] = (k == 0) ? bf[b][i + (size_t)npts * lm]
                                                     : bf[b][i + (size_t)npts * lm] * rp;
            }
        const double qmax = pts[npts - 1];
        const double cx = centers[3 * b], cy = centers[3 * b + 1], cz = centers[3 * b + 2];
        /* points are independent: with OpenMP the z-planes are shared out,
         * each point evaluated exactly as in the serial loop */
#pragma omp parallel for schedule(dynamic, 1) private(S, polys)
        for (int kk = 0; kk < nz; ++kk)
            for (int jj = 0; jj < ny; ++jj)
                for (int ii = 0; ii < nx; ++ii) {
                    const double dx = x[ii] - cx, dy = y[jj] - cy, dz = z[kk] - cz;
                    const double r = sqrt(dx * dx + dy * dy + dz * dz);
                    const int ic = dage_oracle_get_icell(celld, nc, 0.0, qmax, r);
                    double res = 0.0;
                    if (ic > 0) {
                        const double crd = (r - celld[ic - 1]) / steps[b][ic - 1] + first;
                        for (int j = 0; j < nlip; ++j) {  /* eval_polys, lip.F90:551-566 */
                            double v = lipc[j];
                            for (int q = 1; q < nlip; ++q) v = v * crd + lipc[j + nlip * q];
                            polys[j] = v;
                        }
                        dage_oracle_solid_harmonics(lmax, dx, dy, dz, S);
                        const int start = (ic - 1) * (nlip - 1);
                        for (int l = lmin; l <= lmax; ++l) {
                            const double rl = (r > eps) ? pow(r, l) : 1.0;
                            for (int m = -l; m <= l; ++m) {
                                const int lm = l * l + l + m;
                                double fv = 0.0;
                                for (int j = 0; j < nlip; ++j)
                                    fv += fk[start + j + (size_t)npts * lm] * polys[j];
                                res += fv * (S[lm] / rl);
                            }
                        }
                    }
                    const size_t id = ii + (size_t)nx * (jj + (size_t)ny * kk);
                    cube[id] = cube[id] + res;
                }
        free(pts);
        free(celld);
        free(fk);
    }
    return 0;
}

/* ------------------------------------------------------------------- FMM
 * multipole.F90:33-560: the interaction (M2L) and translation (M2M/L2L)
 * matrices for Racah-normalised real solid harmonics.  Vectors are indexed
 * lm = l*l + l + m (0-based); N = (lmax+1)^2. */
static double o_fact(int n) { /* product([(1.d0*i, i=1,n)]), 1 for n <= 0 */
    double p = 1.0;
    for (int i = 1; i <= n; ++i) p *= (double)i;
    return p;
}
static int o_sgn(int n) { return (n % 2 == 0) ? 1 : -1; } /* (-1)**n */
static int o_delta(int i, int j) { return i == j; }
static double o_cc(int l, int m) { /* multipole.F90:512-526 */
    if (m == 0) return 1.0 / o_fact(l);
    if (m > 0) return o_sgn(m) / sqrt(2.0 * o_fact(l - m) * o_fact(l + m));
    return o_sgn(m) * o_cc(l, -m);
}
static double o_cs(int l, int m) { /* multipole.F90:528-542 */
    if (m == 0) return 0.0;
    if (m > 0) return o_sgn(m) / sqrt(2.0 * o_fact(l - m) * o_fact(l + m));
    return -o_sgn(m) * o_cs(l, -m);
}
static double o_cst(int l, int m) { return o_fact(l - m) * o_fact(l + m) * o_cs(l, m); }
static double o_cct(int l, int m) { return o_fact(l - m) * o_fact(l + m) * o_cc(l, m); }
static double o_d(int l, int m) { /* multipole.F90:564-573 */
    return sqrt(o_fact(l + m) * o_fact(l - m) * (2.0 - o_delta(m, 0)));
}
static void o_lm(int idx, int *l, int *m) { /* 0-based lm -> (l, m) */
    int L = 0;
    while ((L + 1) * (L + 1) <= idx) ++L;
    *l = L;
    *m = idx - L * L - L;
}

/* InteractionMatrix_compute_coeffs + eval_pick (multipole.F90:104-183,
 * 268-312): T (N x N, column-major) with L_target = T(dR) q_source, where
 * dR = source centre - target centre. */
int dage_oracle_fmm_interaction(int lmax, const double dR[3], double *T) {
    const int N = (lmax + 1) * (lmax + 1);
    const int L2 = 2 * lmax;
    double *I = (double *)malloc(sizeof(double) * (L2 + 1) * (L2 + 1));
    if (!I) return -1;
    dage_oracle_solid_harmonics(L2, dR[0], dR[1], dR[2], I);
    const double r = sqrt(dR[0] * dR[0] + dR[1] * dR[1] + dR[2] * dR[2]);
    for (int j = 0; j <= L2; ++j)
        for (int k = -j; k <= j; ++k) I[j * j + j + k] /= pow(r, 2 * j + 1);
    const double ij[3] = {1.0, 0.5, 0.25};
    for (int s = 0; s < N; ++s) {
        int j, k;
        o_lm(s, &j, &k);
        for (int i = s; i < N; ++i) {
            int l, m;
            o_lm(i, &l, &m);
            const double c = 2.0 * o_sgn(j) * ij[o_delta(m, 0) + o_delta(k, 0)];
            int M1, M2;
            double c1, c2;
            const int LL = l + j;
            if (m >= 0 && k >= 0) {
                M1 = abs(m + k); c1 = o_cct(LL, m + k);
                M2 = abs(m - k); c2 = o_sgn(k) * o_cct(LL, m - k);
            } else if (m >= 0 && k < 0) {
                M1 = -abs(m + k); c1 = o_cst(LL, m + k);
                M2 = -abs(m - k); c2 = -o_sgn(k) * o_cst(LL, m - k);
                c1 = -c1 * o_sgn(k); c2 = -c2 * o_sgn(k);
            } else if (m < 0 && k >= 0) {
                M1 = -abs(m + k); c1 = o_cst(LL, m + k);
                M2 = -abs(m - k); c2 = o_sgn(k) * o_cst(LL, m - k);
                c1 = -c1 * o_sgn(m); c2 = -c2 * o_sgn(m);
            } else {
                M1 = abs(m + k); c1 = -o_cct(LL, m + k);
                M2 = abs(m - k); c2 = o_sgn(k) * o_cct(LL, m - k);
                c1 = c1 * o_sgn(m + k); c2 = c2 * o_sgn(m + k);
            }
            c1 *= c; c2 *= c;
            c1 = o_sgn(m) * c1 / o_d(l, m); c2 = o_sgn(m) * c2 / o_d(l, m);
            c1 = o_sgn(k) * c1 / o_d(j, k); c2 = o_sgn(k) * c2 / o_d(j, k);
            const double w = c1 * I[LL * LL + LL + M1] + c2 * I[LL * LL + LL + M2];
            T[i + (size_t)N * s] = w;
            T[s + (size_t)N * i] = o_sgn(l - j) * w;
        }
    }
    free(I);
    return 0;
}

/* TranslationMatrix_compute_coeffs + eval (multipole.F90:349-470): packed
 * unit-lower-triangular W (N(N+1)/2, column-major packed) for dR = from - to. */
int dage_oracle_fmm_translation(int lmax, const double dR[3], double *W) {
    const int N = (lmax + 1) * (lmax + 1);
    double *S = (double *)malloc(sizeof(double) * N);
    if (!S) return -1;
    dage_oracle_solid_harmonics(lmax, dR[0], dR[1], dR[2], S);
    const double ij[2] = {1.0, 0.5};
    for (int s = 0; s < N; ++s) {
        int j, k;
        o_lm(s, &j, &k);
        for (int i = s; i < N; ++i) {
            int l, m;
            o_lm(i, &l, &m);
            const int r = i + (2 * N - s - 1) * s / 2;
            const int d = l - j;
            int ok1 = 0, ok2 = 0, M1 = 0, M2 = 0;
            double c1 = 0.0, c2 = 0.0;
            if (m >= 0 && k >= 0) {
                if (abs(m - k) <= d) { ok1 = 1; M1 = abs(m - k); c1 = o_cc(d, m - k) * ij[o_delta(0, k)]; }
                if (abs(m + k) <= d) { ok2 = 1; M2 = abs(m + k); c2 = o_sgn(k) * o_cc(d, m + k) * ij[o_delta(0, k)]; }
            } else if (m >= 0 && k < 0) {
                if (abs(m - k) <= d) { ok1 = 1; M1 = -abs(m - k); c1 = o_cs(d, m - k) * o_sgn(k); }
                if (abs(m + k) <= d) { ok2 = 1; M2 = -abs(m + k); c2 = -o_cs(d, m + k); }
            } else if (m < 0 && k >= 0) {
                if (abs(m - k) <= d) { ok1 = 1; M1 = -abs(m - k); c1 = -o_cs(d, m - k) * o_sgn(m) * ij[o_delta(0, k)]; }
                if (abs(m + k) <= d) { ok2 = 1; M2 = -abs(m + k); c2 = -o_cs(d, m + k) * ij[o_delta(0, k)] * o_sgn(m + k); }
            } else {
                if (abs(m - k) <= d) { ok1 = 1; M1 = abs(m - k); c1 = o_cc(d, m - k) * o_sgn(m - k); }
                if (abs(m + k) <= d) { ok2 = 1; M2 = abs(m + k); c2 = -o_cc(d, m + k) * o_sgn(m); }
            }
            const double sc = o_d(l, m) * o_sgn(m - k) / o_d(j, k);
            c1 *= sc;
            c2 *= sc;
            double w = 0.0;
            if (ok1) w = c1 * S[d * d + d + M1];
            if (ok2) w = w + c2 * S[d * d + d + M2];
            W[r] = w;
        }
    }
    free(S);
    return 0;
}

/* dtpmv('L', 'N'|'T', 'U', N, W, q, 1) as TranslationMatrix_apply calls it
 * (multipole.F90:475-492): q := W q or W^T q, unit diagonal. */
void dage_oracle_fmm_translate(int lmax, const double *W, int transpose, double *q) {
    const int N = (lmax + 1) * (lmax + 1);
    double *out = (double *)malloc(sizeof(double) * N);
    if (!out) return;
    for (int i = 0; i < N; ++i) out[i] = q[i];
    for (int s = 0; s < N; ++s)
        for (int i = s + 1; i < N; ++i) {
            const double w = W[i + (2 * N - s - 1) * s / 2];
            if (transpose) out[s] += w * q[i];
            else out[i] += w * q[s];
        }
    for (int i = 0; i < N; ++i) q[i] = out[i];
    free(out);
}

/* Box multipoles as the GPU evaluator forms them (integrator.cu:1122-1196):
 * q_lm = sum_p rho_p wx wy wz S_lm(r_p - centre) over the box sub-grid. */
void dage_oracle_box_multipoles(int lmax, const double *rho, int ld_x, int ld_y,
                                const double *x, const double *wx, int nx,
                                const double *y, const double *wy, int ny,
                                const double *z, const double *wz, int nz,
                                const double centre[3], double *q) {
    const int N = (lmax + 1) * (lmax + 1);
    double S[(24 + 1) * (24 + 1)];
    for (int i = 0; i < N; ++i) q[i] = 0.0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double v = rho[i + (size_t)ld_x * (j + (size_t)ld_y * k)] * wx[i] * wy[j] * wz[k];
                dage_oracle_solid_harmonics(lmax, x[i] - centre[0], y[j] - centre[1], z[k] - centre[2], S);
                for (int n = 0; n < N; ++n) q[n] += S[n] * v;
            }
}

/* Local-expansion evaluation (gbfmm_coulomb3d.cu:81-280, .F90:1701-1745):
 * out(p) += sum_lm L_lm S_lm(r_p - centre) over an nx*ny*nz block. */
void dage_oracle_eval_local_expansion(int lmax, const double *L, const double *x, int nx,
                                      const double *y, int ny, const double *z, int nz,
                                      const double centre[3], double *out, int ld_x, int ld_y) {
    const int N = (lmax + 1) * (lmax + 1);
    double S[(24 + 1) * (24 + 1)];
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                dage_oracle_solid_harmonics(lmax, x[i] - centre[0], y[j] - centre[1], z[k] - centre[2], S);
                double v = 0.0;
                for (int n = 0; n < N; ++n) v += S[n] * L[n];
                out[i + (size_t)ld_x * (j + (size_t)ld_y * k)] += v;
            }
}

/* OpenMP threads of the loops above (no-op without OpenMP); returns the
 * count in use. */
#ifdef _OPENMP
#include <omp.h>
#endif
int dage_oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ===================================================== bubble radial transforms
 * Radial parts of the Coulomb and Helmholtz operators acting on bubbles
 * (coulomb.F90:222-336, helmholtz3d.F90:519-657) with the LIP cumulative
 * integrals they use (lip.F90:364-485, evaluators.F90:1263-1320) and the
 * modified spherical Bessel functions, evaluated in quadruple precision as the
 * reference does (REAL128, helmholtz3d.F90:879-1000, 1133-1210). */
#include <quadmath.h>

/* x**n for integer n as gfortran evaluates it (binary powering; 1/x^|n|). */
static double o_powi(double x, int n) { return __builtin_powi(x, n); }

/* LIPBasis_integral_coeffs (lip.F90:364-383): antiderivative coefficients,
 * column j (0-based) the coefficient of x^(nlip-j), column nlip zero. */
static void lip_integral_coeffs(int nlip, const double *c, double *ic) {
    for (int r = 0; r < nlip; ++r) ic[r + nlip * nlip] = 0.0;
    for (int i = 0; i < nlip; ++i)  /* int_coeffs(:,n-i) = coeffs(:,n-i)/(i+1) */
        for (int r = 0; r < nlip; ++r)
            ic[r + nlip * (nlip - 1 - i)] = c[r + nlip * (nlip - 1 - i)] / (i + 1);
}

int dage_oracle_lip_inout_integrals(int nlip, int grid_type, double *ints_in,
                                    double *ints_out) {
    double c[64], ic[72], first, last, tmp1[8], tmp2[8];
    if (nlip < 2 || nlip > 8 || dage_oracle_lip_coeffs(nlip, grid_type, c)) return -1;
    dage_oracle_lip_first_last(nlip, grid_type, &first, &last);
    lip_integral_coeffs(nlip, c, ic);
    /* inward: int_{x_j}^{last} chi_i, x_j = first + j (lip.F90:421-446) */
    for (int r = 0; r < nlip; ++r) tmp2[r] = 0.0;
    for (int i = 0; i < nlip; ++i)
        for (int r = 0; r < nlip; ++r) tmp2[r] = (tmp2[r] + ic[r + nlip * i]) * last;
    double x = first;
    for (int j = 0; j < nlip - 1; ++j) {
        for (int r = 0; r < nlip; ++r) tmp1[r] = 0.0;
        for (int i = 0; i < nlip; ++i)
            for (int r = 0; r < nlip; ++r) tmp1[r] = (tmp1[r] + ic[r + nlip * i]) * x;
        for (int r = 0; r < nlip; ++r) ints_in[r + nlip * j] = tmp2[r] - tmp1[r];
        x = x + 1;
    }
    /* outward: int_{first}^{x_j} chi_i, x_j = first + j + 1 (lip.F90:457-485) */
    for (int r = 0; r < nlip; ++r) tmp1[r] = 0.0;
    for (int i = 0; i < nlip; ++i)
        for (int r = 0; r < nlip; ++r) tmp1[r] = (tmp1[r] + ic[r + nlip * i]) * first;
    x = first;
    for (int j = 0; j < nlip - 1; ++j) {
        x = x + 1.0;
        for (int r = 0; r < nlip; ++r) tmp2[r] = 0.0;
        for (int i = 0; i < nlip; ++i)
            for (int r = 0; r < nlip; ++r) tmp2[r] = (tmp2[r] + ic[r + nlip * i]) * x;
        for (int r = 0; r < nlip; ++r) ints_out[r + nlip * j] = tmp2[r] - tmp1[r];
    }
    return 0;
}

int dage_oracle_integrate_inout(int ncell, int nlip, const double *steps, const double *f,
                                double *inw, double *outw) {
    double ii[64], io[64];
    if (dage_oracle_lip_inout_integrals(nlip, 1, ii, io)) return -1;
    const int n = nlip - 1, sz = ncell * n + 1;
    /* inwards (evaluators.F90:1283-1302): from the last cell down */
    if (inw) {
        inw[sz - 1] = 0.0;
        int start = sz - 1;
        for (int c = ncell - 1; c >= 0; --c) {
            start -= n;
            const double carry = inw[start + n];
            for (int j = 0; j < n; ++j) {
                double s = 0.0;
                for (int i = 0; i < nlip; ++i) s += f[start + i] * ii[i + nlip * j];
                inw[start + j] = s * steps[c] + carry;
            }
        }
    }
    /* outwards (evaluators.F90:1304-1320): from the first cell up */
    if (outw) {
        outw[0] = 0.0;
        int start = 0;
        for (int c = 0; c < ncell; ++c) {
            const double carry = outw[start];
            for (int j = 0; j < n; ++j) {
                double s = 0.0;
                for (int i = 0; i < nlip; ++i) s += f[start + i] * io[i + nlip * j];
                outw[start + 1 + j] = s * steps[c] + carry;
            }
            start += n;
        }
    }
    return 0;
}

/* Bubbles_rpow (bubbles.F90:3363-3378) */
static void o_rpow(const double *r, int npts, int n, double *out) {
    for (int i = 0; i < npts; ++i) out[i] = o_powi(r[i], n);
    if (n < 0) out[0] = 0.0;
}

static __float128 q_powi(__float128 x, int n) {
    /* binary powering, as gfortran's REAL(16)**INTEGER */
    __float128 y = 1.0Q;
    int m = n < 0 ? -n : n;
    while (m) {
        if (m & 1) y *= x;
        m >>= 1;
        if (m) x *= x;
    }
    return n < 0 ? 1.0Q / y : y;
}

/* FirstModSphBesselCollection_evaluate_small (helmholtz3d.F90:935-980) */
static __float128 q_bessel_i_series(__float128 z, int n, int maxterms) {
    __float128 prefactor = q_powi(z / 1.0Q, n);
    __float128 divider = 1.0Q;
    for (int i = 1; i <= 2 * n + 1; i += 2) divider = divider * i;
    prefactor = prefactor / divider;
    __float128 addition = 1.0Q, reslt = 1.0Q, fact = 1.0Q, a = 0.5Q * (z * z), b = 1.0Q;
    divider = 1.0Q;
    const __float128 tol = (__float128)1e-15;
    int i = 1;
    while (addition > tol * reslt && i < maxterms) {
        b = b * a;
        fact = fact * i;
        divider = divider * (2 * n + 2 * i + 1);
        addition = b / (fact * divider);
        reslt = reslt + addition;
        ++i;
    }
    return prefactor * reslt;
}

void dage_oracle_bessel_i(int lmax, double xd, double *out) {
    const __float128 x = xd;
    const __float128 eps = (__float128)2.220446049250313e-16;
    __float128 r[64];
    if (fabsq(x) > 0.0Q) {
        r[lmax] = q_bessel_i_series(x, lmax, 300);
        if (lmax >= 1) r[lmax - 1] = q_bessel_i_series(x, lmax - 1, 300);
        for (int l = lmax - 2; l >= 0; --l)
            r[l] = (2 * l + 3) * r[l + 1] / x * 1.0Q + r[l + 2] * 1.0Q;
    } else {
        for (int l = 0; l <= lmax; ++l) r[l] = 0.0Q;
        if (fabsq(x) < eps) r[0] = 1.0Q;
    }
    for (int l = 0; l <= lmax; ++l) out[l] = (double)r[l];
}

void dage_oracle_bessel_k(int lmax, double xd, double *out) {
    const __float128 x = xd;
    const __float128 pi = (__float128)o_pi();
    __float128 r[64];
    if (fabsq(x) > 0.0Q) {
        r[0] = expq(-x) / x * pi / 2;
        if (lmax >= 1) r[1] = expq(-x) * (x + 1) * 1.0Q / (x * x) * pi / 2;
        for (int l = 2; l <= lmax; ++l)
            r[l] = (2 * l - 1) * r[l - 1] * 1.0Q / x + r[l - 2] * 1.0Q;
    } else {
        for (int l = 0; l <= lmax; ++l) r[l] = 0.0Q;
    }
    for (int l = 0; l <= lmax; ++l) out[l] = (double)r[l];
}

int dage_oracle_bubble_coulomb(int lmax, int k, int ncell, int nlip, const double *steps,
                               const double *bf, double *out) {
    const int npts = ncell * (nlip - 1) + 1, nlm = (lmax + 1) * (lmax + 1);
    double *r = malloc(sizeof(double) * npts), *celld = malloc(sizeof(double) * ncell);
    double *w = malloc(sizeof(double) * 8 * npts);
    if (!r || !celld || !w || lmax > 62) { free(r); free(celld); free(w); return -1; }
    dage_oracle_grid1d(0.0, ncell, nlip, steps, 1, r, celld);
    double *p1 = w, *p2 = w + npts, *p3 = w + 2 * npts, *p4 = w + 3 * npts;
    double *g = w + 4 * npts, *o = w + 5 * npts, *in = w + 6 * npts;
    const double fourpi = 4.0 * o_pi();
    int rc = 0;
    for (int l = 0; l <= lmax && !rc; ++l) {
        o_rpow(r, npts, -(l + 1), p1);
        o_rpow(r, npts, l + 2 + k, p2);
        o_rpow(r, npts, l, p3);
        o_rpow(r, npts, -(l + 1) + 2 + k, p4);
        for (int m = -l; m <= l; ++m) {
            const int lm = l * l + l + m;
            const double *f = bf + (size_t)npts * lm;
            for (int i = 0; i < npts; ++i) g[i] = f[i] * p2[i];
            rc |= dage_oracle_integrate_inout(ncell, nlip, steps, g, NULL, o);
            for (int i = 0; i < npts; ++i) g[i] = f[i] * p4[i];
            rc |= dage_oracle_integrate_inout(ncell, nlip, steps, g, in, NULL);
            const double pre = fourpi / (2 * l + 1);
            double *v = out + (size_t)npts * lm;
            for (int i = 0; i < npts; ++i) v[i] = pre * (p1[i] * o[i] + p3[i] * in[i]);
        }
    }
    (void)nlm;
    free(r); free(celld); free(w);
    return rc;
}

/* Exponentially weighted inward / outward LIP integrals of one cell in cell
 * units (nodes t = 0..n): ii[r + nlip*j] = int_j^n exp(-mu (t-j)) L_r(t) dt,
 * io[r + nlip*j] = int_0^(j+1) exp(-mu (j+1-t)) L_r(t) dt, by composite
 * 20-point Gauss-Legendre quadrature (4 panels per node interval) in
 * quadruple precision: an evaluation independent of the closed-form moments
 * the library uses. */
void dage_oracle_exp_inout(int nlip, double mu, double *ii, double *io) {
    double c[256], first, last, gx[20], gw[20];
    dage_oracle_lip_coeffs(nlip, 1, c);
    dage_oracle_lip_first_last(nlip, 1, &first, &last);
    dage_oracle_gauleg(20, gx, gw);
    const int n = nlip - 1, panels = 4;
    for (int r = 0; r < nlip; ++r) {
        for (int j = 0; j < n; ++j) {
            __float128 sin_ = 0.0Q, sout = 0.0Q;
            for (int q = 0; q < n; ++q)
                for (int pnl = 0; pnl < panels; ++pnl) {
                    const __float128 a = q + (__float128)pnl / panels;
                    const __float128 hw = 0.5Q / panels;
                    for (int g = 0; g < 20; ++g) {
                        const __float128 t = a + hw * (1.0Q + gx[g]);
                        __float128 L = 0.0Q;
                        const __float128 X = t + first;
                        for (int kk = 0; kk < nlip; ++kk) L = L * X + c[r + nlip * kk];
                        const __float128 wl = hw * gw[g] * L;
                        if (q >= j) sin_ += wl * expq(-(__float128)mu * (t - j));
                        if (q <= j) sout += wl * expq(-(__float128)mu * (j + 1 - t));
                    }
                }
            ii[r + nlip * j] = (double)sin_;
            io[r + nlip * j] = (double)sout;
        }
    }
}

/* The exponentially fitted form of the radial Helmholtz convolution, used
 * when a cell has kappa L > 4 (there the LIP interpolant of k_l(ks) f or
 * i_l(ks) f over the cell carries errors of order exp(kappa L)): with
 * i_l(x) = e^x I(x), k_l(x) = e^-x K(x),
 *   V = 8k [K(kr) int_0^r e^-k(r-s) u ds + I(kr) int_r^R e^-k(s-r) v ds],
 * u = s^(2+k) I(ks) f, v = s^(2+k) K(ks) f interpolated cell by cell, the
 * exponentials integrated exactly (dage_oracle_exp_inout), carries between
 * cells scaled by e^(-k L). */
static int o_helmholtz_fitted(int lmax, int k, int ncell, int nlip, const double *steps,
                              double kappa, const double *r, const double *bf, double *out) {
    const int npts = ncell * (nlip - 1) + 1, n = nlip - 1, tab = nlip * n;
    double *Is = malloc(sizeof(double) * (size_t)npts * (lmax + 1));
    double *Ks = malloc(sizeof(double) * (size_t)npts * (lmax + 1));
    double *W = malloc(sizeof(double) * (size_t)ncell * 2 * tab);
    double *u = malloc(sizeof(double) * npts), *v = malloc(sizeof(double) * npts);
    double *outw = malloc(sizeof(double) * npts), *inw = malloc(sizeof(double) * npts);
    double *D = malloc(sizeof(double) * ncell), *C = malloc(sizeof(double) * ncell);
    if (!Is || !Ks || !W || !u || !v || !outw || !inw || !D || !C) {
        free(Is); free(Ks); free(W); free(u); free(v); free(outw); free(inw); free(D); free(C);
        return -1;
    }
    for (int i = 0; i < npts; ++i) {
        const __float128 x = (__float128)kappa * r[i];
        double tmp[64];
        dage_oracle_bessel_i(lmax, kappa * r[i], tmp);
        for (int l = 0; l <= lmax; ++l) Is[i + (size_t)npts * l] = (double)(tmp[l] * expq(-x));
        if (i == 0) {
            for (int l = 0; l <= lmax; ++l) Ks[(size_t)npts * l] = 0.0;
        } else {
            const __float128 pi = (__float128)o_pi();
            __float128 kq[64];
            kq[0] = pi / 2 / x;
            if (lmax >= 1) kq[1] = (x + 1) / (x * x) * pi / 2;
            for (int l = 2; l <= lmax; ++l) kq[l] = (2 * l - 1) * kq[l - 1] / x + kq[l - 2];
            for (int l = 0; l <= lmax; ++l) Ks[i + (size_t)npts * l] = (double)kq[l];
        }
    }
    for (int c = 0; c < ncell; ++c)
        dage_oracle_exp_inout(nlip, kappa * steps[c], W + (size_t)c * 2 * tab,
                              W + (size_t)c * 2 * tab + tab);
    double *rp = malloc(sizeof(double) * npts);
    o_rpow(r, npts, 2 + k, rp);
    for (int l = 0; l <= lmax; ++l) {
        const double *I = Is + (size_t)npts * l, *K = Ks + (size_t)npts * l;
        for (int m = -l; m <= l; ++m) {
            const int lm = l * l + l + m;
            const double *f = bf + (size_t)npts * lm;
            for (int i = 0; i < npts; ++i) {
                u[i] = rp[i] * I[i] * f[i];
                v[i] = rp[i] * K[i] * f[i];
            }
            v[0] = (l == 0) ? 3.0 * v[1] - 3.0 * v[2] + 1.0 * v[3] : 0.0;
            /* outward: partials within each cell, then carries D_c */
            outw[0] = 0.0;
            for (int c = 0; c < ncell; ++c)
                for (int j = 0; j < n; ++j) {
                    double sum = 0.0;
                    for (int i = 0; i <= n; ++i)
                        sum += u[c * n + i] * W[(size_t)c * 2 * tab + tab + i + nlip * j];
                    outw[c * n + 1 + j] = sum * steps[c];
                }
            double Dc = 0.0;
            for (int c = 0; c < ncell; ++c) {
                D[c] = Dc;
                Dc = exp(-kappa * n * steps[c]) * Dc + outw[c * n + n];
            }
            for (int i = 1; i < npts; ++i) {
                const int c = (i - 1) / n, j = i - c * n;
                outw[i] = exp(-kappa * j * steps[c]) * D[c] + outw[i];
            }
            /* inward */
            for (int c = 0; c < ncell; ++c)
                for (int j = 0; j < n; ++j) {
                    double sum = 0.0;
                    for (int i = 0; i <= n; ++i)
                        sum += v[c * n + i] * W[(size_t)c * 2 * tab + i + nlip * j];
                    inw[c * n + j] = sum * steps[c];
                }
            inw[npts - 1] = 0.0;
            double Cc = 0.0;
            for (int c = ncell - 1; c >= 0; --c) {
                C[c] = Cc;
                Cc = exp(-kappa * n * steps[c]) * Cc + inw[c * n];
            }
            double *o = out + (size_t)npts * lm;
            for (int i = 0; i < npts; ++i) {
                double in_i = 0.0;
                if (i < npts - 1) {
                    const int c = i / n, j = i - c * n;
                    in_i = inw[i] + exp(-kappa * (n - j) * steps[c]) * C[c];
                }
                o[i] = (K[i] * outw[i] + I[i] * in_i) * 8.0 * kappa;
            }
        }
    }
    free(Is); free(Ks); free(W); free(u); free(v); free(outw); free(inw); free(D); free(C);
    free(rp);
    return 0;
}

int dage_oracle_bubble_helmholtz(int lmax, int k, int ncell, int nlip, const double *steps,
                                 double energy, const double *bf, double *out) {
    const int npts = ncell * (nlip - 1) + 1, nlm = (lmax + 1) * (lmax + 1);
    if (lmax > 62) return -1;
    double *r = malloc(sizeof(double) * npts), *celld = malloc(sizeof(double) * ncell);
    double *bi = malloc(sizeof(double) * (size_t)npts * (lmax + 1));
    double *bk = malloc(sizeof(double) * (size_t)npts * (lmax + 1));
    double *w = malloc(sizeof(double) * 5 * npts);
    if (!r || !celld || !bi || !bk || !w) {
        free(r); free(celld); free(bi); free(bk); free(w);
        return -1;
    }
    dage_oracle_grid1d(0.0, ncell, nlip, steps, 1, r, celld);
    const double kappa = sqrt(-2.0 * energy);
    double kl = 0.0;
    for (int c = 0; c < ncell; ++c)
        if (kappa * (nlip - 1) * steps[c] > kl) kl = kappa * (nlip - 1) * steps[c];
    if (kl > 4.0) {
        const int rc = o_helmholtz_fitted(lmax, k, ncell, nlip, steps, kappa, r, bf, out);
        free(r); free(celld); free(bi); free(bk); free(w);
        return rc;
    }
    double tmp[64];
    for (int i = 0; i < npts; ++i) {
        dage_oracle_bessel_i(lmax, kappa * r[i], tmp);
        for (int l = 0; l <= lmax; ++l) bi[i + (size_t)npts * l] = tmp[l];
        dage_oracle_bessel_k(lmax, kappa * r[i], tmp);
        for (int l = 0; l <= lmax; ++l) bk[i + (size_t)npts * l] = tmp[l];
    }
    double *rp = w, *t1 = w + npts, *t2 = w + 2 * npts, *o = w + 3 * npts, *in = w + 4 * npts;
    o_rpow(r, npts, 2 + k, rp);
    int rc = 0;
    for (int l = 0; l <= lmax && !rc; ++l) {
        const double *I = bi + (size_t)npts * l, *K = bk + (size_t)npts * l;
        for (int m = -l; m <= l; ++m) {
            const int lm = l * l + l + m;
            const double *f = bf + (size_t)npts * lm;
            for (int i = 0; i < npts; ++i) {
                t1[i] = rp[i] * I[i] * f[i];
                t2[i] = rp[i] * K[i] * f[i];
            }
            /* extrapolate_first_nlip7(3, equidistant) for l = 0
             * (evaluators.F90:1588-1640) */
            if (l == 0) t2[0] = 3.0 * t2[1] - 3.0 * t2[2] + 1.0 * t2[3];
            else t2[0] = 0.0;
            rc |= dage_oracle_integrate_inout(ncell, nlip, steps, t1, NULL, o);
            rc |= dage_oracle_integrate_inout(ncell, nlip, steps, t2, in, NULL);
            double *v = out + (size_t)npts * lm;
            for (int i = 0; i < npts; ++i) v[i] = K[i] * o[i] + I[i] * in[i];
        }
    }
    for (size_t e = 0; e < (size_t)npts * nlm; ++e) out[e] = out[e] * 8.0 * kappa;
    free(r); free(celld); free(bi); free(bk); free(w);
    return rc;
}

/* ================================================================ products
 * YProduct (harmonic.F90:884-1000) restated as the projection
 * c(lm1, lm2 -> LM) = (2L+1)/(4 pi) int Y1 Y2 Y_LM dOmega with Y = S/r^l
 * (Racah), by Gauss-Legendre (cos theta) x uniform (phi) quadrature exact
 * for these degrees; coef is dense, (n1, n2, nL) column-major with
 * n = (lmax+1)^2 and L = l1max + l2max. */
int dage_oracle_yproduct(int l1max, int l2max, double *coef) {
    const int L = l1max + l2max;
    const int n1 = (l1max + 1) * (l1max + 1), n2 = (l2max + 1) * (l2max + 1);
    const int nL = (L + 1) * (L + 1), nt = L + 1, np = 2 * L + 2;
    if (L > 30) return -1;
    double *xt = malloc(sizeof(double) * nt), *wt = malloc(sizeof(double) * nt);
    double *S = malloc(sizeof(double) * nL);
    double *Y = malloc(sizeof(double) * (size_t)nt * np * nL);
    if (!xt || !wt || !S || !Y) { free(xt); free(wt); free(S); free(Y); return -1; }
    dage_oracle_gauleg(nt, xt, wt);
    const double pi = o_pi();
    for (int a = 0; a < nt; ++a)
        for (int b = 0; b < np; ++b) {
            const double ct = xt[a], st = sqrt(1.0 - ct * ct), ph = 2.0 * pi * b / np;
            dage_oracle_solid_harmonics(L, st * cos(ph), st * sin(ph), ct,
                                        Y + ((size_t)a * np + b) * nL);
        }
    for (int o = 0; o < nL; ++o) {
        const int lo = (int)sqrt((double)o);
        for (int j = 0; j < n2; ++j)
            for (int i = 0; i < n1; ++i) {
                double s = 0.0;
                for (int a = 0; a < nt; ++a)
                    for (int b = 0; b < np; ++b) {
                        const double *y = Y + ((size_t)a * np + b) * nL;
                        s += wt[a] * (2.0 * pi / np) * y[i] * y[j] * y[o];
                    }
                double c = s * (2 * lo + 1) / (4.0 * pi);
                if (fabs(c) < 1e-10) c = 0.0;  /* quadrature noise of zero entries */
                coef[i + (size_t)n1 * (j + (size_t)n2 * o)] = c;
            }
    }
    free(xt); free(wt); free(S); free(Y);
    return 0;
}

/* Bubble product on one bubble's radial points (bubbles.F90:3703-3756):
 * out(:, LM) = sum over (l1 m1), (l2 m2) in that order of c f1 f2. */
int dage_oracle_bubble_product(int l1max, int l2max, int npts, const double *f1,
                               const double *f2, double *out) {
    const int L = l1max + l2max;
    const int n1 = (l1max + 1) * (l1max + 1), n2 = (l2max + 1) * (l2max + 1);
    const int nL = (L + 1) * (L + 1);
    double *coef = malloc(sizeof(double) * (size_t)n1 * n2 * nL);
    if (!coef || dage_oracle_yproduct(l1max, l2max, coef)) { free(coef); return -1; }
    memset(out, 0, sizeof(double) * (size_t)npts * nL);
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j)
            for (int o = 0; o < nL; ++o) {
                const double c = coef[i + (size_t)n1 * (j + (size_t)n2 * o)];
                if (c == 0.0) continue;
                for (int p = 0; p < npts; ++p)
                    out[p + (size_t)npts * o] =
                        out[p + (size_t)npts * o] + c * f1[p + (size_t)npts * i] * f2[p + (size_t)npts * j];
            }
    free(coef);
    return 0;
}

/* ------------------------------------------------- GBFMM far field, direct
 * What the GBFMM far field approximates (gbfmm_coulomb3d.F90:650-821 with
 * 1/r, gbfmm_helmholtz3d.F90:730-964 with exp(-kappa r)/r): at every point x
 * a box owns (its first cut points per axis), the sum over all boxes c of
 * the finest level that are not nearest neighbours of x's box of
 * rho(y) w_c(y) G(|x - y|) over the points y of c's whole sub-grid, with
 * w_c the product of c's sub-axis integration weights.  kappa = 0 gives 1/r.
 * Axis a, box b: first point lo[a*nb + b], sub-grid points n[a*nb + b],
 * owned points cut[a*nb + b], weights w[(a*nb + b)*wld + i].  pot +=. */
void dage_oracle_gbfmm_far_direct(int nb, const int *lo, const int *n, const int *cut,
                                  const double *w, int wld, const double *x, const double *y,
                                  const double *z, int ldx, int ldy, const double *rho,
                                  double kappa, double *pot) {
    const int nbox = nb * nb * nb;
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < nbox; ++t) {
        const int tb[3] = {t % nb, (t / nb) % nb, t / (nb * nb)};
        for (int k = 0; k < cut[2 * nb + tb[2]]; ++k)
            for (int j = 0; j < cut[nb + tb[1]]; ++j)
                for (int i = 0; i < cut[tb[0]]; ++i) {
                    const int gi = lo[tb[0]] + i, gj = lo[nb + tb[1]] + j,
                              gk = lo[2 * nb + tb[2]] + k;
                    const double px = x[gi], py = y[gj], pz = z[gk];
                    double acc = 0.0;
                    for (int c = 0; c < nbox; ++c) {
                        const int cb[3] = {c % nb, (c / nb) % nb, c / (nb * nb)};
                        if (abs(cb[0] - tb[0]) <= 1 && abs(cb[1] - tb[1]) <= 1 &&
                            abs(cb[2] - tb[2]) <= 1)
                            continue;
                        const int x0 = lo[cb[0]], y0 = lo[nb + cb[1]], z0 = lo[2 * nb + cb[2]];
                        const int nx = n[cb[0]], ny = n[nb + cb[1]], nz = n[2 * nb + cb[2]];
                        const double *wx = w + (size_t)cb[0] * wld;
                        const double *wy = w + (size_t)(nb + cb[1]) * wld;
                        const double *wz = w + (size_t)(2 * nb + cb[2]) * wld;
                        for (int c3 = 0; c3 < nz; ++c3) {
                            const double dz = pz - z[z0 + c3];
                            for (int c2 = 0; c2 < ny; ++c2) {
                                const double dy = py - y[y0 + c2];
                                const double wyz = wy[c2] * wz[c3];
                                const double *rr = rho + (x0 + (size_t)ldx * ((y0 + c2) + (size_t)ldy * (z0 + c3)));
                                for (int c1 = 0; c1 < nx; ++c1) {
                                    const double dx = px - x[x0 + c1];
                                    const double r = sqrt(dx * dx + dy * dy + dz * dz);
                                    const double g = kappa > 0.0 ? exp(-kappa * r) / r : 1.0 / r;
                                    acc += rr[c1] * wx[c1] * wyz * g;
                                }
                            }
                        }
                    }
                    pot[gi + (size_t)ldx * (gj + (size_t)ldy * gk)] += acc;
                }
    }
}
