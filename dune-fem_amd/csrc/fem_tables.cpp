// Host construction of the 1D quadrature and basis tables used by the kernels.
//
// Quadrature: m-point Gauss-Legendre on [0,1] (the rule CubeQuadrature picks
// for order 2m-1, dune/fem/quadrature/femquadratures_inline.hh:59-93),
// recomputed here by Newton iteration on P_m in long double.
// Modal basis: orthonormal shifted Legendre sqrt(2n+1) P_n(2x-1)
// (LegendrePolynomials, dune/fem/space/shapefunctionset/legendrepolynomials.hh:24-60).
// Nodal Q2 basis: Lagrange on {0, 1/2, 1} (dune/fem/space/lagrange/shapefunctionset.hh).
#include <cmath>

#include "fem_common.hpp"

namespace femamd {

static void legendre_pm(int m, long double t, long double &p, long double &dp)
{
    long double p0 = 1.0L, p1 = t;
    if (m == 0) { p = 1.0L; dp = 0.0L; return; }
    for (int k = 2; k <= m; ++k) {
        long double pk = ((2.0L * k - 1.0L) * t * p1 - (k - 1.0L) * p0) / k;
        p0 = p1;
        p1 = pk;
    }
    p = p1;
    dp = m * (t * p1 - p0) / (t * t - 1.0L);
}

static void gauss_ld(int m, long double *x, long double *w)
{
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < m; ++i) {
        long double t = std::cos(pi * (i + 0.75L) / (m + 0.5L)), p, dp;
        for (int it = 0; it < 100; ++it) {
            legendre_pm(m, t, p, dp);
            long double dt = p / dp;
            t -= dt;
            if (std::fabs(dt) < 1e-30L) break;
        }
        legendre_pm(m, t, p, dp);
        x[i] = (1.0L - t) * 0.5L;  // ascending
        w[i] = 1.0L / ((1.0L - t * t) * dp * dp);
    }
    if (m % 2 == 1) x[m / 2] = 0.5L;
}

void gauss_legendre01(int m, double *x, double *w)
{
    long double xl[64], wl[64];
    gauss_ld(m, xl, wl);
    for (int i = 0; i < m; ++i) { x[i] = (double)xl[i]; w[i] = (double)wl[i]; }
}

static long double leg_ld(int n, long double x, long double *dval)
{
    long double t = 2.0L * x - 1.0L;
    long double P[KMAX + 2], dP[KMAX + 2];
    P[0] = 1.0L; dP[0] = 0.0L;
    if (n >= 1) { P[1] = t; dP[1] = 1.0L; }
    for (int k = 2; k <= n; ++k) {
        P[k] = ((2.0L * k - 1.0L) * t * P[k - 1] - (k - 1.0L) * P[k - 2]) / k;
        dP[k] = dP[k - 2] + (2.0L * k - 1.0L) * P[k - 1];
    }
    long double s = std::sqrt(2.0L * n + 1.0L);
    if (dval) *dval = s * 2.0L * dP[n];
    return s * P[n];
}

double legendre01(int n, double x) { return (double)leg_ld(n, x, nullptr); }
double legendre01_d(int n, double x)
{
    long double d;
    leg_ld(n, x, &d);
    return (double)d;
}

double lagrange2(int i, double x)
{
    return i == 0 ? 2.0 * (x - 0.5) * (x - 1.0) : i == 1 ? -4.0 * x * (x - 1.0) : 2.0 * x * (x - 0.5);
}
double lagrange2_d(int i, double x)
{
    return i == 0 ? 4.0 * x - 3.0 : i == 1 ? -8.0 * x + 4.0 : 4.0 * x - 1.0;
}

// Lagrange polynomial of node r over nodes xs, value and derivative at x
static void lagr_ld(int m, const long double *xs, int r, long double x, long double &v, long double &d)
{
    v = 1.0L;
    d = 0.0L;
    for (int k = 0; k < m; ++k) {
        if (k == r) continue;
        long double f = (x - xs[k]) / (xs[r] - xs[k]);
        d = d * f + v / (xs[r] - xs[k]);
        v *= f;
    }
}

Tab1D<double> make_tab1d(int k)
{
    Tab1D<double> t;
    std::memset(&t, 0, sizeof(t));
    long double xs[KMAX], ws[KMAX];
    gauss_ld(k, xs, ws);
    for (int q = 0; q < k; ++q) {
        t.x[q] = (double)xs[q];
        t.w[q] = (double)ws[q];
        for (int a = 0; a < k; ++a) {
            long double d;
            t.B[q][a] = (double)leg_ld(a, xs[q], &d);
            t.Bd[q][a] = (double)d;
        }
        for (int r = 0; r < k; ++r) {
            long double v, d;
            lagr_ld(k, xs, r, xs[q], v, d);
            t.Dt[q][r] = (double)d;
        }
    }
    // even/odd halves of Dt and Dt^T, from the long double values
    {
        long double D[KMAX][KMAX];
        for (int q = 0; q < k; ++q)
            for (int r = 0; r < k; ++r) {
                long double v, d;
                lagr_ld(k, xs, r, xs[q], v, d);
                D[q][r] = d;
            }
        const int h = k / 2, c = k / 2;
        for (int tr = 0; tr < 2; ++tr) {
            auto at = [&](int q, int r) { return tr ? D[r][q] : D[q][r]; };
            double(*Pp)[KMAX / 2 + 1] = tr ? t.DTp : t.Dp;
            double(*Pm)[KMAX / 2] = tr ? t.DTm : t.Dm;
            for (int q = 0; q < h; ++q) {
                for (int r = 0; r < h; ++r) {
                    Pp[q][r] = (double)(0.5L * (at(q, r) + at(q, k - 1 - r)));
                    Pm[q][r] = (double)(0.5L * (at(q, r) - at(q, k - 1 - r)));
                }
                if (k % 2) Pp[q][h] = (double)at(q, c);
            }
            if (k % 2)
                for (int r = 0; r < h; ++r) Pm[h][r] = (double)at(c, r);
        }
    }
    for (int a = 0; a < k; ++a)
        for (int b = 0; b < k; ++b) {
            long double s = 0;
            for (int q = 0; q < k; ++q) {
                long double da, db;
                leg_ld(a, xs[q], &da);
                leg_ld(b, xs[q], &db);
                s += ws[q] * da * db;
            }
            t.S[a][b] = (double)s;
        }
    for (int s = 0; s < 2; ++s) {
        for (int q = 0; q < k; ++q) {
            long double v, d;
            lagr_ld(k, xs, q, (long double)s, v, d);
            t.E[s][q] = (double)v;
            t.dE[s][q] = (double)d;
        }
        for (int a = 0; a < k; ++a) {
            long double d;
            t.P[s][a] = (double)leg_ld(a, (long double)s, &d);
            t.dP[s][a] = (double)d;
        }
    }
    return t;
}

Tab1D<float> to_float(const Tab1D<double> &t)
{
    Tab1D<float> f;
    const double *src = reinterpret_cast<const double *>(&t);
    float *dst = reinterpret_cast<float *>(&f);
    static_assert(sizeof(Tab1D<double>) == 2 * sizeof(Tab1D<float>), "layout");
    for (size_t i = 0; i < sizeof(Tab1D<double>) / sizeof(double); ++i) dst[i] = (float)src[i];
    return f;
}

}  // namespace femamd
