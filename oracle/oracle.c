/*
 * oracle.c -- CPU restatement of dune-fem's matrix-free Galerkin apply and
 * Krylov solvers.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* quadrature                                                               */
/* ------------------------------------------------------------------------ */

/* Legendre P_m and derivative on [-1,1] by the three-term recurrence */
static void legendre_pm(int m, long double t, long double *p, long double *dp)
{
    long double p0 = 1.0L, p1 = t;
    if (m == 0) { *p = 1.0L; *dp = 0.0L; return; }
    for (int k = 2; k <= m; ++k) {
        long double pk = ((2.0L * k - 1.0L) * t * p1 - (k - 1.0L) * p0) / k;
        p0 = p1; p1 = pk;
    }
    *p = p1;
    /* P'_m = m (t P_m - P_{m-1}) / (t^2 - 1) */
    *dp = m * (t * p1 - p0) / (t * t - 1.0L);
}

int ora_gauss(int m, double *x, double *w)
{
    if (m < 1 || m > 40) return -1;
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < m; ++i) {
        /* root i of P_m in descending t; Chebyshev-like first guess */
        long double t = cosl(pi * (i + 0.75L) / (m + 0.5L));
        long double p, dp;
        for (int it = 0; it < 100; ++it) {
            legendre_pm(m, t, &p, &dp);
            long double dt = p / dp;
            t -= dt;
            if (fabsl(dt) < 1e-30L) break;
        }
        legendre_pm(m, t, &p, &dp);
        long double wt = 2.0L / ((1.0L - t * t) * dp * dp);
        /* t descends with i, so x = (1-t)/2 ascends */
        x[i] = (double)((1.0L - t) * 0.5L);
        w[i] = (double)(wt * 0.5L);
    }
    if (m % 2 == 1) x[m / 2] = 0.5;
    return 0;
}

int ora_gauss_m_for_order(int order)
{
    if (order <= 0) order = 1;
    int m = 1;
    while (2 * m - 1 < order) ++m;
    return m;
}

/* ------------------------------------------------------------------------ */
/* 1D shape functions                                                       */
/* ------------------------------------------------------------------------ */

double ora_legendre(int n, double x)
{
    /* P_n(t), t = 2x-1, by recurrence, scaled by sqrt(2n+1) */
    double t = 2.0 * x - 1.0, p0 = 1.0, p1 = t;
    if (n == 0) return 1.0;
    for (int k = 2; k <= n; ++k) {
        double pk = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1; p1 = pk;
    }
    return sqrt(2.0 * n + 1.0) * p1;
}

double ora_legendre_d(int n, double x)
{
    /* d/dx P_n(2x-1) = 2 P'_n(t); P'_n via P'_k = P'_{k-2} + (2k-1) P_{k-1} */
    if (n == 0) return 0.0;
    double t = 2.0 * x - 1.0;
    double P[64], dP[64];
    P[0] = 1.0; P[1] = t; dP[0] = 0.0; dP[1] = 1.0;
    for (int k = 2; k <= n; ++k) {
        P[k] = ((2.0 * k - 1.0) * t * P[k - 1] - (k - 1.0) * P[k - 2]) / k;
        dP[k] = dP[k - 2] + (2.0 * k - 1.0) * P[k - 1];
    }
    return sqrt(2.0 * n + 1.0) * 2.0 * dP[n];
}

double ora_lagrange2(int i, double x)
{
    switch (i) {
    case 0: return 2.0 * (x - 0.5) * (x - 1.0);
    case 1: return -4.0 * x * (x - 1.0);
    default: return 2.0 * x * (x - 0.5);
    }
}

double ora_lagrange2_d(int i, double x)
{
    switch (i) {
    case 0: return 4.0 * x - 3.0;
    case 1: return -8.0 * x + 4.0;
    default: return 4.0 * x - 1.0;
    }
}

/* ------------------------------------------------------------------------ */
/* DG SIPG operator                                                          */
/* ------------------------------------------------------------------------ */

typedef struct {
    int p, k, nb, m, nq, nqf;
    double g[16], gw[16];
    /* volume tables: V[q*nb+i], G[(d*nq+q)*nb+i] (reference gradients) */
    double *V, *G;
    /* face tables for local face lf = 2*d + s: FV[(lf*nqf+q)*nb+i],
     * FG[((lf*3+dd)*nqf+q)*nb+i] */
    double *FV, *FG;
    double wvol[1024], wface[256];
} dg_tables;

static void dg_tables_init(dg_tables *T, int p)
{
    T->p = p; T->k = p + 1; T->nb = T->k * T->k * T->k;
    /* volume order 2p, face order 2p+1 (galerkin.hh:131-132) */
    T->m = ora_gauss_m_for_order(2 * p);
    int mf = ora_gauss_m_for_order(2 * p + 1);
    /* both give p+1 points; keep one rule */
    (void)mf;
    int m = T->m;
    ora_gauss(m, T->g, T->gw);
    T->nq = m * m * m; T->nqf = m * m;
    int nb = T->nb, nq = T->nq, nqf = T->nqf, k = T->k;
    T->V = (double *)malloc(sizeof(double) * nq * nb);
    T->G = (double *)malloc(sizeof(double) * 3 * nq * nb);
    T->FV = (double *)malloc(sizeof(double) * 6 * nqf * nb);
    T->FG = (double *)malloc(sizeof(double) * 18 * nqf * nb);
    for (int q = 0; q < nq; ++q) {
        int qq[3] = { q % m, (q / m) % m, q / (m * m) };
        double xi[3] = { T->g[qq[0]], T->g[qq[1]], T->g[qq[2]] };
        T->wvol[q] = T->gw[qq[0]] * T->gw[qq[1]] * T->gw[qq[2]];
        for (int i = 0; i < nb; ++i) {
            int a = i / (k * k), bb = (i / k) % k, c = i % k;
            double L[3] = { ora_legendre(a, xi[0]), ora_legendre(bb, xi[1]), ora_legendre(c, xi[2]) };
            double D[3] = { ora_legendre_d(a, xi[0]), ora_legendre_d(bb, xi[1]), ora_legendre_d(c, xi[2]) };
            T->V[q * nb + i] = L[0] * L[1] * L[2];
            T->G[(0 * nq + q) * nb + i] = D[0] * L[1] * L[2];
            T->G[(1 * nq + q) * nb + i] = L[0] * D[1] * L[2];
            T->G[(2 * nq + q) * nb + i] = L[0] * L[1] * D[2];
        }
    }
    for (int q = 0; q < nqf; ++q) T->wface[q] = T->gw[q % m] * T->gw[q / m];
    for (int d = 0; d < 3; ++d) {
        int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        for (int s = 0; s < 2; ++s) {
            int lf = 2 * d + s;
            for (int q = 0; q < nqf; ++q) {
                double xi[3];
                xi[d] = (double)s; xi[t1] = T->g[q % m]; xi[t2] = T->g[q / m];
                for (int i = 0; i < nb; ++i) {
                    int a = i / (k * k), bb = (i / k) % k, c = i % k;
                    double L[3] = { ora_legendre(a, xi[0]), ora_legendre(bb, xi[1]), ora_legendre(c, xi[2]) };
                    double D[3] = { ora_legendre_d(a, xi[0]), ora_legendre_d(bb, xi[1]), ora_legendre_d(c, xi[2]) };
                    T->FV[(lf * nqf + q) * nb + i] = L[0] * L[1] * L[2];
                    T->FG[((lf * 3 + 0) * nqf + q) * nb + i] = D[0] * L[1] * L[2];
                    T->FG[((lf * 3 + 1) * nqf + q) * nb + i] = L[0] * D[1] * L[2];
                    T->FG[((lf * 3 + 2) * nqf + q) * nb + i] = L[0] * L[1] * D[2];
                }
            }
        }
    }
}

static void dg_tables_free(dg_tables *T)
{
    free(T->V); free(T->G); free(T->FV); free(T->FG);
}

int64_t ora_dg_ndofs(const ora_dg_t *op)
{
    int k = op->p + 1;
    return (int64_t)op->n[0] * op->n[1] * op->n[2] * k * k * k;
}

static inline int64_t dg_face_index(const ora_dg_t *op, int d, const int c[3], int s)
{
    int N0 = op->n[0] + (d == 0), N1 = op->n[1] + (d == 1);
    int cc[3] = { c[0], c[1], c[2] };
    cc[d] += s;
    return (int64_t)cc[0] + (int64_t)N0 * (cc[1] + (int64_t)N1 * cc[2]);
}

static inline double dg_kappa_vol(const ora_dg_t *op, int64_t e, int q, int nq)
{
    switch (op->kappa_mode) {
    case ORA_KAPPA_QP: return op->kvol[e * nq + q];
    case ORA_KAPPA_CELL: return op->kcell[e];
    default: return op->kappa_const;
    }
}

/* evaluate value and physical gradient at face points of local face lf */
static void dg_face_eval(const dg_tables *T, int lf, const double *ul, const double h[3],
                         double *val, double *grad /* [3][nqf] */)
{
    int nb = T->nb, nqf = T->nqf;
    for (int q = 0; q < nqf; ++q) {
        double v = 0, g0 = 0, g1 = 0, g2 = 0;
        const double *fv = T->FV + (lf * nqf + q) * nb;
        const double *f0 = T->FG + ((lf * 3 + 0) * nqf + q) * nb;
        const double *f1 = T->FG + ((lf * 3 + 1) * nqf + q) * nb;
        const double *f2 = T->FG + ((lf * 3 + 2) * nqf + q) * nb;
        for (int i = 0; i < nb; ++i) {
            v += fv[i] * ul[i]; g0 += f0[i] * ul[i]; g1 += f1[i] * ul[i]; g2 += f2[i] * ul[i];
        }
        val[q] = v;
        grad[0 * nqf + q] = g0 / h[0]; grad[1 * nqf + q] = g1 / h[1]; grad[2 * nqf + q] = g2 / h[2];
    }
}

/* wl += sum_q c0[q] phi_i + c1[q] . grad phi_i at face points (c1 physical) */
static void dg_face_axpy(const dg_tables *T, int lf, const double *c0, const double *c1,
                         const double h[3], double *wl)
{
    int nb = T->nb, nqf = T->nqf;
    for (int q = 0; q < nqf; ++q) {
        const double *fv = T->FV + (lf * nqf + q) * nb;
        const double *f0 = T->FG + ((lf * 3 + 0) * nqf + q) * nb;
        const double *f1 = T->FG + ((lf * 3 + 1) * nqf + q) * nb;
        const double *f2 = T->FG + ((lf * 3 + 2) * nqf + q) * nb;
        double a0 = c0[q], a1 = c1[0 * nqf + q] / h[0], a2 = c1[1 * nqf + q] / h[1], a3 = c1[2 * nqf + q] / h[2];
        for (int i = 0; i < nb; ++i) wl[i] += a0 * fv[i] + a1 * f0[i] + a2 * f1[i] + a3 * f2[i];
    }
}

/* interior integral addInteriorIntegral (galerkin.hh:333-360) */
static void dg_interior(const ora_dg_t *op, const dg_tables *T, int64_t e, const double *ul,
                        const double h[3], double *wl)
{
    int nb = T->nb, nq = T->nq;
    double detJ = h[0] * h[1] * h[2];
    for (int q = 0; q < nq; ++q) {
        double u = 0, g[3] = { 0, 0, 0 };
        const double *v = T->V + q * nb;
        const double *G0 = T->G + (0 * nq + q) * nb, *G1 = T->G + (1 * nq + q) * nb, *G2 = T->G + (2 * nq + q) * nb;
        for (int i = 0; i < nb; ++i) { u += v[i] * ul[i]; g[0] += G0[i] * ul[i]; g[1] += G1[i] * ul[i]; g[2] += G2[i] * ul[i]; }
        for (int d = 0; d < 3; ++d) g[d] /= h[d];
        double kap = dg_kappa_vol(op, e, q, nq);
        double W = T->wvol[q] * detJ;
        /* integrand interior: (source, flux) = (c u, kappa grad u - b u) */
        double s = W * (op->c + (op->cvol ? op->cvol[e * nq + q] : 0.0)) * u;
        double F[3];
        for (int d = 0; d < 3; ++d) F[d] = W * (kap * g[d] - op->b[d] * u) / h[d];
        for (int i = 0; i < nb; ++i) wl[i] += s * v[i] + F[0] * G0[i] + F[1] * G1[i] + F[2] * G2[i];
    }
}

typedef struct { int64_t e; double *w; } dg_defer;

static void dg_apply_range(const ora_dg_t *op, const dg_tables *T, const double *u, double *w,
                           int64_t e0, int64_t e1, double *defbuf, int64_t *ndef, int64_t *defelem)
{
    int nb = T->nb, nqf = T->nqf, m = T->m;
    double h[3] = { 1.0 / op->n[0], 1.0 / op->n[1], 1.0 / op->n[2] };
    double *wl = (double *)malloc(sizeof(double) * nb);
    double *wo = (double *)malloc(sizeof(double) * nb);
    double vin[256], gin[768], vout[256], gout[768], c0[256], c1[768], d0[256], d1[768];
    int64_t nx = op->n[0], ny = op->n[1];
    for (int64_t e = e0; e < e1; ++e) {
        int c[3] = { (int)(e % nx), (int)((e / nx) % ny), (int)(e / (nx * ny)) };
        const double *ul = u + e * nb;
        memset(wl, 0, sizeof(double) * nb);
        dg_interior(op, T, e, ul, h, wl);
        /* intersections: faces 2d+s, s=0 lower, s=1 upper */
        for (int d = 0; d < 3; ++d) {
            int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
            double area = h[t1] * h[t2];
            for (int s = 0; s < 2; ++s) {
                int lf = 2 * d + s;
                double nrm = s ? 1.0 : -1.0;
                int nbc = c[d] + (s ? 1 : -1);
                int64_t fidx = dg_face_index(op, d, c, s);
                if (nbc < 0 || nbc >= op->n[d]) {
                    /* boundary integral addBoundaryIntegral (galerkin.hh:415-435) */
                    dg_face_eval(T, lf, ul, h, vin, gin);
                    double bn = op->b[d] * nrm;
                    for (int q = 0; q < nqf; ++q) {
                        double kap = op->kappa_mode == ORA_KAPPA_QP ? op->kface[d][fidx * nqf + q]
                                   : op->kappa_mode == ORA_KAPPA_CELL ? op->kcell[e] : op->kappa_const;
                        double W = T->wface[q] * area;
                        double dn = gin[d * nqf + q] * nrm;
                        double pen = kap * op->beta / h[d];
                        c0[q] = W * (pen * vin[q] - kap * dn + (bn > 0 ? bn : 0.0) * vin[q]);
                        for (int dd = 0; dd < 3; ++dd) c1[dd * nqf + q] = 0.0;
                        c1[d * nqf + q] = W * (-kap * vin[q] * nrm);
                    }
                    dg_face_axpy(T, lf, c0, c1, h, wl);
                    continue;
                }
                int cn[3] = { c[0], c[1], c[2] };
                cn[d] = nbc;
                int64_t en = cn[0] + nx * (cn[1] + ny * (int64_t)cn[2]);
                if (!(e < en)) continue;   /* skeleton once per pair (galerkin.hh:879) */
                /* addSkeletonIntegral (galerkin.hh:476-537): inside = e (s = 1 side),
                 * outside = en (its lower face), n = +e_d */
                int lfo = 2 * d + 0;
                const double *uo = u + en * nb;
                dg_face_eval(T, lf, ul, h, vin, gin);
                dg_face_eval(T, lfo, uo, h, vout, gout);
                double bn = op->b[d];
                double hbp = bn > 0 ? bn : 0.0, hbm = -bn > 0 ? -bn : 0.0;
                for (int q = 0; q < nqf; ++q) {
                    double kin, kout, kpen;
                    if (op->kappa_mode == ORA_KAPPA_QP) {
                        kin = kout = kpen = op->kface[d][fidx * nqf + q];
                    } else if (op->kappa_mode == ORA_KAPPA_CELL) {
                        kin = op->kcell[e]; kout = op->kcell[en]; kpen = 0.5 * (kin + kout);
                    } else {
                        kin = kout = kpen = op->kappa_const;
                    }
                    double W = T->wface[q] * area;
                    double jump = vin[q] - vout[q];
                    double sigma = 0.5 * (kin * gin[d * nqf + q] + kout * gout[d * nqf + q]);
                    double adv = hbp * vin[q] - hbm * vout[q];
                    double t0 = W * (kpen * op->beta / h[d] * jump - sigma + adv);
                    for (int dd = 0; dd < 3; ++dd) { c1[dd * nqf + q] = 0.0; d1[dd * nqf + q] = 0.0; }
                    c0[q] = t0;
                    c1[d * nqf + q] = W * (-0.5 * kin * jump);
                    d0[q] = -t0;
                    d1[d * nqf + q] = W * (-0.5 * kout * jump);
                }
                (void)m;
                dg_face_axpy(T, lf, c0, c1, h, wl);
                memset(wo, 0, sizeof(double) * nb);
                dg_face_axpy(T, lfo, d0, d1, h, wo);
                if (en < e1) {
                    for (int i = 0; i < nb; ++i) w[en * nb + i] += wo[i];
                } else {
                    int64_t k2 = (*ndef)++;
                    defelem[k2] = en;
                    memcpy(defbuf + k2 * nb, wo, sizeof(double) * nb);
                }
            }
        }
        for (int i = 0; i < nb; ++i) w[e * nb + i] += wl[i];
    }
    free(wl); free(wo);
}

void ora_dg_apply(const ora_dg_t *op, const double *u, double *w, int nthreads)
{
    dg_tables T;
    dg_tables_init(&T, op->p);
    int64_t E = (int64_t)op->n[0] * op->n[1] * op->n[2];
    int64_t N = E * T.nb;
    memset(w, 0, sizeof(double) * N);   /* w.clear() (galerkin.hh:1463) */
    if (nthreads < 1) nthreads = 1;
    if ((int64_t)nthreads > E) nthreads = (int)E;
    /* contiguous element chunks as ThreadIterator (threaditerator.hh:150-175);
     * contributions to elements of a later chunk are deferred and added after
     * the parallel region, which takes the place of the reference's locks */
    int64_t maxdef = 3 * (int64_t)op->n[0] * op->n[1] + 3 * (int64_t)op->n[0] * op->n[2] + 3 * (int64_t)op->n[1] * op->n[2] + 16;
    int64_t *ndef = (int64_t *)calloc(nthreads, sizeof(int64_t));
    double **defbuf = (double **)calloc(nthreads, sizeof(double *));
    int64_t **defelem = (int64_t **)calloc(nthreads, sizeof(int64_t *));
    int64_t chunkmax = (E + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        int64_t cap = nthreads == 1 ? 1 : (chunkmax < maxdef ? chunkmax * 3 : maxdef);
        defbuf[t] = (double *)malloc(sizeof(double) * T.nb * cap);
        defelem[t] = (int64_t *)malloc(sizeof(int64_t) * cap);
    }
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
#endif
    for (int t = 0; t < nthreads; ++t) {
        int64_t e0 = E * t / nthreads, e1 = E * (t + 1) / nthreads;
        dg_apply_range(op, &T, u, w, e0, e1, defbuf[t], &ndef[t], defelem[t]);
    }
    for (int t = 0; t < nthreads; ++t) {
        for (int64_t k = 0; k < ndef[t]; ++k) {
            int64_t en = defelem[t][k];
            for (int i = 0; i < T.nb; ++i) w[en * T.nb + i] += defbuf[t][k * T.nb + i];
        }
        free(defbuf[t]); free(defelem[t]);
    }
    free(ndef); free(defbuf); free(defelem);
    dg_tables_free(&T);
}

/* ---- sum-factorised apply (the CPU counterpart of the GPU algorithm) -------
 * The same operator as ora_dg_apply, evaluated with 1D contractions per
 * direction instead of the reference's dense nq x nb tables: u and its three
 * reference gradients at the volume points by K x K passes, fluxes, and the
 * transposed passes back; faces "owner computes": every element evaluates
 * its own and its neighbour's traces (value and normal derivative) on its six
 * faces and writes only its own DoFs (no locks, no deferred writes).  Used
 * only to time a sum-factorised CPU apply beside the reference algorithm. */
#define SF_KMAX 8

/* out = M along dimension dim (0: a outermost, 2: c innermost) of a K^3 tensor */
static void sf_pass(int K, const double *M /* [K][K]: out index, in index */, int dim, const double *in,
                    double *out)
{
    const int s[3] = { K * K, K, 1 };
    for (int a = 0; a < K; ++a)
        for (int b = 0; b < K; ++b)
            for (int c = 0; c < K; ++c) {
                const int idx[3] = { a, b, c };
                int base = 0;
                for (int d = 0; d < 3; ++d)
                    if (d != dim) base += idx[d] * s[d];
                double acc = 0;
                for (int k = 0; k < K; ++k) acc += M[idx[dim] * K + k] * in[base + k * s[dim]];
                out[a * K * K + b * K + c] = acc;
            }
}

typedef struct {
    int K;
    double B[SF_KMAX * SF_KMAX], Bt[SF_KMAX * SF_KMAX];     /* B[q][a] = phi_a(x_q) and transpose */
    double D[SF_KMAX * SF_KMAX], Dt[SF_KMAX * SF_KMAX];     /* phi_a'(x_q) and transpose */
    double P[2][SF_KMAX], dP[2][SF_KMAX], w[SF_KMAX];
} sf_tab;

static void sf_tab_init(sf_tab *t, int p)
{
    int K = p + 1;
    double x[SF_KMAX];
    t->K = K;
    ora_gauss(K, x, t->w);
    for (int q = 0; q < K; ++q)
        for (int a = 0; a < K; ++a) {
            t->B[q * K + a] = ora_legendre(a, x[q]);
            t->Bt[a * K + q] = t->B[q * K + a];
            t->D[q * K + a] = ora_legendre_d(a, x[q]);
            t->Dt[a * K + q] = t->D[q * K + a];
        }
    for (int s = 0; s < 2; ++s)
        for (int a = 0; a < K; ++a) {
            t->P[s][a] = ora_legendre(a, (double)s);
            t->dP[s][a] = ora_legendre_d(a, (double)s);
        }
}

/* value and d/dxi_d (reference) of the modal tensor u at the K^2 points of
 * face side s normal to d; point (q1, q2) over the tangential dims t1 < t2 */
static void sf_trace(const sf_tab *t, const double *u, int d, int s, double *val, double *der)
{
    int K = t->K;
    int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    const int st[3] = { K * K, K, 1 };
    double v2[SF_KMAX * SF_KMAX], d2[SF_KMAX * SF_KMAX];
    /* normal contraction -> [a_t1][a_t2] */
    for (int i = 0; i < K; ++i)
        for (int j = 0; j < K; ++j) {
            double a0 = 0, a1 = 0;
            for (int m = 0; m < K; ++m) {
                double um = u[m * st[d] + i * st[t1] + j * st[t2]];
                a0 += t->P[s][m] * um;
                a1 += t->dP[s][m] * um;
            }
            v2[i * K + j] = a0;
            d2[i * K + j] = a1;
        }
    /* tangential contractions -> [q1][q2] stored q1 + K q2 */
    for (int q1 = 0; q1 < K; ++q1)
        for (int q2 = 0; q2 < K; ++q2) {
            double a0 = 0, a1 = 0;
            for (int i = 0; i < K; ++i) {
                double bi = t->B[q1 * K + i];
                for (int j = 0; j < K; ++j) {
                    double bij = bi * t->B[q2 * K + j];
                    a0 += bij * v2[i * K + j];
                    a1 += bij * d2[i * K + j];
                }
            }
            val[q1 + K * q2] = a0;
            der[q1 + K * q2] = a1;
        }
}

/* r += sum_q c0(q) phi(q) + c1(q) dphi/dxi_d(q) over the face points */
static void sf_lift(const sf_tab *t, int d, int s, const double *c0, const double *c1, double *r)
{
    int K = t->K;
    int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    const int st[3] = { K * K, K, 1 };
    double a0[SF_KMAX * SF_KMAX], a1[SF_KMAX * SF_KMAX];
    for (int i = 0; i < K; ++i)
        for (int j = 0; j < K; ++j) {
            double x0 = 0, x1 = 0;
            for (int q1 = 0; q1 < K; ++q1)
                for (int q2 = 0; q2 < K; ++q2) {
                    double bb = t->B[q1 * K + i] * t->B[q2 * K + j];
                    x0 += bb * c0[q1 + K * q2];
                    x1 += bb * c1[q1 + K * q2];
                }
            a0[i * K + j] = x0;
            a1[i * K + j] = x1;
        }
    for (int m = 0; m < K; ++m)
        for (int i = 0; i < K; ++i)
            for (int j = 0; j < K; ++j)
                r[m * st[d] + i * st[t1] + j * st[t2]] += t->P[s][m] * a0[i * K + j] + t->dP[s][m] * a1[i * K + j];
}

void ora_dg_apply_sf(const ora_dg_t *op, const double *u, double *w, int nthreads)
{
    sf_tab T;
    sf_tab_init(&T, op->p);
    const int K = T.K, K3 = K * K * K;
    const int64_t E = (int64_t)op->n[0] * op->n[1] * op->n[2];
    const int64_t nx = op->n[0], ny = op->n[1];
    const double h[3] = { 1.0 / op->n[0], 1.0 / op->n[1], 1.0 / op->n[2] };
    const double detJ = h[0] * h[1] * h[2];
    if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (int64_t e = 0; e < E; ++e) {
        double t0[SF_KMAX * SF_KMAX * SF_KMAX], t1b[SF_KMAX * SF_KMAX * SF_KMAX];
        double uq[SF_KMAX * SF_KMAX * SF_KMAX], g[3][SF_KMAX * SF_KMAX * SF_KMAX];
        double r[SF_KMAX * SF_KMAX * SF_KMAX];
        const double *ue = u + e * K3;
        const int c[3] = { (int)(e % nx), (int)((e / nx) % ny), (int)(e / (nx * ny)) };
        /* u and reference gradients at the volume points: tensor (qa, qb, qc) = (x, y, z) */
        for (int d = -1; d < 3; ++d) {
            sf_pass(K, d == 0 ? T.D : T.B, 0, ue, t0);
            sf_pass(K, d == 1 ? T.D : T.B, 1, t0, t1b);
            sf_pass(K, d == 2 ? T.D : T.B, 2, t1b, d < 0 ? uq : g[d]);
        }
        /* fluxes: W |K| (kappa g_d / h_d - b_d u) / h_d against dphi/dxi_d, W |K| c u against phi */
        double F[3][SF_KMAX * SF_KMAX * SF_KMAX], S0[SF_KMAX * SF_KMAX * SF_KMAX];
        for (int qa = 0; qa < K; ++qa)
            for (int qb = 0; qb < K; ++qb)
                for (int qc = 0; qc < K; ++qc) {
                    int i = qa * K * K + qb * K + qc;
                    double W = T.w[qa] * T.w[qb] * T.w[qc] * detJ;
                    double kap = op->kappa_mode == ORA_KAPPA_QP ? op->kvol[e * K3 + qa + K * qb + K * K * qc]
                               : op->kappa_mode == ORA_KAPPA_CELL ? op->kcell[e] : op->kappa_const;
                    for (int d = 0; d < 3; ++d) F[d][i] = W * (kap * g[d][i] / h[d] - op->b[d] * uq[i]) / h[d];
                    const double cv = op->cvol ? op->cvol[e * K3 + qa + K * qb + K * K * qc] : 0.0;
                    S0[i] = W * (op->c + cv) * uq[i];
                }
        memset(r, 0, sizeof(double) * K3);
        for (int d = -1; d < 3; ++d) {
            const double *src = d < 0 ? S0 : F[d];
            sf_pass(K, d == 0 ? T.Dt : T.Bt, 0, src, t0);
            sf_pass(K, d == 1 ? T.Dt : T.Bt, 1, t0, t1b);
            sf_pass(K, d == 2 ? T.Dt : T.Bt, 2, t1b, t0);
            for (int i = 0; i < K3; ++i) r[i] += t0[i];
        }
        /* faces */
        for (int d = 0; d < 3; ++d) {
            int td1 = (d == 0) ? 1 : 0, td2 = (d == 2) ? 1 : 2;
            double area = h[td1] * h[td2];
            for (int s = 0; s < 2; ++s) {
                double nrm = s ? 1.0 : -1.0;
                int nbc = c[d] + (s ? 1 : -1);
                int interior = nbc >= 0 && nbc < op->n[d];
                int64_t fidx = dg_face_index(op, d, c, s);
                double vin[SF_KMAX * SF_KMAX], din[SF_KMAX * SF_KMAX], vout[SF_KMAX * SF_KMAX], dout[SF_KMAX * SF_KMAX];
                double c0[SF_KMAX * SF_KMAX], c1[SF_KMAX * SF_KMAX];
                sf_trace(&T, ue, d, s, vin, din);
                int64_t en = e;
                if (interior) {
                    int cn[3] = { c[0], c[1], c[2] };
                    cn[d] = nbc;
                    en = cn[0] + nx * (cn[1] + ny * (int64_t)cn[2]);
                    sf_trace(&T, u + en * K3, d, 1 - s, vout, dout);
                }
                double bn = op->b[d] * nrm, bp = bn > 0 ? bn : 0.0, bm = bn < 0 ? -bn : 0.0;
                for (int q = 0; q < K * K; ++q) {
                    int q1 = q % K, q2 = q / K;
                    double Wf = T.w[q1] * T.w[q2] * area;
                    double kin, kout, kp;
                    if (op->kappa_mode == ORA_KAPPA_QP) kin = kout = kp = op->kface[d][fidx * K * K + q];
                    else if (op->kappa_mode == ORA_KAPPA_CELL) {
                        kin = op->kcell[e];
                        kout = interior ? op->kcell[en] : kin;
                        kp = interior ? 0.5 * (kin + kout) : kin;
                    } else kin = kout = kp = op->kappa_const;
                    double dnin = din[q] * nrm / h[d];
                    double uo = interior ? vout[q] : 0.0, dno = interior ? dout[q] * nrm / h[d] : 0.0;
                    double wgt = interior ? 0.5 : 1.0;
                    double jump = vin[q] - uo;
                    double sigma = wgt * (kin * dnin + kout * dno);
                    c0[q] = Wf * (kp * op->beta / h[d] * jump - sigma + bp * vin[q] - bm * uo);
                    c1[q] = Wf * (-wgt * kin * jump) * nrm / h[d];
                }
                sf_lift(&T, d, s, c0, c1, r);
            }
        }
        memcpy(w + e * K3, r, sizeof(double) * K3);
    }
}

void ora_dg_rhs(const ora_dg_t *op, const double *fq, double *F)
{
    dg_tables T;
    dg_tables_init(&T, op->p);
    int64_t E = (int64_t)op->n[0] * op->n[1] * op->n[2];
    double detJ = 1.0 / ((double)op->n[0] * op->n[1] * op->n[2]);
    for (int64_t e = 0; e < E; ++e) {
        for (int i = 0; i < T.nb; ++i) F[e * T.nb + i] = 0.0;
        for (int q = 0; q < T.nq; ++q) {
            double W = T.wvol[q] * detJ * fq[e * T.nq + q];
            for (int i = 0; i < T.nb; ++i) F[e * T.nb + i] += W * T.V[q * T.nb + i];
        }
    }
    dg_tables_free(&T);
}

/* F += boundary-data part of the load vector for u = g on the boundary:
 * the Nitsche terms kappa beta/h g v - kappa g grad v . n and the inflow
 * term max(-b.n, 0) g v of dune/fempy/test/parallelTest.py:127-129 and
 * pydemo/advectiondiffusion.py:43-49 (the g-dependent part of the
 * boundary integrand, moved to the right-hand side); g at the face
 * quadrature points [faces_d][m^2] in the kface layout */
void ora_dg_rhs_bnd(const ora_dg_t *op, const double *const gface[3], double *F)
{
    dg_tables T;
    dg_tables_init(&T, op->p);
    int nb = T.nb, nqf = T.nqf;
    double h[3] = { 1.0 / op->n[0], 1.0 / op->n[1], 1.0 / op->n[2] };
    int64_t E = (int64_t)op->n[0] * op->n[1] * op->n[2];
    int64_t nx = op->n[0], ny = op->n[1];
    double c0[256], c1[768];
    for (int64_t e = 0; e < E; ++e) {
        int c[3] = { (int)(e % nx), (int)((e / nx) % ny), (int)(e / (nx * ny)) };
        for (int d = 0; d < 3; ++d) {
            int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
            double area = h[t1] * h[t2];
            for (int s = 0; s < 2; ++s) {
                int nbc = c[d] + (s ? 1 : -1);
                if (nbc >= 0 && nbc < op->n[d]) continue;
                int lf = 2 * d + s;
                double nrm = s ? 1.0 : -1.0;
                int64_t fidx = dg_face_index(op, d, c, s);
                double bn = op->b[d] * nrm;
                for (int q = 0; q < nqf; ++q) {
                    double kap = op->kappa_mode == ORA_KAPPA_QP ? op->kface[d][fidx * nqf + q]
                               : op->kappa_mode == ORA_KAPPA_CELL ? op->kcell[e] : op->kappa_const;
                    double W = T.wface[q] * area;
                    double g = gface[d][fidx * nqf + q];
                    c0[q] = W * (kap * op->beta / h[d] + (bn < 0 ? -bn : 0.0)) * g;
                    for (int dd = 0; dd < 3; ++dd) c1[dd * nqf + q] = 0.0;
                    c1[d * nqf + q] = -W * kap * g * nrm;
                }
                dg_face_axpy(&T, lf, c0, c1, h, F + e * nb);
            }
        }
    }
    dg_tables_free(&T);
}

/* element self block: contributions to element e from u supported on e only */
void ora_dg_block_diag(const ora_dg_t *op, double *blocks)
{
    int64_t E = (int64_t)op->n[0] * op->n[1] * op->n[2];
    int k = op->p + 1, nb = k * k * k;
    /* a 3x3x3-coloured sweep: apply A to u = sum over a colour class of unit
     * vectors; a neighbour of a coloured element is never coloured, so the
     * element block columns read off directly */
    int64_t N = E * nb;
    double *u = (double *)calloc(N, sizeof(double));
    double *w = (double *)malloc(sizeof(double) * N);
    int64_t nx = op->n[0], ny = op->n[1];
    for (int col = 0; col < 8; ++col) {
        int cx = col & 1, cy = (col >> 1) & 1, cz = (col >> 2) & 1;
        for (int j = 0; j < nb; ++j) {
            memset(u, 0, sizeof(double) * N);
            for (int64_t e = 0; e < E; ++e) {
                int c0 = (int)(e % nx), c1 = (int)((e / nx) % ny), c2 = (int)(e / (nx * ny));
                if ((c0 & 1) == cx && (c1 & 1) == cy && (c2 & 1) == cz) u[e * nb + j] = 1.0;
            }
            ora_dg_apply(op, u, w, omp_get_max_threads());
            for (int64_t e = 0; e < E; ++e) {
                int c0 = (int)(e % nx), c1 = (int)((e / nx) % ny), c2 = (int)(e / (nx * ny));
                if ((c0 & 1) == cx && (c1 & 1) == cy && (c2 & 1) == cz)
                    for (int i = 0; i < nb; ++i) blocks[(e * nb + i) * nb + j] = w[e * nb + i];
            }
        }
    }
    free(u); free(w);
}

void ora_dg_apply_cb(void *op, const double *x, double *y)
{
    ora_dg_apply((const ora_dg_t *)op, x, y, 1);
}

/* ------------------------------------------------------------------------ */
/* CG Q2 operator                                                            */
/* ------------------------------------------------------------------------ */

int64_t ora_cg_ndofs(const ora_cg_t *op)
{
    return (int64_t)(2 * op->n[0] + 1) * (2 * op->n[1] + 1);
}

typedef struct {
    double g[3], gw[3];
    double V[9][9], G[2][9][9], W[9];
} cg_tables;

static void cg_tables_init(cg_tables *T)
{
    ora_gauss(3, T->g, T->gw);
    for (int q = 0; q < 9; ++q) {
        int qx = q % 3, qy = q / 3;
        T->W[q] = T->gw[qx] * T->gw[qy];
        for (int i = 0; i < 9; ++i) {
            int ix = i % 3, iy = i / 3;
            T->V[q][i] = ora_lagrange2(ix, T->g[qx]) * ora_lagrange2(iy, T->g[qy]);
            T->G[0][q][i] = ora_lagrange2_d(ix, T->g[qx]) * ora_lagrange2(iy, T->g[qy]);
            T->G[1][q][i] = ora_lagrange2(ix, T->g[qx]) * ora_lagrange2_d(iy, T->g[qy]);
        }
    }
}

static inline int64_t cg_node(const ora_cg_t *op, int64_t ex, int64_t ey, int i)
{
    int64_t NX = 2 * op->n[0] + 1;
    return (2 * ey + i / 3) * NX + 2 * ex + i % 3;
}

static inline int cg_is_boundary(const ora_cg_t *op, int64_t node)
{
    int64_t NX = 2 * op->n[0] + 1, NY = 2 * op->n[1] + 1;
    int64_t i = node % NX, j = node / NX;
    return i == 0 || j == 0 || i == NX - 1 || j == NY - 1;
}

void ora_cg_apply(const ora_cg_t *op, const double *u, double *w)
{
    cg_tables T;
    cg_tables_init(&T);
    int64_t N = ora_cg_ndofs(op);
    memset(w, 0, sizeof(double) * N);
    double hx = 1.0 / op->n[0], hy = 1.0 / op->n[1], detJ = hx * hy;
    for (int64_t ey = 0; ey < op->n[1]; ++ey)
        for (int64_t ex = 0; ex < op->n[0]; ++ex) {
            int64_t e = ex + op->n[0] * ey;
            double ul[9], wl[9] = { 0 };
            for (int i = 0; i < 9; ++i) ul[i] = u[cg_node(op, ex, ey, i)];
            for (int q = 0; q < 9; ++q) {
                double uq = 0, gx = 0, gy = 0;
                for (int i = 0; i < 9; ++i) { uq += T.V[q][i] * ul[i]; gx += T.G[0][q][i] * ul[i]; gy += T.G[1][q][i] * ul[i]; }
                gx /= hx; gy /= hy;
                double kap = op->kappa_mode == ORA_KAPPA_QP ? op->kvol[e * 9 + q] : op->kappa_const;
                double Wt = T.W[q] * detJ;
                for (int i = 0; i < 9; ++i)
                    wl[i] += Wt * (kap * (gx * T.G[0][q][i] / hx + gy * T.G[1][q][i] / hy) +
                                   (op->c + (op->cvol ? op->cvol[e * 9 + q] : 0.0)) * uq * T.V[q][i]);
            }
            for (int i = 0; i < 9; ++i) w[cg_node(op, ex, ey, i)] += wl[i];
        }
    /* subConstraints: w_D = u_D (dirichletwrapper.hh:101-105) */
    for (int64_t k = 0; k < N; ++k)
        if (cg_is_boundary(op, k)) w[k] = u[k];
}

void ora_cg_eval_qp(const ora_cg_t *op, const double *u, double *out)
{
    cg_tables T;
    cg_tables_init(&T);
    for (int64_t ey = 0; ey < op->n[1]; ++ey)
        for (int64_t ex = 0; ex < op->n[0]; ++ex) {
            int64_t e = ex + op->n[0] * ey;
            for (int q = 0; q < 9; ++q) {
                double uq = 0;
                for (int i = 0; i < 9; ++i) uq += T.V[q][i] * u[cg_node(op, ex, ey, i)];
                out[e * 9 + q] = uq;
            }
        }
}

void ora_cg_rhs(const ora_cg_t *op, const double *fq, double *F)
{
    cg_tables T;
    cg_tables_init(&T);
    int64_t N = ora_cg_ndofs(op);
    memset(F, 0, sizeof(double) * N);
    double detJ = 1.0 / ((double)op->n[0] * op->n[1]);
    for (int64_t ey = 0; ey < op->n[1]; ++ey)
        for (int64_t ex = 0; ex < op->n[0]; ++ex) {
            int64_t e = ex + op->n[0] * ey;
            for (int q = 0; q < 9; ++q) {
                double Wt = T.W[q] * detJ * fq[e * 9 + q];
                for (int i = 0; i < 9; ++i) F[cg_node(op, ex, ey, i)] += Wt * T.V[q][i];
            }
        }
    for (int64_t k = 0; k < N; ++k)
        if (cg_is_boundary(op, k)) F[k] = 0.0;
}

void ora_cg_diag(const ora_cg_t *op, double *d)
{
    cg_tables T;
    cg_tables_init(&T);
    int64_t N = ora_cg_ndofs(op);
    memset(d, 0, sizeof(double) * N);
    double hx = 1.0 / op->n[0], hy = 1.0 / op->n[1], detJ = hx * hy;
    for (int64_t ey = 0; ey < op->n[1]; ++ey)
        for (int64_t ex = 0; ex < op->n[0]; ++ex) {
            int64_t e = ex + op->n[0] * ey;
            for (int i = 0; i < 9; ++i) {
                double a = 0;
                for (int q = 0; q < 9; ++q) {
                    double kap = op->kappa_mode == ORA_KAPPA_QP ? op->kvol[e * 9 + q] : op->kappa_const;
                    double gx = T.G[0][q][i] / hx, gy = T.G[1][q][i] / hy;
                    double r = op->c + (op->cvol ? op->cvol[e * 9 + q] : 0.0);
                    a += T.W[q] * detJ * (kap * (gx * gx + gy * gy) + r * T.V[q][i] * T.V[q][i]);
                }
                d[cg_node(op, ex, ey, i)] += a;
            }
        }
    for (int64_t k = 0; k < N; ++k)
        if (cg_is_boundary(op, k)) d[k] = 1.0;
}

void ora_cg_apply_cb(void *op, const double *x, double *y)
{
    ora_cg_apply((const ora_cg_t *)op, x, y);
}

void ora_diag_cb(void *ctx, const double *x, double *y)
{
    const ora_diag_t *D = (const ora_diag_t *)ctx;
    for (int64_t i = 0; i < D->n; ++i) y[i] = D->d[i] * x[i];
}

void ora_block_cb(void *ctx, const double *x, double *y)
{
    const ora_block_t *B = (const ora_block_t *)ctx;
    int nb = B->nb;
    for (int64_t e = 0; e < B->nblocks; ++e) {
        const double *M = B->inv + e * nb * nb;
        for (int i = 0; i < nb; ++i) {
            double s = 0;
            for (int j = 0; j < nb; ++j) s += M[i * nb + j] * x[e * nb + j];
            y[e * nb + i] = s;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* BLAS-1 helpers                                                            */
/* ------------------------------------------------------------------------ */

static double dot(int64_t n, const double *a, const double *b)
{
    double s = 0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static void axpy(int64_t n, double a, const double *x, double *y)
{
    for (int64_t i = 0; i < n; ++i) y[i] += a * x[i];
}

/* ------------------------------------------------------------------------ */
/* CG: LinearSolver::cg (dune/fem/solver/linear/cg.hh:19-117)                */
/* ------------------------------------------------------------------------ */

int ora_cg(int64_t n, ora_apply_fn op, void *opctx, ora_apply_fn pre, void *prectx,
           double *x, const double *b, double eps, int maxit, int criteria)
{
    double *h = (double *)malloc(sizeof(double) * n);
    double *r = (double *)malloc(sizeof(double) * n);
    double *p = (double *)malloc(sizeof(double) * n);
    double *s = pre ? (double *)malloc(sizeof(double) * n) : p;
    double *q = pre ? (double *)malloc(sizeof(double) * n) : p;
    op(opctx, x, h);                                     /* h = A x */
    for (int64_t i = 0; i < n; ++i) { r[i] = h[i] - b[i]; p[i] = b[i] - h[i]; }
    if (pre) { pre(prectx, p, q); memcpy(s, q, sizeof(double) * n); }
    double prev = 0, res = dot(n, p, q);
    double tol = eps * eps * (criteria == 1 ? dot(n, b, b) : criteria == 2 ? res : 1.0);
    int it;
    for (it = 0; res > tol && it < maxit; ++it) {
        if (it > 0) {
            double beta = res / prev;
            for (int64_t i = 0; i < n; ++i) q[i] *= beta;
            if (pre) axpy(n, 1.0, s, q);
            else axpy(n, -1.0, r, p);
        }
        op(opctx, q, h);
        double qh = dot(n, q, h);
        double alpha = res / qh;
        axpy(n, alpha, q, x);
        if (pre) {
            axpy(n, -alpha, h, p);
            pre(prectx, p, s);
            prev = res;
            res = dot(n, p, s);
        } else {
            axpy(n, alpha, h, r);
            prev = res;
            res = dot(n, r, r);
        }
    }
    free(h); free(r); free(p);
    if (pre) { free(s); free(q); }
    return it < maxit ? it : -it;
}

/* BiCGStab, right preconditioned, as LinearSolver::bicgstab
 * (dune/fem/solver/linear/bicgstab.hh:56-222): the iterate is y with
 * x = M^-1 y; the initial guess enters as y and the result is mapped back */
int ora_bicgstab(int64_t n, ora_apply_fn op, void *opctx, ora_apply_fn pre, void *prectx,
                 double *x, const double *b, double eps, int maxit, int criteria)
{
    double *r = (double *)malloc(sizeof(double) * n);
    double *rs = (double *)malloc(sizeof(double) * n);
    double *p = (double *)malloc(sizeof(double) * n);
    double *s = (double *)malloc(sizeof(double) * n);
    double *t = (double *)malloc(sizeof(double) * n);
    double *z = pre ? (double *)malloc(sizeof(double) * n) : NULL;
    double tol = eps;
    if (criteria == 1) tol *= sqrt(dot(n, b, b));
    if (pre) { pre(prectx, x, z); op(opctx, z, r); }
    else op(opctx, x, r);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
    memcpy(p, r, sizeof(double) * n);
    memcpy(rs, r, sizeof(double) * n);
    double nu = dot(n, r, rs);
    if (criteria == 2) tol *= sqrt(nu);
    int it = 0;
    while (1) {
        if (pre) { pre(prectx, p, z); op(opctx, z, t); }
        else op(opctx, p, t);
        double alpha = nu / dot(n, t, rs);
        for (int64_t i = 0; i < n; ++i) s[i] = r[i] - alpha * t[i];
        if (pre) { pre(prectx, s, z); op(opctx, z, r); }
        else op(opctx, s, r);
        double g0 = dot(n, r, s), g1 = dot(n, r, r), g2 = dot(n, s, s), g3 = dot(n, s, rs), g4 = dot(n, r, rs);
        double omega = g0 / g1;
        double res = sqrt(g2 - omega * (2.0 * g0 - omega * g1));
        double beta = (g3 - omega * g4) * alpha / (omega * nu);
        nu = g3 - omega * g4;
        for (int64_t i = 0; i < n; ++i) x[i] += alpha * p[i] + omega * s[i];
        ++it;
        if (res < tol || it >= maxit) break;
        for (int64_t i = 0; i < n; ++i) {
            r[i] = s[i] - omega * r[i];
            p[i] = beta * (p[i] - omega * t[i]) + r[i];
        }
    }
    if (pre) { pre(prectx, x, z); memcpy(x, z, sizeof(double) * n); }
    free(r); free(rs); free(p); free(s); free(t); free(z);
    return it >= maxit ? -it : it;
}

/* ------------------------------------------------------------------------ */
/* GMRES / FGMRES                                                            */
/* ------------------------------------------------------------------------ */

static void rot(double *x, double *y, double c, double s)
{
    double a = *x, b = *y;
    *x = c * a + s * b;
    *y = c * b - s * a;
}

static int gmres_impl(int flexible, int64_t n, ora_apply_fn op, void *opctx, ora_apply_fn pre,
                      void *prectx, double *u, const double *b, int m, double tol, int maxit,
                      int criteria)
{
    double *V = (double *)malloc(sizeof(double) * n * (m + 2));
    double *Z = flexible && pre ? (double *)malloc(sizeof(double) * n * m) : NULL;
    double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
    double *g = (double *)calloc(6 * m + 2, sizeof(double));
    double *sn = g + (m + 1), *cs = sn + m, *y = cs + m;
    double *tmp = (double *)malloc(sizeof(double) * n);
#define HH(i, j) H[(size_t)(i) * m + (j)]
    double _tol = tol;
    if (criteria == 1) _tol *= sqrt(dot(n, b, b));
    int iterations = 0;
    while (1) {
        double *v0 = V;
        op(opctx, u, v0);
        for (int64_t i = 0; i < n; ++i) v0[i] -= b[i];
        double res = sqrt(dot(n, v0, v0));
        if (criteria == 2 && iterations == 0) _tol *= res;
        if (res <= _tol * (1 + 1e-15)) break;
        g[0] = -res;
        for (int i = 1; i <= m; ++i) g[i] = 0.0;
        for (int64_t i = 0; i < n; ++i) v0[i] *= 1.0 / res;
        int j;
        for (j = 0; j < m; ++j) {
            double *vj = V + (int64_t)j * n, *vjp = V + (int64_t)(j + 1) * n;
            if (pre) {
                double *z = flexible ? Z + (int64_t)j * n : V + (int64_t)(m + 1) * n;
                pre(prectx, vj, z);
                op(opctx, z, vjp);
            } else {
                op(opctx, vj, vjp);
            }
            /* classical Gram-Schmidt: all projections from the same vector */
            for (int l = 0; l <= j; ++l) HH(l, j) = dot(n, vjp, V + (int64_t)l * n);
            for (int l = 0; l <= j; ++l) axpy(n, -HH(l, j), V + (int64_t)l * n, vjp);
            HH(j + 1, j) = sqrt(dot(n, vjp, vjp));
            for (int64_t i = 0; i < n; ++i) vjp[i] *= 1.0 / HH(j + 1, j);
            for (int i = 0; i < j; ++i) rot(&HH(i + 1, j), &HH(i, j), cs[i], sn[i]);
            double a = HH(j, j), bb = HH(j + 1, j);
            double nrm = sqrt(a * a + bb * bb);
            cs[j] = a / nrm;
            sn[j] = -bb / nrm;
            rot(&HH(j + 1, j), &HH(j, j), cs[j], sn[j]);
            rot(&g[j + 1], &g[j], cs[j], sn[j]);
            ++iterations;
            if (fabs(g[j + 1]) < _tol || iterations >= maxit) break;
        }
        int last = iterations % m;
        if (last == 0) last = m;
        for (int i = last - 1; i >= 0; --i) {
            double s = 0;
            for (int k = i + 1; k < last; ++k) s += HH(i, k) * y[k];
            y[i] = (g[i] - s) / HH(i, i);
        }
        if (pre && !flexible) {
            memset(tmp, 0, sizeof(double) * n);
            for (int i = 0; i < last; ++i) axpy(n, y[i], V + (int64_t)i * n, tmp);
            double *z = V + (int64_t)(m + 1) * n;
            pre(prectx, tmp, z);
            axpy(n, 1.0, z, u);
        } else if (pre && flexible) {
            for (int i = 0; i < last; ++i) axpy(n, y[i], Z + (int64_t)i * n, u);
        } else {
            for (int i = 0; i < last; ++i) axpy(n, y[i], V + (int64_t)i * n, u);
        }
        if (fabs(g[last]) < _tol) break;
        if (iterations >= maxit) break;
    }
#undef HH
    free(V); free(Z); free(H); free(g); free(tmp);
    return iterations < maxit ? iterations : -iterations;
}

int ora_gmres(int64_t n, ora_apply_fn op, void *opctx, ora_apply_fn pre, void *prectx,
              double *x, const double *b, int restart, double eps, int maxit, int criteria)
{
    return gmres_impl(0, n, op, opctx, pre, prectx, x, b, restart, eps, maxit, criteria);
}

int ora_fgmres(int64_t n, ora_apply_fn op, void *opctx, ora_apply_fn pre, void *prectx,
               double *x, const double *b, int restart, double eps, int maxit, int criteria)
{
    return gmres_impl(1, n, op, opctx, pre, prectx, x, b, restart, eps, maxit, criteria);
}
