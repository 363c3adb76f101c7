/*
 * CPU oracle for the stream-collide path of 2dRadiation.
 *
 * TEST INFRASTRUCTURE ONLY (see rt_oracle.h).  Each function cites the
 * reference file:line it restates.  The arithmetic keeps the reference's
 * operation order so that an FP64 GPU kernel compiled with contraction off
 * can be compared against it to round-off.
 *
 * Pins (tests/test_oracle_*.py): Hastings polynomials vs libm error bounds and
 * exact special values; stencil closed forms; Lagrange identities of the cubic
 * tables (partition of unity, exactness on cubics); Fourier round trip;
 * bilinear exactness on affine fields; collide closed forms (kappa=0, u=0);
 * metric closed forms (Minkowski exact, Kerr-Schild / Schwarzschild lapse,
 * shift and inverse); tetrad orthonormality; RK45 straight lines in flat space
 * and time reversal in curved space; mirror symmetry of the flat transport.
 * Parity unpinned at the ulp level: the 3x3 inverse (the reference calls
 * Eigen, which is not vendored) and libm rounding inside RK45.
 */
#include "rt_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358979323846
#define OR_PI_2 1.57079632679489661923

int32_t or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void or_set_threads(int32_t n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* Hastings polynomials, order 9 (src/ControlFlow.hh:15, src/Utility.hh)     */
/* ------------------------------------------------------------------------ */
double or_myatan(double z) /* src/Utility.hh:179-188 */
{
    const double a1 = 0.9998660, a3 = -0.3302995, a5 = 0.1801410, a7 = -0.0851330, a9 = 0.0208351;
    double zz = z * z;
    return z * (a1 + zz * (a3 + zz * (a5 + zz * (a7 + zz * a9))));
}
double or_myatan2(double y, double x) /* src/Utility.hh:220-254 */
{
    const double pi = OR_PI, pi_2 = OR_PI_2;
    if (x != 0.0) {
        if (fabsf((float)x) > fabsf((float)y)) {
            const double z = y / x;
            if (x > 0.0) return or_myatan(z);
            else if (y >= 0.0) return or_myatan(z) + pi;
            else return or_myatan(z) - pi;
        } else {
            const double z = x / y;
            if (y > 0.0) return -or_myatan(z) + pi_2;
            else return -or_myatan(z) - pi_2;
        }
    } else {
        if (y > 0.0) return pi_2;
        else if (y < 0.0) return -pi_2;
    }
    return 0.0;
}
double or_mysin(double x) /* src/Utility.hh:272-323 */
{
    const double pi = OR_PI, pi_2 = OR_PI_2;
    if (-pi_2 <= x && x <= pi_2) {
    } else if (pi_2 <= x && x <= 3.0 * pi_2) x = -(x - pi);
    else if (3.0 * pi_2 <= x && x <= 5.0 * pi_2) x = x - 2.0 * pi;
    else return NAN; /* the reference exits with an error here */
    x *= 2.0 / pi;
    const double a1 = 1.57079631847, a3 = -0.64596371106, a5 = 0.07968967928,
                 a7 = -0.00467376557, a9 = 0.00015148419;
    double xx = x * x;
    return x * (a1 + xx * (a3 + xx * (a5 + xx * (a7 + xx * a9))));
}
double or_mycos(double xd) /* src/Utility.hh:340-359: evaluated in float */
{
    const float pi = (float)OR_PI, pi_2 = (float)OR_PI_2;
    float x = (float)xd;
    if (x >= pi) {
        float t = x - pi_2;
        return (double)(float)(-or_mysin((double)t));
    }
    float t = x + pi_2;
    return (double)(float)or_mysin((double)t);
}
double or_phi(double x, double y) /* src/DataTypes.hh:166-170 */
{
    double p = or_myatan2(y, x);
    return (p < 0) ? p + 2.0 * OR_PI : p;
}

/* ------------------------------------------------------------------------ */
/* Grid (src/Grid.cpp:4-37, 97-161)                                         */
/* ------------------------------------------------------------------------ */
void or_grid_make(int32_t nx_, int32_t ny_, double sx, double sy, double ex, double ey,
                  double cfl, or_grid *g)
{
    const int halo = 1;
    g->nx = nx_ + 2 * halo;
    g->ny = ny_ + 2 * halo;
    g->dx = (ex - sx) / (g->nx - 1.0 - 2.0 * halo);
    g->dy = (ey - sy) / (g->ny - 1.0 - 2.0 * halo);
    g->dt = cfl * (g->dx < g->dy ? g->dx : g->dy);
    g->startx = sx - g->dx;
    g->starty = sy - g->dy;
    g->endx = ex + g->dx;
    g->endy = ey + g->dy;
}
static inline int or_outside(const or_grid *g, double x, double y) /* Grid.cpp:152-156 */
{
    return x > g->endx || x < g->startx || y > g->endy || y < g->starty;
}
static inline int or_inside_bh(double r2lim, double x, double y) /* Spacetimes.cpp:38-42 */
{
    return r2lim >= 0.0 && (x * x + y * y) <= r2lim;
}

/* ------------------------------------------------------------------------ */
/* Stencil (src/Stencil.cpp:177-297) and cubic tables (:4-162)              */
/* ------------------------------------------------------------------------ */
void or_stencil(int32_t n, double *w, double *phi, double *cx, double *cy)
{
    for (int d = 0; d < n; d++) {
        w[d] = 1.0 / n;
        phi[d] = 2.0 * OR_PI * (d + 0.5) / n;
        cx[d] = or_mycos(phi[d]);
        cy[d] = or_mysin(phi[d]);
    }
    /* SortDirections (:299-332) leaves an already increasing phi unchanged. */
}
static void or_nearest(int32_t nDir, const double *sphi, double phi, int *index)
{
    double minDist = 10;
    *index = -1;
    for (int d = 0; d < nDir; d++) {
        double delta = fabs(phi - sphi[d]);
        double alt = 2.0 * OR_PI - delta;
        double dist = delta < alt ? delta : alt;
        if (dist < minDist) { minDist = dist; *index = d; }
    }
}
void or_interp_grid(int32_t nDir, const double *sphi, int32_t nGrid, int32_t *Nc, double *Wc)
{
    for (int i = 0; i < nGrid; i++) {
        double phi = 2.0 * OR_PI * i / nGrid;
        int index;
        or_nearest(nDir, sphi, phi, &index);
        double phiMin = fmod(sphi[index] + OR_PI - phi, 2.0 * OR_PI);
        int32_t *nb = Nc + 4 * i;
        if (phiMin < OR_PI) {
            for (int k = 0; k < 4; k++) nb[k] = ((index - 1 + k) % nDir + nDir) % nDir;
        } else {
            for (int k = 0; k < 4; k++) nb[k] = ((index - 2 + k) % nDir + nDir) % nDir;
        }
        /* LagrangeWeightsCubic (:109-122): angles rotated so that phi -> pi */
        for (int a = 0; a < 4; a++) {
            double weight = 1;
            double phi_a = fmod(sphi[nb[a]] + OR_PI - phi + 2.0 * OR_PI, 2.0 * OR_PI);
            for (int b = 0; b < 4; b++) {
                double phi_b = fmod(sphi[nb[b]] + OR_PI - phi + 2.0 * OR_PI, 2.0 * OR_PI);
                if (a != b) weight *= (OR_PI - phi_b) / (phi_a - phi_b);
            }
            Wc[4 * i + a] = weight;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Fourier harmonics (src/FourierHarmonics.cpp:5-62)                         */
/* ------------------------------------------------------------------------ */
double or_fourier_basis(int32_t k, double phi)
{
    const double sqrt2 = 1.4142135623730951; /* sqrt(2) */
    if (k == 0) return 1.0;
    else if (k % 2 != 0) return sqrt2 * cos(phi * (k + 1.0) / 2.0);
    else return sqrt2 * sin(phi * k / 2.0);
}
static void or_fourier_coefficients(int32_t nF, const double *phiS, const double *wS,
                                    const double *data, double *coef)
{
    for (int i = 0; i < nF; i++) coef[i] = 0;
    for (int d = 0; d < nF; d++) {
        double c = data[d] * wS[d];
        for (int i = 0; i < nF; i++) coef[i] += c * or_fourier_basis(i, phiS[d]);
    }
}
static inline double or_fourier_value(const double *basis, const double *coef, int nF)
{
    double r = 0;
    for (int i = 0; i < nF; i++) r += coef[i] * basis[i];
    return r;
}

/* ------------------------------------------------------------------------ */
/* Metric (src/Metric.cpp:126-336, src/Spacetimes.cpp:1-111)                 */
/* ------------------------------------------------------------------------ */
void or_metric_function(int32_t kind, double m, double a, double x, double y, double *g)
{
    for (int k = 0; k < 9; k++) g[k] = 0.0;
    g[0] = -1; g[4] = 1; g[8] = 1;
    if (kind == 1) { /* SchwarzSchild::MetricFunction, Spacetimes.cpp:48-64 */
        double rs = 2.0 * m;
        double r2 = x * x + y * y;
        double r = sqrt(r2);
        if (r > 2.0 * m) {
            g[0] = -1.0 + rs / r;
            g[4] = x * x / (r2 - rs * r) + y * y / r2;
            g[8] = x * x / r2 + y * y / (r2 - rs * r);
            g[5] = g[7] = rs * x * y / (r2 * (r - rs));
        }
    } else if (kind == 2) { /* KerrSchild::MetricFunction, Spacetimes.cpp:87-111 */
        double a2 = a * a;
        double R2 = x * x + y * y;
        if (R2 > a2) {
            double r2 = R2 - a2;
            double r = sqrt(r2);
            double rho2 = r2 + a2;
            double H = m / r;
            double l[3] = {1, (r * x + a * y) / rho2, (r * y - a * x) / rho2};
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    if (i == 0 && j == 0) g[0] = -1.0 + 2.0 * H * l[0] * l[0];
                    else g[3 * i + j] = (double)(i == j) + 2.0 * H * l[i] * l[j];
                }
        }
    }
}
/* Adjugate inverse with cyclic cofactors (stand-in for Eigen's fixed-size
 * 3x3 inverse called at src/DataTypes.hh:417-425). */
static inline double cof3(const double *m, int i, int j)
{
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return m[3 * i1 + j1] * m[3 * i2 + j2] - m[3 * i1 + j2] * m[3 * i2 + j1];
}
void or_invert3(const double *m, double *inv)
{
    double c0 = cof3(m, 0, 0), c1 = cof3(m, 1, 0), c2 = cof3(m, 2, 0);
    double det = (c0 * m[0] + c1 * m[3]) + c2 * m[6];
    double id = 1.0 / det;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            inv[3 * i + j] = cof3(m, j, i) * id;
}
static void or_metric_deriv(int32_t kind, double m, double a, double x, double y, int k,
                            int inverse, double *out)
{ /* MetricDeriv<k> / InverseMetricDeriv<k>, Metric.cpp:189-219 */
    double dk = 1e-8;
    double gm2[9], gm1[9], gp1[9], gp2[9], t[9];
    double ex = (k == 1), ey = (k == 2);
    or_metric_function(kind, m, a, x - 2 * dk * ex, y - 2 * dk * ey, gm2);
    or_metric_function(kind, m, a, x - 1 * dk * ex, y - 1 * dk * ey, gm1);
    or_metric_function(kind, m, a, x + 1 * dk * ex, y + 1 * dk * ey, gp1);
    or_metric_function(kind, m, a, x + 2 * dk * ex, y + 2 * dk * ey, gp2);
    if (inverse) {
        or_invert3(gm2, t); memcpy(gm2, t, sizeof t);
        or_invert3(gm1, t); memcpy(gm1, t, sizeof t);
        or_invert3(gp1, t); memcpy(gp1, t, sizeof t);
        or_invert3(gp2, t); memcpy(gp2, t, sizeof t);
    }
    for (int q = 0; q < 9; q++)
        out[q] = (1.0 / 12.0 * gm2[q] - 1 / 12.0 * gp2[q] - 2.0 / 3.0 * gm1[q] + 2.0 / 3.0 * gp1[q]) / dk;
}
void or_metric_build(const or_grid *g, int32_t kind, double m, double a, double *out)
{
    const long N = (long)g->nx * g->ny;
#define A(q) (out + (long)(q) * N)
#pragma omp parallel for schedule(static)
    for (long ij = 0; ij < N; ij++) {
        long i = ij % g->nx, j = ij / g->nx;
        double x = g->startx + i * g->dx, y = g->starty + j * g->dy;
        double gl[9], gu[9], dxl[9], dyl[9], dxu[9], dyu[9];
        or_metric_function(kind, m, a, x, y, gl);
        or_invert3(gl, gu);
        or_metric_deriv(kind, m, a, x, y, 1, 0, dxl);
        or_metric_deriv(kind, m, a, x, y, 2, 0, dyl);
        or_metric_deriv(kind, m, a, x, y, 1, 1, dxu);
        or_metric_deriv(kind, m, a, x, y, 2, 1, dyu);
        A(OR_G00L)[ij] = gl[0]; A(OR_G01L)[ij] = gl[1]; A(OR_G02L)[ij] = gl[2];
        A(OR_G11L)[ij] = gl[4]; A(OR_G12L)[ij] = gl[5]; A(OR_G22L)[ij] = gl[8];
        A(OR_G00U)[ij] = gu[0]; A(OR_G01U)[ij] = gu[1]; A(OR_G02U)[ij] = gu[2];
        A(OR_G11U)[ij] = gu[4]; A(OR_G12U)[ij] = gu[5]; A(OR_G22U)[ij] = gu[8];
        /* InitializeAdmComponentsOnGrid, Metric.cpp:276-336 */
        double al = 1.0 / sqrt(-gu[0]);
        double b1u = al * al * gu[1], b2u = al * al * gu[2];
        A(OR_ALPHA)[ij] = al; A(OR_B1U)[ij] = b1u; A(OR_B2U)[ij] = b2u;
        A(OR_B1L)[ij] = gl[1]; A(OR_B2L)[ij] = gl[2];
        double ga11 = gl[4], ga12 = gl[5], ga22 = gl[8];
        A(OR_GA11L)[ij] = ga11; A(OR_GA12L)[ij] = ga12; A(OR_GA22L)[ij] = ga22;
        double gu11 = gu[4] + b1u * b1u / (al * al);
        double gu12 = gu[5] + b1u * b2u / (al * al);
        double gu22 = gu[8] + b2u * b2u / (al * al);
        A(OR_GA11U)[ij] = gu11; A(OR_GA12U)[ij] = gu12; A(OR_GA22U)[ij] = gu22;
        double da1 = dxu[0] / (2.0 * pow(-gu[0], 3.0 / 2.0));
        double da2 = dyu[0] / (2.0 * pow(-gu[0], 3.0 / 2.0));
        A(OR_DA1)[ij] = da1; A(OR_DA2)[ij] = da2;
        double d1b1 = al * al * dxu[1] + 2.0 * gu[1] * al * da1;
        double d2b1 = al * al * dyu[1] + 2.0 * gu[1] * al * da2;
        double d1b2 = al * al * dxu[2] + 2.0 * gu[2] * al * da1;
        double d2b2 = al * al * dyu[2] + 2.0 * gu[2] * al * da2;
        A(OR_D1B1U)[ij] = d1b1; A(OR_D1B2U)[ij] = d1b2; A(OR_D2B1U)[ij] = d2b1; A(OR_D2B2U)[ij] = d2b2;
        double dg[2][2][2] = {{{dxl[4], dxl[5]}, {dxl[5], dxl[8]}}, {{dyl[4], dyl[5]}, {dyl[5], dyl[8]}}};
        A(OR_D1GA11L)[ij] = dxl[4]; A(OR_D1GA12L)[ij] = dxl[5]; A(OR_D1GA22L)[ij] = dxl[8];
        A(OR_D2GA11L)[ij] = dyl[4]; A(OR_D2GA12L)[ij] = dyl[5]; A(OR_D2GA22L)[ij] = dyl[8];
        double bu[2] = {b1u, b2u};
        double dbu[2][2] = {{d1b1, d1b2}, {d2b1, d2b2}}; /* [l][k] = d_l beta^k */
        double gam[2][2] = {{ga11, ga12}, {ga12, ga22}};
        double K[2][2];
        for (int p = 0; p < 2; p++)
            for (int q = 0; q < 2; q++) {
                K[p][q] = -0.0;
                for (int c = 0; c < 2; c++)
                    K[p][q] += 2.0 * gam[p][c] * dbu[q][c] + dg[c][p][q] * bu[c];
                K[p][q] /= 2.0 * al;
            }
        A(OR_K11)[ij] = K[0][0]; A(OR_K12)[ij] = K[0][1]; A(OR_K22)[ij] = K[1][1];
    }
    /* InitializeBoostedTetradOnGrid, Metric.cpp:152-187 (serial in the reference) */
#pragma omp parallel for schedule(static)
    for (long ij = 0; ij < N; ij++) {
        double al = A(OR_ALPHA)[ij];
        double aa = 1.0 / sqrt(A(OR_GA11U)[ij]);
        double bb = aa * aa * A(OR_GA12U)[ij];
        double m11 = 1 / aa, m21 = bb / aa, m22 = 1;
        double g2 = m22 * A(OR_GA22L)[ij] * m22;
        double t[9] = {1.0 / al, 0, 0,
                       -A(OR_B1U)[ij] / al, m11, 0,
                       -A(OR_B2U)[ij] / al, m21, 1 / sqrt(g2)};
        double ti[9];
        or_invert3(t, ti);
        for (int q = 0; q < 9; q++) { A(OR_T00 + q)[ij] = t[q]; A(OR_TI00 + q)[ij] = ti[q]; }
    }
#undef A
}

/* Off-grid getters: bilinear interpolation inside the domain, flat values
 * outside (Metric.cpp:338-359, 374-800).  Index clamping only affects terms
 * whose bilinear weight is exactly zero. */
typedef struct { const or_grid *g; const double *M; long N; } or_mv;
#define MA(mv, q) ((mv)->M + (long)(q) * (mv)->N)
static inline double or_interp(const or_mv *mv, int q, double fi, double fj)
{
    const or_grid *g = mv->g;
    long i0 = (long)floor(fi), j0 = (long)floor(fj);
    long i1 = i0 + 1 < g->nx ? i0 + 1 : g->nx - 1;
    long j1 = j0 + 1 < g->ny ? j0 + 1 : g->ny - 1;
    double x = fi - i0, y = fj - j0;
    const double *a = MA(mv, q);
    double f00 = a[i0 + j0 * g->nx], f01 = a[i0 + j1 * g->nx];
    double f10 = a[i1 + j0 * g->nx], f11 = a[i1 + j1 * g->nx];
    double f0 = f00 * (1.0 - x) + f10 * x;
    double f1 = f01 * (1.0 - x) + f11 * x;
    return f0 * (1.0 - y) + f1 * y;
}
typedef struct {
    double alpha, b1, b2, da1, da2, db[2][2], K[2][2], gu[2][2], dg[2][2][2];
} or_local;
static void or_local_metric(const or_mv *mv, double x, double y, or_local *L)
{
    const or_grid *g = mv->g;
    if (or_outside(g, x, y)) {
        memset(L, 0, sizeof *L);
        L->alpha = 1.0;
        L->gu[0][0] = 1; L->gu[1][1] = 1;
        return;
    }
    double fi = (x - g->startx) / g->dx, fj = (y - g->starty) / g->dy;
    L->alpha = or_interp(mv, OR_ALPHA, fi, fj);
    L->b1 = or_interp(mv, OR_B1U, fi, fj);
    L->b2 = or_interp(mv, OR_B2U, fi, fj);
    L->da1 = or_interp(mv, OR_DA1, fi, fj);
    L->da2 = or_interp(mv, OR_DA2, fi, fj);
    L->db[0][0] = or_interp(mv, OR_D1B1U, fi, fj);
    L->db[0][1] = or_interp(mv, OR_D1B2U, fi, fj);
    L->db[1][0] = or_interp(mv, OR_D2B1U, fi, fj);
    L->db[1][1] = or_interp(mv, OR_D2B2U, fi, fj);
    L->K[0][0] = or_interp(mv, OR_K11, fi, fj);
    L->K[1][1] = or_interp(mv, OR_K22, fi, fj);
    L->K[0][1] = L->K[1][0] = or_interp(mv, OR_K12, fi, fj);
    L->gu[0][0] = or_interp(mv, OR_GA11U, fi, fj);
    L->gu[1][1] = or_interp(mv, OR_GA22U, fi, fj);
    L->gu[0][1] = L->gu[1][0] = or_interp(mv, OR_GA12U, fi, fj);
    L->dg[0][0][0] = or_interp(mv, OR_D1GA11L, fi, fj);
    L->dg[0][1][1] = or_interp(mv, OR_D1GA22L, fi, fj);
    L->dg[0][0][1] = L->dg[0][1][0] = or_interp(mv, OR_D1GA12L, fi, fj);
    L->dg[1][0][0] = or_interp(mv, OR_D2GA11L, fi, fj);
    L->dg[1][1][1] = or_interp(mv, OR_D2GA22L, fi, fj);
    L->dg[1][0][1] = L->dg[1][1][0] = or_interp(mv, OR_D2GA12L, fi, fj);
}
/* Dx, Dv, Dnu: GeodesicEquationSolver.cpp:5-55 (Vincent, Gourgoulhon, Novak 2012) */
static void or_rhs(const or_mv *mv, double x, double y, double vx, double vy, double nu,
                   double *kx, double *kv, double *knu)
{
    or_local L;
    or_local_metric(mv, x, y, &L);
    double v[2] = {vx, vy}, dA[2] = {L.da1, L.da2}, bu[2] = {L.b1, L.b2};
    for (int k = 0; k < 2; k++) kx[k] = L.alpha * v[k] - bu[k];
    for (int k = 0; k < 2; k++) {
        double dv = 0.0;
        for (int i = 0; i < 2; i++) {
            dv += v[k] * dA[i] * v[i] - L.gu[k][i] * dA[i] - v[i] * L.db[i][k];
            for (int j = 0; j < 2; j++) {
                dv += 2.0 * L.alpha * L.gu[i][k] * L.K[i][j] * v[j] - L.alpha * v[k] * L.K[i][j] * v[i] * v[j];
                for (int l = 0; l < 2; l++)
                    dv -= L.alpha * 0.5 * L.gu[k][l] * (L.dg[i][l][j] + L.dg[j][i][l] - L.dg[l][i][j]) * v[i] * v[j];
            }
        }
        kv[k] = dv;
    }
    double dnu = 0;
    for (int i = 0; i < 2; i++) {
        dnu -= v[i] * dA[i];
        for (int j = 0; j < 2; j++) dnu += L.alpha * L.K[i][j] * v[i] * v[j];
    }
    *knu = dnu * nu;
}
/* RK45_GeodesicEquation<-1>, GeodesicEquationSolver.cpp:78-188.  The stage-6
 * coefficient 11/44 is the reference's (textbook Fehlberg uses 11/40). */
double or_rk45(const or_grid *g, const double *metric, double bh_r2, double tol,
               double dt, double *px, double *py, double *pvx, double *pvy)
{
    or_mv mv = {g, metric, (long)g->nx * g->ny};
    const double td = -1.0;
    double x[2] = {*px, *py}, v[2] = {*pvx, *pvy};
    double k1x[2], k2x[2], k3x[2], k4x[2], k5x[2], k6x[2];
    double k1v[2], k2v[2], k3v[2], k4v[2], k5v[2], k6v[2];
    double k1n, k2n, k3n, k4n, k5n, k6n;
    double xt[2], vt[2], nt;
    double nu = 1, total = 0, ddt = g->dt;
    /* The reference loops while total < dt; near the horizon the step size
     * can collapse so that this never ends.  Like the GPU tables, stop after
     * 100000 step attempts and keep the point reached. */
    int guard = 0;
    while (total < dt && guard++ < 100000) {
        ddt = ddt < dt - total ? ddt : dt - total;
        or_rhs(&mv, x[0], x[1], v[0], v[1], nu, k1x, k1v, &k1n);
        for (int d = 0; d < 2; d++) {
            xt[d] = x[d] + td * ddt * ((1.0 / 4.0) * k1x[d]);
            vt[d] = v[d] + td * ddt * ((1.0 / 4.0) * k1v[d]);
        }
        nt = nu + td * ddt * ((1.0 / 4.0) * k1n);
        or_rhs(&mv, xt[0], xt[1], vt[0], vt[1], nt, k2x, k2v, &k2n);
        for (int d = 0; d < 2; d++) {
            xt[d] = x[d] + td * ddt * ((3.0 / 32.0) * k1x[d] + (9.0 / 32.0) * k2x[d]);
            vt[d] = v[d] + td * ddt * ((3.0 / 32.0) * k1v[d] + (9.0 / 32.0) * k2v[d]);
        }
        nt = nu + td * ddt * ((3.0 / 32.0) * k1n + (9.0 / 32.0) * k2n);
        or_rhs(&mv, xt[0], xt[1], vt[0], vt[1], nt, k3x, k3v, &k3n);
        for (int d = 0; d < 2; d++) {
            xt[d] = x[d] + td * ddt * ((1932.0 / 2197.0) * k1x[d] - (7200.0 / 2197.0) * k2x[d] + (7296.0 / 2197.0) * k3x[d]);
            vt[d] = v[d] + td * ddt * ((1932.0 / 2197.0) * k1v[d] - (7200.0 / 2197.0) * k2v[d] + (7296.0 / 2197.0) * k3v[d]);
        }
        nt = nu + td * ddt * ((1932.0 / 2197.0) * k1n - (7200.0 / 2197.0) * k2n + (7296.0 / 2197.0) * k3n);
        or_rhs(&mv, xt[0], xt[1], vt[0], vt[1], nt, k4x, k4v, &k4n);
        for (int d = 0; d < 2; d++) {
            xt[d] = x[d] + td * ddt * ((439.0 / 216.0) * k1x[d] - (8.0) * k2x[d] + (3680.0 / 513.0) * k3x[d] - (845.0 / 4104.0) * k4x[d]);
            vt[d] = v[d] + td * ddt * ((439.0 / 216.0) * k1v[d] - (8.0) * k2v[d] + (3680.0 / 513.0) * k3v[d] - (845.0 / 4104.0) * k4v[d]);
        }
        nt = nu + td * ddt * ((439.0 / 216.0) * k1n - (8.0) * k2n + (3680.0 / 513.0) * k3n - (845.0 / 4104.0) * k4n);
        or_rhs(&mv, xt[0], xt[1], vt[0], vt[1], nt, k5x, k5v, &k5n);
        for (int d = 0; d < 2; d++) {
            xt[d] = x[d] + td * ddt * (-(8.0 / 27.0) * k1x[d] + (2.0) * k2x[d] - (3544.0 / 2565.0) * k3x[d] + (1859.0 / 4104.0) * k4x[d] - (11.0 / 44.0) * k5x[d]);
            vt[d] = v[d] + td * ddt * (-(8.0 / 27.0) * k1v[d] + (2.0) * k2v[d] - (3544.0 / 2565.0) * k3v[d] + (1859.0 / 4104.0) * k4v[d] - (11.0 / 44.0) * k5v[d]);
        }
        nt = nu + td * ddt * (-(8.0 / 27.0) * k1n + (2.0) * k2n - (3544.0 / 2565.0) * k3n + (1859.0 / 4104.0) * k4n - (11.0 / 44.0) * k5n);
        or_rhs(&mv, xt[0], xt[1], vt[0], vt[1], nt, k6x, k6v, &k6n);
        double dx[2], dv[2], dn;
        for (int d = 0; d < 2; d++) {
            dx[d] = (16.0 / 135.0) * k1x[d] + (6656.0 / 12825.0) * k3x[d] + (28561.0 / 56430.0) * k4x[d] - (9.0 / 50.0) * k5x[d] + (2.0 / 55.0) * k6x[d];
            dv[d] = (16.0 / 135.0) * k1v[d] + (6656.0 / 12825.0) * k3v[d] + (28561.0 / 56430.0) * k4v[d] - (9.0 / 50.0) * k5v[d] + (2.0 / 55.0) * k6v[d];
        }
        dn = (16.0 / 135.0) * k1n + (6656.0 / 12825.0) * k3n + (28561.0 / 56430.0) * k4n - (9.0 / 50.0) * k5n + (2.0 / 55.0) * k6n;
        double TE = 0;
        for (int d = 0; d < 2; d++) {
            double e = ddt * ((1.0 / 360.0) * k1v[d] - (128.0 / 4275.0) * k3v[d] - (2197.0 / 75240.0) * k4v[d] + (1.0 / 50.0) * k5v[d] + (2.0 / 55.0) * k6v[d]);
            TE += e * e;
        }
        TE = sqrt(TE);
        if (TE > tol) {
            ddt *= 0.95 * pow(tol / TE, 1.0 / 4.0);
            continue;
        }
        for (int d = 0; d < 2; d++) {
            x[d] += td * ddt * dx[d];
            v[d] += td * ddt * dv[d];
        }
        nu += td * ddt * dn;
        total += ddt;
        ddt *= 0.95 * pow(tol / (TE + 1e-50), 1.0 / 4.0);
        if (or_inside_bh(bh_r2, x[0], x[1]) || or_outside(g, x[0], x[1])) break;
    }
    *px = x[0]; *py = x[1]; *pvx = v[0]; *pvy = v[1];
    return nu;
}
/* UpdateFourierCoefficients, Radiation.cpp:191-237 */
void or_geodesic_tables(const or_grid *g, const double *metric, int32_t nF, double bh_r2,
                        double tol, double *cS, double *cX, double *cY, double *cCx, double *cCy)
{
    or_mv mv = {g, metric, (long)g->nx * g->ny};
    double wS[64], phiS[64], cxS[64], cyS[64];
    or_stencil(nF, wS, phiS, cxS, cyS);
#pragma omp parallel for collapse(2) schedule(dynamic, 50)
    for (int j = 1; j < g->ny - 1; j++)
        for (int i = 1; i < g->nx - 1; i++) {
            long ij = i + (long)j * g->nx;
            double dS[64], dX[64], dY[64], dCx[64], dCy[64];
            double x0 = g->startx + i * g->dx, y0 = g->starty + j * g->dy;
            double alpha = MA(&mv, OR_ALPHA)[ij];
            for (int d = 0; d < nF; d++) {
                double s = 1, x = x0, y = y0;
                double ux = cxS[d] * alpha, uy = cyS[d] * alpha;
                /* Vec2ObservedByEulObs<IF,LF> at the cell centre (SpecialMath.cpp:159-164) */
                double vx, vy;
                {
                    double fi = (x - g->startx) / g->dx, fj = (y - g->starty) / g->dy;
                    int out = or_outside(g, x, y);
                    double al = out ? 1.0 : or_interp(&mv, OR_ALPHA, fi, fj);
                    double v1 = ux / al, v2 = uy / al;
                    double T11 = out ? 1 : or_interp(&mv, OR_T11, fi, fj);
                    double T12 = out ? 0 : or_interp(&mv, OR_T12, fi, fj);
                    double T21 = out ? 0 : or_interp(&mv, OR_T21, fi, fj);
                    double T22 = out ? 1 : or_interp(&mv, OR_T22, fi, fj);
                    vx = T11 * v1 + T12 * v2;
                    vy = T21 * v1 + T22 * v2;
                }
                if (!or_inside_bh(bh_r2, x, y))
                    s *= or_rk45(g, metric, bh_r2, tol, g->dt, &x, &y, &vx, &vy);
                else
                    vx = vy = 0.0;
                /* TransformLFtoIF with the inverse tetrad at the departure point */
                double fi = (x - g->startx) / g->dx, fj = (y - g->starty) / g->dy;
                int out = or_outside(g, x, y);
                double I11 = out ? 1 : or_interp(&mv, OR_TI11, fi, fj);
                double I12 = out ? 0 : or_interp(&mv, OR_TI12, fi, fj);
                double I21 = out ? 0 : or_interp(&mv, OR_TI21, fi, fj);
                double I22 = out ? 1 : or_interp(&mv, OR_TI22, fi, fj);
                dS[d] = 1.0 / s;
                dX[d] = x;
                dY[d] = y;
                dCx[d] = I11 * vx + I12 * vy;
                dCy[d] = I21 * vx + I22 * vy;
            }
            or_fourier_coefficients(nF, phiS, wS, dS, cS + ij * nF);
            or_fourier_coefficients(nF, phiS, wS, dX, cX + ij * nF);
            or_fourier_coefficients(nF, phiS, wS, dY, cY + ij * nF);
            or_fourier_coefficients(nF, phiS, wS, dCx, cCx + ij * nF);
            or_fourier_coefficients(nF, phiS, wS, dCy, cCy + ij * nF);
        }
}

/* ------------------------------------------------------------------------ */
/* Stream / moments / collide (src/Radiation.cpp)                           */
/* ------------------------------------------------------------------------ */
static inline double or_bilinear(double x, double y, double f00, double f01, double f10, double f11)
{ /* src/Interpolation.cpp:27-37 */
    double f0 = f00 * (1.0 - x) + f10 * x;
    double f1 = f01 * (1.0 - x) + f11 * x;
    return f0 * (1.0 - y) + f1 * y;
}
/* StreamFlatFixed, Radiation.cpp:380-414 */
void or_stream_flat(const or_sim *s, const double *I, double *Inew)
{
    const or_grid *g = &s->grid;
    const int D = s->nDir, nx = g->nx;
#pragma omp parallel for collapse(2) schedule(dynamic, 50)
    for (int j = 1; j < g->ny - 1; j++)
        for (int i = 1; i < g->nx - 1; i++) {
            long ij = i + (long)j * nx;
            for (int d = 0; d < D; d++) {
                double xT = g->startx + i * g->dx, yT = g->starty + j * g->dy;
                xT -= s->cx[d] * g->dt;
                yT -= s->cy[d] * g->dt;
                double iT = (xT - g->startx) / g->dx, jT = (yT - g->starty) / g->dy;
                long i0 = (long)floor(iT), j0 = (long)floor(jT), i1 = i0 + 1, j1 = j0 + 1;
                Inew[ij * D + d] = or_bilinear(iT - i0, jT - j0, I[(i0 + j0 * nx) * D + d],
                                               I[(i0 + j1 * nx) * D + d], I[(i1 + j0 * nx) * D + d],
                                               I[(i1 + j1 * nx) * D + d]);
            }
        }
}
/* IntensityAt, Radiation.cpp:322-343 (fixed streaming: rotation angle 0,
 * R(0) is the identity because MySin(0)=0 and MyCos(0)=1 exactly). */
static inline double or_intensity_at(const or_sim *s, const double *I, long ij, double vx, double vy)
{
    const int D = s->nDir;
    double angle = or_phi(vx, vy);
    double nG = (double)s->nGrid;
    double dd = nG * angle / (2.0 * OR_PI);
    if (dd < 1e-8) dd = 1e-8;
    if (dd > nG - 1e-8) dd = nG - 1e-8;
    long d0 = (long)floor(dd) % s->nGrid;
    long d1 = (d0 + 1) % s->nGrid;
    const int32_t *n0 = s->Nc + 4 * d0, *n1 = s->Nc + 4 * d1;
    const double *w0 = s->Wc + 4 * d0, *w1 = s->Wc + 4 * d1;
    double v0 = 0, v1 = 0;
    for (int k = 0; k < 4; k++) {
        v0 += w0[k] * I[ij * D + n0[k]];
        v1 += w1[k] * I[ij * D + n1[k]];
    }
    double t = dd - d0;
    double r = v0 * (1.0 - t) + v1 * t;
    return r > 0.0 ? r : 0.0; /* std::max(0.0, r) */
}
/* StreamCurvedFixed, Radiation.cpp:454-512 */
void or_stream_curved(const or_sim *s, const double *I, double *Inew)
{
    const or_grid *g = &s->grid;
    const int D = s->nDir, nx = g->nx, nF = s->nCoeff;
    double *basis = (double *)malloc(sizeof(double) * D * nF);
    for (int d = 0; d < D; d++)
        for (int k = 0; k < nF; k++) basis[d * nF + k] = or_fourier_basis(k, s->phi[d]);
#pragma omp parallel for collapse(2) schedule(dynamic, 50)
    for (int j = 1; j < g->ny - 1; j++)
        for (int i = 1; i < g->nx - 1; i++) {
            long ij = i + (long)j * nx;
            double xc = g->startx + i * g->dx, yc = g->starty + j * g->dy;
            if (or_inside_bh(s->bh_r2, xc, yc)) {
                for (int d = 0; d < D; d++) Inew[ij * D + d] = 0;
                continue;
            }
            double alpha = s->alpha[ij];
            for (int d = 0; d < D; d++) {
                const double *B = basis + d * nF;
                double sh = or_fourier_value(B, s->cS + ij * nF, nF);
                double xT = or_fourier_value(B, s->cX + ij * nF, nF);
                double yT = or_fourier_value(B, s->cY + ij * nF, nF);
                double vx = or_fourier_value(B, s->cCx + ij * nF, nF);
                double vy = or_fourier_value(B, s->cCy + ij * nF, nF);
                if (or_inside_bh(s->bh_r2, xT, yT)) { Inew[ij * D + d] = 0; continue; }
                double iT = (xT - g->startx) / g->dx, jT = (yT - g->starty) / g->dy;
                long i0 = (long)floor(iT), j0 = (long)floor(jT), i1 = i0 + 1, j1 = j0 + 1;
                long k00 = i0 + j0 * nx, k01 = i0 + j1 * nx, k10 = i1 + j0 * nx, k11 = i1 + j1 * nx;
                double r00 = alpha / s->alpha[k00], r01 = alpha / s->alpha[k01];
                double r10 = alpha / s->alpha[k10], r11 = alpha / s->alpha[k11];
                double f00 = (r00 * (r00 * r00)) * or_intensity_at(s, I, k00, vx, vy);
                double f01 = (r01 * (r01 * r01)) * or_intensity_at(s, I, k01, vx, vy);
                double f10 = (r10 * (r10 * r10)) * or_intensity_at(s, I, k10, vx, vy);
                double f11 = (r11 * (r11 * r11)) * or_intensity_at(s, I, k11, vx, vy);
                Inew[ij * D + d] = (sh * (sh * sh)) * or_bilinear(iT - i0, jT - j0, f00, f01, f10, f11);
            }
        }
    free(basis);
}
/* ComputeMomentsIF, Radiation.cpp:239-269 */
void or_moments(const or_sim *s, const double *I, double *mom)
{
    const long N = (long)s->grid.nx * s->grid.ny;
    const int D = s->nDir;
    double *E = mom, *Fx = mom + N, *Fy = mom + 2 * N, *Pxx = mom + 3 * N, *Pxy = mom + 4 * N, *Pyy = mom + 5 * N;
#pragma omp parallel for schedule(dynamic, 50)
    for (long ij = 0; ij < N; ij++) {
        double e = 0, fx = 0, fy = 0, pxx = 0, pxy = 0, pyy = 0;
        /* RotationMatrix(rotationAngle[ij]) * C(d) (:257) */
        const double rs = s->rot ? or_mysin(s->rot[ij]) : 0.0, rc = s->rot ? or_mycos(s->rot[ij]) : 1.0;
        for (int d = 0; d < D; d++) {
            if (s->w[d] == 0.0) continue;
            double nx_ = s->cx[d], ny_ = s->cy[d];
            if (s->rot) {
                nx_ = s->cx[d] * rc - s->cy[d] * rs;
                ny_ = s->cx[d] * rs + s->cy[d] * rc;
            }
            double c = s->w[d] * I[ij * D + d];
            e += c;
            fx += c * nx_;
            fy += c * ny_;
            pxx += c * nx_ * nx_;
            pxy += c * nx_ * ny_;
            pyy += c * ny_ * ny_;
        }
        E[ij] = e; Fx[ij] = fx; Fy[ij] = fy; Pxx[ij] = pxx; Pxy[ij] = pxy; Pyy[ij] = pyy;
    }
}
/* ComputeMomentsLF, Radiation.cpp:270-301: tetrad transform of the IF
 * energy-momentum tensor (TransformIFtoLF, SpecialMath.cpp:85-94, same loop
 * order) and F = sqrt(|Norm2(F, gamma_ll)|) (SpecialMath.cpp:20-27); cells
 * with x^2 + y^2 <= r2 (InsideBH) get 0.  tetrad: [9][N] e[a][A] row-major,
 * gamma: [3][N] (11, 12, 22), out: [7][N] (E, Fx, Fy, Pxx, Pxy, Pyy, F). */
void or_moments_lf(const or_grid *g, const double *mom, const double *tetrad, const double *gamma, double r2,
                   double *out)
{
    const long N = (long)g->nx * g->ny;
    for (long ij = 0; ij < N; ij++) {
        const double x = g->startx + (double)(ij % g->nx) * g->dx, y = g->starty + (double)(ij / g->nx) * g->dy;
        if (or_inside_bh(r2, x, y)) {
            for (int k = 0; k < 7; k++) out[k * N + ij] = 0.0;
            continue;
        }
        const double E = mom[ij], Fx = mom[N + ij], Fy = mom[2 * N + ij];
        const double T[3][3] = {{E, Fx, Fy}, {Fx, mom[3 * N + ij], mom[4 * N + ij]},
                                {Fy, mom[4 * N + ij], mom[5 * N + ij]}};
        double R[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int A = 0; A < 3; A++)
                    for (int B = 0; B < 3; B++)
                        R[a][b] += T[A][B] * tetrad[(3 * a + A) * N + ij] * tetrad[(3 * b + B) * N + ij];
        const double gm[3][3] = {{0, 0, 0}, {0, gamma[ij], gamma[N + ij]}, {0, gamma[N + ij], gamma[2 * N + ij]}};
        const double v[3] = {0, R[0][1], R[0][2]};
        double n2 = 0;
        for (int i = 1; i < 3; i++)
            for (int j = 1; j < 3; j++) n2 += gm[i][j] * v[i] * v[j];
        out[ij] = R[0][0];
        out[N + ij] = R[0][1];
        out[2 * N + ij] = R[0][2];
        out[3 * N + ij] = R[1][1];
        out[4 * N + ij] = R[1][2];
        out[5 * N + ij] = R[2][2];
        out[6 * N + ij] = sqrt(fabs(n2));
    }
}
/* Collide, Radiation.cpp:641-724.  Writes Inew at interior active cells,
 * updates the moment arrays in place, and records cycle counts. */
/* The fluid-frame moments of a cycle (src/Radiation.cpp:664-672) from the
 * inertial-frame moments m = (E, Fx, Fy, Pxx, Pxy, Pyy) and the fluid
 * velocity u: out = (E_FF, Fx_FF, Fy_FF) as the reference forms them. */
static inline void fluid_frame_moments(double u1, double u2, double lorentz, const double *m, double *out)
{
    double pE = m[0], pFx = m[1], pFy = m[2], pPxx = m[3], pPxy = m[4], pPyy = m[5];
    double uF = u1 * pFx + u2 * pFy;
    double uxP = u1 * pPxx + u2 * pPxy;
    double uyP = u1 * pPxy + u2 * pPyy;
    double uuP = u1 * uxP + u2 * uyP;
    out[0] = lorentz * lorentz * (pE - 2.0 * uF + uuP);
    out[1] = lorentz * lorentz * (pFx - lorentz * pE * u1 - uxP + u1 * lorentz / (1.0 + lorentz) * ((2.0 * lorentz + 1.0) * uF - lorentz * uuP));
    out[2] = lorentz * lorentz * (pFy - lorentz * pE * u2 - uyP + u2 * lorentz / (1.0 + lorentz) * ((2.0 * lorentz + 1.0) * uF - lorentz * uuP));
}
/* Doppler factor A and the fluid-frame direction nFF of the inertial-frame
 * direction (n1, n2) (src/Radiation.cpp:689-698): out = (nFF1, nFF2, A). */
static inline void fluid_frame_dir(double u1, double u2, double lorentz, double n1, double n2, double *out)
{
    double un = u1 * n1 + u2 * n2;
    double A = lorentz * (1.0 - un);
    /* Tensor2 nFF = (nIF - (1 - (lorentz*un)/(lorentz+1)) * lorentz * uIF) / A */
    double fac = (1 - (lorentz * un) / (lorentz + 1)) * lorentz;
    out[0] = (n1 - fac * u1) / A;
    out[1] = (n2 - fac * u2) / A;
    out[2] = A;
}
void or_fluid_frame(double u1, double u2, const double m[6], double out[3])
{
    fluid_frame_moments(u1, u2, 1.0 / sqrt(1.0 - (u1 * u1 + u2 * u2)), m, out);
}
void or_fluid_dir(double u1, double u2, double n1, double n2, double out[3])
{
    fluid_frame_dir(u1, u2, 1.0 / sqrt(1.0 - (u1 * u1 + u2 * u2)), n1, n2, out);
}

/* Collide of one active cell (src/Radiation.cpp:650-724): Icell and
 * Inew_cell hold its nDir intensities; m holds E, Fx, Fy, Pxx, Pxy, Pyy of
 * Icell on entry and the moments of the last cycle on return.  Returns the
 * cycle count. */
int or_collide_cell(const or_sim *s, long ij, const double *Icell, double *Inew_cell, double m[6])
{
    const or_grid *g = &s->grid;
    const int D = s->nDir;
    double ux = s->ux[ij], uy = s->uy[ij];
    double u1 = s->t11[ij] * ux + s->t12[ij] * uy;
    double u2 = s->t21[ij] * ux + s->t22[ij] * uy;
    double alpha = s->alpha[ij];
    double uu = u1 * u1 + u2 * u2;
    double lorentz = 1.0 / sqrt(1.0 - uu);
    const double rs = s->rot ? or_mysin(s->rot[ij]) : 0.0, rc = s->rot ? or_mycos(s->rot[ij]) : 1.0;
    int cycles = 0;
    double diff = 1;
    while (fabs(diff) > 1e-4 && (cycles < 100)) {
        cycles++;
        double pE = m[0], pFx = m[1], pFy = m[2], pPxx = m[3], pPxy = m[4], pPyy = m[5];
        double ff[3];
        fluid_frame_moments(u1, u2, lorentz, m, ff);
        double fE = ff[0], fFx = ff[1], fFy = ff[2];
        double e = 0, fx = 0, fy = 0, pxx = 0, pxy = 0, pyy = 0;
        for (int d = 0; d < D; d++) {
            double n1 = s->cx[d], n2 = s->cy[d];
            if (s->rot) { /* nIF = RotationMatrix(rotationAngle[ij]) * C(d) (:691) */
                n1 = s->cx[d] * rc - s->cy[d] * rs;
                n2 = s->cx[d] * rs + s->cy[d] * rc;
            }
            double fd[3];
            fluid_frame_dir(u1, u2, lorentz, n1, n2, fd);
            double nf1 = fd[0], nf2 = fd[1], A = fd[2];
            double nF = nf1 * fFx + nf2 * fFy;
            double M = s->kappa0[ij] * fE + 3.0 * s->kappa1[ij] * nF;
            double v = (Icell[d] + alpha * g->dt * (s->eta[ij] + M) / (A * A)) /
                       (1.0 + alpha * g->dt * (s->kappaA[ij] + s->kappa0[ij]) * A);
            Inew_cell[d] = v;
            double c = s->w[d] * v;
            e += c;
            fx += c * n1;
            fy += c * n2;
            pxx += c * n1 * n1;
            pxy += c * n1 * n2;
            pyy += c * n2 * n2;
        }
        m[0] = e; m[1] = fx; m[2] = fy; m[3] = pxx; m[4] = pxy; m[5] = pyy;
        double dE = (pE - e) / e, dFx = (pFx - fx) / fx, dFy = (pFy - fy) / fy;
        double dPxx = (pPxx - pxx) / pxx, dPxy = (pPxy - pxy) / pxy, dPyy = (pPyy - pyy) / pyy;
        diff = sqrt(dE * dE + dFx * dFx + dFy * dFy + dPxx * dPxx + dPxy * dPxy + dPyy * dPyy);
    }
    return cycles;
}

void or_collide(const or_sim *s, const double *I, double *Inew, double *mom, int32_t *count)
{
    const or_grid *g = &s->grid;
    const long N = (long)g->nx * g->ny;
    const int D = s->nDir;
#pragma omp parallel for collapse(2) schedule(dynamic, 50)
    for (int j = 1; j < g->ny - 1; j++)
        for (int i = 1; i < g->nx - 1; i++) {
            double xc = g->startx + i * g->dx, yc = g->starty + j * g->dy;
            if (or_inside_bh(s->bh_r2, xc, yc)) continue;
            long ij = i + (long)j * g->nx;
            double m[6];
            for (int q = 0; q < 6; q++) m[q] = mom[q * N + ij];
            const int cycles = or_collide_cell(s, ij, I + ij * D, Inew + ij * D, m);
            for (int q = 0; q < 6; q++) mom[q * N + ij] = m[q];
            if (count) count[ij] = cycles;
        }
}
/* RunSimulation step loop, Radiation.cpp:784-832 (no output frames). */
int or_simulate(const or_sim *s, int32_t nSteps, double *I, double *moments,
                int32_t *maxIt, double *sumIt, int32_t *itcount)
{
    const long N = (long)s->grid.nx * s->grid.ny;
    const int D = s->nDir;
    double *Inew = (double *)calloc((size_t)N * D, sizeof(double)); /* halo stays 0 */
    double *mom = (double *)calloc((size_t)6 * N, sizeof(double));
    int32_t *count = (int32_t *)calloc((size_t)N, sizeof(int32_t));
    if (!Inew || !mom || !count) { free(Inew); free(mom); free(count); return -1; }
    int32_t mx = 0;
    double sum = 0;
    double *cur = I, *scr = Inew; /* std::swap(I, Inew) by pointer */
    for (int n = 0; n < nSteps; n++) {
        if (s->streaming == 2) or_stream_curved(s, cur, scr);
        else or_stream_flat(s, cur, scr);
        /* after the stream swap: I = streamed (scr), Inew = state (cur) */
        or_moments(s, scr, mom);
        memset(count, 0, sizeof(int32_t) * N);
        or_collide(s, scr, cur, mom, count);
        for (long ij = 0; ij < N; ij++) {
            if (count[ij] > mx) mx = count[ij];
            sum += count[ij];
        }
        if (itcount) memcpy(itcount + (long)n * N, count, sizeof(int32_t) * N);
        /* second swap: state stays in cur, scratch stays scr */
    }
    if (moments) memcpy(moments, mom, sizeof(double) * 6 * N);
    if (maxIt) *maxIt = mx;
    if (sumIt) *sumIt = sum;
    free(Inew); free(mom); free(count);
    return 0;
}
