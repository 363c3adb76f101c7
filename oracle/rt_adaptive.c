/*
 * Adaptive streaming in the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see rt_oracle.h).  Restates, in the reference's
 * operation order:
 *   - Stencil(nOrder, nGhost) with ghost directions (src/Stencil.cpp:177-297,
 *     SortDirections :299-332), MyAcos (src/Utility.hh:365-471);
 *   - UpdateRotationMatrizes and AverageF (src/Radiation.cpp:345-378);
 *   - StreamFlatAdaptive (:415-453) and StreamCurvedAdaptive (:513-575) with
 *     IntensityAt at a rotated neighbour (:322-343) and RotationMatrix
 *     (src/DataTypes.hh:620-635);
 *   - the RunSimulation loop for the two adaptive streaming types (:784-832).
 * The rotated ComputeMomentsIF and Collide are or_moments / or_collide with
 * or_sim.rot set (rt_oracle.c).
 *
 * Pins (tests/test_oracle_adaptive.py): ghost stencils against closed forms
 * (fallback bucket for nOrder = 4, ghost count and spread per bucket for the
 * reference's own stencils), rotation update against exact atan2 within the
 * Hastings error, zero rotation against fixed streaming (bit for bit), a
 * uniform rotation by whole stencil steps against relabelled fixed streaming,
 * rotated moments of single-direction fields, and curved adaptive streaming
 * on exact flat-space tables against flat adaptive streaming.
 */
#include "rt_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.14159265358979323846
#define OR_PI_2 1.57079632679489661923

/* MyAsin<4> (src/Utility.hh:386-395), Hastings (1955) arcsine. */
static double or_myasin4(double x)
{
    double sign = (double)((0.0 < x) - (x < 0.0));
    x = fabs(x);
    const double a0 = 1.57078786, a1 = -0.21412453, a2 = 0.08466649, a3 = -0.03575663, a4 = 0.00864884;
    double p = a0 + x * (a1 + x * (a2 + x * (a3 + x * a4)));
    return sign * (OR_PI_2 - sqrt(1.0 - x) * p);
}
/* MyAcos(double) -> MyAcos<4>(float) (src/Utility.hh:453-471): float in, float out. */
double or_myacos(double xd)
{
    float x = (float)xd;
    float r = (x >= 0) ? (float)(OR_PI_2 - or_myasin4(x)) : (float)(OR_PI_2 + or_myasin4(-x));
    return (double)r;
}

void or_rotate(double angle, double x, double y, double *ox, double *oy)
{ /* RotationMatrix * Tensor2, src/DataTypes.hh:630-635 */
    double s = or_mysin(angle), c = or_mycos(angle);
    *ox = x * c - y * s;
    *oy = x * s + y * c;
}
static inline double or_inverse_angle(double angle)
{ /* RotationMatrix::Inverse, src/DataTypes.hh:625-628 */
    return fmod(-angle + 2.0 * OR_PI, 2.0 * OR_PI);
}
static inline double or_min(double a, double b) { return (b < a) ? b : a; } /* std::min */

int or_stencil_ghost(int32_t nOrder, int32_t nGhost, double *w, double *phi, double *cx, double *cy)
{
    if (nOrder < 1 || nGhost < 0) return -1;
    const int nDir = nOrder + nGhost;
    /* Fourier directions (src/Stencil.cpp:193-200) */
    for (int d = 0; d < nOrder; d++) {
        w[d] = 1.0 / nOrder;
        phi[d] = 2.0 * OR_PI * (d + 0.5) / nOrder;
        cx[d] = or_mycos(phi[d]);
        cy[d] = or_mysin(phi[d]);
    }
    if (nGhost > 0) {
        if (nOrder % 2 != 0) return -1; /* :204-205 */
        const double deltaPhi = OR_PI / 8.0 + 1e-8; /* src/Stencil.h:78 */
        double *bs = (double *)malloc(sizeof(double) * (nOrder + 1));
        double *be = (double *)malloc(sizeof(double) * (nOrder + 1));
        int *bc = (int *)malloc(sizeof(int) * (nOrder + 1));
        double *pg = (double *)malloc(sizeof(double) * nGhost);
        int nb = 0;
        /* buckets between neighbouring directions within deltaPhi of pi (:207-225) */
        for (int d = 0; d < nOrder; d++) {
            double phi0 = phi[d], phi1 = phi[(d + 1) % nOrder];
            double delta0 = fabs(phi0 - OR_PI), delta1 = fabs(phi1 - OR_PI);
            double dist0 = or_min(delta0, 2.0 * OR_PI - delta0);
            double dist1 = or_min(delta1, 2.0 * OR_PI - delta1);
            if (dist0 <= deltaPhi && dist1 <= deltaPhi) {
                bs[nb] = phi0;
                be[nb] = phi1;
                bc[nb] = 0;
                nb++;
            }
        }
        if (nb == 0) { /* :226-232 */
            bs[0] = OR_PI - phi[0];
            be[0] = OR_PI + phi[0];
            bc[0] = 0;
            nb = 1;
        }
        const double minAngle = bs[0], maxAngle = be[nb - 1];
        /* ghost counts per bucket from the arccos distribution (:238-257) */
        for (int i = 0; i < nGhost; i++) {
            double x = (i + 0.5) / nGhost;
            double D = or_myacos(1.0 - 2.0 * x) / OR_PI;
            double angle = minAngle + (maxAngle - minAngle) * D;
            for (int j = 0; j < nb; j++) {
                double center = (bs[j] + be[j]) / 2.0;
                double size = be[j] - bs[j];
                double delta = fabs(angle - center);
                double dist = or_min(delta, 2.0 * OR_PI - delta);
                if (dist < size / 2.0) bc[j]++;
            }
        }
        int total = 0;
        for (int j = 0; j < nb; j++) total += bc[j];
        /* the reference writes past (or leaves unset) its ghost array when the
         * counts do not add up; such stencils are rejected here */
        int rc = (total == nGhost) ? 0 : -2;
        if (rc == 0) {
            int index = 0; /* :260-266 */
            for (int i = 0; i < nb; i++)
                for (int j = 0; j < bc[i]; j++) {
                    pg[index] = bs[i] + (be[i] - bs[i]) * (j + 0.5) / bc[i] + OR_PI;
                    index++;
                }
            for (int d = 0; d < nGhost; d++) { /* :268-277 */
                const int k = nOrder + d;
                w[k] = 0;
                phi[k] = fmod(pg[d] + 2.0 * OR_PI, 2.0 * OR_PI);
                cx[k] = or_mycos(phi[k]);
                cy[k] = or_mysin(phi[k]);
            }
        }
        free(bs);
        free(be);
        free(bc);
        free(pg);
        if (rc != 0) return rc;
    }
    /* SortDirections (:299-332): ascending phi (all angles are distinct) */
    for (int a = 1; a < nDir; a++) {
        double tw = w[a], tp = phi[a], tx = cx[a], ty = cy[a];
        int b = a - 1;
        while (b >= 0 && phi[b] > tp) {
            w[b + 1] = w[b];
            phi[b + 1] = phi[b];
            cx[b + 1] = cx[b];
            cy[b + 1] = cy[b];
            b--;
        }
        w[b + 1] = tw;
        phi[b + 1] = tp;
        cx[b + 1] = tx;
        cy[b + 1] = ty;
    }
    return 0;
}

/* UpdateRotationMatrizes with AverageF (src/Radiation.cpp:345-378): interior
 * cells only; mom is [6][N] (E, Fx, Fy, ...). */
void or_update_rotation(const or_sim *s, const double *mom, const double *rot, double *rotNew)
{
    const or_grid *g = &s->grid;
    const long N = (long)g->nx * g->ny;
    const double *E = mom, *Fx = mom + N, *Fy = mom + 2 * N;
    for (int j = 1; j < g->ny - 1; j++)
        for (int i = 1; i < g->nx - 1; i++) {
            long ij = i + (long)j * g->nx;
            double fx = 0.0, fy = 0.0;
            for (int b = -1; b <= 1; b++)
                for (int a = -1; a <= 1; a++) {
                    long k = (i + a) + (long)(j + b) * g->nx;
                    fx += Fx[k];
                    fy += Fy[k];
                }
            fx /= 9.0;
            fy /= 9.0;
            double norm = sqrt(fx * fx + fy * fy);
            if (E[ij] > 1e-16 && norm / E[ij] > 0.01) rotNew[ij] = or_phi(fx, fy);
            else rotNew[ij] = rot[ij];
        }
}

static inline double or_bilinear_a(double x, double y, double f00, double f01, double f10, double f11)
{ /* src/Interpolation.cpp:27-37 */
    double f0 = f00 * (1.0 - x) + f10 * x;
    double f1 = f01 * (1.0 - x) + f11 * x;
    return f0 * (1.0 - y) + f1 * y;
}
/* IntensityAt (src/Radiation.cpp:322-343) of cell k for a velocity given in
 * the inertial frame: rotate by the inverse of the cell's rotation angle,
 * then cubic-in-angle interpolation of the cell's intensities, clamped at 0. */
static double or_intensity_at_rot(const or_sim *s, const double *I, long k, double vx, double vy, double rotk)
{
    const int D = s->nDir;
    double x, y;
    or_rotate(or_inverse_angle(rotk), vx, vy, &x, &y);
    double angle = or_phi(x, y);
    double nG = (double)s->nGrid;
    double dd = nG * angle / (2.0 * OR_PI);
    if (dd < 1e-8) dd = 1e-8;
    if (dd > nG - 1e-8) dd = nG - 1e-8;
    long d0 = (long)floor(dd) % s->nGrid;
    long d1 = (d0 + 1) % s->nGrid;
    const int32_t *n0 = s->Nc + 4 * d0, *n1 = s->Nc + 4 * d1;
    const double *w0 = s->Wc + 4 * d0, *w1 = s->Wc + 4 * d1;
    double v0 = 0, v1 = 0;
    for (int q = 0; q < 4; q++) {
        v0 += w0[q] * I[k * D + n0[q]];
        v1 += w1[q] * I[k * D + n1[q]];
    }
    double t = dd - d0;
    double r = v0 * (1.0 - t) + v1 * t;
    return r > 0.0 ? r : 0.0;
}
static inline int or_bh(double r2lim, double x, double y) { return r2lim >= 0.0 && (x * x + y * y) <= r2lim; }

/* StreamFlatAdaptive (src/Radiation.cpp:415-453): cell directions rotated by
 * s->rotNew, neighbour intensities read in the frames of s->rot. */
void or_stream_flat_adaptive(const or_sim *s, const double *I, double *Inew)
{
    const or_grid *g = &s->grid;
    const int D = s->nDir, nx = g->nx;
#pragma omp parallel for collapse(2) schedule(dynamic, 50)
    for (int j = 1; j < g->ny - 1; j++)
        for (int i = 1; i < g->nx - 1; i++) {
            long ij = i + (long)j * nx;
            for (int d = 0; d < D; d++) {
                double ux, uy;
                or_rotate(s->rotNew[ij], s->cx[d], s->cy[d], &ux, &uy);
                double xT = g->startx + i * g->dx, yT = g->starty + j * g->dy;
                xT -= ux * g->dt;
                yT -= uy * g->dt;
                double iT = (xT - g->startx) / g->dx, jT = (yT - g->starty) / g->dy;
                long i0 = (long)floor(iT), j0 = (long)floor(jT), i1 = i0 + 1, j1 = j0 + 1;
                long k00 = i0 + j0 * nx, k01 = i0 + j1 * nx, k10 = i1 + j0 * nx, k11 = i1 + j1 * nx;
                double f00 = or_intensity_at_rot(s, I, k00, ux, uy, s->rot[k00]);
                double f01 = or_intensity_at_rot(s, I, k01, ux, uy, s->rot[k01]);
                double f10 = or_intensity_at_rot(s, I, k10, ux, uy, s->rot[k10]);
                double f11 = or_intensity_at_rot(s, I, k11, ux, uy, s->rot[k11]);
                Inew[ij * D + d] = or_bilinear_a(iT - i0, jT - j0, f00, f01, f10, f11);
            }
        }
}

/* StreamCurvedAdaptive (src/Radiation.cpp:513-575): the Fourier series are
 * evaluated at the rotated direction's angle (GetFrequencyShift,
 * GetTempCoordinate, GetTemp2VelocityIF, :303-320). */
void or_stream_curved_adaptive(const or_sim *s, const double *I, double *Inew)
{
    const or_grid *g = &s->grid;
    const int D = s->nDir, nx = g->nx, nF = s->nCoeff;
#pragma omp parallel for collapse(2) schedule(dynamic, 50)
    for (int j = 1; j < g->ny - 1; j++)
        for (int i = 1; i < g->nx - 1; i++) {
            long ij = i + (long)j * nx;
            double xc = g->startx + i * g->dx, yc = g->starty + j * g->dy;
            if (or_bh(s->bh_r2, xc, yc)) {
                for (int d = 0; d < D; d++) Inew[ij * D + d] = 0;
                continue;
            }
            double alpha = s->alpha[ij];
            for (int d = 0; d < D; d++) {
                double ux, uy;
                or_rotate(s->rotNew[ij], s->cx[d], s->cy[d], &ux, &uy);
                double angle = or_phi(ux, uy);
                double B[16];
                for (int k = 0; k < nF; k++) B[k] = or_fourier_basis(k, angle);
                double q[5];
                const double *tab[5] = {s->cS, s->cX, s->cY, s->cCx, s->cCy};
                for (int m = 0; m < 5; m++) { /* Fourier::GetValue, src/FourierHarmonics.cpp:56-62 */
                    double r = 0;
                    for (int k = 0; k < nF; k++) r += tab[m][ij * nF + k] * B[k];
                    q[m] = r;
                }
                double sh = q[0], xT = q[1], yT = q[2], vx = q[3], vy = q[4];
                if (or_bh(s->bh_r2, xT, yT)) {
                    Inew[ij * D + d] = 0;
                    continue;
                }
                double iT = (xT - g->startx) / g->dx, jT = (yT - g->starty) / g->dy;
                long i0 = (long)floor(iT), j0 = (long)floor(jT), i1 = i0 + 1, j1 = j0 + 1;
                long k00 = i0 + j0 * nx, k01 = i0 + j1 * nx, k10 = i1 + j0 * nx, k11 = i1 + j1 * nx;
                double r00 = alpha / s->alpha[k00], r01 = alpha / s->alpha[k01];
                double r10 = alpha / s->alpha[k10], r11 = alpha / s->alpha[k11];
                double f00 = (r00 * (r00 * r00)) * or_intensity_at_rot(s, I, k00, vx, vy, s->rot[k00]);
                double f01 = (r01 * (r01 * r01)) * or_intensity_at_rot(s, I, k01, vx, vy, s->rot[k01]);
                double f10 = (r10 * (r10 * r10)) * or_intensity_at_rot(s, I, k10, vx, vy, s->rot[k10]);
                double f11 = (r11 * (r11 * r11)) * or_intensity_at_rot(s, I, k11, vx, vy, s->rot[k11]);
                Inew[ij * D + d] = (sh * (sh * sh)) * or_bilinear_a(iT - i0, jT - j0, f00, f01, f10, f11);
            }
        }
}

/* RunSimulation loop for FlatAdaptive (streaming 1) and CurvedAdaptive (3)
 * (src/Radiation.cpp:784-832): rotation update, adaptive stream, swap of I
 * and of the angle buffers, rotated moments, rotated collide, swap. */
int or_simulate_adaptive(const or_sim *s, int32_t nSteps, double *I, double *mom, double *rot, double *rotNew,
                         int32_t *maxIt, double *sumIt, int32_t *itcount)
{
    if (s->streaming != 1 && s->streaming != 3) return -1;
    const long N = (long)s->grid.nx * s->grid.ny;
    const int D = s->nDir;
    double *Inew = (double *)calloc((size_t)N * D, sizeof(double));
    int32_t *count = (int32_t *)calloc((size_t)N, sizeof(int32_t));
    if (!Inew || !count) {
        free(Inew);
        free(count);
        return -1;
    }
    or_sim t = *s;
    double *cr = rot, *cn = rotNew;
    int32_t mx = 0;
    double sum = 0;
    for (int n = 0; n < nSteps; n++) {
        or_update_rotation(&t, mom, cr, cn);
        t.rot = cr;
        t.rotNew = cn;
        if (s->streaming == 3) or_stream_curved_adaptive(&t, I, Inew);
        else or_stream_flat_adaptive(&t, I, Inew);
        double *tmp = cr;
        cr = cn;
        cn = tmp;
        t.rot = cr;
        t.rotNew = cn;
        or_moments(&t, Inew, mom);
        memset(count, 0, sizeof(int32_t) * N);
        or_collide(&t, Inew, I, mom, count);
        for (long ij = 0; ij < N; ij++) {
            if (count[ij] > mx) mx = count[ij];
            sum += count[ij];
        }
        if (itcount) memcpy(itcount + (long)n * N, count, sizeof(int32_t) * N);
    }
    if (cr != rot) { /* hand back rotationAngle in rot and rotationAngleNew in rotNew */
        for (long ij = 0; ij < N; ij++) {
            double a = rot[ij];
            rot[ij] = rotNew[ij];
            rotNew[ij] = a;
        }
    }
    if (maxIt) *maxIt = mx;
    if (sumIt) *sumIt = sum;
    free(Inew);
    free(count);
    return 0;
}
