/*
 * The reference's per-cell lambda iteration of Collide
 * (src/Radiation.cpp:641-724), shared by the reference-order collide
 * (kernels_ref.hip) and the adaptive collide's moving cells
 * (kernels_adaptive.hip).  Both files compile without FP contraction, so the
 * arithmetic is the reference's, operation by operation.
 */
#ifndef RT_REF_COLLIDE_H
#define RT_REF_COLLIDE_H
#include "rt_internal.h"
#include "rt_math.h"

/* Collides active cell ij: reads its streamed intensities I[d][ij] and the
 * moments mom[X][ij] of them, writes Inew[d][ij] and the moments of Inew
 * into mom; returns the cycle count.  rot (adaptive streaming, else null):
 * directions RotationMatrix(rot[ij]) C(d) (:691). */
template <typename T>
__device__ inline int ref_collide_cell(long ij, const T *__restrict__ I, T *__restrict__ Inew, double *__restrict__ mom,
                                       const double *__restrict__ f, const double *__restrict__ alpha,
                                       const double *__restrict__ t11, const double *__restrict__ t12,
                                       const double *__restrict__ t21, const double *__restrict__ t22,
                                       const double *__restrict__ w, const double *__restrict__ cx,
                                       const double *__restrict__ cy, int N, int D, double dt,
                                       const double *__restrict__ rot)
{
    double *E = mom, *Fx = mom + N, *Fy = mom + 2L * N, *Pxx = mom + 3L * N, *Pxy = mom + 4L * N,
           *Pyy = mom + 5L * N;
    const double eta = f[ij], k0 = f[(long)N + ij], k1 = f[2L * N + ij], kA = f[3L * N + ij];
    double ux = f[4L * N + ij], uy = f[5L * N + ij];
    double u1 = t11[ij] * ux + t12[ij] * uy;
    double u2 = t21[ij] * ux + t22[ij] * uy;
    double al = alpha[ij];
    double uu = u1 * u1 + u2 * u2;
    double lorentz = 1.0 / sqrt(1.0 - uu);
    /* nIF = RotationMatrix(rotationAngle[ij]) * C(d) (:691) */
    const double rs = rot ? rtm::my_sin(rot[ij]) : 0.0, rc = rot ? rtm::my_cos(rot[ij]) : 1.0;
    int cycles = 0;
    double diff = 1;
    /* the inputs of the last cycle: Inew is written once, after the loop,
     * with that cycle's arithmetic */
    double lastE = 0, lastFx = 0, lastFy = 0;
    while (fabs(diff) > 1e-4 && (cycles < 100)) {
        cycles++;
        double pE = E[ij], pFx = Fx[ij], pFy = Fy[ij], pPxx = Pxx[ij], pPxy = Pxy[ij], pPyy = Pyy[ij];
        double uF = u1 * pFx + u2 * pFy;
        double uxP = u1 * pPxx + u2 * pPxy;
        double uyP = u1 * pPxy + u2 * pPyy;
        double uuP = u1 * uxP + u2 * uyP;
        double fE = lorentz * lorentz * (pE - 2.0 * uF + uuP);
        double fFx = lorentz * lorentz *
                     (pFx - lorentz * pE * u1 - uxP +
                      u1 * lorentz / (1.0 + lorentz) * ((2.0 * lorentz + 1.0) * uF - lorentz * uuP));
        double fFy = lorentz * lorentz *
                     (pFy - lorentz * pE * u2 - uyP +
                      u2 * lorentz / (1.0 + lorentz) * ((2.0 * lorentz + 1.0) * uF - lorentz * uuP));
        lastE = fE;
        lastFx = fFx;
        lastFy = fFy;
        double e = 0, fx = 0, fy = 0, pxx = 0, pxy = 0, pyy = 0;
        if (u1 == 0.0 && u2 == 0.0) {
            /* Fluid at rest: lorentz = 1, u.n = 0, A = 1, fac = 1 and
             * nFF = n exactly, so dropping the factors and divisions by 1
             * below leaves every rounded value of the general form unchanged
             * (up to the sign of zeros). */
            const double adt = al * dt, den = 1.0 + al * dt * (kA + k0);
            for (int d = 0; d < D; d++) {
                const long idx = (long)d * N + ij;
                double n1 = cx[d], n2 = cy[d];
                if (rot) {
                    n1 = cx[d] * rc - cy[d] * rs;
                    n2 = cx[d] * rs + cy[d] * rc;
                }
                const double nF = n1 * fFx + n2 * fFy;
                const double M = k0 * fE + 3.0 * k1 * nF;
                const double v = ((double)I[idx] + adt * (eta + M)) / den;
                const double c = w[d] * v;
                e += c;
                fx += c * n1;
                fy += c * n2;
                pxx += c * n1 * n1;
                pxy += c * n1 * n2;
                pyy += c * n2 * n2;
            }
        } else
        for (int d = 0; d < D; d++) {
            long idx = (long)d * N + ij;
            double n1 = cx[d], n2 = cy[d];
            if (rot) {
                n1 = cx[d] * rc - cy[d] * rs;
                n2 = cx[d] * rs + cy[d] * rc;
            }
            double un = u1 * n1 + u2 * n2;
            double A = lorentz * (1.0 - un);
            double fac = (1 - (lorentz * un) / (lorentz + 1)) * lorentz;
            double nf1 = (n1 - fac * u1) / A, nf2 = (n2 - fac * u2) / A;
            double nF = nf1 * fFx + nf2 * fFy;
            double M = k0 * fE + 3.0 * k1 * nF;
            double v = ((double)I[idx] + al * dt * (eta + M) / (A * A)) / (1.0 + al * dt * (kA + k0) * A);
            double c = w[d] * v;
            e += c;
            fx += c * n1;
            fy += c * n2;
            pxx += c * n1 * n1;
            pxy += c * n1 * n2;
            pyy += c * n2 * n2;
        }
        E[ij] = e;
        Fx[ij] = fx;
        Fy[ij] = fy;
        Pxx[ij] = pxx;
        Pxy[ij] = pxy;
        Pyy[ij] = pyy;
        double dE = (pE - e) / e, dFx = (pFx - fx) / fx, dFy = (pFy - fy) / fy;
        double dPxx = (pPxx - pxx) / pxx, dPxy = (pPxy - pxy) / pxy, dPyy = (pPyy - pyy) / pyy;
        diff = sqrt(dE * dE + dFx * dFx + dFy * dFy + dPxx * dPxx + dPxy * dPxy + dPyy * dPyy);
    }
    const double fE = lastE, fFx = lastFx, fFy = lastFy;
    if (u1 == 0.0 && u2 == 0.0) {
        const double adt = al * dt, den = 1.0 + al * dt * (kA + k0);
        for (int d = 0; d < D; d++) {
            const long idx = (long)d * N + ij;
            double n1 = cx[d], n2 = cy[d];
            if (rot) {
                n1 = cx[d] * rc - cy[d] * rs;
                n2 = cx[d] * rs + cy[d] * rc;
            }
            const double nF = n1 * fFx + n2 * fFy;
            const double M = k0 * fE + 3.0 * k1 * nF;
            Inew[idx] = (T)(((double)I[idx] + adt * (eta + M)) / den);
        }
    } else {
        for (int d = 0; d < D; d++) {
            long idx = (long)d * N + ij;
            double n1 = cx[d], n2 = cy[d];
            if (rot) {
                n1 = cx[d] * rc - cy[d] * rs;
                n2 = cx[d] * rs + cy[d] * rc;
            }
            double un = u1 * n1 + u2 * n2;
            double A = lorentz * (1.0 - un);
            double fac = (1 - (lorentz * un) / (lorentz + 1)) * lorentz;
            double nf1 = (n1 - fac * u1) / A, nf2 = (n2 - fac * u2) / A;
            double nF = nf1 * fFx + nf2 * fFy;
            double M = k0 * fE + 3.0 * k1 * nF;
            Inew[idx] = (T)(((double)I[idx] + al * dt * (eta + M) / (A * A)) / (1.0 + al * dt * (kA + k0) * A));
        }
    }
    return cycles;
}
#endif
