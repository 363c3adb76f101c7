/*
 * Exact collide of the FP64 step (compiled with -ffp-contract=off).
 *
 * ComputeMomentsIF (src/Radiation.cpp:239-269) followed by Collide
 * (src/Radiation.cpp:641-736) for fixed stencils, one cell per lane, with the
 * cell's streamed intensities held in registers across the lambda cycles.
 * Every rounded value is the reference's: moments summed over d in order,
 * c = w I, then c n1, (c n1) n1 and so on, IEEE division and square root,
 * and the stopping test sqrt(sum ((prev - new) / new)^2) > 1e-4 with its
 * NaN semantics (a moment that is exactly 0 in two consecutive cycles stops
 * the loop).  Those NaN stops are what makes the cycle count of a cell
 * depend on the exact rounding of the moment sums: at an isotropic start the
 * y flux is a sum of rounding noise that comes out exactly 0 in some cells
 * and not in others.
 *
 * The fused FP64 step streams into a buffer of true-unit intensities J
 * (kernels_fused.hip, stream-only form) and this kernel turns J into the next
 * state; for curved streaming the state is stored as alpha^-3 I.
 */
#include "rt_internal.h"

namespace {

typedef const __attribute__((address_space(4))) DirTables CDT;

struct CollideArgs {
    const double *J;        /* [D][N] streamed intensities, true units */
    double *out;            /* [D][N] next state (active cells written) */
    double *mom;            /* [6][N] moments after the last cycle (want_moments) */
    int32_t *count;         /* [N] cycles per cell, 0 for inactive cells */
    const double *cell;     /* [RT_CELL_FIELDS][N] FP64 per-cell inputs */
    const double *alpha;    /* [N] */
    const uint8_t *mask;    /* [N] */
    const DirTables *dt;
    unsigned long long *stats; /* [2] sum, max of the cycle counts */
    double *rec_m;          /* adjoint records: moments of J [6][N], cycles [N] */
    uint8_t *rec_k;
    int N;
    double dt_step;
    int want_moments;
    int scale;              /* store alpha^-3 I (curved fused state) */
};

template <int D, bool REC>
__global__ __launch_bounds__(128) void k_collide_exact(CollideArgs a)
{
    const int ij = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = a.N;
    CDT *dtab = (CDT *)a.dt;
    int cycles = 0;
    const bool active = ij < N && (a.mask[ij] & MASK_ACTIVE);
    if (active) {
        double J[D];
#pragma unroll
        for (int d = 0; d < D; d++) J[d] = a.J[(size_t)d * N + ij];
        const double *cf = a.cell + ij;
        const double eta = cf[(size_t)CF_ETA * N], k0 = cf[(size_t)CF_K0 * N], k1 = cf[(size_t)CF_K1 * N],
                     kA = cf[(size_t)CF_KA * N];
        /* uIF = tetrad * uLF, formed once at set-up in the reference's order */
        const double u1 = cf[(size_t)CF_UX * N], u2 = cf[(size_t)CF_UY * N];
        const double al = a.alpha[ij], dt = a.dt_step;

        /* ComputeMomentsIF: c = w I, then the six sums in d order */
        double E = 0, Fx = 0, Fy = 0, Pxx = 0, Pxy = 0, Pyy = 0;
#pragma unroll
        for (int d = 0; d < D; d++) {
            const double w = dtab->w6[d][0];
            if (w == 0.0) continue;
            const double n1 = dtab->cx[d], n2 = dtab->cy[d];
            const double c = w * J[d];
            E += c;
            Fx += c * n1;
            Fy += c * n2;
            Pxx += c * n1 * n1;
            Pxy += c * n1 * n2;
            Pyy += c * n2 * n2;
        }
        if (REC) {
            const double m0[6] = {E, Fx, Fy, Pxx, Pxy, Pyy};
#pragma unroll
            for (int X = 0; X < 3; X++) a.rec_m[(size_t)X * N + ij] = m0[X]; /* E, Fx, Fy: what the adjoint reads */
        }

        const double uu = u1 * u1 + u2 * u2;
        const double lorentz = 1.0 / sqrt(1.0 - uu);
        const bool rest = u1 == 0.0 && u2 == 0.0;
        double diff = 1;
        double lastE = 0, lastFx = 0, lastFy = 0;
        while (fabs(diff) > 1e-4 && (cycles < 100)) {
            cycles++;
            const double pE = E, pFx = Fx, pFy = Fy, pPxx = Pxx, pPxy = Pxy, pPyy = Pyy;
            const double uF = u1 * pFx + u2 * pFy;
            const double uxP = u1 * pPxx + u2 * pPxy;
            const double uyP = u1 * pPxy + u2 * pPyy;
            const double uuP = u1 * uxP + u2 * uyP;
            const double fE = lorentz * lorentz * (pE - 2.0 * uF + uuP);
            const double fFx = lorentz * lorentz *
                               (pFx - lorentz * pE * u1 - uxP +
                                u1 * lorentz / (1.0 + lorentz) * ((2.0 * lorentz + 1.0) * uF - lorentz * uuP));
            const double fFy = lorentz * lorentz *
                               (pFy - lorentz * pE * u2 - uyP +
                                u2 * lorentz / (1.0 + lorentz) * ((2.0 * lorentz + 1.0) * uF - lorentz * uuP));
            lastE = fE;
            lastFx = fFx;
            lastFy = fFy;
            double e = 0, fx = 0, fy = 0, pxx = 0, pxy = 0, pyy = 0;
            if (rest) {
                /* fluid at rest: lorentz = 1, u.n = 0, A = 1, fac = 1 and
                 * nFF = n, so the factors and divisions by 1 of the general
                 * form leave every rounded value unchanged (up to the sign
                 * of zeros) */
                const double adt = al * dt, den = 1.0 + al * dt * (kA + k0);
#pragma unroll
                for (int d = 0; d < D; d++) {
                    const double n1 = dtab->cx[d], n2 = dtab->cy[d], w = dtab->w6[d][0];
                    const double nF = n1 * fFx + n2 * fFy;
                    const double M = k0 * fE + 3.0 * k1 * nF;
                    const double v = (J[d] + adt * (eta + M)) / den;
                    const double c = w * v;
                    e += c;
                    fx += c * n1;
                    fy += c * n2;
                    pxx += c * n1 * n1;
                    pxy += c * n1 * n2;
                    pyy += c * n2 * n2;
                }
            } else {
#pragma unroll
                for (int d = 0; d < D; d++) {
                    const double n1 = dtab->cx[d], n2 = dtab->cy[d], w = dtab->w6[d][0];
                    const double un = u1 * n1 + u2 * n2;
                    const double A = lorentz * (1.0 - un);
                    const double fac = (1 - (lorentz * un) / (lorentz + 1)) * lorentz;
                    const double nf1 = (n1 - fac * u1) / A, nf2 = (n2 - fac * u2) / A;
                    const double nF = nf1 * fFx + nf2 * fFy;
                    const double M = k0 * fE + 3.0 * k1 * nF;
                    const double v = (J[d] + al * dt * (eta + M) / (A * A)) / (1.0 + al * dt * (kA + k0) * A);
                    const double c = w * v;
                    e += c;
                    fx += c * n1;
                    fy += c * n2;
                    pxx += c * n1 * n1;
                    pxy += c * n1 * n2;
                    pyy += c * n2 * n2;
                }
            }
            E = e;
            Fx = fx;
            Fy = fy;
            Pxx = pxx;
            Pxy = pxy;
            Pyy = pyy;
            const double dE = (pE - e) / e, dFx = (pFx - fx) / fx, dFy = (pFy - fy) / fy;
            const double dPxx = (pPxx - pxx) / pxx, dPxy = (pPxy - pxy) / pxy, dPyy = (pPyy - pyy) / pyy;
            diff = sqrt(dE * dE + dFx * dFx + dFy * dFy + dPxx * dPxx + dPxy * dPxy + dPyy * dPyy);
        }

        /* Inew of the last cycle, from that cycle's inputs in its operation
         * order (the reference writes it every cycle; the last write stays) */
        const double fE = lastE, fFx = lastFx, fFy = lastFy;
        const double sc = cf[(size_t)CF_IA3 * N];
        const bool scale = a.scale && sc > 0.0 && sc < INFINITY;
        if (rest) {
            const double adt = al * dt, den = 1.0 + al * dt * (kA + k0);
#pragma unroll
            for (int d = 0; d < D; d++) {
                const double n1 = dtab->cx[d], n2 = dtab->cy[d];
                const double nF = n1 * fFx + n2 * fFy;
                const double M = k0 * fE + 3.0 * k1 * nF;
                const double v = (J[d] + adt * (eta + M)) / den;
                a.out[(size_t)d * N + ij] = scale ? v * sc : v;
            }
        } else {
#pragma unroll
            for (int d = 0; d < D; d++) {
                const double n1 = dtab->cx[d], n2 = dtab->cy[d];
                const double un = u1 * n1 + u2 * n2;
                const double A = lorentz * (1.0 - un);
                const double fac = (1 - (lorentz * un) / (lorentz + 1)) * lorentz;
                const double nf1 = (n1 - fac * u1) / A, nf2 = (n2 - fac * u2) / A;
                const double nF = nf1 * fFx + nf2 * fFy;
                const double M = k0 * fE + 3.0 * k1 * nF;
                const double v = (J[d] + al * dt * (eta + M) / (A * A)) / (1.0 + al * dt * (kA + k0) * A);
                a.out[(size_t)d * N + ij] = scale ? v * sc : v;
            }
        }
        if (a.want_moments) {
            const double m1[6] = {E, Fx, Fy, Pxx, Pxy, Pyy};
#pragma unroll
            for (int X = 0; X < 6; X++) a.mom[(size_t)X * N + ij] = m1[X];
        }
        if (REC) a.rec_k[ij] = (uint8_t)cycles;
    }
    if (ij < N && a.count) a.count[ij] = cycles;

    /* iteration statistics: one atomic pair per wavefront */
    unsigned long long s = (unsigned long long)cycles, m = s;
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        const unsigned long long o = __shfl_xor(m, off);
        m = m > o ? m : o;
    }
    if ((threadIdx.x & 63) == 0 && m > 0) {
        unsigned long long *st = rt_stat_slot(a.stats, (unsigned)(ij >> 6));
        atomicAdd(&st[0], s);
        atomicMax(&st[1], m);
    }
}

template <bool REC>
int launch_d(rt_ctx *c, const CollideArgs &a)
{
    const unsigned nb = (unsigned)((c->N + 127) / 128);
    switch (c->D) {
    case 16: k_collide_exact<16, REC><<<nb, 128, 0, c->stream>>>(a); break;
    case 20: k_collide_exact<20, REC><<<nb, 128, 0, c->stream>>>(a); break;
    case 24: k_collide_exact<24, REC><<<nb, 128, 0, c->stream>>>(a); break;
    case 32: k_collide_exact<32, REC><<<nb, 128, 0, c->stream>>>(a); break;
    case 48: k_collide_exact<48, REC><<<nb, 128, 0, c->stream>>>(a); break;
    case 64: k_collide_exact<64, REC><<<nb, 128, 0, c->stream>>>(a); break;
    default: return rt_fail(c, RT_ERR_UNSUPPORTED, "exact collide: unsupported nDir");
    }
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}

} // namespace

int launch_collide_exact(rt_ctx *c, const void *J, void *out, bool want_moments, const RecordPtrs *rec)
{
    if (c->fp32) return rt_fail(c, RT_ERR_STATE, "exact collide: FP64 contexts only");
    CollideArgs a;
    a.J = (const double *)J;
    a.out = (double *)out;
    a.mom = c->d_mom;
    a.count = c->d_count;
    a.cell = (const double *)c->d_cell;
    a.alpha = c->d_alpha;
    a.mask = c->d_mask;
    a.dt = c->d_dirt;
    a.stats = c->d_stats;
    a.rec_m = rec ? rec->m : nullptr;
    a.rec_k = rec ? rec->k : nullptr;
    a.N = c->N;
    a.dt_step = c->desc.dt;
    a.want_moments = want_moments ? 1 : 0;
    a.scale = c->scaled ? 1 : 0;
    return rec ? launch_d<true>(c, a) : launch_d<false>(c, a);
}
