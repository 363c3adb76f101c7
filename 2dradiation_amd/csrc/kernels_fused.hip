/*
 * Fused stream + moments + collide step for gfx950.
 *
 * One workgroup = 4 wavefronts walks a 64-cell-wide column strip of the grid
 * down a chunk of rows.  Three input rows (all directions) live in an LDS ring
 * [slot][direction][66] (64 cells plus one halo cell each side, directions
 * 0..3 repeated after D-1); the row two ahead is loaded straight from HBM
 * into the slot that just died (global_load_lds, no register staging) while
 * the current row is computed.  Lane = cell column, wavefront w = direction
 * block [w*D/4, (w+1)*D/4).
 *
 * Stream (reference src/Radiation.cpp:380-414 flat, :454-512 curved):
 *   flat:   bilinear of the same direction at the departure point x - c dt.
 *   curved: the five geodesic Fourier series (S, X, Y, Cx, Cy) are evaluated
 *           at phi_d with the per-direction basis from scalar registers; the
 *           departure velocity angle (Hastings atan2) picks the cubic angular
 *           window; the lerp between the two cubic stencils is folded into 5
 *           combined taps; each of the 4 bilinear corners is clamped >= 0.
 *           The state is held as alpha^-3 I while stepping, so the reference's
 *           (alpha/alpha_k)^3 factor needs no per-corner work.
 * Moments (src/Radiation.cpp:239-269): per wavefront partial sums in FP64,
 *   combined across the 4 wavefronts through LDS in a fixed order.
 * Collide (src/Radiation.cpp:641-736): for cells with zero fluid velocity
 *   every lambda cycle is affine in the moments,
 *     Inew_d = q I_d + b + c (n_d . F_prev),  q = 1/(1 + a dt (kA+k0)),
 *     b = a dt q (eta + k0 E_prev),  c = 3 k1 a dt q,
 *   so the cycle runs on the six moments alone using stencil sums
 *   S0[X] = sum w X, Sx[X] = sum w nx X, Sy[X] = sum w ny X, with the
 *   reference's stopping rule; the intensities are formed once at the end.
 *   Rows containing a moving cell run the reference's per-direction cycle
 *   with a workgroup-wide moment exchange per cycle.
 *
 * FP32-mixed accuracy: tables that every cell shares (angular weights, the
 * Fourier basis) are the one FP32 error that does not average out along a
 * ray, so the angular weight rows sum exactly to 1 and 0 in FP32 and the
 * ray factor s enters as s - 1.
 */
#include "fused_common.h"

namespace {
using namespace rtf;

template <typename T>
struct KP {
    const T *Iin;
    T *Iout;
    const T *cell;              /* [RT_CELL_FIELDS][N] */
    const uint8_t *mask;
    const double *coef;         /* [25][N] FP64 */
    const unsigned long long *excl; /* [N] excised-departure bits */
    double *mom;                /* [6][N] */
    unsigned long long *stats;  /* [2] */
    const DirTables *dt;
    int nx, ny;
    long N;
    int strips, chunks, rows_per_block, nblocks;
    double startx, starty, dx, dy, dtime, r2;
    int want_moments;
    /* adjoint recording (k_fused<..., REC = true>): moments of the streamed
     * intensities [6][N], lambda cycle counts [N], corner clamp bits [4][N] */
    double *rec_m;
    uint8_t *rec_k;
    unsigned long long *rec_mask;
};

/* LDS of one workgroup.  Intensity rows are SoA [direction][66 columns]
 * with 4 wrap-around direction rows (D..D+3 repeat 0..3), so the 5 angular
 * taps of a corner are 5 rows of one column at constant offsets; lanes
 * (= columns) hit consecutive banks.  Rows and FP64 coefficients arrive by
 * direct global-to-LDS loads, without passing through registers. */
template <typename T, int D, bool CURVED>
struct Smem {
    static constexpr int RW = 66;
    static constexpr int NC = CURVED ? 26 : 1; /* 25 Fourier coefficients + exclusion bits */
    T I[3][D + 4][RW];
    double C[NC][64];
    double X[RT_FUSED_WAVES][6][64];
    float wrow[10][12] __attribute__((aligned(16)));
};

/* Reference lambda cycle for rows that contain a moving cell
 * (src/Radiation.cpp:661-722), with a workgroup moment exchange per cycle.
 * Every wavefront evaluates the same per-cell decisions, so the trip count is
 * uniform across the workgroup. */
template <typename T, int DW>
__device__ inline void collide_moving(T (&sv)[DW], double (&mOut)[6], int &cycles, const double (&mI)[6],
                                      bool active, double eta, double k0, double k1, double kA, double u1,
                                      double u2, double adt, int wv, int lane, CDirTables *dtab,
                                      double (*X)[6][64])
{
    const double uu = u1 * u1 + u2 * u2;
    const double lor = 1.0 / sqrt(1.0 - uu);
    double prev[6];
#pragma unroll
    for (int q = 0; q < 6; q++) prev[q] = mI[q];
    double diff = 1;
    T fin[DW];
#pragma unroll
    for (int k = 0; k < DW; k++) fin[k] = sv[k];
    while (true) {
        const bool iter = active && fabs(diff) > 1e-4 && cycles < 100;
        if (!__any(iter)) break;
        double part[6] = {0, 0, 0, 0, 0, 0};
        if (iter) {
            cycles++;
            const double uF = u1 * prev[1] + u2 * prev[2];
            const double uxP = u1 * prev[3] + u2 * prev[4];
            const double uyP = u1 * prev[4] + u2 * prev[5];
            const double uuP = u1 * uxP + u2 * uyP;
            const double fE = lor * lor * (prev[0] - 2.0 * uF + uuP);
            const double fFx = lor * lor * (prev[1] - lor * prev[0] * u1 - uxP +
                                            u1 * lor / (1.0 + lor) * ((2.0 * lor + 1.0) * uF - lor * uuP));
            const double fFy = lor * lor * (prev[2] - lor * prev[0] * u2 - uyP +
                                            u2 * lor / (1.0 + lor) * ((2.0 * lor + 1.0) * uF - lor * uuP));
#pragma unroll
            for (int k = 0; k < DW; k++) {
                const int d = wv * DW + k;
                const double n1 = dtab->cx[d], n2 = dtab->cy[d];
                const double un = u1 * n1 + u2 * n2;
                const double A = lor * (1.0 - un);
                const double fac = (1 - (lor * un) / (lor + 1)) * lor;
                const double nF = ((n1 - fac * u1) / A) * fFx + ((n2 - fac * u2) / A) * fFy;
                const double M = k0 * fE + 3.0 * k1 * nF;
                const double v = ((double)sv[k] + adt * (eta + M) / (A * A)) / (1.0 + adt * (kA + k0) * A);
                fin[k] = (T)v;
#pragma unroll
                for (int q = 0; q < 6; q++) part[q] += dtab->w6[d][q] * v;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 6; q++) X[wv][q][lane] = part[q];
        __syncthreads();
        if (iter) {
            double ss = 0;
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const double nw = ((X[0][q][lane] + X[1][q][lane]) + X[2][q][lane]) + X[3][q][lane];
                const double r = (prev[q] - nw) / nw;
                ss += r * r;
                prev[q] = nw;
            }
            diff = sqrt(ss);
        }
    }
#pragma unroll
    for (int q = 0; q < 6; q++) mOut[q] = prev[q];
#pragma unroll
    for (int k = 0; k < DW; k++) sv[k] = fin[k];
}

/* two workgroups per CU where their LDS fits in 160 KB */
template <typename T, int D, bool CURVED>
constexpr int fused_min_blocks()
{
    return sizeof(Smem<T, D, CURVED>) <= 80 * 1024 ? 2 : 1;
}

template <typename T, int D, bool CURVED, bool REC>
__global__ __launch_bounds__(256, (fused_min_blocks<T, D, CURVED>())) void k_fused(KP<T> p)
{
    constexpr int DW = D / RT_FUSED_WAVES;
    constexpr int RW = Smem<T, D, CURVED>::RW;
    constexpr int NC = Smem<T, D, CURVED>::NC;
    static_assert(DW <= 32, "exclusion bits of one wavefront must fit 32 bits");
    __shared__ Smem<T, D, CURVED> sm;

    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    CDirTables *dtab0 = (CDirTables *)p.dt;
    const long N = p.N;
    const int nx = p.nx;

    /* XCD-aware tile order: consecutive workgroups are dealt round-robin over
     * the 8 XCDs, so give each XCD a contiguous run of tiles (vertically
     * adjacent chunks of a strip share their halo rows in that XCD's L2). */
    int bid = blockIdx.x;
    int tile = bid;
    if ((p.nblocks & 7) == 0) tile = (bid & 7) * (p.nblocks >> 3) + (bid >> 3);
    const int strip = tile / p.chunks;
    const int chunk = tile % p.chunks;
    const int col0 = 64 * strip;                /* grid column of LDS column 0 */
    const int gi = col0 + 1 + lane;             /* this lane's cell column */
    const bool inside = gi <= nx - 2;
    const int rowA = 1 + chunk * p.rows_per_block;
    const int rowB = min(rowA + p.rows_per_block, p.ny - 1); /* exclusive */
    if (rowA >= rowB) return;
    const int ncols = min(RW, nx - col0);       /* valid LDS columns */

    if constexpr (CURVED && sizeof(T) == 4) {
        if (threadIdx.x < 120) (&sm.wrow[0][0])[threadIdx.x] = (&dtab0->wrowf[0][0])[threadIdx.x];
    }

    /* ---- asynchronous staging (direct global -> LDS) ---- */
    auto stage_row = [&](int row) {
        const int slot = row % 3;
        const int ndw = (int)(ncols * sizeof(T) / 4);
        for (int r = wv; r < D + 4; r += RT_FUSED_WAVES) {
            const int d = r < D ? r : r - D;
            const T *src = p.Iin + (size_t)d * N + (size_t)row * nx + col0;
            lds_dma_dwords(src, &sm.I[slot][r][0], ndw, lane);
        }
    };
    auto stage_coef = [&](int row) {
        if constexpr (CURVED) {
            const int n = (int)(min(64, nx - 1 - (col0 + 1)) * 2); /* interior cells of the strip */
            const size_t off = (size_t)row * nx + col0 + 1;
            for (int e = wv; e < NC; e += RT_FUSED_WAVES) {
                const void *src = e < 25 ? (const void *)(p.coef + (size_t)e * N + off) : (const void *)(p.excl + off);
                lds_dma_dwords(src, &sm.C[e][0], n, lane);
            }
        }
    };

    /* prologue: rows rowA-1, rowA, rowA+1 and the coefficients of rowA */
    for (int r = rowA - 1; r <= rowA + 1; r++) stage_row(r);
    stage_coef(rowA);

    unsigned long long it_sum = 0, it_max = 0;
    const double dtime = p.dtime;
    int sl = (rowA - 1) % 3; /* LDS slot of row j-1 */

    for (int j = rowA; j < rowB; j++) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads(); /* slots j-1..j+1 and the coefficients of row j are in LDS */
        const bool more = (j + 1 < rowB);
        const int slm = sl, sl0 = sl == 2 ? 0 : sl + 1, slp = sl0 == 2 ? 0 : sl0 + 1;
        sl = sl0;

        const unsigned ij = (unsigned)j * (unsigned)nx + (unsigned)gi;
        const uint8_t mk = inside ? p.mask[ij] : 0;
        const bool active = (mk & MASK_ACTIVE) != 0;
        /* per-cell inputs, consumed by the collide below */
        T f_eta = 0, f_k0 = 0, f_k1 = 0, f_kA = 0, f_ux = 0, f_uy = 0, f_adt = 0, f_ia3 = 1;
        if (active) {
            f_eta = (p.cell + (size_t)CF_ETA * N)[ij];
            f_k0 = (p.cell + (size_t)CF_K0 * N)[ij];
            f_k1 = (p.cell + (size_t)CF_K1 * N)[ij];
            f_kA = (p.cell + (size_t)CF_KA * N)[ij];
            f_ux = (p.cell + (size_t)CF_UX * N)[ij];
            f_uy = (p.cell + (size_t)CF_UY * N)[ij];
            f_adt = (p.cell + (size_t)CF_ADT * N)[ij];
            if (CURVED) f_ia3 = (p.cell + (size_t)CF_IA3 * N)[ij];
        }
        /* An opaque copy of the table pointer per row keeps the per-direction
         * constants from being hoisted out of the row loop. */
        CDirTables *dtab = dtab0;
        asm volatile("" : "+s"(dtab));

        /* this lane's column in the three row slots */
        const T *colm = &sm.I[slm][0][lane];
        const T *col0p = &sm.I[sl0][0][lane];
        const T *colp = &sm.I[slp][0][lane];

        /* ---- stream ---- */
        T sv[DW];
        double acc[6] = {0, 0, 0, 0, 0, 0};
        unsigned long long mbits = 0; /* REC: bit 4k+c set when corner c of direction k is not clamped */
        if constexpr (CURVED) {
            /* geodesic Fourier coefficients of this cell: S as S - 1, X and Y
             * relative to the cell centre in cell units */
            T cS[5], cX[5], cY[5], cU[5], cV[5];
            cell_series<T>(sm.C, lane, p.startx + gi * p.dx, p.starty + j * p.dy, 1.0 / p.dx, 1.0 / p.dy, cS, cX, cY,
                           cU, cV);
            /* directions of this wavefront whose departure point is excised */
            const uint32_t exw = (uint32_t)(__double_as_longlong(sm.C[25][lane]) >> (wv * DW));
            const T angu = (sizeof(T) == 4) ? (T)dtab->ang_uf : (T)dtab->ang_u;
            const int quarter = dtab->quarter, nGrid = dtab->nGrid;
            const T ddmax = T(nGrid) - T(1e-8);
            /* Software pipeline over the directions of this wavefront: the
             * LDS taps of direction k+1 are in flight while direction k is
             * combined.  prep() computes the departure point, angular window
             * and weights and issues the 20 tap loads; fin() clamps,
             * interpolates and applies the ray factor.  Everything is
             * computed for every lane (no branches); lanes that do not gather
             * select 0.  The clamps keep every LDS index in range for any
             * input, NaN included. */
            struct Prep {
                T v[5][4];
                T Cw[5];
                T fx, fy, s;
                bool go;
            };
            auto prep = [&](int k, Prep &P) {
                const int d = wv * DW + k;
                DirGeom<T> G;
                curved_geometry<T, D>(cS, cX, cY, cU, cV, dtab, dtab0, &sm.wrow[0][0], d, angu, quarter, nGrid, ddmax,
                                      G);
                P.s = G.s;
                P.go = active && !((exw >> k) & 1u);
                P.fx = G.fx;
                P.fy = G.fy;
#pragma unroll
                for (int m = 0; m < 5; m++) P.Cw[m] = G.Cw[m];
                const T *ra = G.lo_j ? colm : col0p;
                const T *rb = G.lo_j ? col0p : colp;
                const int off = (int)__umul24((unsigned)G.s0, (unsigned)RW) + (G.lo_i ? 0 : 1);
                const T *pa = ra + off;
                const T *pb = rb + off;
#pragma unroll
                for (int m = 0; m < 5; m++) {
                    P.v[m][0] = pa[m * RW];
                    P.v[m][1] = pa[m * RW + 1];
                    P.v[m][2] = pb[m * RW];
                    P.v[m][3] = pb[m * RW + 1];
                }
            };
            auto fin = [&](int k, const Prep &P) {
                const int d = wv * DW + k;
                T v00 = 0, v10 = 0, v01 = 0, v11 = 0;
#pragma unroll
                for (int m = 0; m < 5; m++) {
                    v00 += P.Cw[m] * P.v[m][0];
                    v10 += P.Cw[m] * P.v[m][1];
                    v01 += P.Cw[m] * P.v[m][2];
                    v11 += P.Cw[m] * P.v[m][3];
                }
                /* state is alpha^-3 I: the (alpha/alpha_k)^3 factors cancel.  s is
                 * carried as s - 1, so s^3 x = x + (s^3 - 1) x rounds once, at the end */
                const T bil = rtm::bilinear<T>(P.fx, P.fy, tmax0(v00), tmax0(v01), tmax0(v10), tmax0(v11));
                if constexpr (REC) {
                    const unsigned bits = (unsigned)(v00 > T(0)) | ((unsigned)(v01 > T(0)) << 1) |
                                          ((unsigned)(v10 > T(0)) << 2) | ((unsigned)(v11 > T(0)) << 3);
                    mbits |= (unsigned long long)bits << (4 * k);
                }
                const T s3m1 = P.s * (T(3) + P.s * (T(3) + P.s));
                const T val = P.go ? bil + bil * s3m1 : T(0);
                sv[k] = val;
                /* materialise this direction's value here: otherwise the
                 * compiler sinks every gather to the end of the row and
                 * spills the taps of all directions */
                asm volatile("" ::"v"(val));
                const double vd = (double)val;
#pragma unroll
                for (int X = 0; X < 6; X++) acc[X] += dtab->w6[d][X] * vd;
            };
            Prep P0, P1;
            prep(0, P0);
#pragma unroll
            for (int k = 0; k < DW; k += 2) {
                if (k + 1 < DW) prep(k + 1, P1);
                fin(k, P0);
                if (k + 1 < DW) {
                    if (k + 2 < DW) prep(k + 2, P0);
                    fin(k + 1, P1);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < DW; k++) {
                const int d = wv * DW + k;
                T fx, fy;
                int oi, oj;
                if constexpr (sizeof(T) == 4) {
                    oi = dtab->flat_di[d];
                    oj = dtab->flat_dj[d];
                    fx = dtab->flat_fxf[d];
                    fy = dtab->flat_fyf[d];
                } else {
                    /* reference arithmetic of StreamFlatFixed for this cell */
                    double xT = p.startx + gi * p.dx, yT = p.starty + j * p.dy;
                    xT -= dtab->cx[d] * dtime;
                    yT -= dtab->cy[d] * dtime;
                    double iT = (xT - p.startx) / p.dx, jT = (yT - p.starty) / p.dy;
                    double i0 = floor(iT), j0 = floor(jT);
                    oi = (int)i0 - gi;
                    oj = (int)j0 - j;
                    oi = oi < -1 ? -1 : (oi > 0 ? 0 : oi);
                    oj = oj < -1 ? -1 : (oj > 0 ? 0 : oj);
                    fx = (T)(iT - (double)(gi + oi));
                    fy = (T)(jT - (double)(j + oj));
                }
                const T *ra = oj < 0 ? colm : col0p;
                const T *rb = oj < 0 ? col0p : colp;
                const int off = d * RW + 1 + oi;
                const T *pa = ra + off;
                const T *pb = rb + off;
                T val = rtm::bilinear<T>(fx, fy, pa[0], pb[0], pa[1], pb[1]);
                val = active ? val : T(0);
                sv[k] = val;
                const double vd = (double)val;
#pragma unroll
                for (int X = 0; X < 6; X++) acc[X] += dtab->w6[d][X] * vd;
            }
        }
#pragma unroll
        for (int X = 0; X < 6; X++) sm.X[wv][X][lane] = acc[X];
        __syncthreads(); /* partial moments visible; every wave is done with slot (j-1)%3 and C */
        if (more) {
            stage_row(j + 2);
            stage_coef(j + 1);
        }

        /* moments of the streamed intensities (true units: alpha^3 x state) */
        const double a3 = CURVED ? 1.0 / (double)f_ia3 : 1.0;
        double mI[6];
#pragma unroll
        for (int X = 0; X < 6; X++)
            mI[X] = a3 * (((sm.X[0][X][lane] + sm.X[1][X][lane]) + sm.X[2][X][lane]) + sm.X[3][X][lane]);

        /* ---- collide ---- */
        const double eta = (double)f_eta, k0 = (double)f_k0, k1 = (double)f_k1, kA = (double)f_kA;
        const double u1 = (double)f_ux, u2 = (double)f_uy, adt = (double)f_adt;
        const bool moving = active && (u1 != 0.0 || u2 != 0.0);
        double mOut[6];
        int cycles = 0;
        if (!__any(moving)) {
            /* zero fluid velocity: lambda iteration on the moments */
            const double q = 1.0 / (1.0 + adt * (kA + k0));
            const double cc = 3.0 * k1 * adt * q;
            double prev[6];
#pragma unroll
            for (int X = 0; X < 6; X++) prev[X] = mI[X];
            double b = 0, fFx = 0, fFy = 0;
            double ss = 1;
            if (active) {
                /* stop when sqrt(sum ((prev - new) / new)^2) <= 1e-4; the
                 * quotients use a refined reciprocal, and new == 0 keeps the
                 * IEEE quotient (inf, or NaN for 0/0, which stops the loop) */
                while (ss > 1e-8 && cycles < 100) {
                    cycles++;
                    b = adt * q * (eta + k0 * prev[0]);
                    fFx = prev[1];
                    fFy = prev[2];
                    double acc2 = 0;
#pragma unroll
                    for (int X = 0; X < 6; X++) {
                        const double nw = q * mI[X] + b * dtab->S[0][X] + cc * (fFx * dtab->S[1][X] + fFy * dtab->S[2][X]);
                        const double r0 = __builtin_amdgcn_rcp(nw);
                        const double r1 = fma(r0, fma(-nw, r0, 1.0), r0);
                        const double r = (prev[X] - nw) * (nw == 0.0 ? r0 : r1);
                        acc2 += r * r;
                        prev[X] = nw;
                    }
                    ss = acc2;
                }
            }
#pragma unroll
            for (int X = 0; X < 6; X++) mOut[X] = prev[X];
            /* Inew = J + (q-1) J + b + c n.F: the small factor q-1 carries the
             * rounding, which keeps FP32 storage unbiased; b and c are scaled
             * to the state's alpha^-3 units */
            const double is = CURVED ? (double)f_ia3 : 1.0;
            const T qm1 = (T)(q - 1.0), bT = (T)(b * is), cxT = (T)(cc * fFx * is), cyT = (T)(cc * fFy * is);
#pragma unroll
            for (int k = 0; k < DW; k++) {
                const int d = wv * DW + k;
                const T nxd = (sizeof(T) == 4) ? (T)dtab->cxf[d] : (T)dtab->cx[d];
                const T nyd = (sizeof(T) == 4) ? (T)dtab->cyf[d] : (T)dtab->cy[d];
                sv[k] = sv[k] + (qm1 * sv[k] + (bT + (cxT * nxd + cyT * nyd)));
            }
        } else {
            if (CURVED) {
#pragma unroll
                for (int k = 0; k < DW; k++) sv[k] = (T)((double)sv[k] * a3);
            }
            collide_moving<T, DW>(sv, mOut, cycles, mI, active, eta, k0, k1, kA, u1, u2, adt, wv, lane, dtab,
                                  sm.X);
            if (CURVED) {
#pragma unroll
                for (int k = 0; k < DW; k++) sv[k] = (T)((double)sv[k] * (double)f_ia3);
            }
        }

        /* ---- write back ---- */
        if constexpr (REC) {
            if (active) {
                if (wv == 0) {
#pragma unroll
                    for (int X = 0; X < 6; X++) (p.rec_m + (size_t)X * N)[ij] = mI[X];
                    p.rec_k[ij] = (uint8_t)cycles;
                }
                if (CURVED) (p.rec_mask + (size_t)wv * N)[ij] = mbits;
            }
        }
        if (active) {
#pragma unroll
            for (int k = 0; k < DW; k++) (p.Iout + (size_t)(wv * DW + k) * N)[ij] = sv[k];
            if (wv == 0) {
                it_sum += (unsigned long long)cycles;
                it_max = it_max > (unsigned long long)cycles ? it_max : (unsigned long long)cycles;
                if (p.want_moments) {
#pragma unroll
                    for (int X = 0; X < 6; X++) (p.mom + (size_t)X * N)[ij] = mOut[X];
                }
            }
        }
    }

    if (wv == 0) {
        for (int off = 32; off > 0; off >>= 1) {
            it_sum += __shfl_xor(it_sum, off);
            unsigned long long o = __shfl_xor(it_max, off);
            it_max = it_max > o ? it_max : o;
        }
        if (lane == 0) {
            atomicAdd(&p.stats[0], it_sum);
            atomicMax(&p.stats[1], it_max);
        }
    }
}

template <typename T, int D, bool CURVED, bool REC>
int launch_one(rt_ctx *c, const void *in, void *out, bool want_moments, const RecordPtrs *rec)
{
    const rt_desc &d = c->desc;
    KP<T> p;
    p.Iin = (const T *)in;
    p.Iout = (T *)out;
    p.cell = (const T *)c->d_cell;
    p.mask = c->d_mask;
    p.coef = c->d_coef;
    p.excl = c->d_excl;
    p.mom = c->d_mom;
    p.stats = c->d_stats;
    p.dt = c->d_dirt;
    p.nx = d.nx;
    p.ny = d.ny;
    p.N = (long)d.nx * d.ny;
    p.strips = (d.nx - 2 + 63) / 64;
    const int rows = d.ny - 2;
    /* aim for about 1024 workgroups (4 per CU) with at least 4 rows each */
    int chunks = (1024 + p.strips - 1) / p.strips;
    int H = (rows + chunks - 1) / chunks;
    H = H < 4 ? 4 : (H > 64 ? 64 : H);
    chunks = (rows + H - 1) / H;
    p.chunks = chunks;
    p.rows_per_block = H;
    p.nblocks = p.strips * chunks;
    p.startx = d.startx;
    p.starty = d.starty;
    p.dx = d.dx;
    p.dy = d.dy;
    p.dtime = d.dt;
    p.r2 = d.excision_r2;
    p.want_moments = want_moments ? 1 : 0;
    p.rec_m = rec ? rec->m : nullptr;
    p.rec_k = rec ? rec->k : nullptr;
    p.rec_mask = rec ? rec->mask : nullptr;
    k_fused<T, D, CURVED, REC><<<p.nblocks, 256, 0, c->stream>>>(p);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}

template <typename T, bool CURVED, bool REC>
int dispatch_d(rt_ctx *c, const void *in, void *out, bool wm, const RecordPtrs *rec)
{
    switch (c->D) {
    case 16: return launch_one<T, 16, CURVED, REC>(c, in, out, wm, rec);
    case 20: return launch_one<T, 20, CURVED, REC>(c, in, out, wm, rec);
    case 24: return launch_one<T, 24, CURVED, REC>(c, in, out, wm, rec);
    case 32: return launch_one<T, 32, CURVED, REC>(c, in, out, wm, rec);
    case 48: return launch_one<T, 48, CURVED, REC>(c, in, out, wm, rec);
    case 64: return launch_one<T, 64, CURVED, REC>(c, in, out, wm, rec);
    default: return rt_fail(c, RT_ERR_UNSUPPORTED, "fused kernel: unsupported nDir");
    }
}

} // namespace

bool fused_supported(const rt_ctx *c)
{
    const int D = c->D;
    if (!(D == 16 || D == 20 || D == 24 || D == 32 || D == 48 || D == 64)) return false;
    const rt_desc &d = c->desc;
    if (d.streaming == RT_CURVED_FIXED) {
        if (c->nCoeff != 5 || !c->fused_tables_ok) return false;
        if (!(c->max_disp < 0.999)) return false;
    } else {
        /* departure offset c*dt must stay within one cell */
        if (!(d.dt < 0.999 * d.dx && d.dt < 0.999 * d.dy)) return false;
    }
    return true;
}

int launch_fused_step(rt_ctx *c, const void *in, void *out, bool want_moments)
{
    const bool curved = c->desc.streaming == RT_CURVED_FIXED;
    if (c->fp32)
        return curved ? dispatch_d<float, true, false>(c, in, out, want_moments, nullptr)
                      : dispatch_d<float, false, false>(c, in, out, want_moments, nullptr);
    return curved ? dispatch_d<double, true, false>(c, in, out, want_moments, nullptr)
                  : dispatch_d<double, false, false>(c, in, out, want_moments, nullptr);
}

int launch_fused_step_rec(rt_ctx *c, const void *in, void *out, const RecordPtrs &rec)
{
    const bool curved = c->desc.streaming == RT_CURVED_FIXED;
    if (c->fp32)
        return curved ? dispatch_d<float, true, true>(c, in, out, false, &rec)
                      : dispatch_d<float, false, true>(c, in, out, false, &rec);
    return curved ? dispatch_d<double, true, true>(c, in, out, false, &rec)
                  : dispatch_d<double, false, true>(c, in, out, false, &rec);
}
