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
#include "rt_internal.h"

namespace {

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
};

template <typename T>
__device__ inline T fast_div(T a, T b)
{
    return a / b;
}
template <>
__device__ inline float fast_div<float>(float a, float b)
{
    return a * __builtin_amdgcn_rcpf(b);
}

template <typename T>
__device__ inline T tmax0(T v)
{
    return v > T(0) ? v : T(0);
}

/* Per-direction tables read with wave-uniform indices: a pointer into the
 * constant address space lets the compiler use scalar loads. */
typedef const __attribute__((address_space(4))) DirTables CDirTables;

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

/* Copies n dwords from global src to LDS dst with one dword per lane per
 * instruction; lanes past n do nothing.  dst must be wave-uniform. */
__device__ inline void lds_dma_dwords(const void *src, void *dst, int n, int lane)
{
    const uint32_t *g = (const uint32_t *)src;
    uint32_t *l = (uint32_t *)dst;
    for (int b = 0; b < n; b += 64) {
        if (b + lane < n) __builtin_amdgcn_global_load_lds(g + b + lane, l + b, 4, 0, 0);
    }
}

/* Fine angular index nGrid*phi/(2 pi) of Tensor2::Phi (src/DataTypes.hh:166-170)
 * with the Hastings atan2 (src/Utility.hh:220-254), branch-free.  The octant
 * offsets are exact multiples of nGrid/4, so only the polynomial term rounds. */
template <typename T>
__device__ inline T angle_index_bf(T vx, T vy, T u, int quarter)
{
    const bool swap = !(fabsf((float)vx) > fabsf((float)vy));
    const T num = swap ? vx : vy, den = swap ? vy : vx;
    const T p = rtm::hastings_atan(fast_div(num, den));
    /* octant offset in quarters: |vx| > |vy|: 0 (or 4 when p < 0) for vx > 0,
     * 2 for vx < 0; otherwise 1 for vy > 0, 3 for vy < 0 (and the
     * polynomial enters with a minus sign) */
    const int px = (int)(vx > T(0)), pneg = (int)(p < T(0)), py = (int)(vy > T(0));
    const int kx = 2 - 2 * px + 4 * (px & pneg);
    const int ky = 3 - 2 * py;
    const int k = swap ? ky : kx;
    const T up = u * p;
    const T dd = T(k * quarter) + (swap ? -up : up);
    return (vx == T(0) && vy == T(0)) ? T(0) : dd;
}

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

template <typename T, int D, bool CURVED>
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
        if constexpr (CURVED) {
            /* geodesic Fourier coefficients of this cell: S as S - 1, X and Y
             * relative to the cell centre in cell units */
            T cS[5], cX[5], cY[5], cU[5], cV[5];
            {
                const double xc = p.startx + gi * p.dx, yc = p.starty + j * p.dy;
                const double idx = 1.0 / p.dx, idy = 1.0 / p.dy;
#pragma unroll
                for (int k = 0; k < 5; k++) {
                    cS[k] = (T)(k == 0 ? sm.C[0][lane] - 1.0 : sm.C[k][lane]);
                    cX[k] = (T)((k == 0 ? sm.C[5][lane] - xc : sm.C[5 + k][lane]) * idx);
                    cY[k] = (T)((k == 0 ? sm.C[10][lane] - yc : sm.C[10 + k][lane]) * idy);
                    cU[k] = (T)sm.C[15 + k][lane];
                    cV[k] = (T)sm.C[20 + k][lane];
                }
            }
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
                T B[5];
#pragma unroll
                for (int q = 0; q < 5; q++) B[q] = (sizeof(T) == 4) ? (T)dtab->basisf[d][q] : (T)dtab->basis[d][q];
                T s = 0, di = 0, dj = 0, vx = 0, vy = 0;
#pragma unroll
                for (int q = 0; q < 5; q++) {
                    s += cS[q] * B[q];
                    di += cX[q] * B[q];
                    dj += cY[q] * B[q];
                    vx += cU[q] * B[q];
                    vy += cV[q] * B[q];
                }
                P.s = s;
                P.go = active && !((exw >> k) & 1u);
                const T fi = fmin(fmax(floor(di), T(-1)), T(0));
                const T fj = fmin(fmax(floor(dj), T(-1)), T(0));
                P.fx = di - fi;
                P.fy = dj - fj;
                T dd = angle_index_bf<T>(vx, vy, angu, quarter);
                dd = fmin(fmax(dd, T(1e-8)), ddmax);
                int g0 = (int)dd;
                const T t = dd - T(g0);
                g0 = (unsigned)g0 < (unsigned)nGrid ? g0 : 0; /* dd == nGrid after rounding is angle 0 */
                const int q10 = (int)(__umul24((unsigned)g0, 6554u) >> 16); /* g0 / 10 for g0 < 16384 */
                const int r = g0 - 10 * q10;
                int woff;
                if constexpr (sizeof(T) == 4) {
                    const float4 *wr = (const float4 *)((const char *)&sm.wrow[0][0] + __umul24((unsigned)r, 48u));
                    const float4 w0 = wr[0], w1 = wr[1], w2 = wr[2];
                    P.Cw[0] = w0.x + t * w1.y;
                    P.Cw[1] = w0.y + t * w1.z;
                    P.Cw[2] = w0.z + t * w1.w;
                    P.Cw[3] = w0.w + t * w2.x;
                    P.Cw[4] = w1.x + t * w2.y;
                    woff = __float_as_int(w2.z);
                } else {
                    /* FP64 rows come through the vector L1 (faster than LDS
                     * for 8-byte elements) */
#pragma unroll
                    for (int m = 0; m < 5; m++) P.Cw[m] = dtab0->wA[r][m] + t * dtab0->wD[r][m];
                    woff = dtab0->woff[r];
                }
                int s0 = q10 + woff;
                s0 = s0 < 0 ? s0 + D : (s0 >= D ? s0 - D : s0);
                const bool lo_j = fj < T(0);
                const T *ra = lo_j ? colm : col0p;
                const T *rb = lo_j ? col0p : colp;
                const int off = (int)__umul24((unsigned)s0, (unsigned)RW) + (fi < T(0) ? 0 : 1);
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

template <typename T, int D, bool CURVED>
int launch_one(rt_ctx *c, const void *in, void *out, bool want_moments)
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
    k_fused<T, D, CURVED><<<p.nblocks, 256, 0, c->stream>>>(p);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}

template <typename T, bool CURVED>
int dispatch_d(rt_ctx *c, const void *in, void *out, bool wm)
{
    switch (c->D) {
    case 16: return launch_one<T, 16, CURVED>(c, in, out, wm);
    case 20: return launch_one<T, 20, CURVED>(c, in, out, wm);
    case 24: return launch_one<T, 24, CURVED>(c, in, out, wm);
    case 32: return launch_one<T, 32, CURVED>(c, in, out, wm);
    case 48: return launch_one<T, 48, CURVED>(c, in, out, wm);
    case 64: return launch_one<T, 64, CURVED>(c, in, out, wm);
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
    if (c->fp32) return curved ? dispatch_d<float, true>(c, in, out, want_moments) : dispatch_d<float, false>(c, in, out, want_moments);
    return curved ? dispatch_d<double, true>(c, in, out, want_moments) : dispatch_d<double, false>(c, in, out, want_moments);
}
