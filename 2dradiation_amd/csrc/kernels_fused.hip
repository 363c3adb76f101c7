/*
 * Fused stream + moments + collide step for gfx950.
 *
 * One workgroup = 4 wavefronts walks a 64-cell-wide column strip of the grid
 * down a chunk of rows.  Three input rows (all directions) live in an LDS ring
 * [slot][d][66] (64 cells plus one halo cell each side); the row two ahead is
 * prefetched into registers with coalesced SoA loads while the current row is
 * computed, then written into the slot that just died.  Lane = cell column,
 * wavefront w = direction block [w*D/4, (w+1)*D/4).
 *
 * Stream (reference src/Radiation.cpp:380-414 flat, :454-512 curved):
 *   flat:   bilinear of the same direction at the departure point x - c dt.
 *   curved: the five geodesic Fourier series (S, X, Y, Cx, Cy) are evaluated
 *           at phi_d with the per-direction basis from scalar registers; the
 *           departure velocity angle (Hastings atan2) picks the cubic angular
 *           window; the lerp between the two cubic stencils is folded into 5
 *           combined taps; each of the 4 bilinear corners is clamped >= 0; the
 *           (alpha/alpha_k)^3 factor is applied as alpha^3 * (I_k alpha_k^-3)
 *           with the alpha_k^-3 scaling done once while staging into LDS.
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

/* Fine angular index nGrid*phi/(2 pi) of Tensor2::Phi (src/DataTypes.hh:166-170)
 * with the Hastings atan2 (src/Utility.hh:220-254).  The octant offsets are
 * exact multiples of nGrid/4, so only the polynomial term is rounded. */
template <typename T>
__device__ inline T angle_index(T vx, T vy, T u, int quarter, int nGrid)
{
    if (vx != T(0)) {
        if (fabsf((float)vx) > fabsf((float)vy)) {
            T p = rtm::hastings_atan(fast_div(vy, vx));
            if (vx > T(0)) return p >= T(0) ? u * p : u * p + T(nGrid);
            return u * p + T(2 * quarter);
        }
        T p = rtm::hastings_atan(fast_div(vx, vy));
        if (vy > T(0)) return T(quarter) - u * p;
        return T(3 * quarter) - u * p;
    }
    if (vy > T(0)) return T(quarter);
    if (vy < T(0)) return T(3 * quarter);
    return T(0);
}

template <typename T>
__device__ inline T tmax0(T v)
{
    return v > T(0) ? v : T(0);
}

template <typename T, int D, bool CURVED>
struct Smem {
    T I[3][D][66];
    T C[CURVED ? 2 : 1][CURVED ? 25 : 1][64];
    double X[RT_FUSED_WAVES][6][64];
    T wA[10][5], wD[10][5];
    int woff[10];
};

template <typename T, int D, bool CURVED>
__global__ __launch_bounds__(256) void k_fused(KP<T> p)
{
    constexpr int DW = D / RT_FUSED_WAVES;
    __shared__ Smem<T, D, CURVED> sm;

    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const DirTables *__restrict__ dtab = p.dt;
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
    const int rowA = 1 + chunk * p.rows_per_block;
    const int rowB = min(rowA + p.rows_per_block, p.ny - 1); /* exclusive */
    if (rowA >= rowB) return;

    if (CURVED) {
        if (threadIdx.x < 50) {
            int r = threadIdx.x / 5, m = threadIdx.x % 5;
            if constexpr (sizeof(T) == 4) {
                sm.wA[r][m] = dtab->wAf[r][m];
                sm.wD[r][m] = dtab->wDf[r][m];
            } else {
                sm.wA[r][m] = dtab->wA[r][m];
                sm.wD[r][m] = dtab->wD[r][m];
            }
        }
        if (threadIdx.x < 10) sm.woff[threadIdx.x] = dtab->woff[threadIdx.x];
    }

    /* ---- staging helpers ---- */
    T preg[DW], preg2[DW];
    double creg[7];
    auto load_row = [&](int row) {
        const long rbase = (long)row * nx;
        const int c = col0 + lane, c2 = col0 + 64 + lane;
        T s1 = T(1), s2 = T(1);
        if (CURVED) {
            if (c < nx) s1 = p.cell[CF_IA3 * N + rbase + c];
            if (lane < 2 && c2 < nx) s2 = p.cell[CF_IA3 * N + rbase + c2];
        }
#pragma unroll
        for (int k = 0; k < DW; k++) {
            const int d = wv * DW + k;
            const T *src = p.Iin + (long)d * N + rbase;
            preg[k] = (c < nx) ? src[c] * s1 : T(0);
            preg2[k] = (lane < 2 && c2 < nx) ? src[c2] * s2 : T(0);
        }
    };
    auto store_row = [&](int row) {
        const int slot = row % 3;
#pragma unroll
        for (int k = 0; k < DW; k++) {
            const int d = wv * DW + k;
            sm.I[slot][d][lane] = preg[k];
            if (lane < 2) sm.I[slot][d][64 + lane] = preg2[k];
        }
    };
    auto load_coef = [&](int row) {
        if (!CURVED) return;
        const long ij = (long)row * nx + gi;
        const bool ok = gi <= nx - 2;
#pragma unroll
        for (int q = 0; q < 7; q++) {
            const int e = wv + 4 * q;
            creg[q] = (e < 25 && ok) ? p.coef[(long)e * N + ij] : 0.0;
        }
    };
    auto store_coef = [&](int row) {
        if (!CURVED) return;
        const double xc = p.startx + gi * p.dx, yc = p.starty + row * p.dy;
        const double idx = 1.0 / p.dx, idy = 1.0 / p.dy;
#pragma unroll
        for (int q = 0; q < 7; q++) {
            const int e = wv + 4 * q;
            if (e < 25) {
                double v = creg[q];
                const int ser = e / 5, k = e % 5;
                if (ser == 1) v = (k == 0 ? v - xc : v) * idx;
                if (ser == 2) v = (k == 0 ? v - yc : v) * idy;
                sm.C[row & 1][e][lane] = (T)v;
            }
        }
    };

    /* prologue: rows rowA-1, rowA, rowA+1 and the coefficients of rowA */
    for (int r = rowA - 1; r <= rowA + 1; r++) {
        load_row(r);
        store_row(r);
    }
    load_coef(rowA);
    store_coef(rowA);

    unsigned long long it_sum = 0, it_max = 0;
    const double *w6 = &dtab->w6[0][0];

    for (int j = rowA; j < rowB; j++) {
        __syncthreads(); /* slots j-1..j+1 and coefficients of row j are visible */
        const bool more = (j + 1 < rowB);
        if (more) {
            load_row(j + 2);
            load_coef(j + 1);
        }

        const long ij = (long)j * nx + gi;
        const bool inside = gi <= nx - 2;
        const uint8_t mk = inside ? p.mask[ij] : 0;
        const bool active = (mk & MASK_ACTIVE) != 0;

        /* ---- stream ---- */
        T sv[DW];
        double acc[6] = {0, 0, 0, 0, 0, 0};
        if (CURVED) {
            T cS[5], cX[5], cY[5], cU[5], cV[5];
#pragma unroll
            for (int k = 0; k < 5; k++) {
                cS[k] = sm.C[j & 1][k][lane];
                cX[k] = sm.C[j & 1][5 + k][lane];
                cY[k] = sm.C[j & 1][10 + k][lane];
                cU[k] = sm.C[j & 1][15 + k][lane];
                cV[k] = sm.C[j & 1][20 + k][lane];
            }
            const T a3 = active ? T(1) / p.cell[CF_IA3 * N + ij] : T(0);
            const bool near = (mk & MASK_NEAR_BH) != 0;
            const double xc = p.startx + gi * p.dx, yc = p.starty + j * p.dy;
            const T angu = (sizeof(T) == 4) ? (T)dtab->ang_uf : (T)dtab->ang_u;
            const int quarter = dtab->quarter, nGrid = dtab->nGrid;
#pragma unroll
            for (int k = 0; k < DW; k++) {
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
                T val = T(0);
                bool go = active;
                if (go && near) {
                    const double xd = xc + (double)di * p.dx, yd = yc + (double)dj * p.dy;
                    if (xd * xd + yd * yd <= p.r2) go = false;
                }
                if (go) {
                    T fi = floor(di), fj = floor(dj);
                    fi = fi < T(-1) ? T(-1) : (fi > T(0) ? T(0) : fi);
                    fj = fj < T(-1) ? T(-1) : (fj > T(0) ? T(0) : fj);
                    const T fx = di - fi, fy = dj - fj;
                    T dd = angle_index<T>(vx, vy, angu, quarter, nGrid);
                    dd = dd < T(1e-8) ? T(1e-8) : dd;
                    dd = dd > T(nGrid) - T(1e-8) ? T(nGrid) - T(1e-8) : dd;
                    int g0 = (int)dd;
                    const T t = dd - T(g0);
                    if (g0 >= nGrid) g0 -= nGrid;
                    const int r = g0 % 10;
                    int s0 = g0 / 10 + sm.woff[r];
                    s0 = s0 < 0 ? s0 + D : (s0 >= D ? s0 - D : s0);
                    T Cw[5];
                    int dirs[5];
#pragma unroll
                    for (int m = 0; m < 5; m++) {
                        Cw[m] = sm.wA[r][m] + t * sm.wD[r][m];
                        int dm = s0 + m;
                        dirs[m] = dm >= D ? dm - D : dm;
                    }
                    const int sa = (j + (int)fj + 3) % 3, sb = (j + (int)fj + 4) % 3;
                    const int ca = lane + 1 + (int)fi, cb = ca + 1;
                    T v00 = 0, v01 = 0, v10 = 0, v11 = 0;
#pragma unroll
                    for (int m = 0; m < 5; m++) {
                        v00 += Cw[m] * sm.I[sa][dirs[m]][ca];
                        v01 += Cw[m] * sm.I[sb][dirs[m]][ca];
                        v10 += Cw[m] * sm.I[sa][dirs[m]][cb];
                        v11 += Cw[m] * sm.I[sb][dirs[m]][cb];
                    }
                    val = (s * s * s) * a3 *
                          rtm::bilinear<T>(fx, fy, tmax0(v00), tmax0(v01), tmax0(v10), tmax0(v11));
                }
                sv[k] = val;
                const double vd = (double)val;
#pragma unroll
                for (int X = 0; X < 6; X++) acc[X] += w6[d * 6 + X] * vd;
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
                    xT -= dtab->cx[d] * p.dtime;
                    yT -= dtab->cy[d] * p.dtime;
                    double iT = (xT - p.startx) / p.dx, jT = (yT - p.starty) / p.dy;
                    double i0 = floor(iT), j0 = floor(jT);
                    oi = (int)i0 - gi;
                    oj = (int)j0 - j;
                    oi = oi < -1 ? -1 : (oi > 0 ? 0 : oi);
                    oj = oj < -1 ? -1 : (oj > 0 ? 0 : oj);
                    fx = (T)(iT - (double)(gi + oi));
                    fy = (T)(jT - (double)(j + oj));
                }
                const int sa = (j + oj + 3) % 3, sb = (j + oj + 4) % 3;
                const int ca = lane + 1 + oi, cb = ca + 1;
                T val = T(0);
                if (active)
                    val = rtm::bilinear<T>(fx, fy, sm.I[sa][d][ca], sm.I[sb][d][ca], sm.I[sa][d][cb], sm.I[sb][d][cb]);
                sv[k] = val;
                const double vd = (double)val;
#pragma unroll
                for (int X = 0; X < 6; X++) acc[X] += w6[d * 6 + X] * vd;
            }
        }
#pragma unroll
        for (int X = 0; X < 6; X++) sm.X[wv][X][lane] = acc[X];
        __syncthreads();

        double mI[6];
#pragma unroll
        for (int X = 0; X < 6; X++)
            mI[X] = ((sm.X[0][X][lane] + sm.X[1][X][lane]) + sm.X[2][X][lane]) + sm.X[3][X][lane];

        /* ---- collide ---- */
        double eta = 0, k0 = 0, k1 = 0, kA = 0, u1 = 0, u2 = 0, adt = 0;
        if (active) {
            eta = (double)p.cell[CF_ETA * N + ij];
            k0 = (double)p.cell[CF_K0 * N + ij];
            k1 = (double)p.cell[CF_K1 * N + ij];
            kA = (double)p.cell[CF_KA * N + ij];
            u1 = (double)p.cell[CF_UX * N + ij];
            u2 = (double)p.cell[CF_UY * N + ij];
            adt = (double)p.cell[CF_ADT * N + ij];
        }
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
            double diff = 1;
            if (active) {
                while (fabs(diff) > 1e-4 && cycles < 100) {
                    cycles++;
                    b = adt * q * (eta + k0 * prev[0]);
                    fFx = prev[1];
                    fFy = prev[2];
                    double nw[6], ss = 0;
#pragma unroll
                    for (int X = 0; X < 6; X++) {
                        nw[X] = q * mI[X] + b * dtab->S[0][X] + cc * (fFx * dtab->S[1][X] + fFy * dtab->S[2][X]);
                        const double r = (prev[X] - nw[X]) / nw[X];
                        ss += r * r;
                    }
                    diff = sqrt(ss);
#pragma unroll
                    for (int X = 0; X < 6; X++) prev[X] = nw[X];
                }
            }
#pragma unroll
            for (int X = 0; X < 6; X++) mOut[X] = prev[X];
            const T qT = (T)q, bT = (T)b, cxT = (T)(cc * fFx), cyT = (T)(cc * fFy);
#pragma unroll
            for (int k = 0; k < DW; k++) {
                const int d = wv * DW + k;
                const T nxd = (sizeof(T) == 4) ? (T)dtab->cxf[d] : (T)dtab->cx[d];
                const T nyd = (sizeof(T) == 4) ? (T)dtab->cyf[d] : (T)dtab->cy[d];
                sv[k] = qT * sv[k] + (bT + (cxT * nxd + cyT * nyd));
            }
        } else {
            /* moving fluid somewhere in this row: the reference's per-direction
             * cycle (src/Radiation.cpp:661-722) with a moment exchange per
             * cycle.  Every wavefront evaluates the same per-cell decisions, so
             * the loop trip count is uniform across the workgroup. */
            const double uu = u1 * u1 + u2 * u2;
            const double lor = 1.0 / sqrt(1.0 - uu);
            double prev[6];
#pragma unroll
            for (int X = 0; X < 6; X++) prev[X] = mI[X];
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
                        for (int X = 0; X < 6; X++) part[X] += w6[d * 6 + X] * v;
                    }
                }
                __syncthreads();
#pragma unroll
                for (int X = 0; X < 6; X++) sm.X[wv][X][lane] = part[X];
                __syncthreads();
                if (iter) {
                    double ss = 0;
#pragma unroll
                    for (int X = 0; X < 6; X++) {
                        const double nw = ((sm.X[0][X][lane] + sm.X[1][X][lane]) + sm.X[2][X][lane]) + sm.X[3][X][lane];
                        const double r = (prev[X] - nw) / nw;
                        ss += r * r;
                        prev[X] = nw;
                    }
                    diff = sqrt(ss);
                }
            }
#pragma unroll
            for (int X = 0; X < 6; X++) mOut[X] = prev[X];
#pragma unroll
            for (int k = 0; k < DW; k++) sv[k] = fin[k];
        }

        /* ---- write back ---- */
        if (active) {
#pragma unroll
            for (int k = 0; k < DW; k++) p.Iout[(long)(wv * DW + k) * N + ij] = sv[k];
            if (wv == 0) {
                it_sum += (unsigned long long)cycles;
                it_max = it_max > (unsigned long long)cycles ? it_max : (unsigned long long)cycles;
                if (p.want_moments) {
#pragma unroll
                    for (int X = 0; X < 6; X++) p.mom[(long)X * N + ij] = mOut[X];
                }
            }
        }
        if (more) {
            __syncthreads(); /* every wavefront is done with slot (j-1)%3 and the exchange buffer */
            store_row(j + 2);
            store_coef(j + 1);
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
