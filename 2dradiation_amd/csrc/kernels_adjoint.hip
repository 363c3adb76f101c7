/*
 * Adjoint of the fused stream-collide step for gfx950 (gradients of a loss
 * on the final state with respect to the per-cell emissivity eta and
 * absorption kappa0), plus the observer-ring sampling and its transpose.
 *
 * The forward step being differentiated (fluid velocity u = 0) is
 *   J = Stream(I)   (src/Radiation.cpp:380-414 flat, :454-512 curved, with the
 *                    clamp >= 0 of IntensityAt, :342)
 *   v_0 = (E, Fx, Fy)(J),  v_k = q m(J) + a q (eta + k0 v_{k-1}[0]) S0
 *                              + cc (v_{k-1}[1] Sx + v_{k-1}[2] Sy)   (k < K)
 *   Inew_d = q J_d + a q (eta + k0 E_{K-1}) + cc (Fx_{K-1} nx_d + Fy_{K-1} ny_d)
 * (Collide, src/Radiation.cpp:641-736, whose lambda cycle is affine in the
 * moments for u = 0), with K recorded per cell by the forward pass.
 *
 * One workgroup = 4 wavefronts walks a 64-column strip down a chunk of rows
 * like the forward kernel.  Per row: the adjoint of the collide is formed
 * per cell (forward-mode tangents for eta and kappa0, reverse mode through
 * the 3x3 moment recursion for J); then every (cell, direction) recomputes
 * its gather geometry from the tables and scatters its share into an LDS
 * accumulation ring of three rows with LDS atomics.  A row is complete once
 * the row below it has scattered; it is then added to the global adjoint
 * with one global atomic per element (neighbouring strips and chunks share
 * their edge rows and columns) and its slot is cleared for reuse.
 */
#include "fused_common.h"

namespace {
using namespace rtf;

template <typename T>
struct AP {
    T *lam_in;         /* [D][N] adjoint of the state after the step (state units); its
                        * active cells are cleared once read, so that the buffer can
                        * take the next step's scatter */
    T *lam_out;        /* [D][N] adjoint of the state before the step (accumulated) */
    const T *Inext;    /* [D][N] state after the step */
    const double *rec_m;
    const uint8_t *rec_k;
    const unsigned long long *rec_mask;
    int rec_waves; /* wavefronts of the recording launch: 4, or 8 (two words per adjoint wavefront) */
    const T *cell;
    const uint8_t *mask;
    const T *coefT;
    const unsigned long long *excl;
    const uint32_t *arec; /* FP32 curved: angular records of the forward (k_angle_rec) */
    const DirTables *dt;
    double *grad_eta, *grad_k0;
    int nx, ny;
    long N;
    int chunks, rows_per_block, nblocks;
    double startx, starty, dx, dy;
};

template <typename T, int D, bool CURVED>
struct ASmem {
    static constexpr int RW = 66;
    /* adjoint accumulation, slot = row % 3; FP64 in both precisions
     * (ds_add_f64: the FP32 LDS add, ds_add_f32, measured about 5x slower
     * here) */
    double A[3][RT_RING_ROWS(D)][RW];
    T C[CURVED ? 25 : 1][64];               /* coefficient series (d_coefT form) */
    unsigned long long EX[CURVED ? 64 : 1]; /* excised-departure bits */
    double X[RT_FUSED_WAVES][4][64];
    float wrow[11][12] __attribute__((aligned(16)));
};

template <typename T, int D, bool CURVED>
constexpr int adj_min_blocks()
{
    return sizeof(ASmem<T, D, CURVED>) <= 80 * 1024 ? 2 : 1;
}

template <typename T, int D, bool CURVED>
__global__ __launch_bounds__(256, (adj_min_blocks<T, D, CURVED>())) void k_adjoint(AP<T> p)
{
    constexpr int DW = D / RT_FUSED_WAVES;
    constexpr int RW = ASmem<T, D, CURVED>::RW;
    constexpr int NA = RT_RING_ROWS(D) * RW;
    __shared__ ASmem<T, D, CURVED> sm;

    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    CDirTables *dtab0 = (CDirTables *)p.dt;
    const long N = p.N;
    const int nx = p.nx;

    int bid = blockIdx.x;
    int tile = bid;
    if ((p.nblocks & 7) == 0) tile = (bid & 7) * (p.nblocks >> 3) + (bid >> 3);
    const int strip = tile / p.chunks;
    const int chunk = tile % p.chunks;
    const int col0 = 64 * strip;
    const int gi = col0 + 1 + lane;
    const bool inside = gi <= nx - 2;
    const int rowA = 1 + chunk * p.rows_per_block;
    const int rowB = min(rowA + p.rows_per_block, p.ny - 1);
    if (rowA >= rowB) return;

    for (int e = threadIdx.x; e < 3 * NA; e += 256) (&sm.A[0][0][0])[e] = 0.0;
    /* FP32 curved: A operands of the matrix-core series (as in the forward) */
    constexpr int NG = (DW + 3) / 4;
    float bas4[(CURVED && sizeof(T) == 4) ? NG : 1][4];
    if constexpr (CURVED && sizeof(T) == 4) basis_groups<DW, NG>(dtab0, wv, lane, bas4);
    if constexpr (CURVED && sizeof(T) == 4) {
        if (threadIdx.x < 132) (&sm.wrow[0][0])[threadIdx.x] = (&dtab0->wrowf[0][0])[threadIdx.x];
    }
    auto stage_coef = [&](int row) {
        if constexpr (CURVED) {
            const int ncell = min(64, nx - 2 - col0);
            const size_t off = (size_t)row * nx + col0 + 1;
            constexpr int E = (int)(sizeof(T) / 4);
            /* FP32 takes its windows from the records: S, X and Y only */
            constexpr int NS = sizeof(T) == 4 ? 15 : 25;
#pragma unroll 1
            for (int e = wv; e < NS + 1; e += RT_FUSED_WAVES) {
                const uint32_t *src = e < NS ? (const uint32_t *)(p.coefT + (size_t)e * N + off)
                                             : (const uint32_t *)(p.excl + off);
                uint32_t *dst = e < NS ? (uint32_t *)&sm.C[e][0] : (uint32_t *)&sm.EX[0];
                const int n = (e < NS ? E : 2) * ncell;
                if (ncell == 64) {
                    /* full strip: whole unmasked instructions */
                    for (int b = 0; b < n; b += 64) __builtin_amdgcn_global_load_lds(src + b + lane, dst + b, 4, 0, 0);
                } else {
                    lds_dma_dwords(src, dst, n, lane);
                }
            }
        }
    };
    /* adds the finished accumulation row `row` into the global adjoint and
     * clears its slot */
    auto flush = [&](int row) {
        /* wavefront = direction, lane = column (64 lanes, then the two
         * right-hand columns); the ring rows that repeat a direction (pad
         * rows 0, 1 and D+2..D+5) are folded into its row first, so every
         * element takes one global atomic per workgroup */
        const int ncol = min(RW, nx - col0);
#pragma unroll 1
        for (int d = wv; d < D; d += RT_FUSED_WAVES) {
            const int rp = d + RT_RING_PAD;
            const int dup = d >= D - 2 ? rp - D : (d < 4 ? rp + D : -1);
            T *dst = p.lam_out + (size_t)d * N + (size_t)row * nx + col0;
            double *src = &sm.A[row % 3][rp][0];
            double *srd = &sm.A[row % 3][dup < 0 ? rp : dup][0];
#pragma unroll
            for (int c0 = 0; c0 < RW; c0 += 64) {
                const int c = c0 + lane;
                if (c < RW) {
                    const double v = src[c] + (dup < 0 ? 0.0 : srd[c]);
                    src[c] = 0.0;
                    srd[c] = 0.0;
                    if (v != 0.0 && c < ncol) unsafeAtomicAdd(dst + c, (T)v);
                }
            }
        }
    };
    /* FP32-mixed: 1/x as v_rcp_f64 refined by one Newton step (x >= 1
     * here); the FP64 kernel divides */
    auto inv = [](double x) -> double {
        if constexpr (sizeof(T) == 4) {
            const double r = __builtin_amdgcn_rcp(x);
            return fma(r, fma(-x, r, 1.0), r);
        } else {
            return 1.0 / x;
        }
    };
    /* Per-cell inputs of a row, loaded one row ahead (issued after the
     * previous row's scatter, so their latency passes behind its flush and
     * barriers): the adjoint after the step (cleared in lam_in as read) and
     * the state after it for this wavefront's directions, the mask, fields,
     * recorded moments, cycle count, clamp bits and angular records. */
    constexpr bool AREC = CURVED && sizeof(T) == 4;
    T r_lam[DW], r_In[DW];
    uint32_t arec[AREC ? DW : 1];
    T r_f[6];
    double r_m[3];
    int r_K = 0;
    uint8_t r_mk = 0;
    unsigned long long r_mlo = 0, r_mhi = 0;
    auto load_row = [&](int row) {
        const unsigned ijr = (unsigned)row * (unsigned)nx + (unsigned)gi;
        r_mk = inside ? p.mask[ijr] : 0;
        const bool act = (r_mk & MASK_ACTIVE) != 0;
#pragma unroll
        for (int k = 0; k < DW; k++) {
            const int d = wv * DW + k;
            r_lam[k] = act ? (p.lam_in + (size_t)d * N)[ijr] : T(0);
            if (act) (p.lam_in + (size_t)d * N)[ijr] = T(0);
            r_In[k] = act ? (p.Inext + (size_t)d * N)[ijr] : T(0);
        }
        if constexpr (AREC) {
#pragma unroll
            for (int k = 0; k < DW; k++) arec[k] = inside ? (p.arec + (size_t)(wv * DW + k) * N)[ijr] : 0u;
        }
        constexpr int FIDX[6] = {CF_ETA, CF_K0, CF_K1, CF_KA, CF_ADT, CF_IA3};
#pragma unroll
        for (int f = 0; f < 6; f++) r_f[f] = act ? (p.cell + (size_t)FIDX[f] * N)[ijr] : T(f == 5 ? 1 : 0);
#pragma unroll
        for (int X = 0; X < 3; X++) r_m[X] = act ? (p.rec_m + (size_t)X * N)[ijr] : 0.0;
        r_K = act ? p.rec_k[ijr] : 0;
        if constexpr (CURVED) {
            r_mlo = r_mhi = 0ull;
            if (act) {
                const int w0 = p.rec_waves == 8 ? 2 * wv : wv;
                r_mlo = D <= 32 ? ((const uint32_t *)p.rec_mask + (size_t)w0 * N)[ijr] : (p.rec_mask + (size_t)w0 * N)[ijr];
                if (p.rec_waves == 8)
                    r_mhi = D <= 32 ? ((const uint32_t *)p.rec_mask + (size_t)(w0 + 1) * N)[ijr]
                                    : (p.rec_mask + (size_t)(w0 + 1) * N)[ijr];
            }
        }
    };
    stage_coef(rowA);
    load_row(rowA);

    for (int j = rowA; j < rowB; j++) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads(); /* coefficients and per-cell inputs of row j in place; accumulation slot of row j+1 clear */
        const bool more = (j + 1 < rowB);
        const unsigned ij = (unsigned)j * (unsigned)nx + (unsigned)gi;
        const uint8_t mk = r_mk;
        const bool active = (mk & MASK_ACTIVE) != 0;
        CDirTables *dtab = dtab0;
        asm volatile("" : "+s"(dtab));

        /* ---- adjoint of the state after the step: partial sums ---- */
        T lam[DW];
        double P0 = 0, Px = 0, Py = 0, PI = 0;
#pragma unroll
        for (int k = 0; k < DW; k++) {
            const int d = wv * DW + k;
            lam[k] = r_lam[k];
            const T In = r_In[k];
            const double l = (double)lam[k];
            P0 += l;
            Px += l * dtab->cx[d];
            Py += l * dtab->cy[d];
            PI += l * (double)In;
        }
        sm.X[wv][0][lane] = P0;
        sm.X[wv][1][lane] = Px;
        sm.X[wv][2][lane] = Py;
        sm.X[wv][3][lane] = PI;
        __syncthreads();
        double Ls[4];
#pragma unroll
        for (int q = 0; q < 4; q++) Ls[q] = ((sm.X[0][q][lane] + sm.X[1][q][lane]) + sm.X[2][q][lane]) + sm.X[3][q][lane];

        /* ---- adjoint of the collide (per cell) ---- */
        double gb[3] = {0, 0, 0};
        double a3 = 1.0, ia3 = 1.0;
        if (active) {
            const double eta = (double)r_f[0];
            const double k0 = (double)r_f[1];
            const double k1 = (double)r_f[2];
            const double kA = (double)r_f[3];
            const double a = (double)r_f[4];
            if (CURVED) {
                ia3 = (double)r_f[5];
                a3 = inv(ia3);
            }
            const double den = 1.0 + a * (kA + k0);
            const double q = inv(den);
            const double cc = 3.0 * k1 * a * q;
            const double dq = -a * q * q, dcc = 3.0 * k1 * a * dq;
            double mI[3];
#pragma unroll
            for (int X = 0; X < 3; X++) mI[X] = r_m[X];
            const int K = r_K;
            /* moments entering the last cycle and their tangents */
            double v[3] = {mI[0], mI[1], mI[2]}, te[3] = {0, 0, 0}, tk[3] = {0, 0, 0};
            for (int cyc = 1; cyc < K; cyc++) {
                const double b = a * q * (eta + k0 * v[0]);
                const double db_de = a * q * (1.0 + k0 * te[0]);
                const double db_dk = a * (dq * (eta + k0 * v[0]) + q * (v[0] + k0 * tk[0]));
                double nv[3], ne[3], nk[3];
#pragma unroll
                for (int X = 0; X < 3; X++) {
                    const double S0 = dtab->S[0][X], Sx = dtab->S[1][X], Sy = dtab->S[2][X];
                    nv[X] = q * mI[X] + b * S0 + cc * (v[1] * Sx + v[2] * Sy);
                    ne[X] = db_de * S0 + cc * (te[1] * Sx + te[2] * Sy);
                    nk[X] = dq * mI[X] + db_dk * S0 + dcc * (v[1] * Sx + v[2] * Sy) + cc * (tk[1] * Sx + tk[2] * Sy);
                }
#pragma unroll
                for (int X = 0; X < 3; X++) {
                    v[X] = nv[X];
                    te[X] = ne[X];
                    tk[X] = nk[X];
                }
            }
            /* adjoint sums in true intensity units */
            const double L0 = ia3 * Ls[0], Lx = ia3 * Ls[1], Ly = ia3 * Ls[2];
            const double b = a * q * (eta + k0 * v[0]);
            /* / q (the FP32-mixed kernel multiplies by den = 1/q) */
            const double num = Ls[3] - b * L0 - cc * (v[1] * Lx + v[2] * Ly);
            const double LJ = sizeof(T) == 4 ? num * den : num / q;
            if (wv == 0) {
                p.grad_eta[ij] += a * q * (1.0 + k0 * te[0]) * L0 + cc * (te[1] * Lx + te[2] * Ly);
                p.grad_k0[ij] += dq * LJ + a * (dq * (eta + k0 * v[0]) + q * v[0] + q * k0 * tk[0]) * L0 +
                                 dcc * (v[1] * Lx + v[2] * Ly) + cc * (tk[1] * Lx + tk[2] * Ly);
            }
            /* reverse through v_k = c0 + A v_{k-1}, A = [a q k0 S0 | cc Sx | cc Sy] */
            double u[3] = {a * q * k0 * L0, cc * Lx, cc * Ly}, Cb[3] = {0, 0, 0};
            for (int cyc = K - 1; cyc >= 1; cyc--) {
                double nu[3];
#pragma unroll
                for (int X = 0; X < 3; X++) Cb[X] += u[X];
                /* A^T u: component Y is column Y of A (the S row of that
                 * column's coefficient) dotted with u */
                nu[0] = a * q * k0 * (dtab->S[0][0] * u[0] + dtab->S[0][1] * u[1] + dtab->S[0][2] * u[2]);
                nu[1] = cc * (dtab->S[1][0] * u[0] + dtab->S[1][1] * u[1] + dtab->S[1][2] * u[2]);
                nu[2] = cc * (dtab->S[2][0] * u[0] + dtab->S[2][1] * u[1] + dtab->S[2][2] * u[2]);
#pragma unroll
                for (int Y = 0; Y < 3; Y++) u[Y] = nu[Y];
            }
#pragma unroll
            for (int X = 0; X < 3; X++) gb[X] = q * Cb[X] + u[X];
            /* J-bar in state units: J' = alpha^-3 J  =>  dL/dJ' = alpha^3 dL/dJ */
#pragma unroll
            for (int k = 0; k < DW; k++) lam[k] = (T)(q * (double)lam[k]);
        }

        /* ---- transpose of the gather (scatter into the LDS ring) ---- */
        const int slm = (j + 2) % 3, sl0 = j % 3, slp = (j + 1) % 3;
        if constexpr (CURVED) {
            T cS[5], cX[5], cY[5], cU[5], cV[5];
#pragma unroll
            for (int k = 0; k < 5; k++) {
                cS[k] = sm.C[k][lane];
                cX[k] = sm.C[5 + k][lane];
                cY[k] = sm.C[10 + k][lane];
                cU[k] = AREC ? T(0) : sm.C[15 + k][lane];
                cV[k] = AREC ? T(0) : sm.C[20 + k][lane];
            }
            const uint32_t exw = (uint32_t)(sm.EX[lane] >> (wv * DW));
            /* the recording wavefronts 2 wv and 2 wv + 1 of an 8-wavefront
             * recording hold this wavefront's first and second DW/2
             * directions */
            const unsigned long long mbits = p.rec_waves == 8 ? (r_mlo | (r_mhi << (2 * DW))) : r_mlo;
            const T angu = (sizeof(T) == 4) ? (T)dtab->ang_uf : (T)dtab->ang_u;
            const int quarter = dtab->quarter, nGrid = dtab->nGrid;
            const T ddmax = T(nGrid) - T(1e-8);
            f32x4 ser[3];
#pragma unroll
            for (int k = 0; k < DW; k++) {
                const int d = wv * DW + k;
                if constexpr (sizeof(T) == 4) {
                    if (k % 4 == 0) {
                        ser[0] = series4_mfma(cS, bas4[k / 4]);
                        ser[1] = series4_mfma(cX, bas4[k / 4]);
                        ser[2] = series4_mfma(cY, bas4[k / 4]);
                    }
                }
                const double jb = (double)lam[k] + a3 * dtab->w6[d][0] *
                                                       (gb[0] + gb[1] * dtab->cx[d] + gb[2] * dtab->cy[d]);
                const bool go = active && !((exw >> k) & 1u) && jb != 0.0;
                if (__any(go)) {
                    DirGeom<T> G;
                    if constexpr (sizeof(T) == 4) {
                        const int i = k % 4;
                        geometry_from_record<D>(ser[0][i], ser[1][i], ser[2][i], arec[k], &sm.wrow[0][0], G);
                    } else {
                        curved_geometry<T, D>(cS, cX, cY, cU, cV, dtab, dtab0, &sm.wrow[0][0], d, angu, quarter,
                                              nGrid, ddmax, G);
                    }
                    if (go) {
                        const T s3 = T(1) + G.s * (T(3) + G.s * (T(3) + G.s));
                        const T f = (T)jb * s3;
                        const unsigned bits = (unsigned)(mbits >> (4 * k)) & 15u;
                        /* corner weights to FP64 once; the 20 products in FP64 */
                        const double w00 = (bits & 1u) ? (double)(f * (T(1) - G.fx) * (T(1) - G.fy)) : 0.0;
                        const double w01 = (bits & 2u) ? (double)(f * (T(1) - G.fx) * G.fy) : 0.0;
                        const double w10 = (bits & 4u) ? (double)(f * G.fx * (T(1) - G.fy)) : 0.0;
                        const double w11 = (bits & 8u) ? (double)(f * G.fx * G.fy) : 0.0;
                        const int ra = G.lo_j ? slm : sl0, rb = G.lo_j ? sl0 : slp;
                        const int ca = lane + (G.lo_i ? 0 : 1);
#pragma unroll
                        for (int m = 0; m < 5; m++) {
                            double *pa = &sm.A[ra][G.s0 + m][ca];
                            double *pb = &sm.A[rb][G.s0 + m][ca];
                            const double cw = (double)G.Cw[m];
                            atomicAdd(pa, w00 * cw);
                            atomicAdd(pa + 1, w10 * cw);
                            atomicAdd(pb, w01 * cw);
                            atomicAdd(pb + 1, w11 * cw);
                        }
                    }
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < DW; k++) {
                const int d = wv * DW + k;
                const double jb = (double)lam[k] + dtab->w6[d][0] * (gb[0] + gb[1] * dtab->cx[d] + gb[2] * dtab->cy[d]);
                if (active && jb != 0.0) {
                    const int oi = dtab->flat_di[d], oj = dtab->flat_dj[d];
                    const T fx = (sizeof(T) == 4) ? (T)dtab->flat_fxf[d] : (T)dtab->flat_fx[d];
                    const T fy = (sizeof(T) == 4) ? (T)dtab->flat_fyf[d] : (T)dtab->flat_fy[d];
                    const T f = (T)jb;
                    const int ra = oj < 0 ? slm : sl0, rb = oj < 0 ? sl0 : slp;
                    const int ca = lane + 1 + oi;
                    const int dr = d + RT_RING_PAD;
                    atomicAdd(&sm.A[ra][dr][ca], (double)(f * (T(1) - fx) * (T(1) - fy)));
                    atomicAdd(&sm.A[ra][dr][ca + 1], (double)(f * fx * (T(1) - fy)));
                    atomicAdd(&sm.A[rb][dr][ca], (double)(f * (T(1) - fx) * fy));
                    atomicAdd(&sm.A[rb][dr][ca + 1], (double)(f * fx * fy));
                }
            }
        }
        if (more) load_row(j + 1);
        __syncthreads(); /* every scatter of row j is in the ring; row j-1 is complete */
        flush(j - 1);
        if (more) stage_coef(j + 1);
    }
    __syncthreads();
    flush(rowB - 1);
    flush(rowB);
}

/* Observer samples: out[p][d] = bilinear of I_d (true units) at (x_p, y_p) for
 * the directions with n_d . (x_p, y_p) > 0, 0 otherwise. */
template <typename T>
__global__ void k_sample(const T *I, const T *cell, int scaled, const double *cx, const double *cy, int D, long N,
                         int nx, double startx, double starty, double dx, double dy, int npts, const double *px,
                         const double *py, double *out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts * D) return;
    const int pnt = t / D, d = t - pnt * D;
    const double x = px[pnt], y = py[pnt];
    double o = 0;
    if (cx[d] * x + cy[d] * y > 0) {
        const double iT = (x - startx) / dx, jT = (y - starty) / dy;
        const long i0 = (long)floor(iT), j0 = (long)floor(jT);
        const double fx = iT - i0, fy = jT - j0;
        const long k[4] = {i0 + j0 * nx, i0 + (j0 + 1) * nx, i0 + 1 + j0 * nx, i0 + 1 + (j0 + 1) * nx};
        const double bw[4] = {(1 - fx) * (1 - fy), (1 - fx) * fy, fx * (1 - fy), fx * fy};
        for (int c = 0; c < 4; c++) {
            double v = (double)I[(size_t)d * N + k[c]];
            if (scaled) v /= (double)cell[(size_t)CF_IA3 * N + k[c]];
            o += bw[c] * v;
        }
    }
    out[t] = o;
}

/* Transpose of k_sample: lam[d][k] += g[p][d] * dI_true/dI_state * weight. */
template <typename T>
__global__ void k_sample_adjoint(T *lam, const T *cell, int scaled, const double *cx, const double *cy, int D, long N,
                                 int nx, double startx, double starty, double dx, double dy, int npts,
                                 const double *px, const double *py, const double *g)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts * D) return;
    const int pnt = t / D, d = t - pnt * D;
    const double x = px[pnt], y = py[pnt];
    if (!(cx[d] * x + cy[d] * y > 0) || g[t] == 0.0) return;
    const double iT = (x - startx) / dx, jT = (y - starty) / dy;
    const long i0 = (long)floor(iT), j0 = (long)floor(jT);
    const double fx = iT - i0, fy = jT - j0;
    const long k[4] = {i0 + j0 * nx, i0 + (j0 + 1) * nx, i0 + 1 + j0 * nx, i0 + 1 + (j0 + 1) * nx};
    const double bw[4] = {(1 - fx) * (1 - fy), (1 - fx) * fy, fx * (1 - fy), fx * fy};
    for (int c = 0; c < 4; c++) {
        double v = g[t] * bw[c];
        if (scaled) v /= (double)cell[(size_t)CF_IA3 * N + k[c]];
        unsafeAtomicAdd(lam + (size_t)d * N + k[c], (T)v);
    }
}

template <typename T, int D, bool CURVED>
int launch_adj(rt_ctx *c, const void *lam_in, void *lam_out, const void *Inext, const RecordPtrs &rec)
{
    const rt_desc &d = c->desc;
    AP<T> p;
    p.lam_in = (T *)const_cast<void *>(lam_in);
    p.lam_out = (T *)lam_out;
    p.Inext = (const T *)Inext;
    p.rec_m = rec.m;
    p.rec_k = rec.k;
    p.rec_mask = rec.mask;
    p.rec_waves = c->rec_waves;
    p.cell = (const T *)c->d_cell;
    p.mask = c->d_mask;
    p.coefT = (const T *)c->d_coefT;
    p.excl = c->d_excl;
    p.arec = c->d_arec;
    p.dt = c->d_dirt;
    p.grad_eta = c->d_grad[0];
    p.grad_k0 = c->d_grad[1];
    p.nx = d.nx;
    p.ny = d.ny;
    p.N = (long)d.nx * d.ny;
    const int strips = (d.nx - 2 + 63) / 64;
    const int rows = d.ny - 2;
    /* one round of resident workgroups (as the forward step) */
    static int per_cu = 0;
    if (per_cu == 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_adjoint<T, D, CURVED>, 256, 0) != hipSuccess || n < 1)
            n = 1;
        per_cu = n;
    }
    const int wgs = per_cu * c->ncu;
    int chunks = (wgs + strips - 1) / strips;
    int H = (rows + chunks - 1) / chunks;
    /* chunks down to 2 rows: on small grids that fills the second resident
     * workgroup per CU the LDS allows, at the price of more halo-row atomics
     * (every element of a chunk's H + 2 completed rows takes one global
     * atomic; FP32 256^2 x 32, 500 steps with their recomputation: 1 row
     * 27.2 ms, 2 rows 22.7, 3 rows 24.8, 4 rows 26.6) */
    H = H < 2 ? 2 : (H > 64 ? 64 : H);
    chunks = (rows + H - 1) / H;
    p.chunks = chunks;
    p.rows_per_block = H;
    p.nblocks = strips * chunks;
    p.startx = d.startx;
    p.starty = d.starty;
    p.dx = d.dx;
    p.dy = d.dy;
    k_adjoint<T, D, CURVED><<<p.nblocks, 256, 0, c->stream>>>(p);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}

template <typename T, bool CURVED>
int adj_dispatch(rt_ctx *c, const void *li, void *lo, const void *In, const RecordPtrs &rec)
{
    switch (c->D) {
    case 16: return launch_adj<T, 16, CURVED>(c, li, lo, In, rec);
    case 20: return launch_adj<T, 20, CURVED>(c, li, lo, In, rec);
    case 24: return launch_adj<T, 24, CURVED>(c, li, lo, In, rec);
    case 32: return launch_adj<T, 32, CURVED>(c, li, lo, In, rec);
    case 48: return launch_adj<T, 48, CURVED>(c, li, lo, In, rec);
    case 64: return launch_adj<T, 64, CURVED>(c, li, lo, In, rec);
    default: return rt_fail(c, RT_ERR_UNSUPPORTED, "adjoint: unsupported nDir");
    }
}

unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }
} // namespace

int launch_adjoint_step(rt_ctx *c, const void *lam_in, void *lam_out, const void *Inext, const RecordPtrs &rec)
{
    const bool curved = c->desc.streaming == RT_CURVED_FIXED;
    if (c->fp32)
        return curved ? adj_dispatch<float, true>(c, lam_in, lam_out, Inext, rec)
                      : adj_dispatch<float, false>(c, lam_in, lam_out, Inext, rec);
    return curved ? adj_dispatch<double, true>(c, lam_in, lam_out, Inext, rec)
                  : adj_dispatch<double, false>(c, lam_in, lam_out, Inext, rec);
}

int launch_sample(rt_ctx *c, const void *I, int scaled, int npts, const double *px, const double *py, double *out)
{
    const rt_desc &d = c->desc;
    const size_t n = (size_t)npts * c->D;
    if (c->fp32)
        k_sample<float><<<nblk(n, 256), 256, 0, c->stream>>>((const float *)I, (const float *)c->d_cell, scaled, c->d_cx,
                                                             c->d_cy, c->D, c->N, d.nx, d.startx, d.starty, d.dx, d.dy,
                                                             npts, px, py, out);
    else
        k_sample<double><<<nblk(n, 256), 256, 0, c->stream>>>((const double *)I, (const double *)c->d_cell, scaled,
                                                              c->d_cx, c->d_cy, c->D, c->N, d.nx, d.startx, d.starty,
                                                              d.dx, d.dy, npts, px, py, out);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}

int launch_sample_adjoint(rt_ctx *c, void *lam, int scaled, int npts, const double *px, const double *py,
                          const double *g)
{
    const rt_desc &d = c->desc;
    const size_t n = (size_t)npts * c->D;
    if (c->fp32)
        k_sample_adjoint<float><<<nblk(n, 256), 256, 0, c->stream>>>((float *)lam, (const float *)c->d_cell, scaled,
                                                                     c->d_cx, c->d_cy, c->D, c->N, d.nx, d.startx,
                                                                     d.starty, d.dx, d.dy, npts, px, py, g);
    else
        k_sample_adjoint<double><<<nblk(n, 256), 256, 0, c->stream>>>((double *)lam, (const double *)c->d_cell,
                                                                      scaled, c->d_cx, c->d_cy, c->D, c->N, d.nx,
                                                                      d.startx, d.starty, d.dx, d.dy, npts, px, py, g);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}
