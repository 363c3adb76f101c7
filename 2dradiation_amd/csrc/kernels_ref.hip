/*
 * Reference-order kernels (compiled with -ffp-contract=off):
 *  - layout: AoS FP64 host intensities <-> SoA [d][cell] device intensities
 *  - per-method stream / moments / collide in FP64 arithmetic with the
 *    reference's operation order (bitwise comparable with a strict-IEEE CPU
 *    build of the same formulas)
 *  - the FP64 geodesic-table kernel (adaptive RK45 backtrace + Fourier fit)
 */
#include <cstring>
#include "rt_internal.h"
#include "ref_collide.h"

namespace {

constexpr int TB = 32; /* transpose tile */

__device__ inline long clampl(long v, long hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

/* AoS [N][D] double -> SoA [D][N] T through a padded LDS tile. */
template <typename T>
__global__ __launch_bounds__(256) void k_transpose_in(const double *__restrict__ aos, T *__restrict__ soa,
                                                      int N, int D)
{
    __shared__ double tile[TB][TB + 1];
    int c0 = blockIdx.x * TB, d0 = blockIdx.y * TB;
    int tx = threadIdx.x % TB, ty = threadIdx.x / TB; /* 32 x 8 */
    for (int r = ty; r < TB; r += 8) {
        int c = c0 + r, d = d0 + tx;
        tile[r][tx] = (c < N && d < D) ? aos[(size_t)c * D + d] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < TB; r += 8) {
        int d = d0 + r, c = c0 + tx;
        if (c < N && d < D) soa[(size_t)d * N + c] = (T)tile[tx][r];
    }
}
template <typename T>
__global__ __launch_bounds__(256) void k_transpose_out(const T *__restrict__ soa, double *__restrict__ aos,
                                                       int N, int D)
{
    __shared__ double tile[TB][TB + 1];
    int c0 = blockIdx.x * TB, d0 = blockIdx.y * TB;
    int tx = threadIdx.x % TB, ty = threadIdx.x / TB;
    for (int r = ty; r < TB; r += 8) {
        int d = d0 + r, c = c0 + tx;
        tile[r][tx] = (c < N && d < D) ? (double)soa[(size_t)d * N + c] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < TB; r += 8) {
        int c = c0 + r, d = d0 + tx;
        if (c < N && d < D) aos[(size_t)c * D + d] = tile[tx][r];
    }
}

__global__ void k_fill(double *p, size_t n, double v)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

/* Halo and excised cells never change (src/Radiation.cpp:413, 511, 648-649,
 * 735): copy them from the state buffer to the scratch buffer so that either
 * buffer can serve as the next state. */
template <typename T>
__global__ void k_copy_fixed(const T *__restrict__ src, T *__restrict__ dst, const uint8_t *__restrict__ mask,
                             int N, int D)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)N * D) return;
    int ij = (int)(i % N);
    if (!(mask[ij] & MASK_ACTIVE)) dst[i] = src[i];
}

/* Curved streaming multiplies every gathered intensity by (alpha/alpha_k)^3
 * (src/Radiation.cpp:497-505).  The fused kernel keeps the state as
 * alpha^-3 I, so the factor becomes the destination's alpha^3 alone. */
template <typename T>
__global__ void k_scale_state(T *__restrict__ I, const T *__restrict__ ia3, int N, int D, int inverse)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)N * D) return;
    const T s = ia3[i % N];
    if (!(s > T(0)) || !(s < T(INFINITY))) return;
    I[i] = inverse ? I[i] / s : I[i] * s;
}

/* Cell mask: active = interior and not InsideBH at the cell centre
 * (src/Radiation.cpp:644-649); near = the departure point may be excised. */
__global__ void k_mask(uint8_t *mask, int nx, int ny, double startx, double starty, double dx, double dy,
                       double r2, double near_r2)
{
    int ij = blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= nx * ny) return;
    int i = ij % nx, j = ij / nx;
    double x = startx + i * dx, y = starty + j * dy;
    double rr = x * x + y * y;
    bool interior = i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1;
    bool inside = r2 >= 0.0 && rr <= r2;
    uint8_t m = 0;
    if (interior && !inside) m |= MASK_ACTIVE;
    if (r2 >= 0.0 && rr <= near_r2) m |= MASK_NEAR_BH;
    mask[ij] = m;
}

/* Per-cell inputs of the fused kernel: fields, the inertial-frame fluid
 * velocity uIF = tetrad * uLF (the reference passes the tetrad, not its
 * inverse, src/Radiation.cpp:652-653), alpha*dt and alpha^-3. */
template <typename T>
__global__ void k_prepare(const double *__restrict__ f, const double *__restrict__ alpha,
                          const double *__restrict__ t11, const double *__restrict__ t12,
                          const double *__restrict__ t21, const double *__restrict__ t22, T *__restrict__ cell,
                          const uint8_t *__restrict__ mask, int N, double dt)
{
    int ij = blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= N) return;
    double ux = f[4 * (size_t)N + ij], uy = f[5 * (size_t)N + ij];
    double u1 = t11[ij] * ux + t12[ij] * uy;
    double u2 = t21[ij] * ux + t22[ij] * uy;
    double a = alpha[ij];
    cell[CF_ETA * (size_t)N + ij] = (T)f[ij];
    cell[CF_K0 * (size_t)N + ij] = (T)f[(size_t)N + ij];
    cell[CF_K1 * (size_t)N + ij] = (T)f[2 * (size_t)N + ij];
    cell[CF_KA * (size_t)N + ij] = (T)f[3 * (size_t)N + ij];
    cell[CF_UX * (size_t)N + ij] = (T)u1;
    cell[CF_UY * (size_t)N + ij] = (T)u2;
    cell[CF_ADT * (size_t)N + ij] = (T)(a * dt);
    cell[CF_IA3 * (size_t)N + ij] = (T)(1.0 / (a * a * a));
    cell[CF_ACT * (size_t)N + ij] = (mask[ij] & MASK_ACTIVE) ? T(1) : T(0);
}

/* ---- reference-order stream (src/Radiation.cpp:380-414) ---- */
template <typename T>
__global__ void k_ref_stream_flat(const T *__restrict__ I, T *__restrict__ Inew, const double *__restrict__ cx,
                                  const double *__restrict__ cy, int nx, int ny, int D, double startx,
                                  double starty, double dx, double dy, double dt)
{
    const int nix = nx - 2, niy = ny - 2;
    const long Ni = (long)nix * niy;
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= Ni * D) return;
    int d = (int)(t / Ni);
    long c = t % Ni;
    int i = 1 + (int)(c % nix), j = 1 + (int)(c / nix);
    const long N = (long)nx * ny;
    double xT = startx + i * dx, yT = starty + j * dy;
    xT -= cx[d] * dt;
    yT -= cy[d] * dt;
    double iT = (xT - startx) / dx, jT = (yT - starty) / dy;
    long i0 = (long)floor(iT), j0 = (long)floor(jT), i1 = i0 + 1, j1 = j0 + 1;
    /* departures beyond the halo are undefined in the reference; keep the
     * gather inside the grid */
    i0 = clampl(i0, nx - 1); i1 = clampl(i1, nx - 1); j0 = clampl(j0, ny - 1); j1 = clampl(j1, ny - 1);
    const T *Id = I + (long)d * N;
    double v = rtm::bilinear<double>(iT - i0, jT - j0, (double)Id[i0 + j0 * nx], (double)Id[i0 + j1 * nx],
                                     (double)Id[i1 + j0 * nx], (double)Id[i1 + j1 * nx]);
    Inew[(long)d * N + i + (long)j * nx] = (T)v;
}

/* IntensityAt (src/Radiation.cpp:322-343); rotation angle 0 is the identity. */
template <typename T>
__device__ inline double ref_intensity_at(const T *__restrict__ I, long N, long k, double vx, double vy,
                                          const int32_t *__restrict__ Nc, const double *__restrict__ Wc,
                                          int nGrid)
{
    double angle = rtm::my_phi(vx, vy);
    double nG = (double)nGrid;
    double dd = nG * angle / (2.0 * rtm::kPi);
    if (!(dd >= 1e-8)) dd = 1e-8; /* NaN angle: keep the table reads in range */
    if (dd > nG - 1e-8) dd = nG - 1e-8;
    /* floor(dd) < nGrid, so the reference's "% nGrid" of d0 is the identity */
    const int d0 = (int)floor(dd);
    const int d1 = (d0 + 1 == nGrid) ? 0 : d0 + 1;
    double v0 = 0, v1 = 0;
    for (int q = 0; q < 4; q++) {
        v0 += Wc[4 * d0 + q] * (double)I[(long)Nc[4 * d0 + q] * N + k];
        v1 += Wc[4 * d1 + q] * (double)I[(long)Nc[4 * d1 + q] * N + k];
    }
    double t = dd - d0;
    double r = v0 * (1.0 - t) + v1 * t;
    return r > 0.0 ? r : 0.0;
}

/* ---- reference-order curved stream (src/Radiation.cpp:454-512) ---- */
template <typename T>
__global__ void k_ref_stream_curved(const T *__restrict__ I, T *__restrict__ Inew,
                                    const double *__restrict__ basis, const double *__restrict__ coef,
                                    const double *__restrict__ alpha, const int32_t *__restrict__ Nc,
                                    const double *__restrict__ Wc, int nGrid, int nF, int nx, int ny, int D,
                                    double startx, double starty, double dx, double dy, double r2)
{
    const int nix = nx - 2, niy = ny - 2;
    const long Ni = (long)nix * niy;
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= Ni * D) return;
    int d = (int)(t / Ni);
    long c = t % Ni;
    int i = 1 + (int)(c % nix), j = 1 + (int)(c / nix);
    const long N = (long)nx * ny;
    long ij = i + (long)j * nx;
    double xc = startx + i * dx, yc = starty + j * dy;
    T *out = Inew + (long)d * N + ij;
    if (r2 >= 0.0 && xc * xc + yc * yc <= r2) {
        *out = (T)0;
        return;
    }
    double al = alpha[ij];
    const double *B = basis + (long)d * nF;
    double q[5];
    for (int s = 0; s < 5; s++) {
        double r = 0;
        for (int k = 0; k < nF; k++) r += coef[((long)s * nF + k) * N + ij] * B[k];
        q[s] = r;
    }
    double sh = q[0], xT = q[1], yT = q[2], vx = q[3], vy = q[4];
    if (r2 >= 0.0 && xT * xT + yT * yT <= r2) {
        *out = (T)0;
        return;
    }
    double iT = (xT - startx) / dx, jT = (yT - starty) / dy;
    if (!(iT >= 0.0)) iT = 0.0;
    if (!(jT >= 0.0)) jT = 0.0;
    if (iT > nx - 1.0) iT = nx - 1.0;
    if (jT > ny - 1.0) jT = ny - 1.0;
    long i0 = (long)floor(iT), j0 = (long)floor(jT), i1 = i0 + 1, j1 = j0 + 1;
    i1 = clampl(i1, nx - 1); j1 = clampl(j1, ny - 1);
    long k00 = i0 + j0 * nx, k01 = i0 + j1 * nx, k10 = i1 + j0 * nx, k11 = i1 + j1 * nx;
    double r00 = al / alpha[k00], r01 = al / alpha[k01], r10 = al / alpha[k10], r11 = al / alpha[k11];
    double f00 = (r00 * (r00 * r00)) * ref_intensity_at(I, N, k00, vx, vy, Nc, Wc, nGrid);
    double f01 = (r01 * (r01 * r01)) * ref_intensity_at(I, N, k01, vx, vy, Nc, Wc, nGrid);
    double f10 = (r10 * (r10 * r10)) * ref_intensity_at(I, N, k10, vx, vy, Nc, Wc, nGrid);
    double f11 = (r11 * (r11 * r11)) * ref_intensity_at(I, N, k11, vx, vy, Nc, Wc, nGrid);
    *out = (T)((sh * (sh * sh)) * rtm::bilinear<double>(iT - i0, jT - j0, f00, f01, f10, f11));
}

/* ---- ComputeMomentsIF (src/Radiation.cpp:239-269), every stored cell ----
 * rot (adaptive streaming, else null): directions RotationMatrix(rot[ij]) C(d)
 * (:257; src/DataTypes.hh:630-635) */
template <typename T>
__global__ void k_ref_moments(const T *__restrict__ I, double *__restrict__ mom, const double *__restrict__ w,
                              const double *__restrict__ cx, const double *__restrict__ cy, int N, int D,
                              const double *__restrict__ rot)
{
    int ij = blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= N) return;
    double e = 0, fx = 0, fy = 0, pxx = 0, pxy = 0, pyy = 0;
    const double rs = rot ? rtm::my_sin(rot[ij]) : 0.0, rc = rot ? rtm::my_cos(rot[ij]) : 1.0;
    /* FP32: 8 directions' loads are issued before their sums (same
     * summation order as one direction at a time; FP64 measured faster
     * one at a time) */
    constexpr int B = sizeof(T) == 4 ? 8 : 1;
    for (int d0 = 0; d0 < D; d0 += B) {
        T v[B];
#pragma unroll
        for (int k = 0; k < B; k++)
            if (d0 + k < D) v[k] = I[(long)(d0 + k) * N + ij];
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int d = d0 + k;
            if (d >= D || w[d] == 0.0) continue;
            double n1 = cx[d], n2 = cy[d];
            if (rot) {
                n1 = cx[d] * rc - cy[d] * rs;
                n2 = cx[d] * rs + cy[d] * rc;
            }
            double c = w[d] * (double)v[k];
            e += c;
            fx += c * n1;
            fy += c * n2;
            pxx += c * n1 * n1;
            pxy += c * n1 * n2;
            pyy += c * n2 * n2;
        }
    }
    mom[ij] = e;
    mom[(long)N + ij] = fx;
    mom[2L * N + ij] = fy;
    mom[3L * N + ij] = pxx;
    mom[4L * N + ij] = pxy;
    mom[5L * N + ij] = pyy;
}

/* ---- Collide (src/Radiation.cpp:641-724), interior active cells ---- */
template <typename T>
__global__ void k_ref_collide(const T *__restrict__ I, T *__restrict__ Inew, double *__restrict__ mom,
                              int32_t *__restrict__ count, const double *__restrict__ f,
                              const double *__restrict__ alpha, const double *__restrict__ t11,
                              const double *__restrict__ t12, const double *__restrict__ t21,
                              const double *__restrict__ t22, const uint8_t *__restrict__ mask,
                              const double *__restrict__ w, const double *__restrict__ cx,
                              const double *__restrict__ cy, int N, int D, double dt,
                              unsigned long long *stats, const double *__restrict__ rot)
{
    int ij = blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= N) return;
    if (!(mask[ij] & MASK_ACTIVE)) {
        count[ij] = 0;
        return;
    }
    const int cycles = ref_collide_cell<T>(ij, I, Inew, mom, f, alpha, t11, t12, t21, t22, w, cx, cy, N, D, dt, rot);
    count[ij] = cycles;
    unsigned long long *st = rt_stat_slot(stats, (unsigned)ij);
    atomicAdd(&st[0], (unsigned long long)cycles);
    atomicMax(&st[1], (unsigned long long)cycles);
}

/* ------------------------------------------------------------------------
 * Geodesic tables (src/Radiation.cpp:191-237; src/GeodesicEquationSolver.cpp)
 * ---------------------------------------------------------------------- */
struct GeoGrid {
    int nx, ny;
    double dx, dy, startx, starty, endx, endy, dt, r2;
};
__device__ inline bool geo_outside(const GeoGrid &g, double x, double y)
{
    return x > g.endx || x < g.startx || y > g.endy || y < g.starty;
}
__device__ inline bool geo_inside_bh(const GeoGrid &g, double x, double y)
{
    return g.r2 >= 0.0 && (x * x + y * y) <= g.r2;
}
/* Bilinear interpolation of metric quantities q0..q0+nq-1 of the per-cell
 * block at fractional grid position (fi,fj) (src/Metric.cpp:338-359).  Upper
 * neighbours are clamped into the grid; they then carry zero weight. */
template <int NQ>
__device__ inline void geo_interp(const double *__restrict__ M, const GeoGrid &g, double fi, double fj,
                                  const int (&qs)[NQ], double (&out)[NQ])
{
    if (!(fi >= 0.0)) fi = 0.0; /* NaN-safe: keeps every gather inside the grid */
    if (!(fj >= 0.0)) fj = 0.0;
    if (fi > g.nx - 1.0) fi = g.nx - 1.0;
    if (fj > g.ny - 1.0) fj = g.ny - 1.0;
    long i0 = (long)floor(fi), j0 = (long)floor(fj);
    long i1 = i0 + 1 < g.nx ? i0 + 1 : g.nx - 1;
    long j1 = j0 + 1 < g.ny ? j0 + 1 : g.ny - 1;
    double x = fi - i0, y = fj - j0;
    const double *p00 = M + (i0 + j0 * g.nx) * RT_NMETRIC, *p01 = M + (i0 + j1 * g.nx) * RT_NMETRIC;
    const double *p10 = M + (i1 + j0 * g.nx) * RT_NMETRIC, *p11 = M + (i1 + j1 * g.nx) * RT_NMETRIC;
#pragma unroll
    for (int k = 0; k < NQ; k++) {
        int q = qs[k];
        out[k] = rtm::bilinear<double>(x, y, p00[q], p01[q], p10[q], p11[q]);
    }
}
struct GeoLocal {
    double alpha, b1, b2, da1, da2, db[2][2], K[2][2], gu[2][2], dg[2][2][2];
};
__device__ inline void geo_local(const double *__restrict__ M, const GeoGrid &g, double x, double y,
                                 GeoLocal &L)
{
    if (geo_outside(g, x, y)) {
        L.alpha = 1.0;
        L.b1 = L.b2 = L.da1 = L.da2 = 0;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                L.db[a][b] = 0;
                L.K[a][b] = 0;
                L.gu[a][b] = (a == b) ? 1.0 : 0.0;
                L.dg[0][a][b] = L.dg[1][a][b] = 0;
            }
        return;
    }
    double fi = (x - g.startx) / g.dx, fj = (y - g.starty) / g.dy;
    const int qs[21] = {RT_M_ALPHA,        RT_M_BETA1_U,      RT_M_BETA2_U,      RT_M_D1_ALPHA,
                        RT_M_D2_ALPHA,     RT_M_D1_BETA1_U,   RT_M_D1_BETA2_U,   RT_M_D2_BETA1_U,
                        RT_M_D2_BETA2_U,   RT_M_K11,          RT_M_K12,          RT_M_K22,
                        RT_M_GAMMA11_UU,   RT_M_GAMMA12_UU,   RT_M_GAMMA22_UU,   RT_M_D1_GAMMA11_LL,
                        RT_M_D1_GAMMA12_LL, RT_M_D1_GAMMA22_LL, RT_M_D2_GAMMA11_LL, RT_M_D2_GAMMA12_LL,
                        RT_M_D2_GAMMA22_LL};
    double v[21];
    geo_interp<21>(M, g, fi, fj, qs, v);
    L.alpha = v[0];
    L.b1 = v[1];
    L.b2 = v[2];
    L.da1 = v[3];
    L.da2 = v[4];
    L.db[0][0] = v[5];
    L.db[0][1] = v[6];
    L.db[1][0] = v[7];
    L.db[1][1] = v[8];
    L.K[0][0] = v[9];
    L.K[0][1] = L.K[1][0] = v[10];
    L.K[1][1] = v[11];
    L.gu[0][0] = v[12];
    L.gu[0][1] = L.gu[1][0] = v[13];
    L.gu[1][1] = v[14];
    L.dg[0][0][0] = v[15];
    L.dg[0][0][1] = L.dg[0][1][0] = v[16];
    L.dg[0][1][1] = v[17];
    L.dg[1][0][0] = v[18];
    L.dg[1][0][1] = L.dg[1][1][0] = v[19];
    L.dg[1][1][1] = v[20];
}
/* Dx, Dv, Dnu (src/GeodesicEquationSolver.cpp:5-55). */
__device__ inline void geo_rhs(const double *__restrict__ M, const GeoGrid &g, double x, double y, double vx,
                               double vy, double nu, double kx[2], double kv[2], double &knu)
{
    GeoLocal L;
    geo_local(M, g, x, y, L);
    double v[2] = {vx, vy}, dA[2] = {L.da1, L.da2}, bu[2] = {L.b1, L.b2};
    for (int k = 0; k < 2; k++) kx[k] = L.alpha * v[k] - bu[k];
    for (int k = 0; k < 2; k++) {
        double dv = 0.0;
        for (int i = 0; i < 2; i++) {
            dv += v[k] * dA[i] * v[i] - L.gu[k][i] * dA[i] - v[i] * L.db[i][k];
            for (int j = 0; j < 2; j++) {
                dv += 2.0 * L.alpha * L.gu[i][k] * L.K[i][j] * v[j] - L.alpha * v[k] * L.K[i][j] * v[i] * v[j];
                for (int l = 0; l < 2; l++)
                    dv -= L.alpha * 0.5 * L.gu[k][l] * (L.dg[i][l][j] + L.dg[j][i][l] - L.dg[l][i][j]) * v[i] *
                          v[j];
            }
        }
        kv[k] = dv;
    }
    double dnu = 0;
    for (int i = 0; i < 2; i++) {
        dnu -= v[i] * dA[i];
        for (int j = 0; j < 2; j++) dnu += L.alpha * L.K[i][j] * v[i] * v[j];
    }
    knu = dnu * nu;
}
/* Adaptive RK45 backward over dt (src/GeodesicEquationSolver.cpp:78-188);
 * the stage-6 coefficient 11/44 is the reference's. */
__device__ double geo_rk45(const double *__restrict__ M, const GeoGrid &g, double tol, double dt, double &px,
                           double &py, double &pvx, double &pvy)
{
    const double td = -1.0;
    double x[2] = {px, py}, v[2] = {pvx, pvy};
    double k1x[2], k2x[2], k3x[2], k4x[2], k5x[2], k6x[2];
    double k1v[2], k2v[2], k3v[2], k4v[2], k5v[2], k6v[2];
    double k1n, k2n, k3n, k4n, k5n, k6n;
    double xt[2], vt[2], nt;
    double nu = 1, total = 0, ddt = g.dt;
    int guard = 0;
    while (total < dt && guard++ < 100000) {
        ddt = ddt < dt - total ? ddt : dt - total;
        geo_rhs(M, g, x[0], x[1], v[0], v[1], nu, k1x, k1v, k1n);
        for (int d = 0; d < 2; d++) {
            xt[d] = x[d] + td * ddt * ((1.0 / 4.0) * k1x[d]);
            vt[d] = v[d] + td * ddt * ((1.0 / 4.0) * k1v[d]);
        }
        nt = nu + td * ddt * ((1.0 / 4.0) * k1n);
        geo_rhs(M, g, xt[0], xt[1], vt[0], vt[1], nt, k2x, k2v, k2n);
        for (int d = 0; d < 2; d++) {
            xt[d] = x[d] + td * ddt * ((3.0 / 32.0) * k1x[d] + (9.0 / 32.0) * k2x[d]);
            vt[d] = v[d] + td * ddt * ((3.0 / 32.0) * k1v[d] + (9.0 / 32.0) * k2v[d]);
        }
        nt = nu + td * ddt * ((3.0 / 32.0) * k1n + (9.0 / 32.0) * k2n);
        geo_rhs(M, g, xt[0], xt[1], vt[0], vt[1], nt, k3x, k3v, k3n);
        for (int d = 0; d < 2; d++) {
            xt[d] = x[d] + td * ddt * ((1932.0 / 2197.0) * k1x[d] - (7200.0 / 2197.0) * k2x[d] + (7296.0 / 2197.0) * k3x[d]);
            vt[d] = v[d] + td * ddt * ((1932.0 / 2197.0) * k1v[d] - (7200.0 / 2197.0) * k2v[d] + (7296.0 / 2197.0) * k3v[d]);
        }
        nt = nu + td * ddt * ((1932.0 / 2197.0) * k1n - (7200.0 / 2197.0) * k2n + (7296.0 / 2197.0) * k3n);
        geo_rhs(M, g, xt[0], xt[1], vt[0], vt[1], nt, k4x, k4v, k4n);
        for (int d = 0; d < 2; d++) {
            xt[d] = x[d] + td * ddt * ((439.0 / 216.0) * k1x[d] - (8.0) * k2x[d] + (3680.0 / 513.0) * k3x[d] - (845.0 / 4104.0) * k4x[d]);
            vt[d] = v[d] + td * ddt * ((439.0 / 216.0) * k1v[d] - (8.0) * k2v[d] + (3680.0 / 513.0) * k3v[d] - (845.0 / 4104.0) * k4v[d]);
        }
        nt = nu + td * ddt * ((439.0 / 216.0) * k1n - (8.0) * k2n + (3680.0 / 513.0) * k3n - (845.0 / 4104.0) * k4n);
        geo_rhs(M, g, xt[0], xt[1], vt[0], vt[1], nt, k5x, k5v, k5n);
        for (int d = 0; d < 2; d++) {
            xt[d] = x[d] + td * ddt * (-(8.0 / 27.0) * k1x[d] + (2.0) * k2x[d] - (3544.0 / 2565.0) * k3x[d] + (1859.0 / 4104.0) * k4x[d] - (11.0 / 44.0) * k5x[d]);
            vt[d] = v[d] + td * ddt * (-(8.0 / 27.0) * k1v[d] + (2.0) * k2v[d] - (3544.0 / 2565.0) * k3v[d] + (1859.0 / 4104.0) * k4v[d] - (11.0 / 44.0) * k5v[d]);
        }
        nt = nu + td * ddt * (-(8.0 / 27.0) * k1n + (2.0) * k2n - (3544.0 / 2565.0) * k3n + (1859.0 / 4104.0) * k4n - (11.0 / 44.0) * k5n);
        geo_rhs(M, g, xt[0], xt[1], vt[0], vt[1], nt, k6x, k6v, k6n);
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
        if (geo_inside_bh(g, x[0], x[1]) || geo_outside(g, x[0], x[1])) break;
    }
    px = x[0];
    py = x[1];
    pvx = v[0];
    pvy = v[1];
    return nu;
}

/* One thread per interior cell: nF backtraces, then the Fourier fit
 * (Fourier::GetCoefficients, src/FourierHarmonics.cpp:33-47) into the SoA
 * table coef[(series*nF + k)*N + ij]. */
struct GeoStencil {
    double w[8], phi[8], cx[8], cy[8];
    double basis[8][8]; /* basis[d][k] at streaming angles, nF <= 8 */
};
__global__ __launch_bounds__(128) void k_geodesic_tables(const double *__restrict__ M, GeoGrid g, GeoStencil s,
                                                         int nF, double tol, double *__restrict__ coef)
{
    const int nix = g.nx - 2, niy = g.ny - 2;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nix * niy) return;
    int i = 1 + c % nix, j = 1 + c / nix;
    long ij = i + (long)j * g.nx;
    const long N = (long)g.nx * g.ny;
    double x0 = g.startx + i * g.dx, y0 = g.starty + j * g.dy;
    double alpha = M[ij * RT_NMETRIC + RT_M_ALPHA];
    double acc[5][8];
    for (int q = 0; q < 5; q++)
        for (int k = 0; k < nF; k++) acc[q][k] = 0;
    for (int d = 0; d < nF; d++) {
        double sh = 1, x = x0, y = y0;
        double ux = s.cx[d] * alpha, uy = s.cy[d] * alpha;
        double vx, vy;
        {
            /* Vec2ObservedByEulObs<IF,LF> at the cell centre (src/SpecialMath.cpp:159-164) */
            const int qs[5] = {RT_M_ALPHA, RT_M_TETRAD11, RT_M_TETRAD12, RT_M_TETRAD21, RT_M_TETRAD22};
            double v[5];
            if (geo_outside(g, x, y)) {
                v[0] = 1; v[1] = 1; v[2] = 0; v[3] = 0; v[4] = 1;
            } else {
                geo_interp<5>(M, g, (x - g.startx) / g.dx, (y - g.starty) / g.dy, qs, v);
            }
            double v1 = ux / v[0], v2 = uy / v[0];
            vx = v[1] * v1 + v[2] * v2;
            vy = v[3] * v1 + v[4] * v2;
        }
        if (!geo_inside_bh(g, x, y))
            sh *= geo_rk45(M, g, tol, g.dt, x, y, vx, vy);
        else
            vx = vy = 0.0;
        const int qi[4] = {RT_M_TETRADINV11, RT_M_TETRADINV12, RT_M_TETRADINV21, RT_M_TETRADINV22};
        double ti[4];
        if (geo_outside(g, x, y)) {
            ti[0] = 1; ti[1] = 0; ti[2] = 0; ti[3] = 1;
        } else {
            geo_interp<4>(M, g, (x - g.startx) / g.dx, (y - g.starty) / g.dy, qi, ti);
        }
        double data[5] = {1.0 / sh, x, y, ti[0] * vx + ti[1] * vy, ti[2] * vx + ti[3] * vy};
        for (int q = 0; q < 5; q++) {
            double cc = data[q] * s.w[d];
            for (int k = 0; k < nF; k++) acc[q][k] += cc * s.basis[d][k];
        }
    }
    for (int q = 0; q < 5; q++)
        for (int k = 0; k < nF; k++) coef[((long)q * nF + k) * N + ij] = acc[q][k];
}

/* Largest |departure - cell centre| in cells over active cells and
 * intensity directions, from the Fourier tables (decides whether the fused
 * kernel's one-cell halo covers every bilinear corner). */
__global__ void k_departure_bound(const double *__restrict__ coef, const double *__restrict__ basis,
                                  const uint8_t *__restrict__ mask, int nx, int ny, int D, int nF, double startx,
                                  double starty, double dx, double dy, double r2, unsigned long long *out,
                                  unsigned long long *excl)
{
    int ij = blockIdx.x * blockDim.x + threadIdx.x;
    const long N = (long)nx * ny;
    double m = 0;
    unsigned long long bits = 0;
    if (ij < N && (mask[ij] & MASK_ACTIVE)) {
        int i = ij % nx, j = ij / nx;
        double xc = startx + i * dx, yc = starty + j * dy;
        for (int d = 0; d < D; d++) {
            double xT = 0, yT = 0;
            for (int k = 0; k < nF; k++) {
                xT += coef[((long)1 * nF + k) * N + ij] * basis[d * nF + k];
                yT += coef[((long)2 * nF + k) * N + ij] * basis[d * nF + k];
            }
            if (r2 >= 0.0 && xT * xT + yT * yT <= r2) { /* excised departure: not gathered */
                if (d < 64) bits |= 1ull << d;
                continue;
            }
            double a = fabs(xT - xc) / dx, b = fabs(yT - yc) / dy;
            if (!(a == a) || !(b == b)) { m = 1e30; break; }
            m = fmax(m, fmax(a, b));
        }
    }
    if (ij < N) excl[ij] = bits;
    /* non-negative doubles order like their bit patterns */
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

} // namespace

/* ------------------------------------------------------------------------ */
/* launchers                                                                 */
/* ------------------------------------------------------------------------ */
static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

int launch_transpose_in(rt_ctx *c, const double *aos, void *soa)
{
    dim3 grid(nblk(c->N, TB), nblk(c->D, TB));
    if (c->fp32)
        k_transpose_in<float><<<grid, 256, 0, c->stream>>>(aos, (float *)soa, c->N, c->D);
    else
        k_transpose_in<double><<<grid, 256, 0, c->stream>>>(aos, (double *)soa, c->N, c->D);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}
int launch_transpose_out(rt_ctx *c, const void *soa, double *aos)
{
    dim3 grid(nblk(c->N, TB), nblk(c->D, TB));
    if (c->fp32)
        k_transpose_out<float><<<grid, 256, 0, c->stream>>>((const float *)soa, aos, c->N, c->D);
    else
        k_transpose_out<double><<<grid, 256, 0, c->stream>>>((const double *)soa, aos, c->N, c->D);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}
int launch_fill(rt_ctx *c, double *p, size_t n, double v)
{
    if (n == 0) return RT_OK;
    k_fill<<<nblk(n, 256), 256, 0, c->stream>>>(p, n, v);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}
int launch_copy_fixed_cells(rt_ctx *c, const void *src, void *dst)
{
    size_t n = (size_t)c->N * c->D;
    if (c->fp32)
        k_copy_fixed<float><<<nblk(n, 256), 256, 0, c->stream>>>((const float *)src, (float *)dst, c->d_mask, c->N, c->D);
    else
        k_copy_fixed<double><<<nblk(n, 256), 256, 0, c->stream>>>((const double *)src, (double *)dst, c->d_mask, c->N, c->D);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}
int launch_scale_state(rt_ctx *c, void *buf, bool inverse)
{
    size_t n = (size_t)c->N * c->D;
    const size_t off = (size_t)CF_IA3 * c->N;
    if (c->fp32)
        k_scale_state<float><<<nblk(n, 256), 256, 0, c->stream>>>((float *)buf, (const float *)c->d_cell + off, c->N, c->D,
                                                                 inverse ? 1 : 0);
    else
        k_scale_state<double><<<nblk(n, 256), 256, 0, c->stream>>>((double *)buf, (const double *)c->d_cell + off, c->N,
                                                                   c->D, inverse ? 1 : 0);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}
int launch_mask(rt_ctx *c)
{
    const rt_desc &d = c->desc;
    double r2 = d.excision_r2;
    double near_r2 = -1;
    if (r2 >= 0) {
        double R = sqrt(r2) + 2.5 * fmax(d.dx, d.dy);
        near_r2 = R * R;
    }
    k_mask<<<nblk(c->N, 256), 256, 0, c->stream>>>(c->d_mask, d.nx, d.ny, d.startx, d.starty, d.dx, d.dy, r2, near_r2);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}
int launch_prepare_cells(rt_ctx *c)
{
    c->z0_dirty = true;
    if (c->fp32)
        k_prepare<float><<<nblk(c->N, 256), 256, 0, c->stream>>>(c->d_f64, c->d_alpha, c->d_t11, c->d_t12, c->d_t21, c->d_t22,
                                                                 (float *)c->d_cell, c->d_mask, c->N, c->desc.dt);
    else
        k_prepare<double><<<nblk(c->N, 256), 256, 0, c->stream>>>(c->d_f64, c->d_alpha, c->d_t11, c->d_t12, c->d_t21, c->d_t22,
                                                                  (double *)c->d_cell, c->d_mask, c->N, c->desc.dt);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}
int launch_ref_stream(rt_ctx *c, const void *in, void *out)
{
    const rt_desc &d = c->desc;
    if (rt_is_adaptive(d)) return launch_stream_adaptive(c, in, out);
    size_t n = (size_t)(d.nx - 2) * (d.ny - 2) * c->D;
    if (d.streaming == RT_FLAT_FIXED) {
        if (c->fp32)
            k_ref_stream_flat<float><<<nblk(n, 256), 256, 0, c->stream>>>((const float *)in, (float *)out, c->d_cx, c->d_cy, d.nx, d.ny, c->D,
                                                                          d.startx, d.starty, d.dx, d.dy, d.dt);
        else
            k_ref_stream_flat<double><<<nblk(n, 256), 256, 0, c->stream>>>((const double *)in, (double *)out, c->d_cx, c->d_cy, d.nx, d.ny,
                                                                           c->D, d.startx, d.starty, d.dx, d.dy, d.dt);
    } else {
        if (c->fp32)
            k_ref_stream_curved<float><<<nblk(n, 256), 256, 0, c->stream>>>(
                (const float *)in, (float *)out, c->d_basis, c->d_coef, c->d_alpha, c->d_Nc, c->d_Wc, c->nGrid, c->nCoeff, d.nx,
                d.ny, c->D, d.startx, d.starty, d.dx, d.dy, d.excision_r2);
        else
            k_ref_stream_curved<double><<<nblk(n, 256), 256, 0, c->stream>>>(
                (const double *)in, (double *)out, c->d_basis, c->d_coef, c->d_alpha, c->d_Nc, c->d_Wc, c->nGrid, c->nCoeff, d.nx,
                d.ny, c->D, d.startx, d.starty, d.dx, d.dy, d.excision_r2);
    }
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}
/* Lab-frame moments of ComputeMomentsLF (src/Radiation.cpp:270-301): the
 * reference's loop order of TransformIFtoLF (src/SpecialMath.cpp:85-94) and
 * Norm2 (:20-27), no contraction (this file is built with
 * -ffp-contract=off). */
__global__ void k_moments_lf(const double *__restrict__ mom, const double *__restrict__ frame,
                             double *__restrict__ lf, long N, int nx, double startx, double starty, double dx,
                             double dy, double r2)
{
    const long ij = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (ij >= N) return;
    const double x = startx + (double)(ij % nx) * dx, y = starty + (double)(ij / nx) * dy;
    if (r2 >= 0.0 && x * x + y * y <= r2) {
        for (int k = 0; k < 7; k++) lf[(size_t)k * N + ij] = 0.0;
        return;
    }
    const double E = mom[ij], Fx = mom[N + ij], Fy = mom[2 * N + ij];
    const double Pxx = mom[3 * N + ij], Pxy = mom[4 * N + ij], Pyy = mom[5 * N + ij];
    const double T[3][3] = {{E, Fx, Fy}, {Fx, Pxx, Pxy}, {Fy, Pxy, Pyy}};
    double e[3][3];
    for (int a = 0; a < 3; a++)
        for (int A = 0; A < 3; A++) e[a][A] = frame[(size_t)(3 * a + A) * N + ij];
    double R[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            for (int A = 0; A < 3; A++)
                for (int B = 0; B < 3; B++) R[a][b] += T[A][B] * e[a][A] * e[b][B];
    const double g[3][3] = {{0, 0, 0},
                            {0, frame[9 * (size_t)N + ij], frame[10 * (size_t)N + ij]},
                            {0, frame[10 * (size_t)N + ij], frame[11 * (size_t)N + ij]}};
    const double v[3] = {0, R[0][1], R[0][2]};
    double n2 = 0;
    for (int i = 1; i < 3; i++)
        for (int j = 1; j < 3; j++) n2 += g[i][j] * v[i] * v[j];
    lf[ij] = R[0][0];
    lf[N + ij] = R[0][1];
    lf[2 * N + ij] = R[0][2];
    lf[3 * N + ij] = R[1][1];
    lf[4 * N + ij] = R[1][2];
    lf[5 * N + ij] = R[2][2];
    lf[6 * N + ij] = sqrt(fabs(n2));
}

/* keepSourceNodesActive (src/Radiation.cpp:790-791, LoadInitialData at the
 * start of every step): the source cells get their initial intensities back;
 * a state held as alpha^-3 I gets them scaled as k_scale_state does. */
template <typename T>
__global__ void k_restore_sources(T *__restrict__ I, const int32_t *__restrict__ idx, const double *__restrict__ src,
                                  int nsrc, int D, long N, const T *__restrict__ ia3)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= (long)nsrc * D) return;
    const int k = (int)(i % nsrc), d = (int)(i / nsrc);
    const long ij = idx[k];
    T x = (T)src[i];
    if (ia3) {
        const T s = ia3[ij];
        if (s > T(0) && s < T(INFINITY)) x = x * s;
    }
    I[(size_t)d * N + ij] = x;
}

int launch_restore_sources(rt_ctx *c, void *state, bool scaled)
{
    if (c->nsrc == 0) return RT_OK;
    const long n = (long)c->nsrc * c->D;
    const size_t off = (size_t)CF_IA3 * c->N;
    if (c->fp32)
        k_restore_sources<float><<<nblk(n, 256), 256, 0, c->stream>>>(
            (float *)state, c->d_src_idx, c->d_src_I, c->nsrc, c->D, c->N, scaled ? (const float *)c->d_cell + off : nullptr);
    else
        k_restore_sources<double><<<nblk(n, 256), 256, 0, c->stream>>>(
            (double *)state, c->d_src_idx, c->d_src_I, c->nsrc, c->D, c->N, scaled ? (const double *)c->d_cell + off : nullptr);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}

int launch_moments_lf(rt_ctx *c)
{
    const rt_desc &d = c->desc;
    k_moments_lf<<<nblk(c->N, 256), 256, 0, c->stream>>>(c->d_mom, c->d_frame, c->d_lf, c->N, d.nx, d.startx,
                                                          d.starty, d.dx, d.dy, d.excision_r2);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}

/* current rotation angles of an adaptive context (null: fixed streaming) */
static const double *rot_or_null(const rt_ctx *c) { return rt_is_adaptive(c->desc) ? c->d_rot[c->rcur] : nullptr; }

int launch_ref_moments(rt_ctx *c, const void *in)
{
    if (c->fp32)
        k_ref_moments<float><<<nblk(c->N, 256), 256, 0, c->stream>>>((const float *)in, c->d_mom, c->d_w, c->d_cx, c->d_cy, c->N, c->D,
                                                                     rot_or_null(c));
    else
        k_ref_moments<double><<<nblk(c->N, 256), 256, 0, c->stream>>>((const double *)in, c->d_mom, c->d_w, c->d_cx, c->d_cy, c->N, c->D,
                                                                      rot_or_null(c));
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}
int launch_ref_collide(rt_ctx *c, const void *in, void *out)
{
    if (c->fp32)
        k_ref_collide<float><<<nblk(c->N, 128), 128, 0, c->stream>>>(
            (const float *)in, (float *)out, c->d_mom, c->d_count, c->d_f64, c->d_alpha, c->d_t11, c->d_t12, c->d_t21, c->d_t22,
            c->d_mask, c->d_w, c->d_cx, c->d_cy, c->N, c->D, c->desc.dt, c->d_stats, rot_or_null(c));
    else
        k_ref_collide<double><<<nblk(c->N, 128), 128, 0, c->stream>>>(
            (const double *)in, (double *)out, c->d_mom, c->d_count, c->d_f64, c->d_alpha, c->d_t11, c->d_t12, c->d_t21,
            c->d_t22, c->d_mask, c->d_w, c->d_cx, c->d_cy, c->N, c->D, c->desc.dt, c->d_stats, rot_or_null(c));
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}
int launch_geodesic_tables(rt_ctx *c, const double *d_metric, double tol)
{
    const rt_desc &d = c->desc;
    GeoGrid g;
    g.nx = d.nx;
    g.ny = d.ny;
    g.dx = d.dx;
    g.dy = d.dy;
    g.startx = d.startx;
    g.starty = d.starty;
    g.endx = d.startx + (d.nx - 1) * d.dx;
    g.endy = d.starty + (d.ny - 1) * d.dy;
    g.dt = d.dt;
    g.r2 = d.excision_r2;
    GeoStencil s;
    for (int k = 0; k < c->nCoeff; k++) {
        s.w[k] = c->sw[k];
        s.phi[k] = c->sphi[k];
        s.cx[k] = c->scx[k];
        s.cy[k] = c->scy[k];
        for (int q = 0; q < c->nCoeff; q++) s.basis[k][q] = rtm::fourier_basis(q, c->sphi[k]);
    }
    size_t n = (size_t)(d.nx - 2) * (d.ny - 2);
    k_geodesic_tables<<<nblk(n, 128), 128, 0, c->stream>>>(d_metric, g, s, c->nCoeff, tol, c->d_coef);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}
int launch_departure_bound(rt_ctx *c, double *max_disp)
{
    const rt_desc &d = c->desc;
    unsigned long long *tmp = c->d_red;
    RT_HIP(c, hipMemsetAsync(tmp, 0, sizeof(unsigned long long), c->stream));
    k_departure_bound<<<nblk(c->N, 256), 256, 0, c->stream>>>(c->d_coef, c->d_basis, c->d_mask, d.nx, d.ny, c->D, c->nCoeff,
                                                              d.startx, d.starty, d.dx, d.dy, d.excision_r2, tmp,
                                                              c->d_excl);
    RT_HIP(c, hipGetLastError());
    unsigned long long bits = 0;
    RT_HIP(c, hipMemcpyAsync(&bits, tmp, sizeof bits, hipMemcpyDeviceToHost, c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    double m;
    memcpy(&m, &bits, sizeof m);
    *max_disp = m;
    return RT_OK;
}
