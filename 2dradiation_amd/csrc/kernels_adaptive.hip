/*
 * Adaptive streaming for gfx950: per-cell rotated stencils
 * (StreamingType FlatAdaptive and CurvedAdaptive).
 *
 * Reference: UpdateRotationMatrizes / AverageF (src/Radiation.cpp:345-378),
 * StreamFlatAdaptive (:415-453), StreamCurvedAdaptive (:513-575), IntensityAt
 * (:322-343), RotationMatrix (src/DataTypes.hh:620-635).  The rotation
 * update keeps the reference's FP64 operation order (this file is built
 * without FP contraction; its threshold is discontinuous).  The FP64 stream
 * follows the reference's order with a*b+c contracted into FMAs (every value
 * it rounds feeds a continuous function).  FP32-mixed contexts (FP32 state)
 * stream with FP32 arithmetic in k_stream_adaptive_f32.
 *
 * Layout: intensities SoA [D][N] as everywhere in the library; rotation
 * angles [N] in two buffers that swap like the reference's rotationAngle and
 * rotationAngleNew; per step one small kernel writes the sines and cosines
 * the stream needs, trig [4][N] = (sin, cos) of the new angle of every cell
 * and (sin, cos) of the inverse of its current angle, so the stream evaluates
 * no rotation polynomial of its own.
 *
 * Stream kernel: a workgroup is 64 consecutive interior cells (one per lane)
 * times 4 waves; wave w handles directions w, w+4, ...  A lane keeps its
 * cell's rotation and, for curved streaming, its 5 x nCoeff Fourier
 * coefficients in registers for all of its directions; stores of one
 * direction are coalesced over the 64 cells.
 */
#include "fused_common.h"
#include "ref_collide.h"

namespace {

constexpr double kTwoPi = 2.0 * rtm::kPi;

/* UpdateRotationMatrizes (interior cells, UPDATE) and the rotation sines and
 * cosines of every stored cell.  rotNew of halo cells is left as it is, as
 * in the reference. */
template <bool UPDATE>
__global__ __launch_bounds__(256) void k_rotation(const double *__restrict__ mom, const double *__restrict__ rot,
                                                  double *__restrict__ rotNew, double *__restrict__ trig,
                                                  float *__restrict__ trigf, int nx, int ny)
{
    const long N = (long)nx * ny;
    const long ij = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (ij >= N) return;
    const int i = (int)(ij % nx), j = (int)(ij / nx);
    double rn;
    if (UPDATE && i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1) {
        /* AverageF (:345-359): rows b = -1..1, columns a = -1..1 */
        const double *Fx = mom + N, *Fy = mom + 2 * N;
        double fx = 0.0, fy = 0.0;
        for (int b = -1; b <= 1; b++)
            for (int a = -1; a <= 1; a++) {
                const long k = (i + a) + (long)(j + b) * nx;
                fx += Fx[k];
                fy += Fy[k];
            }
        fx /= 9.0;
        fy /= 9.0;
        const double norm = sqrt(fx * fx + fy * fy);
        const double E = mom[ij];
        rn = (E > 1e-16 && norm / E > 0.01) ? rtm::my_phi(fx, fy) : rot[ij];
        rotNew[ij] = rn;
    } else {
        rn = rotNew[ij];
    }
    trig[ij] = rtm::my_sin(rn);
    trig[N + ij] = rtm::my_cos(rn);
    /* RotationMatrix::Inverse (src/DataTypes.hh:625-628) */
    const double inv = fmod(-rot[ij] + kTwoPi, kTwoPi);
    trig[2 * N + ij] = rtm::my_sin(inv);
    trig[3 * N + ij] = rtm::my_cos(inv);
    if (trigf)
        for (int q = 0; q < 4; q++) trigf[q * N + ij] = (float)trig[q * N + ij];
}

/* IntensityAt (src/Radiation.cpp:322-343) of cell k: the inertial-frame
 * velocity (vx, vy) rotated by the inverse of the cell's angle (sine si,
 * cosine ci), cubic angular interpolation on the fine grid, clamp at 0.
 * IX is the offset type: 32-bit when every offset into I fits. */
template <typename T, typename IX>
__device__ inline double intensity_at_rot(const T *__restrict__ I, IX N, IX k, double vx, double vy, double si,
                                          double ci, const int32_t *__restrict__ Nc, const double *__restrict__ Wc,
                                          int nGrid)
{
#pragma clang fp contract(fast)
    const double x = vx * ci - vy * si, y = vx * si + vy * ci;
    const double angle = rtm::my_phi(x, y);
    const double nG = (double)nGrid;
    double dd = nG * angle / kTwoPi;
    /* InterpolationGrid::d clamps to [1e-8, nGrid - 1e-8]; a NaN angle
     * (non-finite state) also goes to the lower end, keeping the table
     * reads in range */
    if (!(dd >= 1e-8)) dd = 1e-8;
    if (dd > nG - 1e-8) dd = nG - 1e-8;
    /* floor(dd) < nGrid, so the reference's "% nGrid" of d0 is the identity */
    const int d0 = (int)floor(dd);
    const int d1 = (d0 + 1 == nGrid) ? 0 : d0 + 1;
    const int4 n0 = *(const int4 *)(Nc + 4 * d0), n1 = *(const int4 *)(Nc + 4 * d1);
    const double2 w0a = *(const double2 *)(Wc + 4 * d0), w0b = *(const double2 *)(Wc + 4 * d0 + 2);
    const double2 w1a = *(const double2 *)(Wc + 4 * d1), w1b = *(const double2 *)(Wc + 4 * d1 + 2);
    double v0 = 0, v1 = 0;
    v0 += w0a.x * (double)I[(IX)n0.x * N + k];
    v1 += w1a.x * (double)I[(IX)n1.x * N + k];
    v0 += w0a.y * (double)I[(IX)n0.y * N + k];
    v1 += w1a.y * (double)I[(IX)n1.y * N + k];
    v0 += w0b.x * (double)I[(IX)n0.z * N + k];
    v1 += w1b.x * (double)I[(IX)n1.z * N + k];
    v0 += w0b.y * (double)I[(IX)n0.w * N + k];
    v1 += w1b.y * (double)I[(IX)n1.w * N + k];
    const double t = dd - d0;
    const double r = v0 * (1.0 - t) + v1 * t;
    return r > 0.0 ? r : 0.0;
}

__device__ inline long clampl(long v, long hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

struct AdaptiveArgs {
    const double *trig;   /* [4][N] */
    const double *cx, *cy;
    const int32_t *Nc;
    const double *Wc;
    int nGrid;
    const double *coef;   /* [5][nF][N] curved */
    const double *alpha;  /* [N] curved */
    int nF;
    int nx, ny, D;
    double startx, starty, dx, dy, dt, r2;
};

/* NF > 0: number of Fourier coefficients known at compile time (held in
 * registers); NF == 0: a.nF coefficients read from memory per direction.
 * IX: offset type into the state (uint32_t when N * D < 2^32). */
template <typename T, bool CURVED, int NF, typename IX>
__global__ __launch_bounds__(256) void k_stream_adaptive(const T *__restrict__ I, T *__restrict__ Inew, AdaptiveArgs a)
{
    /* the stream is continuous in every intermediate it rounds, so
     * contracting a*b+c into one FMA moves its result by a few ulp only */
#pragma clang fp contract(fast)
    const int nix = a.nx - 2;
    const long Ni = (long)nix * (a.ny - 2);
    const long c = blockIdx.x * 64L + threadIdx.x;
    if (c >= Ni) return;
    const int i = 1 + (int)(c % nix), j = 1 + (int)(c / nix);
    const IX N = (IX)a.nx * (IX)a.ny, ij = (IX)i + (IX)j * (IX)a.nx;
    const double rs = a.trig[ij], rc = a.trig[N + ij];
    const double *__restrict__ si = a.trig + 2 * N;
    const double *__restrict__ ci = a.trig + 3 * N;
    const double xc = a.startx + i * a.dx, yc = a.starty + j * a.dy;
    double al = 1.0;
    double cf[5][NF > 0 ? NF : 1];
    if (CURVED) {
        if (a.r2 >= 0.0 && xc * xc + yc * yc <= a.r2) {
            for (int d = threadIdx.y; d < a.D; d += 4) Inew[(IX)d * N + ij] = (T)0;
            return;
        }
        al = a.alpha[ij];
        if (NF > 0)
            for (int s = 0; s < 5; s++)
                for (int k = 0; k < NF; k++) cf[s][k] = a.coef[((size_t)s * NF + k) * (size_t)N + ij];
    }
    for (int d = threadIdx.y; d < a.D; d += 4) {
        /* RotationMatrix(rotationAngleNew[ij]) * C(d) */
        const double cxd = a.cx[d], cyd = a.cy[d];
        const double ux = cxd * rc - cyd * rs, uy = cxd * rs + cyd * rc;
        T *out = Inew + (IX)d * N + ij;
        if (!CURVED) {
            double xT = xc, yT = yc;
            xT -= ux * a.dt;
            yT -= uy * a.dt;
            const double iT = (xT - a.startx) / a.dx, jT = (yT - a.starty) / a.dy;
            /* departures beyond the halo are undefined in the reference; keep
             * the gather inside the grid (the floor is taken after clamping
             * to [-1, nx], so the conversion to int is always defined) */
            int i0 = (int)floor(fmin(fmax(iT, -1.0), (double)a.nx));
            int j0 = (int)floor(fmin(fmax(jT, -1.0), (double)a.ny));
            int i1 = i0 + 1, j1 = j0 + 1;
            i0 = (int)clampl(i0, a.nx - 1); i1 = (int)clampl(i1, a.nx - 1);
            j0 = (int)clampl(j0, a.ny - 1); j1 = (int)clampl(j1, a.ny - 1);
            const IX k00 = (IX)(i0 + j0 * a.nx), k01 = (IX)(i0 + j1 * a.nx), k10 = (IX)(i1 + j0 * a.nx),
                     k11 = (IX)(i1 + j1 * a.nx);
            const double f00 = intensity_at_rot(I, N, k00, ux, uy, si[k00], ci[k00], a.Nc, a.Wc, a.nGrid);
            const double f01 = intensity_at_rot(I, N, k01, ux, uy, si[k01], ci[k01], a.Nc, a.Wc, a.nGrid);
            const double f10 = intensity_at_rot(I, N, k10, ux, uy, si[k10], ci[k10], a.Nc, a.Wc, a.nGrid);
            const double f11 = intensity_at_rot(I, N, k11, ux, uy, si[k11], ci[k11], a.Nc, a.Wc, a.nGrid);
            *out = (T)rtm::bilinear<double>(iT - i0, jT - j0, f00, f01, f10, f11);
        } else {
            const double angle = rtm::my_phi(ux, uy);
            const int nF = NF > 0 ? NF : a.nF;
            /* Fourier::Basis (src/FourierHarmonics.cpp:5-15): harmonic m of
             * basis entries 2m-1 (sqrt2 cos m phi) and 2m (sqrt2 sin m phi);
             * one sincos, higher harmonics by the angle-addition recurrence
             * (a few ulp from libm's cos and sin of m phi) */
            double B[8];
            B[0] = 1.0;
            double s1, c1;
            sincos(angle, &s1, &c1);
            double sm = s1, cm = c1;
            for (int m = 1; 2 * m - 1 < nF; m++) {
                B[2 * m - 1] = rtm::kSqrt2 * cm;
                if (2 * m < 8) B[2 * m] = rtm::kSqrt2 * sm;
                const double cn = cm * c1 - sm * s1, sn = sm * c1 + cm * s1;
                cm = cn;
                sm = sn;
            }
            double q[5] = {0, 0, 0, 0, 0};
            for (int k = 0; k < nF; k++) {
                /* Fourier::GetValue (src/FourierHarmonics.cpp:56-62) */
                for (int s = 0; s < 5; s++) {
                    const double cv = NF > 0 ? cf[s][NF > 0 ? k : 0] : a.coef[((size_t)s * nF + k) * (size_t)N + ij];
                    q[s] += cv * B[k];
                }
            }
            const double sh = q[0], xT = q[1], yT = q[2], vx = q[3], vy = q[4];
            if (a.r2 >= 0.0 && xT * xT + yT * yT <= a.r2) {
                *out = (T)0;
                continue;
            }
            double iT = (xT - a.startx) / a.dx, jT = (yT - a.starty) / a.dy;
            if (!(iT >= 0.0)) iT = 0.0;
            if (!(jT >= 0.0)) jT = 0.0;
            if (iT > a.nx - 1.0) iT = a.nx - 1.0;
            if (jT > a.ny - 1.0) jT = a.ny - 1.0;
            const int i0 = (int)floor(iT), j0 = (int)floor(jT);
            const int i1 = (int)clampl(i0 + 1, a.nx - 1), j1 = (int)clampl(j0 + 1, a.ny - 1);
            const IX k00 = (IX)(i0 + j0 * a.nx), k01 = (IX)(i0 + j1 * a.nx), k10 = (IX)(i1 + j0 * a.nx),
                     k11 = (IX)(i1 + j1 * a.nx);
            const double r00 = al / a.alpha[k00], r01 = al / a.alpha[k01], r10 = al / a.alpha[k10],
                         r11 = al / a.alpha[k11];
            const double f00 = (r00 * (r00 * r00)) * intensity_at_rot(I, N, k00, vx, vy, si[k00], ci[k00], a.Nc, a.Wc, a.nGrid);
            const double f01 = (r01 * (r01 * r01)) * intensity_at_rot(I, N, k01, vx, vy, si[k01], ci[k01], a.Nc, a.Wc, a.nGrid);
            const double f10 = (r10 * (r10 * r10)) * intensity_at_rot(I, N, k10, vx, vy, si[k10], ci[k10], a.Nc, a.Wc, a.nGrid);
            const double f11 = (r11 * (r11 * r11)) * intensity_at_rot(I, N, k11, vx, vy, si[k11], ci[k11], a.Nc, a.Wc, a.nGrid);
            *out = (T)((sh * (sh * sh)) * rtm::bilinear<double>(iT - i0, jT - j0, f00, f01, f10, f11));
        }
    }
}

/* ---- FP32 arithmetic (FP32-mixed contexts) ----------------------------
 * The same stream with the per-(cell, direction) work in FP32: rotations
 * from FP32 copies of the sines and cosines, the Hastings angle as one
 * branch-free index (octant offsets exact), one sincosf per direction, the
 * Fourier series in the relative form (S - 1, departure offsets in cells,
 * velocity), the lapse ratio as alpha_c^3 * alpha_k^-3, bilinear and angular
 * interpolation as lerps.  Departure excision is tested in FP64 on absolute
 * coordinates.  Offsets are 32-bit (N * D < 2^32). */
struct AdaptiveArgsF {
    const float *trigf;   /* [4][N] */
    const double *cx, *cy;
    const int32_t *Nc;
    const float *Wcf;     /* [nGrid][4] */
    int nGrid;
    const float *coefT;   /* [25][N]: S - 1, X, Y offsets in cells, vx, vy (5 each) */
    const float *ia3;     /* [N] alpha^-3 */
    int nx, ny, D;
    float dtdx, dtdy;     /* flat departure offset per unit velocity, in cells */
    float fine_u;         /* nGrid / (2 pi) */
    double startx, starty, dx, dy, r2;
};

/* IntensityAt (src/Radiation.cpp:322-343) in FP32: inverse rotation, fine
 * angular index, two 4-tap cubic sums, lerp, clamp at 0. */
__device__ inline float intensity_at_rot_f(const float *__restrict__ I, uint32_t N, uint32_t k, float vx, float vy,
                                           float si, float ci, const int32_t *__restrict__ Nc,
                                           const float *__restrict__ Wcf, int nGrid, float u, float q1, float q2,
                                           float q3, float q4)
{
#pragma clang fp contract(fast)
    const float x = vx * ci - vy * si, y = vx * si + vy * ci;
    float dd = rtf::angle_index_bf<float>(x, y, u, q1, q2, q3, q4);
    if (!(dd >= 0.0f)) dd = 0.0f; /* zero vector (angle 0) and NaN */
    int d0 = (int)dd;
    if (d0 > nGrid - 1) d0 = nGrid - 1;
    const float t = fminf(dd - (float)d0, 1.0f);
    const int d1 = (d0 + 1 == nGrid) ? 0 : d0 + 1;
    const int4 n0 = *(const int4 *)(Nc + 4 * d0), n1 = *(const int4 *)(Nc + 4 * d1);
    const float4 w0 = *(const float4 *)(Wcf + 4 * d0), w1 = *(const float4 *)(Wcf + 4 * d1);
    float v0 = w0.x * I[(uint32_t)n0.x * N + k];
    float v1 = w1.x * I[(uint32_t)n1.x * N + k];
    v0 = fmaf(w0.y, I[(uint32_t)n0.y * N + k], v0);
    v1 = fmaf(w1.y, I[(uint32_t)n1.y * N + k], v1);
    v0 = fmaf(w0.z, I[(uint32_t)n0.z * N + k], v0);
    v1 = fmaf(w1.z, I[(uint32_t)n1.z * N + k], v1);
    v0 = fmaf(w0.w, I[(uint32_t)n0.w * N + k], v0);
    v1 = fmaf(w1.w, I[(uint32_t)n1.w * N + k], v1);
    return rtf::tmax0(fmaf(t, v1 - v0, v0));
}

/* departure offset d (cells) -> base cell and weight, clamped into the grid
 * as the FP64 kernels clamp the departure coordinate */
__device__ inline void split_offset(int i, float d, int n, int &i0, int &i1, float &w)
{
    const float f = floorf(d);
    i0 = i + (int)f;
    w = d - f;
    if (!(d == d) || i0 < 0) {
        i0 = 0;
        w = 0.0f;
    }
    if (i0 >= n - 1) {
        i0 = n - 1;
        w = 0.0f;
    }
    i1 = i0 + 1 < n ? i0 + 1 : n - 1;
}

template <bool CURVED>
__global__ __launch_bounds__(256) void k_stream_adaptive_f32(const float *__restrict__ I, float *__restrict__ Inew,
                                                             AdaptiveArgsF a)
{
#pragma clang fp contract(fast)
    const int nix = a.nx - 2;
    const uint32_t Ni = (uint32_t)nix * (uint32_t)(a.ny - 2);
    const uint32_t c = blockIdx.x * 64u + threadIdx.x;
    if (c >= Ni) return;
    const int i = 1 + (int)(c % nix), j = 1 + (int)(c / nix);
    const uint32_t N = (uint32_t)a.nx * (uint32_t)a.ny, ij = (uint32_t)i + (uint32_t)j * (uint32_t)a.nx;
    const float rs = a.trigf[ij], rc = a.trigf[N + ij];
    const float *__restrict__ si = a.trigf + 2 * N;
    const float *__restrict__ ci = a.trigf + 3 * N;
    const float u = a.fine_u, q1 = 0.25f * a.nGrid, q2 = 0.5f * a.nGrid, q3 = 0.75f * a.nGrid, q4 = (float)a.nGrid;
    const float hp = 1.5707963267948966f, pi = 3.14159265358979323846f, tp = 4.71238898038469f,
                twopi = 6.283185307179586f;
    const double xc = a.startx + i * a.dx, yc = a.starty + j * a.dy;
    float cf[25];
    float a3c = 1.0f;
    if (CURVED) {
        if (a.r2 >= 0.0 && xc * xc + yc * yc <= a.r2) {
            for (int d = threadIdx.y; d < a.D; d += 4) Inew[(uint32_t)d * N + ij] = 0.0f;
            return;
        }
#pragma unroll
        for (int q = 0; q < 25; q++) cf[q] = a.coefT[(size_t)q * N + ij];
        a3c = 1.0f / a.ia3[ij];
    }
    for (int d = threadIdx.y; d < a.D; d += 4) {
        const float cxd = (float)a.cx[d], cyd = (float)a.cy[d];
        const float ux = cxd * rc - cyd * rs, uy = cxd * rs + cyd * rc;
        float di, dj, vx = ux, vy = uy, sh = 1.0f;
        if (!CURVED) {
            di = -ux * a.dtdx;
            dj = -uy * a.dtdy;
        } else {
            const float ang = rtf::angle_index_bf<float>(ux, uy, 1.0f, hp, pi, tp, twopi);
            float s1, c1;
            sincosf(ang, &s1, &c1);
            const float r2 = 1.4142135623730951f;
            const float B[5] = {1.0f, r2 * c1, r2 * s1, r2 * (c1 * c1 - s1 * s1), r2 * (2.0f * s1 * c1)};
            float q[5];
#pragma unroll
            for (int s = 0; s < 5; s++) {
                float acc = cf[5 * s];
#pragma unroll
                for (int k = 1; k < 5; k++) acc = fmaf(cf[5 * s + k], B[k], acc);
                q[s] = acc;
            }
            sh = 1.0f + q[0];
            di = q[1];
            dj = q[2];
            vx = q[3];
            vy = q[4];
            const double xT = xc + a.dx * (double)di, yT = yc + a.dy * (double)dj;
            if (a.r2 >= 0.0 && xT * xT + yT * yT <= a.r2) {
                Inew[(uint32_t)d * N + ij] = 0.0f;
                continue;
            }
        }
        int i0, i1, j0, j1;
        float tx, ty;
        split_offset(i, di, a.nx, i0, i1, tx);
        split_offset(j, dj, a.ny, j0, j1, ty);
        const uint32_t k00 = (uint32_t)(i0 + j0 * a.nx), k01 = (uint32_t)(i0 + j1 * a.nx),
                       k10 = (uint32_t)(i1 + j0 * a.nx), k11 = (uint32_t)(i1 + j1 * a.nx);
        float f00 = intensity_at_rot_f(I, N, k00, vx, vy, si[k00], ci[k00], a.Nc, a.Wcf, a.nGrid, u, q1, q2, q3, q4);
        float f01 = intensity_at_rot_f(I, N, k01, vx, vy, si[k01], ci[k01], a.Nc, a.Wcf, a.nGrid, u, q1, q2, q3, q4);
        float f10 = intensity_at_rot_f(I, N, k10, vx, vy, si[k10], ci[k10], a.Nc, a.Wcf, a.nGrid, u, q1, q2, q3, q4);
        float f11 = intensity_at_rot_f(I, N, k11, vx, vy, si[k11], ci[k11], a.Nc, a.Wcf, a.nGrid, u, q1, q2, q3, q4);
        if (CURVED) {
            /* (alpha_c / alpha_k)^3 = alpha_c^3 * alpha_k^-3 */
            f00 *= a3c * a.ia3[k00];
            f01 *= a3c * a.ia3[k01];
            f10 *= a3c * a.ia3[k10];
            f11 *= a3c * a.ia3[k11];
        }
        const float b = rtf::bilinear_lerp(tx, ty, f00, f01, f10, f11);
        Inew[(uint32_t)d * N + ij] = CURVED ? (sh * (sh * sh)) * b : b;
    }
}

/* Rotation angles re-imposed on source cells (LoadInitialData,
 * src/Radiation.cpp:171, 177); NaN marks a cell whose angle is left alone. */
__global__ void k_restore_source_rot(double *__restrict__ rot, const int32_t *__restrict__ idx,
                                     const double *__restrict__ src, int nsrc)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nsrc) return;
    const double v = src[k];
    if (v == v) rot[idx[k]] = v;
}

inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

template <typename T, bool CURVED, typename IX>
void launch_stream_ix(rt_ctx *c, const void *in, void *out, const AdaptiveArgs &a, unsigned blocks)
{
    const dim3 grid(blocks), block(64, 4);
    if (CURVED && a.nF == 5)
        k_stream_adaptive<T, true, 5, IX><<<grid, block, 0, c->stream>>>((const T *)in, (T *)out, a);
    else if (CURVED)
        k_stream_adaptive<T, true, 0, IX><<<grid, block, 0, c->stream>>>((const T *)in, (T *)out, a);
    else
        k_stream_adaptive<T, false, 0, IX><<<grid, block, 0, c->stream>>>((const T *)in, (T *)out, a);
}
template <typename T, bool CURVED>
void launch_stream_t(rt_ctx *c, const void *in, void *out, const AdaptiveArgs &a, unsigned blocks)
{
    if ((size_t)c->N * c->D < ((size_t)1 << 32)) launch_stream_ix<T, CURVED, uint32_t>(c, in, out, a, blocks);
    else launch_stream_ix<T, CURVED, size_t>(c, in, out, a, blocks);
}

} // namespace

int launch_rotation(rt_ctx *c, bool update)
{
    const rt_desc &d = c->desc;
    const double *rot = c->d_rot[c->rcur];
    double *rotNew = c->d_rot[1 - c->rcur];
    if (update)
        k_rotation<true><<<nblk(c->N, 256), 256, 0, c->stream>>>(c->d_mom, rot, rotNew, c->d_rtrig, c->d_rtrigf, d.nx,
                                                                 d.ny);
    else
        k_rotation<false><<<nblk(c->N, 256), 256, 0, c->stream>>>(c->d_mom, rot, rotNew, c->d_rtrig, c->d_rtrigf, d.nx,
                                                                  d.ny);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}

int launch_stream_adaptive(rt_ctx *c, const void *in, void *out)
{
    const rt_desc &d = c->desc;
    const bool curved = d.streaming == RT_CURVED_ADAPTIVE;
    AdaptiveArgs a;
    a.trig = c->d_rtrig;
    a.cx = c->d_cx;
    a.cy = c->d_cy;
    a.Nc = c->d_Nc;
    a.Wc = c->d_Wc;
    a.nGrid = c->nGrid;
    a.coef = c->d_coef;
    a.alpha = c->d_alpha;
    a.nF = c->nCoeff;
    a.nx = d.nx;
    a.ny = d.ny;
    a.D = c->D;
    a.startx = d.startx;
    a.starty = d.starty;
    a.dx = d.dx;
    a.dy = d.dy;
    a.dt = d.dt;
    a.r2 = d.excision_r2;
    const size_t Ni = (size_t)(d.nx - 2) * (d.ny - 2);
    const unsigned blocks = nblk(Ni, 64);
    if (adaptive_f32_ready(c)) {
        AdaptiveArgsF f;
        f.trigf = c->d_rtrigf;
        f.cx = c->d_cx;
        f.cy = c->d_cy;
        f.Nc = c->d_Nc;
        f.Wcf = c->d_Wcf;
        f.nGrid = c->nGrid;
        f.coefT = (const float *)c->d_coefT;
        f.ia3 = (const float *)c->d_cell + (size_t)CF_IA3 * c->N;
        f.nx = d.nx;
        f.ny = d.ny;
        f.D = c->D;
        f.dtdx = (float)(d.dt / d.dx);
        f.dtdy = (float)(d.dt / d.dy);
        f.fine_u = (float)(c->nGrid / (2.0 * rtm::kPi));
        f.startx = d.startx;
        f.starty = d.starty;
        f.dx = d.dx;
        f.dy = d.dy;
        f.r2 = d.excision_r2;
        const dim3 grid(blocks), block(64, 4);
        if (curved) k_stream_adaptive_f32<true><<<grid, block, 0, c->stream>>>((const float *)in, (float *)out, f);
        else k_stream_adaptive_f32<false><<<grid, block, 0, c->stream>>>((const float *)in, (float *)out, f);
        RT_HIP(c, hipGetLastError());
        return RT_OK;
    }
    if (c->fp32) {
        if (curved) launch_stream_t<float, true>(c, in, out, a, blocks);
        else launch_stream_t<float, false>(c, in, out, a, blocks);
    } else {
        if (curved) launch_stream_t<double, true>(c, in, out, a, blocks);
        else launch_stream_t<double, false>(c, in, out, a, blocks);
    }
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}

/* ---- Collide of a rotated stencil in moment form (FP32-mixed) ----
 * For a cell at rest (u = 0 in the tetrad frame) every cycle of the
 * reference's lambda iteration (src/Radiation.cpp:661-722) is affine in the
 * moments: with q = 1 / (1 + alpha dt (kA + k0)), b = alpha dt q (eta + k0 E)
 * and c = 3 k1 alpha dt q,
 *   v_d = q I_d + b + c n_d.F,
 * so the moments of v are
 *   m'_X = q mI_X + b S0_X + c (Fx S1_X + Fy S2_X)
 * with mI the moments of the streamed intensities and S0_X = sum_d w_d X_d,
 * S1_X = sum_d w_d n_d,x X_d, S2_X = sum_d w_d n_d,y X_d over the cell's
 * rotated directions n_d = R(theta) C(d): power sums of the stencil up to
 * third order, rotated from the unrotated ones as tensors.  This is the
 * fused kernel's moment-space iteration with per-cell stencil sums: the same
 * stopping rule on the moments, then one pass over the directions,
 *   Inew_d = I_d + ((q - 1) I_d + b + C(d).G),  G = c R(theta)^T F,
 * in the storage precision.  One lane is one cell: it iterates (or, for a
 * moving cell, runs the reference cycle, ref_collide_cell, over all
 * directions), then writes the directions of a cell at rest with 8 loads in
 * flight at a time.  Every wavefront of a workgroup (256 consecutive cells)
 * runs its own cells' iterations, none waits for another's. */
struct StencilSums {
    double s[10];
};

template <typename T>
__global__ __launch_bounds__(256) void k_collide_adaptive(const T *__restrict__ I, T *__restrict__ Inew,
                                                          double *__restrict__ mom, int32_t *__restrict__ count,
                                                          const double *__restrict__ f, const double *__restrict__ alpha,
                                                          const double *__restrict__ t11, const double *__restrict__ t12,
                                                          const double *__restrict__ t21, const double *__restrict__ t22,
                                                          const uint8_t *__restrict__ mask, const double *__restrict__ w,
                                                          const double *__restrict__ cx, const double *__restrict__ cy,
                                                          int N, int D, double dt, unsigned long long *stats,
                                                          const double *__restrict__ rot, StencilSums ss)
{
    const int lane = threadIdx.x & 63;
    /* consecutive 256-cell chunks go to one XCD (workgroups are dealt
     * round-robin over the 8 XCDs): a direction row of 64 cells is not
     * 128-byte aligned, and its partial cache lines at both ends are then
     * written back by one L2 instead of two */
    const int nb = gridDim.x;
    const int chunk = (nb & 7) == 0 ? (int)(blockIdx.x & 7) * (nb >> 3) + (int)(blockIdx.x >> 3) : (int)blockIdx.x;
    const long ij = (long)chunk * 256 + threadIdx.x;
    {
        unsigned long long cyc = 0;
        if (ij < N) {
            if (!(mask[ij] & MASK_ACTIVE)) {
                count[ij] = 0;
            } else {
                const double ux = f[4L * N + ij], uy = f[5L * N + ij];
                const double u1 = t11[ij] * ux + t12[ij] * uy;
                const double u2 = t21[ij] * ux + t22[ij] * uy;
                int cycles = 0;
                if (u1 == 0.0 && u2 == 0.0) {
                    const double eta = f[ij], k0 = f[(long)N + ij], k1 = f[2L * N + ij], kA = f[3L * N + ij];
                    const double adt = alpha[ij] * dt;
                    const double q = 1.0 / (1.0 + adt * (kA + k0));
                    const double cc = 3.0 * k1 * adt * q;
                    const double rs = rtm::my_sin(rot[ij]), rc = rtm::my_cos(rot[ij]);
                    /* rotated power sums */
                    const double *u = ss.s;
                    const double c2 = rc * rc, s2 = rs * rs, cs = rc * rs;
                    const double a = rc * u[1] - rs * u[2], b1 = rs * u[1] + rc * u[2];
                    const double txx = c2 * u[3] - 2.0 * cs * u[4] + s2 * u[5];
                    const double txy = cs * u[3] + (c2 - s2) * u[4] - cs * u[5];
                    const double tyy = s2 * u[3] + 2.0 * cs * u[4] + c2 * u[5];
                    const double uxxx = rc * c2 * u[6] - 3.0 * c2 * rs * u[7] + 3.0 * rc * s2 * u[8] - rs * s2 * u[9];
                    const double uxxy = c2 * rs * u[6] + (rc * c2 - 2.0 * rc * s2) * u[7] +
                                        (rs * s2 - 2.0 * c2 * rs) * u[8] + rc * s2 * u[9];
                    const double uxyy = rc * s2 * u[6] + (2.0 * c2 * rs - rs * s2) * u[7] +
                                        (rc * c2 - 2.0 * rc * s2) * u[8] - c2 * rs * u[9];
                    const double uyyy = rs * s2 * u[6] + 3.0 * s2 * rc * u[7] + 3.0 * rs * c2 * u[8] + rc * c2 * u[9];
                    const double S0[6] = {u[0], a, b1, txx, txy, tyy};
                    const double S1[6] = {a, txx, txy, uxxx, uxxy, uxyy};
                    const double S2[6] = {b1, txy, tyy, uxxy, uxyy, uyyy};
                    double mI[6], prev[6];
#pragma unroll
                    for (int X = 0; X < 6; X++) prev[X] = mI[X] = mom[(long)X * N + ij];
                    /* stop when sqrt(sum ((prev - new) / new)^2) <= 1e-4; new == 0
                     * keeps the IEEE quotient (inf, or NaN for 0/0, which stops) */
                    double b = 0, fFx = 0, fFy = 0, sq = 1;
                    while (sq > 1e-8 && cycles < 100) {
                        cycles++;
                        b = adt * q * (eta + k0 * prev[0]);
                        fFx = prev[1];
                        fFy = prev[2];
                        double acc = 0;
#pragma unroll
                        for (int X = 0; X < 6; X++) {
                            const double nw = q * mI[X] + b * S0[X] + cc * (fFx * S1[X] + fFy * S2[X]);
                            const double r = (prev[X] - nw) / nw;
                            acc += r * r;
                            prev[X] = nw;
                        }
                        sq = acc;
                    }
#pragma unroll
                    for (int X = 0; X < 6; X++) mom[(long)X * N + ij] = prev[X];
                    const T qm1 = (T)(q - 1.0), bT = (T)b;
                    const T gx = (T)(cc * (rc * fFx + rs * fFy)), gy = (T)(cc * (rc * fFy - rs * fFx));
                    for (int d0 = 0; d0 < D; d0 += 8) {
                        T v[8];
#pragma unroll
                        for (int k = 0; k < 8; k++)
                            if (d0 + k < D) v[k] = I[(long)(d0 + k) * N + ij];
#pragma unroll
                        for (int k = 0; k < 8; k++) {
                            const int d = d0 + k;
                            if (d < D) {
                                const T t = (T)cx[d] * gx + (T)cy[d] * gy;
                                Inew[(long)d * N + ij] = v[k] + (qm1 * v[k] + (bT + t));
                            }
                        }
                    }
                } else {
                    cycles = ref_collide_cell<T>(ij, I, Inew, mom, f, alpha, t11, t12, t21, t22, w, cx, cy, N, D, dt, rot);
                }
                count[ij] = cycles;
                cyc = (unsigned long long)cycles;
            }
        }
        /* cycle statistics: one atomic pair per wavefront */
        unsigned long long sum = cyc, mx = cyc;
        for (int off = 32; off > 0; off >>= 1) {
            sum += __shfl_xor(sum, off);
            const unsigned long long o = __shfl_xor(mx, off);
            mx = mx > o ? mx : o;
        }
        if (lane == 0 && mx > 0) {
            unsigned long long *st = rt_stat_slot(stats, blockIdx.x * 4u + (threadIdx.x >> 6));
            atomicAdd(&st[0], sum);
            atomicMax(&st[1], mx);
        }
    }
}

int launch_collide_adaptive(rt_ctx *c, const void *in, void *out)
{
    StencilSums ss;
    for (int q = 0; q < 10; q++) ss.s[q] = c->stencil_sums[q];
    const double *rot = c->d_rot[c->rcur];
    const unsigned blocks = (nblk(c->N, 256) + 7) / 8 * 8;
    if (c->fp32)
        k_collide_adaptive<float><<<blocks, 256, 0, c->stream>>>(
            (const float *)in, (float *)out, c->d_mom, c->d_count, c->d_f64, c->d_alpha, c->d_t11, c->d_t12, c->d_t21,
            c->d_t22, c->d_mask, c->d_w, c->d_cx, c->d_cy, c->N, c->D, c->desc.dt, c->d_stats, rot, ss);
    else
        k_collide_adaptive<double><<<blocks, 256, 0, c->stream>>>(
            (const double *)in, (double *)out, c->d_mom, c->d_count, c->d_f64, c->d_alpha, c->d_t11, c->d_t12,
            c->d_t21, c->d_t22, c->d_mask, c->d_w, c->d_cx, c->d_cy, c->N, c->D, c->desc.dt, c->d_stats, rot, ss);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}

/* FP32-mixed contexts stream in FP32 when the FP32 tables exist, the
 * offsets fit 32 bits and (curved) the series have 5 coefficients. */
bool adaptive_f32_ready(const rt_ctx *c)
{
    if (!c->fp32 || !c->d_rtrigf || !c->d_Wcf || (size_t)c->N * c->D >= ((size_t)1 << 32)) return false;
    if (c->desc.streaming == RT_CURVED_ADAPTIVE) return c->nCoeff == 5 && c->d_coefT && c->coefT_valid;
    return true;
}

int launch_restore_source_rot(rt_ctx *c)
{
    if (c->nsrc == 0 || !c->d_src_rot || !c->src_rot_set) return RT_OK;
    k_restore_source_rot<<<nblk(c->nsrc, 256), 256, 0, c->stream>>>(c->d_rot[c->rcur], c->d_src_idx, c->d_src_rot,
                                                                      c->nsrc);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}
