/*
 * Adaptive streaming for gfx950: per-cell rotated stencils
 * (StreamingType FlatAdaptive and CurvedAdaptive).
 *
 * Reference: UpdateRotationMatrizes / AverageF (src/Radiation.cpp:345-378),
 * StreamFlatAdaptive (:415-453), StreamCurvedAdaptive (:513-575), IntensityAt
 * (:322-343), RotationMatrix (src/DataTypes.hh:620-635).  FP64 arithmetic in
 * the reference's operation order; this file is built without FP
 * contraction.  The state may be stored in FP32 (FP32-mixed contexts).
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
#include "rt_internal.h"

namespace {

constexpr double kTwoPi = 2.0 * rtm::kPi;

/* UpdateRotationMatrizes (interior cells, UPDATE) and the rotation sines and
 * cosines of every stored cell.  rotNew of halo cells is left as it is, as
 * in the reference. */
template <bool UPDATE>
__global__ __launch_bounds__(256) void k_rotation(const double *__restrict__ mom, const double *__restrict__ rot,
                                                  double *__restrict__ rotNew, double *__restrict__ trig, int nx,
                                                  int ny)
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

/* Waves per SIMD the stream kernel is compiled for; it sets the register
 * budget (256 VGPRs at 2, 168 at 3). */
#ifndef RT_ADAPTIVE_WAVES
#define RT_ADAPTIVE_WAVES 2
#endif

/* NF > 0: number of Fourier coefficients known at compile time (held in
 * registers); NF == 0: a.nF coefficients read from memory per direction.
 * IX: offset type into the state (uint32_t when N * D < 2^32). */
template <typename T, bool CURVED, int NF, typename IX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RT_ADAPTIVE_WAVES))) void k_stream_adaptive(const T *__restrict__ I, T *__restrict__ Inew, AdaptiveArgs a)
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
        k_rotation<true><<<nblk(c->N, 256), 256, 0, c->stream>>>(c->d_mom, rot, rotNew, c->d_rtrig, d.nx, d.ny);
    else
        k_rotation<false><<<nblk(c->N, 256), 256, 0, c->stream>>>(c->d_mom, rot, rotNew, c->d_rtrig, d.nx, d.ny);
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

int launch_restore_source_rot(rt_ctx *c)
{
    if (c->nsrc == 0 || !c->d_src_rot || !c->src_rot_set) return RT_OK;
    k_restore_source_rot<<<nblk(c->nsrc, 256), 256, 0, c->stream>>>(c->d_rot[c->rcur], c->d_src_idx, c->d_src_rot,
                                                                      c->nsrc);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}
