/*
 * Device helpers shared by the fused forward step and its adjoint: the
 * per-direction geometry of the curved gather (Fourier series at the stencil
 * angle, departure cell, cubic angular window) and the direct-to-LDS copy.
 */
#ifndef RT_FUSED_COMMON_H
#define RT_FUSED_COMMON_H
#include "rt_internal.h"

namespace rtf {

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

/* Bilinear interpolation as three lerps (6 operations; the reference's
 * order, rtm::bilinear, takes 8).  With f >= 0 and x, y in [0, 1] every lerp
 * stays >= 0 after rounding. */
template <typename T>
__device__ inline T bilinear_lerp(T x, T y, T f00, T f01, T f10, T f11)
{
    const T f0 = fma(x, f10 - f00, f00);
    const T f1 = fma(x, f11 - f01, f01);
    return fma(y, f1 - f0, f0);
}

template <typename T>
__device__ inline T tmax0(T v)
{
    return v > T(0) ? v : T(0);
}

/* Per-direction tables read with wave-uniform indices: a pointer into the
 * constant address space lets the compiler use scalar loads. */
typedef const __attribute__((address_space(4))) DirTables CDirTables;

/* One 16-byte direct global-to-LDS load per lane (gfx950): lane l writes
 * dst + 16 l; dst must be wave-uniform and 16-byte aligned, src 16-byte
 * aligned.  (A non-template function: inside a template the host pass
 * rejects the 16-byte form and drops the kernel's host stub.) */
__device__ inline void lds_dma16(const void *src, void *dst)
{
    __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 0);
}

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
 * offsets are exact multiples of nGrid/4, so only the polynomial term rounds.
 * A zero vector (angle 0) gives NaN here (0/0), which the caller's clamp to
 * [1e-8, nGrid - 1e-8] turns into 1e-8, the index of angle 0 after that
 * clamp (IEEE max/min return the non-NaN operand).  q1..q4: k * nGrid/4. */
template <typename T>
__device__ inline T angle_index_bf(T vx, T vy, T u, T q1, T q2, T q3, T q4)
{
    const bool swap = !(fabsf((float)vx) > fabsf((float)vy));
    const T num = swap ? vx : vy, den = swap ? vy : vx;
    const T p = rtm::hastings_atan(fast_div(num, den));
    /* octant offset: |vx| > |vy|: 0 (or 4 quarters when p < 0) for vx > 0,
     * 2 quarters for vx < 0; otherwise 1 quarter for vy > 0, 3 for vy < 0
     * (and the polynomial enters with a minus sign) */
    const T kx = vx > T(0) ? (p < T(0) ? q4 : T(0)) : q2;
    const T ky = vy > T(0) ? q1 : q3;
    const T up = u * p;
    return (swap ? ky : kx) + (swap ? -up : up);
}

/* Geometry of one (cell, direction) of the curved gather: the ray factor as
 * s - 1, the departure cell offset (fi, fj in {-1, 0}) and bilinear weights,
 * and the 5 combined angular taps starting at direction s0 (0 <= s0 < D).
 * Every index is clamped into range for any input, NaN included.
 * Reference: StreamCurvedFixed (src/Radiation.cpp:454-512) and IntensityAt
 * (:322-343).  wrow: the FP32 weight rows in LDS (FP32 only). */
template <typename T>
struct DirGeom {
    T s, fx, fy;
    T Cw[5];
    int s0;
    bool lo_i, lo_j;
};

/* Geometry from the five series values at the direction's angle: s - 1,
 * departure offset (di, dj) in cells, departure velocity (vx, vy). */
template <typename T, int D>
__device__ inline void geometry_from_series(T s, T di, T dj, T vx, T vy, CDirTables *dtab0, const float *wrow,
                                            T angu, int quarter, int nGrid, T ddmax, DirGeom<T> &G)
{
    G.s = s;
    const T fi = fmin(fmax(floor(di), T(-1)), T(0));
    const T fj = fmin(fmax(floor(dj), T(-1)), T(0));
    G.fx = di - fi;
    G.fy = dj - fj;
    G.lo_i = fi < T(0);
    G.lo_j = fj < T(0);
    const T q1 = T(quarter);
    T dd = angle_index_bf<T>(vx, vy, angu, q1, T(2) * q1, T(3) * q1, T(4) * q1);
    dd = fmin(fmax(dd, T(1e-8)), ddmax);
    int g0 = (int)dd;
    /* dd > 0: dd - floor(dd) is exact, and v_fract_f32 computes it in one step */
    const T t = (sizeof(T) == 4) ? (T)__builtin_amdgcn_fractf((float)dd) : dd - T(g0);
    g0 = min(max(g0, 0), nGrid); /* g0 == nGrid (dd rounded up) is angle 0: rows D..D+5 repeat -2..3 */
    const int q10 = (int)(__umul24((unsigned)g0, 6554u) >> 16); /* g0 / 10 for g0 < 16384 */
    /* r = g0 - 10 q10 as a full-rate 24-bit mad (the compiler otherwise
     * forms a quarter-rate 64-bit mad) */
    int r;
    asm("v_mad_i32_i24 %0, %1, -10, %2" : "=v"(r) : "v"(q10), "v"(g0));
    int woff;
    if constexpr (sizeof(T) == 4) {
        /* the rows hold A - D and D (A = (A - D) + D exactly) */
        const float4 *wr = (const float4 *)((const char *)wrow + __umul24((unsigned)r, 48u));
        const float4 w0 = wr[0], w1 = wr[1], w2 = wr[2];
        G.Cw[0] = (w0.x + w1.y) + t * w1.y;
        G.Cw[1] = (w0.y + w1.z) + t * w1.z;
        G.Cw[2] = (w0.z + w1.w) + t * w1.w;
        G.Cw[3] = (w0.w + w2.x) + t * w2.x;
        G.Cw[4] = (w1.x + w2.y) + t * w2.y;
        woff = __float_as_int(w2.z);
    } else {
        /* FP64 rows come through the vector L1 (faster than LDS for 8-byte
         * elements) */
#pragma unroll
        for (int m = 0; m < 5; m++) G.Cw[m] = dtab0->wA[r][m] + t * dtab0->wD[r][m];
        woff = dtab0->woff[r];
    }
    /* LDS row of the window start: woff is in [-2, -1] (checked when the
     * tables are built) and q10 in [0, D], so the window rows lie in
     * [0, D+5] without wrapping */
    G.s0 = q10 + woff + RT_RING_PAD;
}

/* Angular window of a departure velocity (vx, vy) in FP64, as
 * geometry_from_series<double> forms it: first LDS row s0 of the 5 combined
 * taps (window offset and ring pad included), weight-row residue r and
 * fraction t. */
__device__ inline void angle_window64(double vx, double vy, CDirTables *dt, int &s0, int &r, double &t)
{
    const double q1 = double(dt->quarter);
    double dd = angle_index_bf<double>(vx, vy, dt->ang_u, q1, 2.0 * q1, 3.0 * q1, 4.0 * q1);
    dd = fmin(fmax(dd, 1e-8), double(dt->nGrid) - 1e-8);
    int g0 = (int)dd;
    t = dd - double(g0);
    g0 = min(max(g0, 0), dt->nGrid);
    const int q10 = g0 / 10;
    r = g0 - 10 * q10;
    s0 = q10 + dt->woff[r] + RT_RING_PAD;
}

/* Angular record of one (cell, direction): s0 in bits 23..29, t in 19
 * fixed-point bits 4..22 (rounded, at most 1 - 2^-19), r in bits 0..3.  The
 * bits of t sit where an FP32 mantissa keeps them, so 1 + t is one bitwise
 * operation.  r = 10 (all weights 0) marks a departure point inside the
 * excision radius, which gathers nothing. */
__device__ inline uint32_t pack_angle_rec(int s0, int r, double t)
{
    const double tq = fmin(rint(t * 524288.0), 524287.0);
    return ((uint32_t)s0 << 23) | ((uint32_t)tq << 4) | (uint32_t)r;
}

/* FP32 geometry from the S, X and Y series values and the angular record.
 * Records are valid by construction (k_angle_rec, or zero before the first
 * tables): s0 <= D + 1 keeps the taps in the ring and r <= 10 in the weight
 * table.  The departure offsets (di, dj) lie in (-1, 1) (the fused path
 * requires a departure bound below 0.999 cells), so the corner is the
 * cell's own or the one before it, and v_fract_f32 gives the bilinear weight
 * di - floor(di) in one step.  C = (A - D) + (1 + t) D is the FMA of
 * A + t D: the same exact value, rounded once. */
template <int D>
__device__ inline void geometry_from_record(float s, float di, float dj, uint32_t rec, const float *wrow,
                                            DirGeom<float> &G)
{
    G.s = s;
    G.fx = __builtin_amdgcn_fractf(di);
    G.fy = __builtin_amdgcn_fractf(dj);
    G.lo_i = di < 0.0f;
    G.lo_j = dj < 0.0f;
    const unsigned r = rec & 15u;
    const float t1 = __uint_as_float((rec & 0x007FFFF0u) | 0x3F800000u);
    const float4 *wr = (const float4 *)((const char *)wrow + __umul24(r, 48u));
    const float4 w0 = wr[0], w1 = wr[1];
    const float2 w2 = *(const float2 *)(wr + 2);
    G.Cw[0] = fmaf(t1, w1.y, w0.x);
    G.Cw[1] = fmaf(t1, w1.z, w0.y);
    G.Cw[2] = fmaf(t1, w1.w, w0.z);
    G.Cw[3] = fmaf(t1, w2.x, w0.w);
    G.Cw[4] = fmaf(t1, w2.y, w1.x);
    G.s0 = (int)(rec >> 23);
}

template <typename T, int D>
__device__ inline void curved_geometry(const T (&cS)[5], const T (&cX)[5], const T (&cY)[5], const T (&cU)[5],
                                       const T (&cV)[5], CDirTables *dtab, CDirTables *dtab0, const float *wrow,
                                       int d, T angu, int quarter, int nGrid, T ddmax, DirGeom<T> &G)
{
    /* Fourier basis 0 is 1, and fma(c0, 1, 0) = c0 exactly, so the series
     * start from c0 */
    T B[5];
#pragma unroll
    for (int q = 1; q < 5; q++) B[q] = (sizeof(T) == 4) ? (T)dtab->basisf[d][q] : (T)dtab->basis[d][q];
    T s = cS[0], di = cX[0], dj = cY[0], vx = cU[0], vy = cV[0];
#pragma unroll
    for (int q = 1; q < 5; q++) {
        s += cS[q] * B[q];
        di += cX[q] * B[q];
        dj += cY[q] * B[q];
        vx += cU[q] * B[q];
        vy += cV[q] * B[q];
    }
    geometry_from_series<T, D>(s, di, dj, vx, vy, dtab0, wrow, angu, quarter, nGrid, ddmax, G);
}

/* The five FP32 Fourier series of this lane's cell at 4 directions with
 * v_mfma_f32_4x4x1f32 (16 blocks of 4x4, K = 1): lane 4b+j supplies the
 * coefficient of its own cell as B[0][j] of block b and basis value of
 * direction i = lane & 3 as A[i][0]; register i of lane 4b+j receives
 * D[i][j] = C[i][j] + A[i] B[j].  Starting from C = 0 with A = 1 for
 * coefficient 0, the chain gives every series the bits of the FMA chain
 * c0, fma(c_q, B_q, .) for q = 1..4.  bq: basis values 1..4 of direction
 * (lane & 3) of the group. */
typedef float f32x4 __attribute__((ext_vector_type(4)));

/* The A operands of series4_mfma for the NG direction groups of wavefront
 * wv (DW directions): basis values 1..4 of direction 4g + (lane & 3), 0 past
 * the wavefront's last direction. */
template <int DW, int NG>
__device__ inline void basis_groups(CDirTables *dtab, int wv, int lane, float (&bas4)[NG][4])
{
#pragma unroll
    for (int g = 0; g < NG; g++) {
        const int i = 4 * g + (lane & 3);
        const int d = wv * DW + (i < DW ? i : 0);
#pragma unroll
        for (int q = 0; q < 4; q++) bas4[g][q] = i < DW ? dtab->basisf[d][q + 1] : 0.f;
    }
}

__device__ inline f32x4 series4_mfma(const float (&c)[5], const float (&bq)[4])
{
    f32x4 r = {0.f, 0.f, 0.f, 0.f};
    r = __builtin_amdgcn_mfma_f32_4x4x1f32(1.0f, c[0], r, 0, 0, 0);
#pragma unroll
    for (int q = 1; q < 5; q++) r = __builtin_amdgcn_mfma_f32_4x4x1f32(bq[q - 1], c[q], r, 0, 0, 0);
    return r;
}

} // namespace rtf
#endif
