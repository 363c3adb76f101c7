/*
 * Fused stream + moments + collide step for gfx950.
 *
 * One workgroup = 4 wavefronts walks a 64-cell-wide column strip of the grid
 * down a chunk of rows.  Three input rows (all directions) live in an LDS ring
 * [slot][direction][66] (64 cells plus one halo cell each side, direction
 * rows -2..D+3 taken mod D); the row two ahead is loaded straight from HBM
 * into the slot that just died (global_load_lds, no register staging) while
 * the current row is computed.  Lane = cell column, wavefront w = direction
 * block [w*D/4, (w+1)*D/4).
 *
 * Stream (reference src/Radiation.cpp:380-414 flat, :454-512 curved):
 *   flat:   bilinear of the same direction at the departure point x - c dt.
 *   curved: the five geodesic Fourier series (S, X, Y, Cx, Cy) are evaluated
 *           at phi_d (FP32: four directions at a time on the matrix cores,
 *           FP64: FMA chains with the basis in scalar registers); the
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
#include <string>
#include <vector>
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
    const T *coefT;             /* [25][N] in kernel form (S - 1, relative X, Y) */
    const unsigned long long *excl; /* [N] excised-departure bits */
    const uint32_t *arec;       /* FP32 curved: [D][N] angular records (k_angle_rec) */
    double *mom;                /* [6][N] */
    unsigned long long *stats;  /* [2] */
    const DirTables *dt;
    int nx, ny;
    long N;
    int strips, chunks, rows_per_block, nblocks;
    double startx, starty, dx, dy, dtime, r2;
    int want_moments;
    int wide; /* FP32 16-byte staging (Smem) allowed for this launch */
    int cls;  /* k_fused_queue: index of (nDir, streaming) in queue_class */
    /* adjoint recording (k_fused<..., REC = true>): moments E, Fx, Fy of the
     * streamed intensities [3][N] (of a [6][N] buffer), lambda cycle counts
     * [N], corner clamp bits [4 or 8][N] (32-bit words up to 32 directions,
     * else 64-bit) */
    double *rec_m;
    uint8_t *rec_k;
    unsigned long long *rec_mask;
};

/* The reference's first lambda cycle from all-zero streamed intensities
 * (src/Radiation.cpp:661-721): with prev = 0 and u = 0 every direction gets
 * the same v = (0 + a dt eta) / (1 + a dt (kA + k0)), the new moments are
 * the d-ordered sums of c = w v times {1, n1, n2, n1 n1, n1 n2, n2 n2}, and
 * (0 - new) / new is NaN, which stops the loop, wherever a sum is exactly 0.
 * At an isotropic start the y flux is such a sum of rounding noise, exactly
 * 0 in some cells and not in others; the moment recursion cannot see it.
 * Evaluated in the reference's operation order, without contraction. */
template <int D>
__device__ __noinline__ bool zero_start_stops(const double *__restrict__ f64, const double *__restrict__ alpha64,
                                              long ij, long N, double dt, CDirTables *dtab)
{
#pragma clang fp contract(off)
    const double eta = f64[ij], k0 = f64[N + ij], kA = f64[3 * N + ij], al = alpha64[ij];
    const double adt = al * dt, den = 1.0 + al * dt * (kA + k0);
    const double v = (0.0 + adt * eta) / den;
    double e = 0, fx = 0, fy = 0, pxx = 0, pxy = 0, pyy = 0;
#pragma unroll 4
    for (int d = 0; d < D; d++) {
        const double n1 = dtab->cx[d], n2 = dtab->cy[d];
        const double c = dtab->w6[d][0] * v;
        e += c;
        fx += c * n1;
        fy += c * n2;
        pxx += c * n1 * n1;
        pxy += c * n1 * n2;
        pyy += c * n2 * n2;
    }
    return e == 0.0 || fx == 0.0 || fy == 0.0 || pxx == 0.0 || pxy == 0.0 || pyy == 0.0;
}

/* zero_start_stops depends on a cell's fields and lapse only: evaluated once
 * per field or table upload, it is kept as CF_ACT = 2 (active, first cycle
 * from all-zero intensities stops), so the step reads a flag instead of
 * running 6 D FP64 operations per such cell and row. */
template <typename T, int D>
__global__ void k_mark_zero_start(const double *__restrict__ f64, const double *__restrict__ alpha64,
                                  const DirTables *dt, T *__restrict__ cell, long N, double dtime)
{
    const long ij = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (ij >= N) return;
    T *act = cell + (size_t)CF_ACT * N + ij;
    if (*act != T(0) && zero_start_stops<D>(f64, alpha64, ij, N, dtime, (CDirTables *)dt)) *act = T(2);
}

/* LDS of one workgroup.  Intensity rows are SoA [direction][columns]
 * holding directions -2..D+3 (mod D), so the 5 angular taps of a corner are
 * 5 rows of one column at constant offsets, never wrapped; lanes
 * (= columns) hit consecutive banks.  A row covers the strip's 64 cells and
 * one halo cell each side (66 columns).  Rows and per-cell inputs arrive by
 * direct global-to-LDS loads, without passing through registers.
 *
 * FP32 rows and per-cell rows are 68 wide, a whole number of 16-byte
 * chunks: on grids of even width (and N a multiple of 4) they are staged
 * with 16-byte loads from the aligned column at or left of the first one
 * needed, 0 or 2 columns (rows) and 1 or 3 (cells) ahead of it; the readers
 * add that shift.  Otherwise 4-byte loads fill them from column 0. */
template <typename T, int D, bool CURVED, int NW>
struct __attribute__((aligned(16))) Smem {
    static constexpr int RL = 66;                        /* columns used */
    static constexpr int RW = sizeof(T) == 4 ? 68 : 66;  /* row stride */
    static constexpr int CW = sizeof(T) == 4 ? 68 : 64;  /* per-cell row stride */
    /* the deferred collide of the tile body (its CP) in the forms that can
     * hold a row of streamed intensities in registers */
    static constexpr bool HAS_CP = NW == 4 && (sizeof(T) == 8 || D >= 48);
    /* series staged per cell: FP32 takes its angular windows from the
     * records (k_angle_rec) and evaluates only S, X and Y */
    static constexpr int NSER = sizeof(T) == 4 ? 15 : 25;
    T I[3][RT_RING_ROWS(D)][RW];
    T C[CURVED ? NSER : 1][CW];                  /* coefficient series (d_coefT form) */
    unsigned long long EX[(CURVED && sizeof(T) == 8) ? 64 : 1]; /* FP64: excised-departure bits */
    T F[RT_CELL_FIELDS][CW];                     /* per-cell fields */
    double X[NW][6][64];                         /* per-wavefront moment partials */
    T CP[HAS_CP ? 4 : 1][64];                    /* collide of a row: q - 1, b, c Fx, c Fy */
    float wrow[11][12] __attribute__((aligned(16)));
};
/* two workgroups per CU for the FP32 64-direction curved form, three for 32
 * directions (512-byte granules) */
static_assert((sizeof(Smem<float, 64, true, 4>) + 511) / 512 * 2 <= 320, "FP32 D=64 curved LDS exceeds half a CU");
static_assert((sizeof(Smem<float, 32, true, 4>) + 511) / 512 * 3 <= 320, "FP32 D=32 curved LDS exceeds a third of a CU");

/* Eight-wavefront workgroups (D/8 directions each) for FP32 with 32 or 48
 * directions and FP64 with 32, 48 or 64.  The launch takes them when the
 * 4-wavefront workgroups would fill at most half of one round of resident
 * workgroups even with one-row chunks, i.e. for very small grids: twice the
 * wavefronts per workgroup then hide the gather latency.  On grids that
 * fill the device the 4-wavefront form is faster, because every wavefront
 * repeats the per-row collide, except where the LDS of the 4-wavefront form
 * allows only one workgroup per CU (FP64 at 32 to 64 curved directions):
 * there the 8-wavefront form doubles the wavefronts per SIMD at the same
 * LDS and is faster on any grid (FP64 1024^2 x 64: 716 against 821 us per
 * step).  Both forms give the same bits (wave_sum).  The recording variant
 * notes the form it ran (rt_ctx::rec_waves): the adjoint reads its
 * per-wavefront clamp bits accordingly. */
template <typename T, int D>
constexpr bool has_wide()
{
    return (sizeof(T) == 4 && (D == 32 || D == 48)) || (sizeof(T) == 8 && (D == 32 || D == 48 || D == 64));
}

/* Sum of the NW wavefront partials of moment q of this lane's cell.  Four
 * wavefronts: ((X0 + X1) + X2) + X3.  Eight (each half of a 4-wavefront
 * partial): the pairs first, then the same order over the pairs. */
template <int NW>
__device__ inline double wave_sum(const double (*X)[6][64], int q, int lane)
{
    if constexpr (NW == 8) {
        double s = X[0][q][lane] + X[1][q][lane];
#pragma unroll
        for (int w = 2; w < NW; w += 2) s += X[w][q][lane] + X[w + 1][q][lane];
        return s;
    } else {
        double s = X[0][q][lane];
#pragma unroll
        for (int w = 1; w < NW; w++) s += X[w][q][lane];
        return s;
    }
}

/* Reference lambda cycle for rows that contain a moving cell
 * (src/Radiation.cpp:661-722), with a workgroup moment exchange per cycle.
 * Every wavefront evaluates the same per-cell decisions, so the trip count is
 * uniform across the workgroup. */
template <typename T, int DW, int NW>
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
        /* two chains per 4-wavefront partial, as in the stream */
        constexpr int NCH = (has_wide<T, DW * NW>() && NW == 4) ? 2 : 1;
        double part[NCH][6] = {};
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
                const int ch = (NCH == 2 && k >= DW / 2) ? 1 : 0;
#pragma unroll
                for (int q = 0; q < 6; q++) part[ch][q] += dtab->w6[d][q] * v;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 6; q++) X[wv][q][lane] = NCH == 2 ? part[0][q] + part[NCH - 1][q] : part[0][q];
        __syncthreads();
        if (iter) {
            double ss = 0;
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const double nw = wave_sum<NW>(X, q, lane);
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

/* Wavefronts per SIMD the register allocation must allow: as many
 * workgroups per CU as their LDS allows in 160 KB (at most 3), each with
 * NW/4 wavefronts on every SIMD, at most 2 (FP32) or 4 (FP64) in all. */
template <typename T, int D, bool CURVED, int NW>
constexpr int fused_waves_per_simd()
{
    /* LDS is allocated to workgroups in 512-byte granules */
    constexpr size_t lds = (sizeof(Smem<T, D, CURVED, NW>) + 511) / 512 * 512;
    constexpr int by_lds = (int)((160 * 1024) / lds);
    constexpr int wgs = by_lds < 1 ? 1 : (by_lds > 3 ? 3 : by_lds);
    /* FP32 forms stop at 2: with 3 per SIMD their 168 registers spill
     * (256^2 x 32 Schwarzschild: 16.5 us per step at 2 against 19.8 us at
     * 3) */
    constexpr int cap = sizeof(T) == 4 ? 2 : 4;
    return wgs * NW / 4 > cap ? cap : wgs * NW / 4;
}

/* XCD-aware tile order: consecutive workgroups are dealt round-robin over
 * the 8 XCDs, so each XCD gets a contiguous run of tiles.  Tiles number the
 * strips fastest, so a run is a band of rows across the grid: horizontally
 * adjacent strips, running side by side, share the cache lines at their
 * common edge in that XCD's L2 (1024^2 x 64 FP32: 0.304 -> 0.281 ms per
 * step, fetched bytes per step 949 -> 712 MB, against runs of whole
 * strips). */
__device__ __forceinline__ int xcd_tile(int bid, int nblocks)
{
    return (nblocks & 7) == 0 ? (bid & 7) * (nblocks >> 3) + (bid >> 3) : bid;
}

/* One tile of the fused step (fused_tile_body.inc) in the workgroup's LDS
 * sm; PK: KP<T> in the constant address space, so that its fields are reloaded
 * where used rather than held in scalar registers across the tile. */
template <typename T, int D, bool CURVED, bool REC, int NW, bool SONLY, typename PK>
__device__ __forceinline__ void fused_tile(PK &p, int tile, const T *Iin, T *Iout, Smem<T, D, CURVED, NW> &sm)
{
    constexpr bool LEAN = true;
#include "fused_tile_body.inc"
}

/* The single-problem kernel includes the body itself: its parameters stay
 * in the kernel-argument segment, which the compiler reloads freely (through
 * a reference they would take scalar registers for the whole body). */
template <typename T, int D, bool CURVED, bool REC, int NW, bool SONLY>
__global__ __launch_bounds__(64 * NW, (fused_waves_per_simd<T, D, CURVED, NW>())) void k_fused(KP<T> p)
{
    __shared__ Smem<T, D, CURVED, NW> sm;
    const int tile = xcd_tile(blockIdx.x, p.nblocks);
    const T *Iin = p.Iin;
    T *Iout = p.Iout;
    constexpr bool LEAN = false;
#include "fused_tile_body.inc"
}

/* Many independent problems of one (precision, nDir, streaming) class in one
 * launch: workgroups [start[i], start[i+1]) run the tiles of problem i.  The
 * problem table holds each problem's step with `parity` 0; odd parity swaps
 * the two state buffers. */
template <typename T, int D, bool CURVED>
__global__ __launch_bounds__(64 * RT_FUSED_WAVES, (fused_waves_per_simd<T, D, CURVED, RT_FUSED_WAVES>())) void k_fused_batch(
    const KP<T> *__restrict__ probs, const int *__restrict__ start, int nprob, int parity)
{
    __shared__ Smem<T, D, CURVED, RT_FUSED_WAVES> sm;
    int i = 0;
    while (i + 1 < nprob && start[i + 1] <= (int)blockIdx.x) i++;
    i = __builtin_amdgcn_readfirstlane(i);
    const KP<T> &p = probs[i];
    const int tile = xcd_tile((int)blockIdx.x - start[i], p.nblocks);
    const T *Iin = parity ? (const T *)p.Iout : p.Iin;
    T *Iout = parity ? const_cast<T *>(p.Iin) : p.Iout;
    constexpr bool REC = false, SONLY = false;
    constexpr int NW = RT_FUSED_WAVES;
    constexpr bool LEAN = false;
#include "fused_tile_body.inc"
}

/* ---- one persistent launch for a batch of FP32 problems ------------------
 * Every FP32 problem of a batch, whatever its nDir and streaming, in one
 * launch of resident workgroups that take (problem, tile) items from a queue
 * (largest tiles first) until it is empty.  Class c = 2 * (index of nDir in
 * 16, 20, 24, 32, 48, 64) + curved selects the tile body; the LDS is the
 * union of the classes' layouts. */
union QueueSmem {
    Smem<float, 16, false, RT_FUSED_WAVES> f16;
    Smem<float, 16, true, RT_FUSED_WAVES> c16;
    Smem<float, 20, false, RT_FUSED_WAVES> f20;
    Smem<float, 20, true, RT_FUSED_WAVES> c20;
    Smem<float, 24, false, RT_FUSED_WAVES> f24;
    Smem<float, 24, true, RT_FUSED_WAVES> c24;
    Smem<float, 32, false, RT_FUSED_WAVES> f32;
    Smem<float, 32, true, RT_FUSED_WAVES> c32;
    Smem<float, 48, false, RT_FUSED_WAVES> f48;
    Smem<float, 48, true, RT_FUSED_WAVES> c48;
    Smem<float, 64, false, RT_FUSED_WAVES> f64;
    Smem<float, 64, true, RT_FUSED_WAVES> c64;
};

int queue_class(int D, bool curved)
{
    const int k = D == 16 ? 0 : D == 20 ? 1 : D == 24 ? 2 : D == 32 ? 3 : D == 48 ? 4 : D == 64 ? 5 : -1;
    return k < 0 ? -1 : 2 * k + (curved ? 1 : 0);
}

/* Item w = (problem, tile) of the table; *counter starts at 0 (k_queue_reset). */
__global__ __launch_bounds__(64 * RT_FUSED_WAVES, 2) void k_fused_queue(const KP<float> *__restrict__ probs,
                                                                        const int2 *__restrict__ items, int nitems,
                                                                        int *counter, int parity)
{
    __shared__ QueueSmem sm;
    __shared__ int s_item;
    for (;;) {
        __syncthreads(); /* the previous tile is done with the LDS and with s_item */
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        if (it >= nitems) return;
        const int2 w = items[it];
        typedef const __attribute__((address_space(4))) KP<float> CKP;
        CKP &p = ((CKP *)probs)[w.x];
        const float *Iin = parity ? (const float *)p.Iout : p.Iin;
        float *Iout = parity ? const_cast<float *>(p.Iin) : p.Iout;
        constexpr int NW = RT_FUSED_WAVES;
        switch (p.cls) {
        case 0: fused_tile<float, 16, false, false, NW, false>(p, w.y, Iin, Iout, sm.f16); break;
        case 1: fused_tile<float, 16, true, false, NW, false>(p, w.y, Iin, Iout, sm.c16); break;
        case 2: fused_tile<float, 20, false, false, NW, false>(p, w.y, Iin, Iout, sm.f20); break;
        case 3: fused_tile<float, 20, true, false, NW, false>(p, w.y, Iin, Iout, sm.c20); break;
        case 4: fused_tile<float, 24, false, false, NW, false>(p, w.y, Iin, Iout, sm.f24); break;
        case 5: fused_tile<float, 24, true, false, NW, false>(p, w.y, Iin, Iout, sm.c24); break;
        case 6: fused_tile<float, 32, false, false, NW, false>(p, w.y, Iin, Iout, sm.f32); break;
        case 7: fused_tile<float, 32, true, false, NW, false>(p, w.y, Iin, Iout, sm.c32); break;
        case 8: fused_tile<float, 48, false, false, NW, false>(p, w.y, Iin, Iout, sm.f48); break;
        case 9: fused_tile<float, 48, true, false, NW, false>(p, w.y, Iin, Iout, sm.c48); break;
        case 10: fused_tile<float, 64, false, false, NW, false>(p, w.y, Iin, Iout, sm.f64); break;
        default: fused_tile<float, 64, true, false, NW, false>(p, w.y, Iin, Iout, sm.c64); break;
        }
    }
}

__global__ void k_queue_reset(int *counter)
{
    if (threadIdx.x == 0) counter[0] = 0;
}

/* Kernel parameters of one fused step of context c; H = rows per workgroup
 * (0: about `wgs` workgroups with 1 to 64 rows each; on small grids, chunks
 * of 1 or 2 rows fill the device at the price of staging their halo rows
 * again: 100^2 x 64 takes 17.7 us per step against 30.4 us with chunks of
 * at least 4 rows, 200^2 x 32 Kerr-Schild 26.3 against 32.6 us). */
template <typename T>
KP<T> make_kp(const rt_ctx *c, const void *in, void *out, bool want_moments, const RecordPtrs *rec, int H = 0,
              int wgs = 1024)
{
    const rt_desc &d = c->desc;
    KP<T> p;
    p.Iin = (const T *)in;
    p.Iout = (T *)out;
    p.cell = (const T *)c->d_cell;
    p.mask = c->d_mask;
    p.coef = c->d_coef;
    p.coefT = (const T *)c->d_coefT;
    p.excl = c->d_excl;
    p.arec = c->d_arec;
    p.mom = c->d_mom;
    p.stats = c->d_stats;
    p.dt = c->d_dirt;
    p.nx = d.nx;
    p.ny = d.ny;
    p.N = (long)d.nx * d.ny;
    p.strips = (d.nx - 2 + 63) / 64;
    const int rows = d.ny - 2;
    if (H <= 0) {
        const int target = (wgs + p.strips - 1) / p.strips;
        H = (rows + target - 1) / target;
        H = H < 1 ? 1 : (H > 64 ? 64 : H);
    }
    const int chunks = (rows + H - 1) / H;
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
    /* 16-byte staging reads up to 2 columns past a row's last needed one:
     * every buffer the fused step reads a state from (d_I, and the
     * checkpoints and segment states of the adjoint) carries a 256-byte tail
     * guard, or is followed by another state of the same allocation */
    auto a16 = [](const void *q) { return ((uintptr_t)q & 15) == 0; };
    /* (the per-lane DMA offsets of a row are 32-bit: D N sizeof(T) < 2^32) */
    p.wide = (sizeof(T) == 4 && d.nx % 2 == 0 && p.N % 4 == 0 && a16(in) && a16(c->d_cell) && a16(c->d_coefT) &&
              a16(c->d_excl) && (double)c->D * (double)p.N * sizeof(T) < 4294967296.0)
                 ? 1
                 : 0;
    return p;
}

/* Resident workgroups per CU of k_fused<T, D, CURVED, REC, NW, SONLY>. */
template <typename T, int D, bool CURVED, bool REC, int NW, bool SONLY>
int fused_per_cu()
{
    static int per_cu = 0;
    if (per_cu == 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused<T, D, CURVED, REC, NW, SONLY>, 64 * NW, 0) != hipSuccess ||
            n < 1)
            n = 1;
        per_cu = n;
    }
    return per_cu;
}

template <typename T, int D, bool CURVED, bool REC, bool SONLY>
int launch_one(rt_ctx *c, const void *in, void *out, bool want_moments, const RecordPtrs *rec)
{
    /* one round of resident workgroups: each walks a long chunk of rows,
     * which amortises the two halo rows it stages first */
    const int round4 = fused_per_cu<T, D, CURVED, REC, RT_FUSED_WAVES, SONLY>() * c->ncu;
    KP<T> p = make_kp<T>(c, in, out, want_moments, rec, 0, round4);
    if constexpr (has_wide<T, D>()) {
        constexpr bool one_wg = fused_waves_per_simd<T, D, CURVED, RT_FUSED_WAVES>() == 1;
        if (one_wg || 2 * p.nblocks <= round4) {
            KP<T> p8 =
                make_kp<T>(c, in, out, want_moments, rec, 0, fused_per_cu<T, D, CURVED, REC, 8, SONLY>() * c->ncu);
            k_fused<T, D, CURVED, REC, 8, SONLY><<<p8.nblocks, 512, 0, c->stream>>>(p8);
            RT_HIP(c, hipGetLastError());
            if (REC) c->rec_waves = 8;
            return RT_OK;
        }
    }
    if (REC) c->rec_waves = RT_FUSED_WAVES;
    k_fused<T, D, CURVED, REC, RT_FUSED_WAVES, SONLY><<<p.nblocks, 64 * RT_FUSED_WAVES, 0, c->stream>>>(p);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}

template <typename T, bool CURVED, bool REC, bool SONLY>
int dispatch_d(rt_ctx *c, const void *in, void *out, bool wm, const RecordPtrs *rec)
{
    switch (c->D) {
    case 16: return launch_one<T, 16, CURVED, REC, SONLY>(c, in, out, wm, rec);
    case 20: return launch_one<T, 20, CURVED, REC, SONLY>(c, in, out, wm, rec);
    case 24: return launch_one<T, 24, CURVED, REC, SONLY>(c, in, out, wm, rec);
    case 32: return launch_one<T, 32, CURVED, REC, SONLY>(c, in, out, wm, rec);
    case 48: return launch_one<T, 48, CURVED, REC, SONLY>(c, in, out, wm, rec);
    case 64: return launch_one<T, 64, CURVED, REC, SONLY>(c, in, out, wm, rec);
    default: return rt_fail(c, RT_ERR_UNSUPPORTED, "fused kernel: unsupported nDir");
    }
}

} // namespace

bool fused_supported(const rt_ctx *c)
{
    const int D = c->D;
    if (!(D == 16 || D == 20 || D == 24 || D == 32 || D == 48 || D == 64)) return false;
    const rt_desc &d = c->desc;
    if (rt_is_adaptive(d)) return false;
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
        return curved ? dispatch_d<float, true, false, false>(c, in, out, want_moments, nullptr)
                      : dispatch_d<float, false, false, false>(c, in, out, want_moments, nullptr);
    return curved ? dispatch_d<double, true, false, false>(c, in, out, want_moments, nullptr)
                  : dispatch_d<double, false, false, false>(c, in, out, want_moments, nullptr);
}

/* FP64: the fused stream into the true-unit buffer J (d_J), then the exact
 * collide (kernels_collide.hip) into out. */
int launch_exact_step(rt_ctx *c, const void *in, void *out, bool want_moments, const RecordPtrs *rec)
{
    const bool curved = c->desc.streaming == RT_CURVED_FIXED;
    if (!c->d_J) return rt_fail(c, RT_ERR_STATE, "exact FP64 step: no streaming buffer");
    int r;
    if (rec)
        r = curved ? dispatch_d<double, true, true, true>(c, in, c->d_J, false, rec)
                   : dispatch_d<double, false, true, true>(c, in, c->d_J, false, rec);
    else
        r = curved ? dispatch_d<double, true, false, true>(c, in, c->d_J, false, nullptr)
                   : dispatch_d<double, false, false, true>(c, in, c->d_J, false, nullptr);
    if (r != RT_OK) return r;
    return launch_collide_exact(c, c->d_J, out, want_moments, rec);
}

int launch_product_step(rt_ctx *c, const void *in, void *out, bool want_moments)
{
    if (!c->fp32) return launch_exact_step(c, in, out, want_moments, nullptr);
    return launch_fused_step(c, in, out, want_moments);
}

int launch_mark_zero_start(rt_ctx *c)
{
    const unsigned nb = (unsigned)((c->N + 255) / 256);
    auto go = [&](auto tag, auto dtag) -> int {
        using T = decltype(tag);
        constexpr int D = decltype(dtag)::value;
        k_mark_zero_start<T, D><<<nb, 256, 0, c->stream>>>(c->d_f64, c->d_alpha, c->d_dirt, (T *)c->d_cell, (long)c->N,
                                                          c->desc.dt);
        RT_HIP(c, hipGetLastError());
        return RT_OK;
    };
    auto byD = [&](auto tag) -> int {
        switch (c->D) {
        case 16: return go(tag, std::integral_constant<int, 16>{});
        case 20: return go(tag, std::integral_constant<int, 20>{});
        case 24: return go(tag, std::integral_constant<int, 24>{});
        case 32: return go(tag, std::integral_constant<int, 32>{});
        case 48: return go(tag, std::integral_constant<int, 48>{});
        case 64: return go(tag, std::integral_constant<int, 64>{});
        default: return rt_fail(c, RT_ERR_UNSUPPORTED, "fused kernel: unsupported nDir");
        }
    };
    return c->fp32 ? byD(float{}) : byD(double{});
}

int launch_fused_step_rec(rt_ctx *c, const void *in, void *out, const RecordPtrs &rec)
{
    const bool curved = c->desc.streaming == RT_CURVED_FIXED;
    if (!c->fp32) return launch_exact_step(c, in, out, false, &rec);
    return curved ? dispatch_d<float, true, true, false>(c, in, out, false, &rec)
                  : dispatch_d<float, false, true, false>(c, in, out, false, &rec);
}

/* The fused kernel's form of the Fourier tables (computed once per table
 * build): S - 1, X and Y relative to the cell centre in cell units, the
 * velocity series unchanged, in the kernel precision. */
template <typename T>
__global__ void k_coef_rel(const double *__restrict__ coef, T *__restrict__ out, int nx, long N, double startx,
                           double starty, double dx, double dy)
{
    const long ij = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (ij >= N) return;
    const int gi = (int)(ij % nx), j = (int)(ij / nx);
    const double xc = startx + gi * dx, yc = starty + j * dy;
    const double idx = 1.0 / dx, idy = 1.0 / dy;
#pragma unroll
    for (int k = 0; k < 5; k++) {
        out[(size_t)k * N + ij] = (T)(k == 0 ? coef[ij] - 1.0 : coef[(size_t)k * N + ij]);
        out[(size_t)(5 + k) * N + ij] = (T)((k == 0 ? coef[5 * (size_t)N + ij] - xc : coef[(size_t)(5 + k) * N + ij]) * idx);
        out[(size_t)(10 + k) * N + ij] =
            (T)((k == 0 ? coef[10 * (size_t)N + ij] - yc : coef[(size_t)(10 + k) * N + ij]) * idy);
        out[(size_t)(15 + k) * N + ij] = (T)coef[(size_t)(15 + k) * N + ij];
        out[(size_t)(20 + k) * N + ij] = (T)coef[(size_t)(20 + k) * N + ij];
    }
}

/* Angular record of every (cell, direction) of an FP32 fixed curved context:
 * the departure velocity series (15..24 of the FP64 tables) at phi_d and the
 * window the FP64 fused stream takes for it (angle_window64), in the same
 * arithmetic, packed by pack_angle_rec.  The FP32 stream then evaluates only
 * the S, X and Y series.  A departure point inside the excision radius (bit d
 * of excl, k_departure_bound) gets the all-zero weight row 10 instead, so
 * the stream needs no excision test of its own. */
__global__ void k_angle_rec(const double *__restrict__ coef, const DirTables *__restrict__ dt0,
                            const unsigned long long *__restrict__ excl, uint32_t *__restrict__ rec, long N, int D)
{
    const long ij = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (ij >= N) return;
    CDirTables *dt = (CDirTables *)dt0;
    const unsigned long long ex = excl[ij];
    double cU[5], cV[5];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        cU[k] = coef[(size_t)(15 + k) * N + ij];
        cV[k] = coef[(size_t)(20 + k) * N + ij];
    }
    for (int d = 0; d < D; d++) {
        double vx = cU[0], vy = cV[0];
#pragma unroll
        for (int q = 1; q < 5; q++) {
            vx += cU[q] * dt->basis[d][q];
            vy += cV[q] * dt->basis[d][q];
        }
        int s0, r;
        double t;
        angle_window64(vx, vy, dt, s0, r, t);
        if ((ex >> d) & 1ull) {
            s0 = RT_RING_PAD;
            r = 10;
            t = 0.0;
        }
        rec[(size_t)d * N + ij] = pack_angle_rec(s0, r, t);
    }
}

int launch_coef_rel(rt_ctx *c)
{
    if (!rt_is_curved(c->desc) || c->nCoeff != 5 || !c->d_coefT) return RT_OK;
    const rt_desc &d = c->desc;
    const unsigned nb = (unsigned)((c->N + 255) / 256);
    if (c->fp32)
        k_coef_rel<float><<<nb, 256, 0, c->stream>>>(c->d_coef, (float *)c->d_coefT, d.nx, c->N, d.startx, d.starty,
                                                     d.dx, d.dy);
    else
        k_coef_rel<double><<<nb, 256, 0, c->stream>>>(c->d_coef, (double *)c->d_coefT, d.nx, c->N, d.startx, d.starty,
                                                      d.dx, d.dy);
    RT_HIP(c, hipGetLastError());
    if (c->d_arec && c->has_stencil) {
        k_angle_rec<<<nb, 256, 0, c->stream>>>(c->d_coef, c->d_dirt, c->d_excl, c->d_arec, (long)c->N, c->D);
        RT_HIP(c, hipGetLastError());
    }
    return RT_OK;
}

/* ---- batches of independent problems ------------------------------------ */

struct rt_batch {
    std::vector<rt_ctx *> ctx;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    struct Group {
        bool fp32 = false, curved = false;
        int D = 0;
        std::vector<int> members;
        std::vector<int> H;        /* rows per workgroup of each member (0: default) */
        void *d_kp = nullptr;      /* KP<T> table */
        int *d_start = nullptr;    /* first workgroup of each member, then the total */
        int nblocks = 0;
    };
    std::vector<Group> groups;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    /* one persistent launch for all FP32 members (k_fused_queue) instead of
     * one launch per FP32 group; qH: rows per tile (0: the single-problem
     * rule; 64 problems of 64^2..512^2 x 16..64 directions take 191 ms per
     * 200 steps with it, 151 ms with 16 or 32 rows, which the autotuner
     * picks) */
    bool queue = true;
    int qH = 16;
    int q_nitems = 0, q_grid = 0;
    KP<float> *d_qkp = nullptr;
    int2 *d_items = nullptr;
    int *d_counter = nullptr;
    size_t q_items_cap = 0;
};

namespace {
int bfail(rt_batch *b, int code, const std::string &msg)
{
    if (b) b->err = msg;
    return code;
}
#define BHIP(b, call)                                                                      \
    do {                                                                                   \
        hipError_t _e = (call);                                                            \
        if (_e != hipSuccess) return bfail((b), RT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

template <typename T>
int fill_group_table(rt_batch *b, rt_batch::Group &g)
{
    std::vector<KP<T>> kp;
    std::vector<int> start;
    int total = 0;
    for (size_t m = 0; m < g.members.size(); m++) {
        rt_ctx *c = b->ctx[g.members[m]];
        kp.push_back(make_kp<T>(c, c->d_I[c->cur], c->d_I[1 - c->cur], false, nullptr, g.H[m]));
        start.push_back(total);
        total += kp.back().nblocks;
    }
    start.push_back(total);
    g.nblocks = total;
    if (!g.d_kp) BHIP(b, hipMalloc(&g.d_kp, kp.size() * sizeof(KP<T>)));
    if (!g.d_start) BHIP(b, hipMalloc((void **)&g.d_start, start.size() * sizeof(int)));
    BHIP(b, hipMemcpyAsync(g.d_kp, kp.data(), kp.size() * sizeof(KP<T>), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(g.d_start, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return RT_OK;
}

/* Resident workgroups of k_fused_queue on the batch's device. */
int queue_grid(const rt_batch *b)
{
    static int per_cu = 0;
    if (per_cu == 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_queue, 64 * RT_FUSED_WAVES, 0) != hipSuccess || n < 1)
            n = 1;
        per_cu = n;
    }
    return per_cu * b->ctx[0]->ncu;
}

bool has_fp32(const rt_batch *b)
{
    for (const auto &g : b->groups)
        if (g.fp32) return true;
    return false;
}

/* The problem table and the item queue of k_fused_queue: every FP32 member
 * with b->qH rows per tile (0: the single-problem default for a round of
 * resident workgroups), tiles ordered by decreasing work (rows x nDir, curved
 * counted twice), ties by problem and tile. */
int fill_queue_table(rt_batch *b)
{
    if (!b->queue || !has_fp32(b)) return RT_OK;
    b->q_grid = queue_grid(b);
    std::vector<KP<float>> kp;
    struct Item {
        double cost;
        int prob, tile;
    };
    std::vector<Item> items;
    for (const auto &g : b->groups) {
        if (!g.fp32) continue;
        for (int m : g.members) {
            rt_ctx *c = b->ctx[m];
            KP<float> p = make_kp<float>(c, c->d_I[c->cur], c->d_I[1 - c->cur], false, nullptr, b->qH, b->q_grid);
            p.cls = queue_class(c->D, g.curved);
            if (p.cls < 0) return bfail(b, RT_ERR_UNSUPPORTED, "batch: unsupported nDir");
            const int prob = (int)kp.size();
            for (int t = 0; t < p.nblocks; t++) {
                const int chunk = t / p.strips;
                const int rowA = 1 + chunk * p.rows_per_block;
                const int rows = std::min(rowA + p.rows_per_block, p.ny - 1) - rowA;
                items.push_back({(double)rows * c->D * (g.curved ? 2.0 : 1.0), prob, t});
            }
            kp.push_back(p);
        }
    }
    std::stable_sort(items.begin(), items.end(), [](const Item &x, const Item &y) { return x.cost > y.cost; });
    std::vector<int2> w(items.size());
    for (size_t k = 0; k < items.size(); k++) w[k] = make_int2(items[k].prob, items[k].tile);
    b->q_nitems = (int)w.size();
    if (!b->d_qkp) BHIP(b, hipMalloc((void **)&b->d_qkp, kp.size() * sizeof(KP<float>)));
    if (w.size() > b->q_items_cap) {
        if (b->d_items) (void)hipFree(b->d_items);
        b->d_items = nullptr;
        BHIP(b, hipMalloc((void **)&b->d_items, w.size() * sizeof(int2)));
        b->q_items_cap = w.size();
    }
    if (!b->d_counter) BHIP(b, hipMalloc((void **)&b->d_counter, sizeof(int)));
    BHIP(b, hipMemcpyAsync(b->d_qkp, kp.data(), kp.size() * sizeof(KP<float>), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->d_items, w.data(), w.size() * sizeof(int2), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return RT_OK;
}

/* One step of every FP32 member in one persistent launch. */
int launch_queue(rt_batch *b, int parity)
{
    k_queue_reset<<<1, 64, 0, b->stream>>>(b->d_counter);
    BHIP(b, hipGetLastError());
    k_fused_queue<<<b->q_grid, 64 * RT_FUSED_WAVES, 0, b->stream>>>(b->d_qkp, b->d_items, b->q_nitems, b->d_counter,
                                                                   parity);
    BHIP(b, hipGetLastError());
    return RT_OK;
}

template <typename T, bool CURVED>
int launch_group_t(rt_batch *b, const rt_batch::Group &g, int parity)
{
    const KP<T> *kp = (const KP<T> *)g.d_kp;
    const int n = (int)g.members.size();
    switch (g.D) {
    case 16: k_fused_batch<T, 16, CURVED><<<g.nblocks, 64 * RT_FUSED_WAVES, 0, b->stream>>>(kp, g.d_start, n, parity); break;
    case 20: k_fused_batch<T, 20, CURVED><<<g.nblocks, 64 * RT_FUSED_WAVES, 0, b->stream>>>(kp, g.d_start, n, parity); break;
    case 24: k_fused_batch<T, 24, CURVED><<<g.nblocks, 64 * RT_FUSED_WAVES, 0, b->stream>>>(kp, g.d_start, n, parity); break;
    case 32: k_fused_batch<T, 32, CURVED><<<g.nblocks, 64 * RT_FUSED_WAVES, 0, b->stream>>>(kp, g.d_start, n, parity); break;
    case 48: k_fused_batch<T, 48, CURVED><<<g.nblocks, 64 * RT_FUSED_WAVES, 0, b->stream>>>(kp, g.d_start, n, parity); break;
    case 64: k_fused_batch<T, 64, CURVED><<<g.nblocks, 64 * RT_FUSED_WAVES, 0, b->stream>>>(kp, g.d_start, n, parity); break;
    default: return bfail(b, RT_ERR_UNSUPPORTED, "batch: unsupported nDir");
    }
    BHIP(b, hipGetLastError());
    return RT_OK;
}

/* FP64 members step as rt_step does (fused stream, exact collide), one
 * member after another on the batch's stream. */
int launch_group_exact(rt_batch *b, const rt_batch::Group &g, int parity)
{
    for (int m : g.members) {
        rt_ctx *c = b->ctx[m];
        void *in = c->d_I[parity ? 1 - c->cur : c->cur], *out = c->d_I[parity ? c->cur : 1 - c->cur];
        hipStream_t own = c->stream;
        c->stream = b->stream;
        const int r = launch_exact_step(c, in, out, false, nullptr);
        c->stream = own;
        if (r != RT_OK) return bfail(b, r, std::string("batch member: ") + c->err);
    }
    return RT_OK;
}

int launch_group(rt_batch *b, const rt_batch::Group &g, int parity)
{
    if (g.fp32) return g.curved ? launch_group_t<float, true>(b, g, parity) : launch_group_t<float, false>(b, g, parity);
    return launch_group_exact(b, g, parity);
}

/* One step of every problem: the FP32 members in one queue launch (or one
 * launch per FP32 group), the FP64 members as rt_step does. */
int launch_all(rt_batch *b, int parity)
{
    const bool q = b->queue && has_fp32(b);
    if (q)
        if (int r = launch_queue(b, parity)) return r;
    for (const auto &g : b->groups)
        if (!(q && g.fp32))
            if (int r = launch_group(b, g, parity)) return r;
    return RT_OK;
}

/* Brings every member to the fused state convention and refreshes the
 * problem tables (their buffers follow each context's current state). */
int batch_prepare(rt_batch *b)
{
    for (rt_ctx *c : b->ctx) {
        if (c->nsrc > 0) return bfail(b, RT_ERR_UNSUPPORTED, "batch members cannot have source cells (rt_set_sources)");
        if (int r = prepare_fused(c)) return bfail(b, r, std::string("batch member: ") + c->err);
        c->ck_valid = false;
        BHIP(b, hipStreamSynchronize(c->stream));
    }
    for (auto &g : b->groups)
        if (int r = g.fp32 ? fill_group_table<float>(b, g) : fill_group_table<double>(b, g)) return r;
    return fill_queue_table(b);
}

void drop_graph(rt_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

/* A graph of two steps (parity 0 then 1): replays leave every state in its
 * current buffer. */
int capture_graph(rt_batch *b)
{
    drop_graph(b);
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int r = launch_all(b, 0);
    if (r == RT_OK) r = launch_all(b, 1);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (r != RT_OK) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) return bfail(b, RT_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    b->graph = g;
    BHIP(b, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    return RT_OK;
}

/* nSteps steps of every problem: graph replays for the pairs, one direct
 * step for an odd remainder (after which the members' current buffers flip:
 * the FP32 groups read their buffers from the refreshed tables, the FP64
 * members' launches carry them as arguments, so the graph is captured
 * again). */
int batch_run(rt_batch *b, int nSteps)
{
    for (int k = 0; k < nSteps / 2; k++) BHIP(b, hipGraphLaunch(b->exec, b->stream));
    if (nSteps & 1) {
        if (int r = launch_all(b, 0)) return r;
        for (rt_ctx *c : b->ctx) c->cur = 1 - c->cur;
        BHIP(b, hipStreamSynchronize(b->stream));
        bool exact = false;
        for (auto &g : b->groups) {
            if (int r = g.fp32 ? fill_group_table<float>(b, g) : fill_group_table<double>(b, g)) return r;
            exact = exact || !g.fp32;
        }
        if (int r = fill_queue_table(b)) return r;
        if (exact)
            if (int r = capture_graph(b)) return r;
    }
    return RT_OK;
}
} // namespace

extern "C" {

int rt_batch_create(rt_ctx *const *ctxs, int32_t n, rt_batch **out)
{
    if (!out) return RT_ERR_ARG;
    *out = nullptr;
    if (!ctxs || n < 1) return RT_ERR_ARG;
    rt_batch *b = new rt_batch();
    auto fail = [&](int code, const std::string &msg) {
        if (ctxs[0]) rt_fail(ctxs[0], code, "rt_batch_create: " + msg);
        rt_batch_destroy(b);
        return code;
    };
    for (int i = 0; i < n; i++) {
        rt_ctx *c = ctxs[i];
        if (!c) return fail(RT_ERR_ARG, "null context");
        if (i == 0) b->device = c->desc.device;
        if (c->desc.device != b->device) return fail(RT_ERR_ARG, "contexts on different devices");
        if (!c->has_stencil || !c->has_fields || (c->desc.streaming == RT_CURVED_FIXED && !c->has_tables))
            return fail(RT_ERR_STATE, "context " + std::to_string(i) + " is not ready to step");
        if (!fused_supported(c)) return fail(RT_ERR_UNSUPPORTED, "context " + std::to_string(i) + " is not covered by the fused step");
        for (int j = 0; j < i; j++)
            if (ctxs[j] == c) return fail(RT_ERR_ARG, "a context appears twice");
        b->ctx.push_back(c);
        const bool curved = c->desc.streaming == RT_CURVED_FIXED;
        rt_batch::Group *g = nullptr;
        for (auto &gg : b->groups)
            if (gg.fp32 == c->fp32 && gg.curved == curved && gg.D == c->D) g = &gg;
        if (!g) {
            b->groups.emplace_back();
            g = &b->groups.back();
            g->fp32 = c->fp32;
            g->curved = curved;
            g->D = c->D;
        }
        g->members.push_back(i);
        g->H.push_back(0);
    }
    if (hipSetDevice(b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
        return fail(RT_ERR_HIP, "stream creation failed");
    if (int r = batch_prepare(b)) return fail(r, b->err);
    if (int r = capture_graph(b)) return fail(r, b->err);
    *out = b;
    return RT_OK;
}

int rt_batch_destroy(rt_batch *b)
{
    if (!b) return RT_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (auto &g : b->groups) {
        if (g.d_kp) (void)hipFree(g.d_kp);
        if (g.d_start) (void)hipFree(g.d_start);
    }
    if (b->d_qkp) (void)hipFree(b->d_qkp);
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->d_counter) (void)hipFree(b->d_counter);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return RT_OK;
}

const char *rt_batch_last_error(const rt_batch *b) { return b ? b->err.c_str() : "null batch"; }

int rt_batch_step(rt_batch *b, int32_t nSteps)
{
    if (!b) return RT_ERR_ARG;
    if (nSteps < 0) return bfail(b, RT_ERR_ARG, "nSteps must be >= 0");
    if (int r = batch_prepare(b)) return r;
    if (int r = batch_run(b, nSteps)) return r;
    BHIP(b, hipStreamSynchronize(b->stream));
    return RT_OK;
}

int rt_batch_time_steps(rt_batch *b, int32_t nSteps, float *ms)
{
    if (!b || !ms) return RT_ERR_ARG;
    if (nSteps < 0) return bfail(b, RT_ERR_ARG, "nSteps must be >= 0");
    if (int r = batch_prepare(b)) return r;
    BHIP(b, hipEventRecord(b->ev0, b->stream));
    if (int r = batch_run(b, nSteps)) return r;
    BHIP(b, hipEventRecord(b->ev1, b->stream));
    BHIP(b, hipEventSynchronize(b->ev1));
    BHIP(b, hipEventElapsedTime(ms, b->ev0, b->ev1));
    return RT_OK;
}

int rt_batch_autotune(rt_batch *b, int32_t trial_steps, int32_t *chosen)
{
    if (!b) return RT_ERR_ARG;
    if (trial_steps < 1) return bfail(b, RT_ERR_ARG, "trial_steps must be >= 1");
    if (int r = batch_prepare(b)) return r;
    /* the trials advance the states: keep copies and restore them */
    std::vector<void *> saved(b->ctx.size(), nullptr);
    auto release = [&]() {
        for (void *p : saved)
            if (p) (void)hipFree(p);
    };
    for (size_t i = 0; i < b->ctx.size(); i++) {
        rt_ctx *c = b->ctx[i];
        const size_t bytes = (size_t)c->N * c->D * (c->fp32 ? 4 : 8);
        if (hipMalloc(&saved[i], bytes) != hipSuccess) {
            release();
            return bfail(b, RT_ERR_OOM, "autotune: no memory for state copies");
        }
        if (hipMemcpyAsync(saved[i], c->d_I[c->cur], bytes, hipMemcpyDeviceToDevice, b->stream) != hipSuccess) {
            release();
            return bfail(b, RT_ERR_HIP, "autotune: state copy failed");
        }
    }
    static const int candidates[] = {0, 1, 2, 4, 8, 16, 32, 64};
    /* times 2*trial_steps steps of `launch` after one untimed one */
    auto time_it = [&](auto launch, float &ms) -> int {
        int r = launch(0);
        if (r == RT_OK && hipEventRecord(b->ev0, b->stream) != hipSuccess) r = RT_ERR_HIP;
        for (int k = 0; r == RT_OK && k < 2 * trial_steps; k++) r = launch((k + 1) & 1);
        if (r == RT_OK && hipEventRecord(b->ev1, b->stream) != hipSuccess) r = RT_ERR_HIP;
        if (r == RT_OK && hipEventSynchronize(b->ev1) != hipSuccess) r = RT_ERR_HIP;
        if (r == RT_OK && hipEventElapsedTime(&ms, b->ev0, b->ev1) != hipSuccess) r = RT_ERR_HIP;
        return r;
    };
    /* the FP32 members in one queue launch: rows per tile */
    float best_q = 1e30f, sum_groups = 0;
    int best_qH = 0;
    if (has_fp32(b)) {
        b->queue = true;
        for (int H : candidates) {
            if (H == 1) continue; /* one-row tiles stage their halo rows twice over */
            b->qH = H;
            float ms = 0;
            int r = fill_queue_table(b);
            if (r == RT_OK) r = time_it([&](int par) { return launch_queue(b, par); }, ms);
            if (r != RT_OK) {
                release();
                return r == RT_ERR_HIP ? bfail(b, r, "autotune: timing failed") : r;
            }
            if (ms < best_q) {
                best_q = ms;
                best_qH = H;
            }
        }
    }
    int gi = 0;
    for (auto &g : b->groups) {
        float best = 1e30f;
        int bestH = 0;
        for (int H : candidates) {
            for (auto &h : g.H) h = H;
            int r = g.fp32 ? fill_group_table<float>(b, g) : fill_group_table<double>(b, g);
            float ms = 0;
            if (r == RT_OK) r = time_it([&](int par) { return launch_group(b, g, par); }, ms);
            if (r != RT_OK) {
                release();
                return r == RT_ERR_HIP ? bfail(b, r, "autotune: timing failed") : r;
            }
            if (ms < best) {
                best = ms;
                bestH = H;
            }
        }
        for (auto &h : g.H) h = bestH;
        if (chosen) chosen[gi] = bestH;
        if (g.fp32) sum_groups += best;
        gi++;
    }
    /* the queue when it beats one launch per FP32 group */
    b->queue = has_fp32(b) && best_q <= sum_groups;
    b->qH = best_qH;
    if (b->queue && chosen) {
        gi = 0;
        for (auto &g : b->groups) {
            if (g.fp32) chosen[gi] = best_qH;
            gi++;
        }
    }
    for (size_t i = 0; i < b->ctx.size(); i++) {
        rt_ctx *c = b->ctx[i];
        const size_t bytes = (size_t)c->N * c->D * (c->fp32 ? 4 : 8);
        if (hipMemcpyAsync(c->d_I[c->cur], saved[i], bytes, hipMemcpyDeviceToDevice, b->stream) != hipSuccess) {
            release();
            return bfail(b, RT_ERR_HIP, "autotune: state restore failed");
        }
    }
    (void)hipStreamSynchronize(b->stream);
    release();
    if (int r = batch_prepare(b)) return r;
    return capture_graph(b);
}

int rt_batch_set_queue(rt_batch *b, int32_t on)
{
    if (!b) return RT_ERR_ARG;
    b->queue = on != 0;
    if (int r = batch_prepare(b)) return r;
    return capture_graph(b);
}

int rt_batch_uses_queue(const rt_batch *b, int32_t *queue)
{
    if (!b || !queue) return RT_ERR_ARG;
    *queue = (b->queue && has_fp32(b)) ? 1 : 0;
    return RT_OK;
}

int rt_batch_groups(const rt_batch *b, int32_t *ngroups)
{
    if (!b || !ngroups) return RT_ERR_ARG;
    *ngroups = (int32_t)b->groups.size();
    return RT_OK;
}

} /* extern "C" */
