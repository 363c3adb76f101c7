/*
 * C ABI of the stream-collide library (declarations and reference mapping in
 * include/rt2d.h).  Host-side orchestration only; kernels live in
 * kernels_ref.hip and kernels_fused.hip.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>
#include "rt_internal.h"
#include "../cpp/Spacetimes.h"
#include "../cpp/Stencil.h"

int launch_mask(rt_ctx *c);
static int ensure_unscaled(rt_ctx *c);
static int build_tables_on_device(rt_ctx *c, double tol);

/* Staging of rt_sample_outgoing_noalloc: the sample points and their
 * samples, RT_NOALLOC_SCRATCH_DOUBLES doubles on every GPU the library runs
 * on, for the life of the process.  Host code uses each device's address of
 * it (g_scratch), never the name: the name is a host-side stand-in. */
static __device__ __align__(256) double scratch[RT_NOALLOC_SCRATCH_DOUBLES];

namespace {
thread_local std::string g_create_error;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

/* The address of `scratch` on the current device (d), looked up once per
 * device. */
int scratch_address(int d, double **out, std::string &err)
{
    if (d < 0 || d >= kMaxDevices) {
        err = "device ordinal beyond the no-allocation table";
        return RT_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[d]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return RT_ERR_HIP;
        }
        g_scratch[d] = (double *)p;
    }
    *out = g_scratch[d];
    return RT_OK;
}

template <typename P>
int dalloc(rt_ctx *c, P **p, size_t bytes)
{
    if (*p) return RT_OK;
    hipError_t e = hipMalloc((void **)p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return rt_fail(c, RT_ERR_OOM, std::string("device allocation of ") + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    return RT_OK;
}
template <typename P>
void dfree(P *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}
bool all_finite(const double *a, size_t n)
{
    for (size_t i = 0; i < n; i++)
        if (!std::isfinite(a[i])) return false;
    return true;
}
size_t tsize(const rt_ctx *c) { return c->fp32 ? sizeof(float) : sizeof(double); }

#define NEED(c, p)                                                                                 \
    do {                                                                                           \
        if (!(p)) return rt_fail((c), RT_ERR_ARG, "null pointer argument: " #p);                    \
    } while (0)

/* Rounds five weights to multiples of 2^-23 (exact in FP32 for |w| < 2) and
 * puts the remainder of target - sum on the largest one, so the FP32 weights
 * sum to target exactly. */
void exact_sum_f32(const double w[5], double target, float out[5])
{
    const double q = 0x1p-23;
    long long k[5], sum = 0;
    int big = 0;
    for (int m = 0; m < 5; m++) {
        k[m] = std::llround(w[m] / q);
        sum += k[m];
        if (std::fabs(w[m]) > std::fabs(w[big])) big = m;
    }
    k[big] += std::llround(target / q) - sum;
    for (int m = 0; m < 5; m++) out[m] = (float)((double)k[m] * q);
}

/* Builds the per-direction constants of the fused kernel from the stencil and,
 * for the curved path, checks that the cubic tables repeat with period 10 fine
 * angles (true for the uniform stencils Stencil(n,0) builds). */
int build_dir_tables(rt_ctx *c, const double *w, const double *phi, const double *cx, const double *cy,
                     const int32_t *Nc, const double *Wc)
{
    const int D = c->D;
    c->fused_tables_ok = false;
    c->z0_dirty = true;
    if (D > RT_MAX_FUSED_DIR) return RT_OK;
    DirTables *h = new DirTables();
    memset(h, 0, sizeof *h);
    const rt_desc &d = c->desc;
    for (int k = 0; k < D; k++) {
        for (int q = 0; q < c->nCoeff && q < 8; q++) {
            h->basis[k][q] = rtm::fourier_basis(q, phi[k]);
            h->basisf[k][q] = (float)h->basis[k][q];
        }
        const double X[6] = {1.0, cx[k], cy[k], cx[k] * cx[k], cx[k] * cy[k], cy[k] * cy[k]};
        for (int q = 0; q < 6; q++) {
            h->w6[k][q] = w[k] * X[q];
            h->S[0][q] += w[k] * X[q];
            h->S[1][q] += w[k] * cx[k] * X[q];
            h->S[2][q] += w[k] * cy[k] * X[q];
        }
        h->cx[k] = cx[k];
        h->cy[k] = cy[k];
        h->cxf[k] = (float)cx[k];
        h->cyf[k] = (float)cy[k];
        const double di = -cx[k] * d.dt / d.dx, dj = -cy[k] * d.dt / d.dy;
        const double fi = std::floor(di), fj = std::floor(dj);
        h->flat_di[k] = (int)fi;
        h->flat_dj[k] = (int)fj;
        h->flat_fx[k] = di - fi;
        h->flat_fy[k] = dj - fj;
        h->flat_fxf[k] = (float)(di - fi);
        h->flat_fyf[k] = (float)(dj - fj);
    }
    const int nG = c->nGrid;
    h->nGrid = nG;
    h->quarter = nG / 4;
    h->ang_u = nG / (2.0 * rtm::kPi);
    h->ang_uf = (float)h->ang_u;
    bool ok = (nG == 10 * D) && (nG % 4 == 0) && Nc && Wc && D >= 16;
    /* Weights of fine angle g (and of g+1) over directions b-4..b+3 with
     * b = g/10; a window that differs only by taps of (numerically) zero
     * weight, as at fine angles that coincide with a stencil direction,
     * expands to the same vector. */
    auto expand = [&](int g, int b, double out[8]) {
        for (int k = 0; k < 8; k++) out[k] = 0;
        for (int m = 0; m < 4; m++) {
            int rel = ((Nc[4 * g + m] - b) % D + D) % D;
            if (rel > D / 2) rel -= D;
            if (rel < -4 || rel > 3) return false;
            out[rel + 4] += Wc[4 * g + m];
        }
        return true;
    };
    for (int r = 0; ok && r < 10; r++) {
        double W0r[8], W1r[8];
        ok = expand(r, 0, W0r) && expand((r + 1) % nG, 0, W1r);
        for (int g = r; ok && g < nG; g += 10) {
            double W0[8], W1[8];
            const int b = g / 10;
            ok = expand(g, b, W0) && expand((g + 1) % nG, b, W1);
            for (int k = 0; ok && k < 8; k++)
                if (std::fabs(W0[k] - W0r[k]) > 1e-12 || std::fabs(W1[k] - W1r[k]) > 1e-12) ok = false;
        }
        if (!ok) break;
        int pmin = 8, pmax = -1;
        for (int k = 0; k < 8; k++)
            if (std::fabs(W0r[k]) > 1e-13 || std::fabs(W1r[k]) > 1e-13) {
                pmin = std::min(pmin, k);
                pmax = std::max(pmax, k);
            }
        /* the fused kernels' LDS ring holds the windows of offsets -2 and -1 */
        if (pmax < 0 || pmax - pmin > 4 || pmin - 4 < -RT_RING_PAD || pmin - 4 > -1) {
            ok = false;
            break;
        }
        h->woff[r] = pmin - 4;
        for (int m = 0; m < 5; m++) {
            const int k = pmin + m;
            const double a = k < 8 ? W0r[k] : 0.0, b = k < 8 ? W1r[k] : 0.0;
            h->wA[r][m] = a;
            h->wD[r][m] = b - a;
        }
        /* FP32 copies whose exact sums are 1 and 0, like the FP64 weights:
         * every cell shares these rows, so a rounding error in their sum
         * would act as one coherent gain on all rays */
        exact_sum_f32(h->wA[r], 1.0, h->wAf[r]);
        exact_sum_f32(h->wD[r], 0.0, h->wDf[r]);
        /* A - D, exact (both on the 2^-23 grid, difference below 2): the
         * record-driven FP32 stream forms C = (A - D) + (1 + t) D with the
         * fraction t as 1 + t (geometry_from_record) */
        for (int m = 0; m < 5; m++) {
            h->wrowf[r][m] = (float)((double)h->wAf[r][m] - (double)h->wDf[r][m]);
            h->wrowf[r][5 + m] = h->wDf[r][m];
        }
        std::memcpy(&h->wrowf[r][10], &h->woff[r], sizeof(float));
        h->wrowf[r][11] = 0.0f;
    }
    /* row 10: all weights 0, the window of an excised departure (k_angle_rec) */
    for (int m = 0; m < 12; m++) h->wrowf[10][m] = 0.0f;
    c->fused_tables_ok = ok;
    hipError_t e = hipMemcpyAsync(c->d_dirt, h, sizeof *h, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    delete h;
    if (e != hipSuccess) return rt_hip_check(c, e, "upload direction tables");
    return RT_OK;
}

int refresh_departure_bound(rt_ctx *c)
{
    if (c->desc.streaming == RT_CURVED_ADAPTIVE && c->d_coefT && c->has_tables) {
        /* relative-form series of the FP32 adaptive stream */
        TRY(launch_coef_rel(c));
        c->coefT_valid = true;
        return RT_OK;
    }
    if (c->desc.streaming != RT_CURVED_FIXED || !c->has_tables || !c->has_stencil) return RT_OK;
    double m = 0;
    TRY(launch_departure_bound(c, &m));
    c->max_disp = m;
    TRY(launch_coef_rel(c));
    return RT_OK;
}
} // namespace

int rt_fail(rt_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}
int rt_hip_check(rt_ctx *c, hipError_t e, const char *what)
{
    std::string m = std::string("HIP error in ") + what + ": " + hipGetErrorString(e);
    return rt_fail(c, e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP, m);
}

extern "C" {

int rt_abi_version(void) { return RT2D_ABI_VERSION; }

const char *rt_last_error(const rt_ctx *c)
{
    if (!c) return g_create_error.c_str();
    return c->err.c_str();
}

int rt_create(const rt_desc *desc, rt_ctx **out)
{
    if (!desc || !out) return rt_fail(nullptr, RT_ERR_ARG, "rt_create: null argument");
    *out = nullptr;
    const rt_desc &d = *desc;
    if (d.nx < 3 || d.ny < 3) return rt_fail(nullptr, RT_ERR_ARG, "Grid must have at least 2 LP in each Dimension (nx, ny >= 3 with halo).");
    if (!(d.dx > 0 && d.dy > 0 && d.dt > 0) || !std::isfinite(d.dx) || !std::isfinite(d.dy) || !std::isfinite(d.dt))
        return rt_fail(nullptr, RT_ERR_ARG, "dx, dy, dt must be positive and finite");
    if (d.nDir < 4 || d.nDir > 4096) return rt_fail(nullptr, RT_ERR_ARG, "nDir must be in [4, 4096]");
    if (d.nCoeff < 1 || d.nCoeff > 8) return rt_fail(nullptr, RT_ERR_ARG, "nCoeff must be in [1, 8]");
    if (d.precision != RT_FP64 && d.precision != RT_FP32_MIXED) return rt_fail(nullptr, RT_ERR_ARG, "unknown precision");
    if (d.streaming == RT_GEODESIC_FIXED)
        return rt_fail(nullptr, RT_ERR_UNSUPPORTED, "GeodesicFixed streaming (a performance test of the reference) is not implemented");
    if (d.streaming != RT_FLAT_FIXED && d.streaming != RT_CURVED_FIXED && !rt_is_adaptive(d))
        return rt_fail(nullptr, RT_ERR_ARG, "Invalid StreamingType");
    if ((long long)d.nx * d.ny * d.nDir > (1LL << 33)) return rt_fail(nullptr, RT_ERR_ARG, "problem too large");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return rt_fail(nullptr, RT_ERR_HIP, "no HIP device available");
    if (d.device < 0 || d.device >= ndev) return rt_fail(nullptr, RT_ERR_ARG, "device ordinal out of range");
    e = hipSetDevice(d.device);
    if (e != hipSuccess) return rt_fail(nullptr, RT_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));

    rt_ctx *c = new rt_ctx();
    c->desc = d;
    c->N = d.nx * d.ny;
    c->D = d.nDir;
    c->nCoeff = d.nCoeff;
    c->fp32 = d.precision == RT_FP32_MIXED;
    if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, d.device) != hipSuccess || c->ncu < 1)
        c->ncu = 256;
    int r = RT_OK;
    auto fail = [&](int rc) {
        g_create_error = c->err;
        rt_destroy(c);
        return rc;
    };
    {
        std::string serr;
        if (int rs = scratch_address(d.device, &c->d_scratch, serr)) return fail(rt_fail(c, rs, serr));
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(rt_fail(c, RT_ERR_HIP, "stream creation failed"));
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) return fail(rt_fail(c, RT_ERR_HIP, "event creation failed"));
    const size_t N = c->N, D = c->D;
    /* the state buffers carry a 256-byte tail guard: the fused kernel's
     * 16-byte row staging may read up to 2 elements past the last row */
    if ((r = dalloc(c, &c->d_I[0], N * D * tsize(c) + 256)) || (r = dalloc(c, &c->d_I[1], N * D * tsize(c) + 256)) ||
        (r = dalloc(c, &c->d_mom, 6 * N * sizeof(double))) || (r = dalloc(c, &c->d_count, N * sizeof(int32_t))) ||
        (r = dalloc(c, &c->d_stats, RT_STAT_SLOTS * RT_STAT_STRIDE * sizeof(unsigned long long))) || (r = dalloc(c, &c->d_mask, N)) ||
        (r = dalloc(c, &c->d_f64, 6 * N * sizeof(double))) || (r = dalloc(c, &c->d_cell, RT_CELL_FIELDS * N * tsize(c))) ||
        (r = dalloc(c, &c->d_alpha, N * sizeof(double))) || (r = dalloc(c, &c->d_t11, N * sizeof(double))) ||
        (r = dalloc(c, &c->d_t12, N * sizeof(double))) || (r = dalloc(c, &c->d_t21, N * sizeof(double))) ||
        (r = dalloc(c, &c->d_t22, N * sizeof(double))) || (r = dalloc(c, &c->d_dirt, sizeof(DirTables))) ||
        (r = dalloc(c, &c->d_w, D * sizeof(double))) || (r = dalloc(c, &c->d_phi, D * sizeof(double))) ||
        (r = dalloc(c, &c->d_cx, D * sizeof(double))) || (r = dalloc(c, &c->d_cy, D * sizeof(double))) ||
        (r = dalloc(c, &c->d_basis, D * d.nCoeff * sizeof(double))))
        return fail(r);
    if (rt_is_curved(d) && (r = dalloc(c, &c->d_coef, 5 * (size_t)d.nCoeff * N * sizeof(double)))) return fail(r);
    if (d.streaming == RT_CURVED_FIXED && ((r = dalloc(c, &c->d_coefT, 5 * (size_t)d.nCoeff * N * tsize(c))) ||
                                           (r = dalloc(c, &c->d_excl, N * sizeof(unsigned long long)))))
        return fail(r);
    if (d.streaming == RT_CURVED_FIXED && c->fp32 && D <= RT_MAX_FUSED_DIR) {
        if ((r = dalloc(c, &c->d_arec, N * D * sizeof(uint32_t)))) return fail(r);
        /* zero records are valid windows (the fused kernel reads them unchecked) */
        if (hipMemsetAsync(c->d_arec, 0, N * D * sizeof(uint32_t), c->stream) != hipSuccess)
            return fail(rt_fail(c, RT_ERR_HIP, "record clear failed"));
    }
    /* FP64 fixed streaming: the streamed intensities between the fused
     * stream and the exact collide */
    if (!c->fp32 && !rt_is_adaptive(d) && (r = dalloc(c, &c->d_J, N * D * sizeof(double)))) return fail(r);
    if (rt_is_adaptive(d)) {
        if ((r = dalloc(c, &c->d_rot[0], N * sizeof(double))) || (r = dalloc(c, &c->d_rot[1], N * sizeof(double))) ||
            (r = dalloc(c, &c->d_rtrig, 4 * N * sizeof(double))))
            return fail(r);
        if (c->fp32 && (r = dalloc(c, &c->d_rtrigf, 4 * N * sizeof(float)))) return fail(r);
        if (c->fp32 && d.streaming == RT_CURVED_ADAPTIVE && d.nCoeff == 5 &&
            (r = dalloc(c, &c->d_coefT, 5 * (size_t)d.nCoeff * N * sizeof(float))))
            return fail(r);
        /* rotationAngle = rotationAngleNew = 0 (src/Radiation.cpp:55) */
        if (hipMemsetAsync(c->d_rot[0], 0, N * sizeof(double), c->stream) != hipSuccess ||
            hipMemsetAsync(c->d_rot[1], 0, N * sizeof(double), c->stream) != hipSuccess)
            return fail(rt_fail(c, RT_ERR_HIP, "initialisation failed"));
    }
    if ((r = dalloc(c, &c->d_red, sizeof(unsigned long long)))) return fail(r);
    /* zeroed state; flat geometry until rt_set_geometry (alpha = 1, tetrad = I) */
    if (hipMemsetAsync(c->d_I[0], 0, N * D * tsize(c), c->stream) != hipSuccess ||
        hipMemsetAsync(c->d_I[1], 0, N * D * tsize(c), c->stream) != hipSuccess ||
        hipMemsetAsync(c->d_mom, 0, 6 * N * sizeof(double), c->stream) != hipSuccess ||
        hipMemsetAsync(c->d_stats, 0, RT_STAT_SLOTS * RT_STAT_STRIDE * sizeof(unsigned long long), c->stream) != hipSuccess ||
        hipMemsetAsync(c->d_count, 0, N * sizeof(int32_t), c->stream) != hipSuccess ||
        hipMemsetAsync(c->d_f64, 0, 6 * N * sizeof(double), c->stream) != hipSuccess ||
        hipMemsetAsync(c->d_t12, 0, N * sizeof(double), c->stream) != hipSuccess ||
        hipMemsetAsync(c->d_t21, 0, N * sizeof(double), c->stream) != hipSuccess)
        return fail(rt_fail(c, RT_ERR_HIP, "initialisation failed"));
    if ((r = launch_fill(c, c->d_alpha, N, 1.0)) || (r = launch_fill(c, c->d_t11, N, 1.0)) || (r = launch_fill(c, c->d_t22, N, 1.0)) ||
        (r = launch_mask(c)))
        return fail(r);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(rt_fail(c, RT_ERR_HIP, "initialisation failed"));
    *out = c;
    return RT_OK;
}

void rt_destroy(rt_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    dfree(c->d_w); dfree(c->d_phi); dfree(c->d_cx); dfree(c->d_cy); dfree(c->d_Nc); dfree(c->d_Wc);
    dfree(c->d_basis); dfree(c->d_dirt); dfree(c->d_alpha); dfree(c->d_t11); dfree(c->d_t12);
    dfree(c->d_t21); dfree(c->d_t22); dfree(c->d_coef); dfree(c->d_coefT); dfree(c->d_arec); dfree(c->d_excl); dfree(c->d_red);
    dfree(c->d_frame); dfree(c->d_lf); dfree(c->d_src_idx); dfree(c->d_src_I); dfree(c->d_metric);
    dfree(c->d_rot[0]); dfree(c->d_rot[1]); dfree(c->d_rtrig); dfree(c->d_src_rot);
    dfree(c->d_rtrigf); dfree(c->d_Wcf);
    dfree(c->d_ckpt); dfree(c->d_seg); dfree(c->d_rec_m); dfree(c->d_rec_k); dfree(c->d_rec_mask);
    dfree(c->d_lam[0]); dfree(c->d_lam[1]); dfree(c->d_grad[0]); dfree(c->d_grad[1]); dfree(c->d_pts); dfree(c->d_mask); dfree(c->d_f64);
    dfree(c->d_cell); dfree(c->d_I[0]); dfree(c->d_I[1]); dfree(c->d_J); dfree(c->d_mom); dfree(c->d_count);
    dfree(c->d_stats);
    for (hipEvent_t e : {c->ev_seg[0], c->ev_seg[1], c->ev_adj[0], c->ev_adj[1], c->ev_go})
        if (e) (void)hipEventDestroy(e);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int rt_synchronize(rt_ctx *c)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_set_stencil(rt_ctx *c, const double *w, const double *phi, const double *cx, const double *cy, int32_t nGrid,
                   const int32_t *Nc, const double *Wc)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    NEED(c, w); NEED(c, phi); NEED(c, cx); NEED(c, cy);
    const int D = c->D;
    if (!all_finite(w, D) || !all_finite(phi, D) || !all_finite(cx, D) || !all_finite(cy, D))
        return rt_fail(c, RT_ERR_NONFINITE, "stencil arrays contain NaN/Inf");
    if (rt_is_curved(c->desc) || rt_is_adaptive(c->desc)) {
        if (nGrid < 2 || !Nc || !Wc) return rt_fail(c, RT_ERR_ARG, "curved and adaptive streaming need the cubic interpolation tables");
        for (long k = 0; k < 4L * nGrid; k++)
            if (Nc[k] < 0 || Nc[k] >= D) return rt_fail(c, RT_ERR_ARG, "interpolation table index out of range");
        if (!all_finite(Wc, 4L * nGrid)) return rt_fail(c, RT_ERR_NONFINITE, "interpolation weights contain NaN/Inf");
    }
    for (int q = 0; q < 10; q++) c->stencil_sums[q] = 0;
    for (int d = 0; d < D; d++) {
        const double x = cx[d], y = cy[d];
        const double t[10] = {1, x, y, x * x, x * y, y * y, x * x * x, x * x * y, x * y * y, y * y * y};
        for (int q = 0; q < 10; q++) c->stencil_sums[q] += w[d] * t[q];
    }
    RT_HIP(c, hipMemcpyAsync(c->d_w, w, D * sizeof(double), hipMemcpyHostToDevice, c->stream));
    RT_HIP(c, hipMemcpyAsync(c->d_phi, phi, D * sizeof(double), hipMemcpyHostToDevice, c->stream));
    RT_HIP(c, hipMemcpyAsync(c->d_cx, cx, D * sizeof(double), hipMemcpyHostToDevice, c->stream));
    RT_HIP(c, hipMemcpyAsync(c->d_cy, cy, D * sizeof(double), hipMemcpyHostToDevice, c->stream));
    std::vector<double> basis((size_t)D * c->nCoeff);
    for (int d = 0; d < D; d++)
        for (int k = 0; k < c->nCoeff; k++) basis[(size_t)d * c->nCoeff + k] = rtm::fourier_basis(k, phi[d]);
    RT_HIP(c, hipMemcpyAsync(c->d_basis, basis.data(), basis.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (Nc && Wc && nGrid > 0) {
        if (c->nGrid != nGrid) {
            dfree(c->d_Nc);
            dfree(c->d_Wc);
            c->d_Nc = nullptr;
            c->d_Wc = nullptr;
        }
        c->nGrid = nGrid;
        if (!c->d_Nc) TRY(dalloc(c, &c->d_Nc, 4 * (size_t)nGrid * sizeof(int32_t)));
        if (!c->d_Wc) TRY(dalloc(c, &c->d_Wc, 4 * (size_t)nGrid * sizeof(double)));
        RT_HIP(c, hipMemcpyAsync(c->d_Nc, Nc, 4 * (size_t)nGrid * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        RT_HIP(c, hipMemcpyAsync(c->d_Wc, Wc, 4 * (size_t)nGrid * sizeof(double), hipMemcpyHostToDevice, c->stream));
        if (c->fp32 && rt_is_adaptive(c->desc)) {
            std::vector<float> wf(4 * (size_t)nGrid);
            for (size_t k = 0; k < wf.size(); k++) wf[k] = (float)Wc[k];
            dfree(c->d_Wcf);
            c->d_Wcf = nullptr;
            TRY(dalloc(c, &c->d_Wcf, wf.size() * sizeof(float)));
            RT_HIP(c, hipMemcpyAsync(c->d_Wcf, wf.data(), wf.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
            RT_HIP(c, hipStreamSynchronize(c->stream));
        }
    }
    TRY(build_dir_tables(c, w, phi, cx, cy, Nc, Wc));
    c->has_stencil = true;
    TRY(refresh_departure_bound(c));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_set_streaming_stencil(rt_ctx *c, const double *w, const double *phi, const double *cx, const double *cy)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    NEED(c, w); NEED(c, phi); NEED(c, cx); NEED(c, cy);
    for (int k = 0; k < c->nCoeff; k++) {
        c->sw[k] = w[k];
        c->sphi[k] = phi[k];
        c->scx[k] = cx[k];
        c->scy[k] = cy[k];
    }
    c->has_sstencil = true;
    return RT_OK;
}

int rt_set_geometry(rt_ctx *c, const double *alpha, const double *t11, const double *t12, const double *t21, const double *t22)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    c->ck_valid = false;
    NEED(c, alpha); NEED(c, t11); NEED(c, t12); NEED(c, t21); NEED(c, t22);
    const size_t N = c->N;
    const rt_desc &d = c->desc;
    for (size_t k = 0; k < N; k++) {
        /* cells inside the excision radius are never streamed from or collided */
        const double x = d.startx + (double)(k % d.nx) * d.dx, y = d.starty + (double)(k / d.nx) * d.dy;
        if (d.excision_r2 >= 0.0 && x * x + y * y <= d.excision_r2) continue;
        if (!std::isfinite(alpha[k]) || !std::isfinite(t11[k]) || !std::isfinite(t12[k]) || !std::isfinite(t21[k]) ||
            !std::isfinite(t22[k]))
            return rt_fail(c, RT_ERR_NONFINITE, "lapse or tetrad contains NaN/Inf outside the excision region");
        if (!(alpha[k] > 0)) return rt_fail(c, RT_ERR_ARG, "alpha must be positive");
    }
    TRY(ensure_unscaled(c));
    RT_HIP(c, hipMemcpyAsync(c->d_alpha, alpha, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    RT_HIP(c, hipMemcpyAsync(c->d_t11, t11, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    RT_HIP(c, hipMemcpyAsync(c->d_t12, t12, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    RT_HIP(c, hipMemcpyAsync(c->d_t21, t21, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    RT_HIP(c, hipMemcpyAsync(c->d_t22, t22, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->has_geometry = true;
    if (c->has_fields) TRY(launch_prepare_cells(c));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_set_fourier_tables(rt_ctx *c, const double *cS, const double *cX, const double *cY, const double *cCx, const double *cCy)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    if (!rt_is_curved(c->desc)) return rt_fail(c, RT_ERR_STATE, "Fourier tables are only used by curved streaming");
    NEED(c, cS); NEED(c, cX); NEED(c, cY); NEED(c, cCx); NEED(c, cCy);
    const size_t N = c->N, F = c->nCoeff;
    const double *src[5] = {cS, cX, cY, cCx, cCy};
    std::vector<double> soa(5 * F * N);
    for (int q = 0; q < 5; q++) {
        if (!all_finite(src[q], N * F)) return rt_fail(c, RT_ERR_NONFINITE, "Fourier table contains NaN/Inf");
        for (size_t ij = 0; ij < N; ij++)
            for (size_t k = 0; k < F; k++) soa[(q * F + k) * N + ij] = src[q][ij * F + k];
    }
    RT_HIP(c, hipMemcpyAsync(c->d_coef, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->has_tables = true;
    TRY(refresh_departure_bound(c));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_build_fourier_tables(rt_ctx *c, const double *metric, double tol)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    if (!rt_is_curved(c->desc)) return rt_fail(c, RT_ERR_STATE, "Fourier tables are only used by curved streaming");
    NEED(c, metric);
    if (!c->has_sstencil) return rt_fail(c, RT_ERR_STATE, "rt_set_streaming_stencil must come first");
    if (!(tol > 0) || !std::isfinite(tol)) return rt_fail(c, RT_ERR_ARG, "rk45 tolerance must be positive");
    /* the block may hold NaN at a coordinate singularity inside the excision
     * region (Kerr-Schild at r = 0); the tables are validated instead.  The
     * block stays on the device for rt_rebuild_fourier_tables. */
    const size_t bytes = (size_t)c->N * RT_NMETRIC * sizeof(double);
    TRY(dalloc(c, &c->d_metric, bytes));
    RT_HIP(c, hipMemcpyAsync(c->d_metric, metric, bytes, hipMemcpyHostToDevice, c->stream));
    return build_tables_on_device(c, tol);
}

int rt_rebuild_fourier_tables(rt_ctx *c, double tol)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    if (!c->d_metric) return rt_fail(c, RT_ERR_STATE, "rt_build_fourier_tables has not been called");
    if (!(tol > 0) || !std::isfinite(tol)) return rt_fail(c, RT_ERR_ARG, "rk45 tolerance must be positive");
    return build_tables_on_device(c, tol);
}

/* RK45 tables from the device metric block c->d_metric. */
static int build_tables_on_device(rt_ctx *c, double tol)
{
    c->ck_valid = false;
    TRY(launch_fill(c, c->d_coef, 5 * (size_t)c->nCoeff * c->N, 0.0));
    TRY(launch_geodesic_tables(c, c->d_metric, tol));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    c->has_tables = true;
    TRY(refresh_departure_bound(c));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    if (c->has_stencil && !(c->max_disp < 1e29)) {
        c->has_tables = false;
        return rt_fail(c, RT_ERR_NONFINITE, "geodesic tables are not finite at an active cell (metric NaN/Inf near the path?)");
    }
    return RT_OK;
}

int rt_get_fourier_tables(rt_ctx *c, double *cS, double *cX, double *cY, double *cCx, double *cCy)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    if (!c->has_tables) return rt_fail(c, RT_ERR_STATE, "no Fourier tables");
    const size_t N = c->N, F = c->nCoeff;
    std::vector<double> soa(5 * F * N);
    RT_HIP(c, hipMemcpyAsync(soa.data(), c->d_coef, soa.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    double *dst[5] = {cS, cX, cY, cCx, cCy};
    for (int q = 0; q < 5; q++) {
        if (!dst[q]) continue;
        for (size_t ij = 0; ij < N; ij++)
            for (size_t k = 0; k < F; k++) dst[q][ij * F + k] = soa[(q * F + k) * N + ij];
    }
    return RT_OK;
}

int rt_set_fields(rt_ctx *c, const double *eta, const double *kappa0, const double *kappa1, const double *kappaA, const double *ux,
                  const double *uy)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    c->ck_valid = false;
    const double *src[6] = {eta, kappa0, kappa1, kappaA, ux, uy};
    const size_t N = c->N;
    bool any_u = false;
    for (int q = 0; q < 6; q++) {
        if (!src[q]) return rt_fail(c, RT_ERR_ARG, "null field pointer");
        if (!all_finite(src[q], N)) return rt_fail(c, RT_ERR_NONFINITE, "field contains NaN/Inf");
    }
    for (size_t k = 0; k < N; k++) {
        if (ux[k] != 0.0 || uy[k] != 0.0) any_u = true;
        if (ux[k] * ux[k] + uy[k] * uy[k] >= 1.0) return rt_fail(c, RT_ERR_ARG, "fluid velocity must be below the speed of light");
    }
    TRY(ensure_unscaled(c));
    for (int q = 0; q < 6; q++)
        RT_HIP(c, hipMemcpyAsync(c->d_f64 + q * N, src[q], N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->any_u = any_u;
    c->has_fields = true;
    TRY(launch_prepare_cells(c));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_upload_I(rt_ctx *c, const double *I)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    c->ck_valid = false;
    NEED(c, I);
    const size_t n = (size_t)c->N * c->D;
    if (!all_finite(I, n)) return rt_fail(c, RT_ERR_NONFINITE, "intensities contain NaN/Inf");
    double *tmp = nullptr;
    TRY(dalloc(c, &tmp, n * sizeof(double)));
    int r = RT_OK;
    hipError_t e = hipMemcpyAsync(tmp, I, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) r = rt_hip_check(c, e, "upload I");
    if (r == RT_OK) r = launch_transpose_in(c, tmp, c->d_I[c->cur]);
    if (r == RT_OK) {
        /* the scratch buffer starts as the reference's Inew: zero, so that the
         * halo of every streamed state is 0 (its moments feed the rotation
         * update of adaptive streaming); the fused path copies the fixed cells
         * itself before it steps */
        e = hipMemsetAsync(c->d_I[1 - c->cur], 0, n * tsize(c), c->stream);
        if (e != hipSuccess) r = rt_hip_check(c, e, "clear scratch I");
    }
    if (r == RT_OK) {
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) r = rt_hip_check(c, e, "upload I");
    }
    dfree(tmp);
    if (r == RT_OK) c->has_I = true;
    c->scaled = false;
    c->fixed_synced = false;
    return r;
}

int rt_download_I(rt_ctx *c, double *I)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    NEED(c, I);
    const size_t n = (size_t)c->N * c->D;
    TRY(ensure_unscaled(c));
    double *tmp = nullptr;
    TRY(dalloc(c, &tmp, n * sizeof(double)));
    int r = launch_transpose_out(c, c->d_I[c->cur], tmp);
    hipError_t e;
    if (r == RT_OK) {
        e = hipMemcpyAsync(I, tmp, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e != hipSuccess) r = rt_hip_check(c, e, "download I");
    }
    if (r == RT_OK) {
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) r = rt_hip_check(c, e, "download I");
    }
    dfree(tmp);
    return r;
}

int rt_get_moments(rt_ctx *c, double *E, double *Fx, double *Fy, double *Pxx, double *Pxy, double *Pyy)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    double *dst[6] = {E, Fx, Fy, Pxx, Pxy, Pyy};
    for (int q = 0; q < 6; q++)
        if (dst[q]) RT_HIP(c, hipMemcpyAsync(dst[q], c->d_mom + (size_t)q * c->N, c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_set_lab_frame(rt_ctx *c, const double *const tetrad[9], const double *gamma11, const double *gamma12,
                     const double *gamma22)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    NEED(c, tetrad); NEED(c, gamma11); NEED(c, gamma12); NEED(c, gamma22);
    const size_t N = c->N;
    const double *src[12];
    for (int k = 0; k < 9; k++) {
        NEED(c, tetrad[k]);
        src[k] = tetrad[k];
    }
    src[9] = gamma11;
    src[10] = gamma12;
    src[11] = gamma22;
    TRY(dalloc(c, &c->d_frame, 12 * N * sizeof(double)));
    TRY(dalloc(c, &c->d_lf, 7 * N * sizeof(double)));
    for (int k = 0; k < 12; k++)
        RT_HIP(c, hipMemcpyAsync(c->d_frame + (size_t)k * N, src[k], N * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    c->has_frame = true;
    return RT_OK;
}

int rt_set_sources(rt_ctx *c, const uint8_t *isSource, const double *I)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    c->ck_valid = false;
    if (!isSource) {
        c->nsrc = 0;
        return RT_OK;
    }
    NEED(c, I);
    const size_t N = c->N, D = c->D;
    std::vector<int32_t> idx;
    for (size_t ij = 0; ij < N; ij++)
        if (isSource[ij]) idx.push_back((int32_t)ij);
    const int n = (int)idx.size();
    std::vector<double> soa((size_t)n * D);
    for (int k = 0; k < n; k++)
        for (size_t d = 0; d < D; d++) {
            const double v = I[(size_t)idx[k] * D + d];
            if (!std::isfinite(v)) return rt_fail(c, RT_ERR_NONFINITE, "source intensities contain NaN/Inf");
            soa[d * n + k] = v;
        }
    if (n > c->src_cap) {
        dfree(c->d_src_idx);
        dfree(c->d_src_I);
        dfree(c->d_src_rot);
        c->d_src_idx = nullptr;
        c->d_src_I = nullptr;
        c->d_src_rot = nullptr;
        c->src_cap = 0;
        TRY(dalloc(c, &c->d_src_idx, (size_t)n * sizeof(int32_t)));
        TRY(dalloc(c, &c->d_src_I, soa.size() * sizeof(double)));
        if (rt_is_adaptive(c->desc)) TRY(dalloc(c, &c->d_src_rot, (size_t)n * sizeof(double)));
        c->src_cap = n;
    }
    c->src_rot_set = false;
    if (n > 0) {
        RT_HIP(c, hipMemcpyAsync(c->d_src_idx, idx.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        RT_HIP(c, hipMemcpyAsync(c->d_src_I, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        RT_HIP(c, hipStreamSynchronize(c->stream));
    }
    c->nsrc = n;
    return RT_OK;
}

int rt_get_moments_lf(rt_ctx *c, double *E_LF, double *Fx_LF, double *Fy_LF, double *Pxx_LF, double *Pxy_LF,
                      double *Pyy_LF, double *F_LF)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    if (!c->has_frame) return rt_fail(c, RT_ERR_STATE, "rt_set_lab_frame has not been called");
    TRY(launch_moments_lf(c));
    double *dst[7] = {E_LF, Fx_LF, Fy_LF, Pxx_LF, Pxy_LF, Pyy_LF, F_LF};
    for (int q = 0; q < 7; q++)
        if (dst[q])
            RT_HIP(c, hipMemcpyAsync(dst[q], c->d_lf + (size_t)q * c->N, c->N * sizeof(double),
                                     hipMemcpyDeviceToHost, c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_iteration_stats(rt_ctx *c, int32_t *maxCount, double *sumCount)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    unsigned long long s[RT_STAT_SLOTS * RT_STAT_STRIDE];
    RT_HIP(c, hipMemcpyAsync(s, c->d_stats, sizeof s, hipMemcpyDeviceToHost, c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    unsigned long long sum = 0, mx = 0;
    for (int k = 0; k < RT_STAT_SLOTS; k++) {
        sum += s[k * RT_STAT_STRIDE];
        mx = std::max(mx, s[k * RT_STAT_STRIDE + 1]);
    }
    if (maxCount) *maxCount = (int32_t)mx;
    if (sumCount) *sumCount = (double)sum;
    return RT_OK;
}

int rt_get_iteration_counts(rt_ctx *c, int32_t *count)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    NEED(c, count);
    RT_HIP(c, hipMemcpyAsync(count, c->d_count, c->N * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_reset_iteration_stats(rt_ctx *c)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    RT_HIP(c, hipMemsetAsync(c->d_stats, 0, RT_STAT_SLOTS * RT_STAT_STRIDE * sizeof(unsigned long long), c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

/* Brings d_I[cur] back to intensities if the fused curved path left it as
 * alpha^-3 I, and makes the fixed cells of both buffers equal again. */
static int ensure_unscaled(rt_ctx *c)
{
    /* the per-method kernels below write the buffers in their own way */
    c->fixed_synced = false;
    if (!c->scaled) return RT_OK;
    TRY(launch_scale_state(c, c->d_I[c->cur], true));
    TRY(launch_copy_fixed_cells(c, c->d_I[c->cur], c->d_I[1 - c->cur]));
    c->scaled = false;
    return RT_OK;
}

/* State convention of the fused kernel: curved state as alpha^-3 I, and the
 * halo and excised cells equal in both buffers. */
int prepare_fused(rt_ctx *c)
{
    if (c->desc.streaming == RT_CURVED_FIXED && !c->scaled) {
        TRY(launch_scale_state(c, c->d_I[c->cur], false));
        c->scaled = true;
        c->fixed_synced = false;
    }
    if (!c->fixed_synced) {
        TRY(launch_copy_fixed_cells(c, c->d_I[c->cur], c->d_I[1 - c->cur]));
        c->fixed_synced = true;
    }
    if (c->z0_dirty) {
        TRY(launch_mark_zero_start(c));
        c->z0_dirty = false;
    }
    return RT_OK;
}

static int check_ready(rt_ctx *c)
{
    if (!c->has_stencil) return rt_fail(c, RT_ERR_STATE, "rt_set_stencil has not been called");
    if (!c->has_fields) return rt_fail(c, RT_ERR_STATE, "rt_set_fields has not been called");
    if (rt_is_curved(c->desc) && !c->has_tables) return rt_fail(c, RT_ERR_STATE, "curved streaming needs Fourier tables");
    if ((rt_is_curved(c->desc) || rt_is_adaptive(c->desc)) && !c->d_Nc)
        return rt_fail(c, RT_ERR_STATE, "curved and adaptive streaming need the cubic interpolation tables");
    return RT_OK;
}

int rt_uses_fused_path(rt_ctx *c, int32_t *fused)
{
    if (!c || !fused) return rt_fail(c, RT_ERR_ARG, "null argument");
    *fused = fused_supported(c) ? 1 : 0;
    return RT_OK;
}

int rt_step(rt_ctx *c, int32_t nSteps, int32_t want_moments)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    c->ck_valid = false;
    if (nSteps < 0) return rt_fail(c, RT_ERR_ARG, "nSteps must be >= 0");
    TRY(check_ready(c));
    if (nSteps == 0) return RT_OK;
    if (fused_supported(c)) {
        TRY(prepare_fused(c));
        if (want_moments) TRY(launch_fill(c, c->d_mom, 6 * (size_t)c->N, 0.0));
        for (int n = 0; n < nSteps; n++) {
            TRY(launch_restore_sources(c, c->d_I[c->cur], c->scaled));
            TRY(launch_product_step(c, c->d_I[c->cur], c->d_I[1 - c->cur], want_moments && n == nSteps - 1));
            c->cur = 1 - c->cur;
        }
    } else if (rt_is_adaptive(c->desc)) {
        /* RunSimulation (src/Radiation.cpp:784-819): LoadInitialData for the
         * source cells, UpdateRotationMatrizes, adaptive stream (swapping
         * I and the angle buffers), ComputeMomentsIF, Collide */
        c->fixed_synced = false;
        for (int n = 0; n < nSteps; n++) {
            TRY(launch_restore_sources(c, c->d_I[c->cur], false));
            TRY(launch_restore_source_rot(c));
            TRY(launch_rotation(c, true));
            TRY(launch_stream_adaptive(c, c->d_I[c->cur], c->d_I[1 - c->cur]));
            c->cur = 1 - c->cur;
            c->rcur = 1 - c->rcur;
            TRY(launch_ref_moments(c, c->d_I[c->cur]));
            if (c->fp32) TRY(launch_collide_adaptive(c, c->d_I[c->cur], c->d_I[1 - c->cur]));
            else TRY(launch_ref_collide(c, c->d_I[c->cur], c->d_I[1 - c->cur]));
            c->cur = 1 - c->cur;
        }
    } else {
        TRY(ensure_unscaled(c));
        for (int n = 0; n < nSteps; n++) {
            TRY(launch_restore_sources(c, c->d_I[c->cur], false));
            TRY(launch_ref_stream(c, c->d_I[c->cur], c->d_I[1 - c->cur]));
            c->cur = 1 - c->cur;
            TRY(launch_ref_moments(c, c->d_I[c->cur]));
            TRY(launch_ref_collide(c, c->d_I[c->cur], c->d_I[1 - c->cur]));
            c->cur = 1 - c->cur;
        }
    }
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}

int rt_stream(rt_ctx *c)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    c->ck_valid = false;
    TRY(check_ready(c));
    TRY(ensure_unscaled(c));
    if (rt_is_adaptive(c->desc)) {
        TRY(launch_rotation(c, false));
        TRY(launch_stream_adaptive(c, c->d_I[c->cur], c->d_I[1 - c->cur]));
        c->rcur = 1 - c->rcur;
    } else {
        TRY(launch_ref_stream(c, c->d_I[c->cur], c->d_I[1 - c->cur]));
    }
    c->cur = 1 - c->cur;
    return RT_OK;
}

int rt_update_rotation(rt_ctx *c)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    if (!rt_is_adaptive(c->desc)) return rt_fail(c, RT_ERR_STATE, "rotation angles belong to adaptive streaming");
    return launch_rotation(c, true);
}

/* angles the reference's MySin/MyCos accept (src/Utility.hh:272-359) */
static bool angles_ok(const double *a, size_t n)
{
    for (size_t k = 0; k < n; k++)
        if (!(a[k] >= -rtm::kPi2 && a[k] <= 5.0 * rtm::kPi2)) return false;
    return true;
}

int rt_set_rotation(rt_ctx *c, const double *angle, const double *angleNew)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    if (!rt_is_adaptive(c->desc)) return rt_fail(c, RT_ERR_STATE, "rotation angles belong to adaptive streaming");
    NEED(c, angle);
    const size_t N = c->N;
    if (!angles_ok(angle, N) || (angleNew && !angles_ok(angleNew, N)))
        return rt_fail(c, RT_ERR_ARG, "rotation angles must lie in [-pi/2, 5pi/2]");
    RT_HIP(c, hipMemcpyAsync(c->d_rot[c->rcur], angle, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (angleNew) RT_HIP(c, hipMemcpyAsync(c->d_rot[1 - c->rcur], angleNew, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    else RT_HIP(c, hipMemsetAsync(c->d_rot[1 - c->rcur], 0, N * sizeof(double), c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_get_rotation(rt_ctx *c, double *angle, double *angleNew)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    if (!rt_is_adaptive(c->desc)) return rt_fail(c, RT_ERR_STATE, "rotation angles belong to adaptive streaming");
    const size_t N = c->N;
    if (angle) RT_HIP(c, hipMemcpyAsync(angle, c->d_rot[c->rcur], N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (angleNew) RT_HIP(c, hipMemcpyAsync(angleNew, c->d_rot[1 - c->rcur], N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_set_source_rotation(rt_ctx *c, const double *angle)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    if (!rt_is_adaptive(c->desc)) return rt_fail(c, RT_ERR_STATE, "rotation angles belong to adaptive streaming");
    NEED(c, angle);
    if (c->nsrc == 0) return rt_fail(c, RT_ERR_STATE, "rt_set_sources has not named any source cell");
    std::vector<int32_t> idx(c->nsrc);
    RT_HIP(c, hipMemcpyAsync(idx.data(), c->d_src_idx, idx.size() * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<double> v(c->nsrc);
    for (int k = 0; k < c->nsrc; k++) {
        v[k] = angle[idx[k]];
        if (v[k] == v[k] && !angles_ok(&v[k], 1)) return rt_fail(c, RT_ERR_ARG, "rotation angles must lie in [-pi/2, 5pi/2]");
    }
    RT_HIP(c, hipMemcpyAsync(c->d_src_rot, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    c->src_rot_set = true;
    return RT_OK;
}

int rt_set_moments(rt_ctx *c, const double *E, const double *Fx, const double *Fy, const double *Pxx, const double *Pxy,
                   const double *Pyy)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    const double *src[6] = {E, Fx, Fy, Pxx, Pxy, Pyy};
    for (int q = 0; q < 6; q++) {
        if (!src[q]) continue;
        if (!all_finite(src[q], c->N)) return rt_fail(c, RT_ERR_NONFINITE, "moments contain NaN/Inf");
        RT_HIP(c, hipMemcpyAsync(c->d_mom + (size_t)q * c->N, src[q], c->N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}
int rt_compute_moments(rt_ctx *c)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    TRY(check_ready(c));
    TRY(ensure_unscaled(c));
    return launch_ref_moments(c, c->d_I[c->cur]);
}
int rt_collide(rt_ctx *c)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    c->ck_valid = false;
    TRY(check_ready(c));
    TRY(ensure_unscaled(c));
    TRY(launch_ref_collide(c, c->d_I[c->cur], c->d_I[1 - c->cur]));
    c->cur = 1 - c->cur;
    return RT_OK;
}

int rt_time_steps(rt_ctx *c, int32_t nSteps, float *ms)
{
    if (!c || !ms) return rt_fail(c, RT_ERR_ARG, "null argument");
    RT_HIP(c, hipEventRecord(c->ev0, c->stream));
    TRY(rt_step(c, nSteps, 0));
    RT_HIP(c, hipEventRecord(c->ev1, c->stream));
    RT_HIP(c, hipEventSynchronize(c->ev1));
    RT_HIP(c, hipEventElapsedTime(ms, c->ev0, c->ev1));
    return RT_OK;
}

/* ---- adjoint ------------------------------------------------------------ */

static int check_adjoint_ready(rt_ctx *c)
{
    TRY(check_ready(c));
    if (c->any_u) return rt_fail(c, RT_ERR_UNSUPPORTED, "the adjoint needs zero fluid velocity");
    if (c->nsrc > 0) return rt_fail(c, RT_ERR_UNSUPPORTED, "the adjoint does not take source cells (rt_set_sources)");
    if (!fused_supported(c)) return rt_fail(c, RT_ERR_UNSUPPORTED, "the adjoint needs the fused step (stencil or time step not covered)");
    return RT_OK;
}

/* Segment storage of the adjoint: two segments of `every` states (tail
 * guard, as d_I) and their records, and the second stream that recomputes
 * one segment while the other is read. */
static int ensure_segments(rt_ctx *c, int every)
{
    const size_t N = c->N, bytes = N * c->D * tsize(c);
    if (every > c->seg_cap) {
        c->seg_valid = false;
        dfree(c->d_seg); dfree(c->d_rec_m); dfree(c->d_rec_k); dfree(c->d_rec_mask);
        c->seg_cap = 0;
        TRY(dalloc(c, &c->d_seg, 2 * (size_t)every * bytes + 256));
        TRY(dalloc(c, &c->d_rec_m, 2 * (size_t)every * 6 * N * sizeof(double)));
        TRY(dalloc(c, &c->d_rec_k, 2 * (size_t)every * N));
        TRY(dalloc(c, &c->d_rec_mask, 2 * (size_t)every * 8 * N * sizeof(unsigned long long)));
        c->seg_cap = every;
    }
    if (!c->stream2) {
        /* default priority, like `stream` */
        RT_HIP(c, hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
        for (hipEvent_t *e : {&c->ev_seg[0], &c->ev_seg[1], &c->ev_adj[0], &c->ev_adj[1], &c->ev_go})
            RT_HIP(c, hipEventCreateWithFlags(e, hipEventDisableTiming));
    }
    return RT_OK;
}
static char *seg_state(rt_ctx *c, int b, int i)
{
    return (char *)c->d_seg + ((size_t)b * c->ck_every + i) * ((size_t)c->N * c->D * tsize(c));
}
static RecordPtrs seg_record(rt_ctx *c, int b, int i)
{
    const size_t k = (size_t)b * c->ck_every + i, N = c->N;
    RecordPtrs r;
    r.m = c->d_rec_m + k * 6 * N;
    r.k = c->d_rec_k + k * N;
    r.mask = c->d_rec_mask + k * 8 * N;
    return r;
}

int rt_forward_checkpointed(rt_ctx *c, int32_t nSteps, int32_t every)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    c->ck_valid = false;
    c->seg_valid = false;
    if (nSteps < 0 || every < 1) return rt_fail(c, RT_ERR_ARG, "nSteps must be >= 0 and every >= 1");
    TRY(check_adjoint_ready(c));
    const size_t ND = (size_t)c->N * c->D, bytes = ND * tsize(c);
    const int count = (nSteps + every - 1) / every;
    if (count > c->ck_cap) {
        dfree(c->d_ckpt);
        c->ck_cap = 0;
        TRY(dalloc(c, &c->d_ckpt, (size_t)count * bytes + 256)); /* tail guard, as d_I */
        c->ck_cap = count;
    }
    TRY(ensure_segments(c, every));
    /* the adjoint's own buffers are allocated here, with the checkpoints,
     * so that rt_adjoint allocates nothing */
    TRY(dalloc(c, &c->d_lam[0], bytes));
    TRY(dalloc(c, &c->d_lam[1], bytes));
    TRY(dalloc(c, &c->d_grad[0], c->N * sizeof(double)));
    TRY(dalloc(c, &c->d_grad[1], c->N * sizeof(double)));
    TRY(prepare_fused(c));
    c->ck_scaled = c->scaled;
    c->ck_every = every;
    for (int s = 0; s < count; s++) {
        RT_HIP(c, hipMemcpyAsync((char *)c->d_ckpt + (size_t)s * bytes, c->d_I[c->cur], bytes, hipMemcpyDeviceToDevice,
                                 c->stream));
        const int steps = std::min(every, nSteps - s * every);
        if (s < count - 1) {
            for (int n = 0; n < steps; n++) {
                TRY(launch_product_step(c, c->d_I[c->cur], c->d_I[1 - c->cur], false));
                c->cur = 1 - c->cur;
            }
            continue;
        }
        /* the last segment runs with recording into segment buffer 0 (the
         * same arithmetic), so that the adjoint starts on it without
         * recomputing it; its final state comes back to d_I */
        /* fixed cells (halo and excised, equal in every state of this run)
         * of every state of both segment buffers: the recording steps and
         * the adjoint's recomputation write active cells only */
        for (int b = 0; b < 2; b++)
            for (int i = 0; i < std::min(every, nSteps); i++)
                TRY(launch_copy_fixed_cells(c, c->d_I[c->cur], seg_state(c, b, i)));
        const void *prev = c->d_I[c->cur];
        for (int i = 0; i < steps; i++) {
            TRY(launch_fused_step_rec(c, prev, seg_state(c, 0, i), seg_record(c, 0, i)));
            prev = seg_state(c, 0, i);
        }
        if (steps > 0)
            RT_HIP(c, hipMemcpyAsync(c->d_I[c->cur], prev, bytes, hipMemcpyDeviceToDevice, c->stream));
    }
    RT_HIP(c, hipStreamSynchronize(c->stream));
    c->ck_steps = nSteps;
    c->ck_count = count;
    c->ck_valid = true;
    c->seg_valid = count > 0;
    return RT_OK;
}

/* d_pts layout: x [npts], y [npts], then npts*nDir values (samples or their
 * gradients) */
static int upload_points(rt_ctx *c, int32_t npts, const double *x, const double *y, const double *g)
{
    const int need = 2 * npts + npts * c->D;
    if (need > c->pts_cap) {
        dfree(c->d_pts);
        TRY(dalloc(c, &c->d_pts, (size_t)need * sizeof(double)));
        c->pts_cap = need;
    }
    RT_HIP(c, hipMemcpyAsync(c->d_pts, x, npts * sizeof(double), hipMemcpyHostToDevice, c->stream));
    RT_HIP(c, hipMemcpyAsync(c->d_pts + npts, y, npts * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (g)
        RT_HIP(c, hipMemcpyAsync(c->d_pts + 2 * npts, g, (size_t)npts * c->D * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
    return RT_OK;
}

static int check_points(rt_ctx *c, int32_t npts, const double *x, const double *y)
{
    if (npts < 1) return rt_fail(c, RT_ERR_ARG, "npts must be >= 1");
    const rt_desc &d = c->desc;
    for (int p = 0; p < npts; p++) {
        const double i = (x[p] - d.startx) / d.dx, j = (y[p] - d.starty) / d.dy;
        if (!(i >= 0 && j >= 0 && i < d.nx - 1 && j < d.ny - 1))
            return rt_fail(c, RT_ERR_ARG, "sample point outside the grid");
    }
    return RT_OK;
}

int rt_sample_outgoing(rt_ctx *c, int32_t npts, const double *x, const double *y, double *out)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    NEED(c, x); NEED(c, y); NEED(c, out);
    TRY(check_points(c, npts, x, y));
    TRY(upload_points(c, npts, x, y, nullptr));
    double *dout = c->d_pts + 2 * npts;
    TRY(launch_sample(c, c->d_I[c->cur], c->scaled ? 1 : 0, npts, c->d_pts, c->d_pts + npts, dout));
    RT_HIP(c, hipMemcpyAsync(out, dout, (size_t)npts * c->D * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_sample_outgoing_noalloc(rt_ctx *c, int32_t npts, const double *x, const double *y, double *out)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    NEED(c, x); NEED(c, y); NEED(c, out);
    TRY(check_points(c, npts, x, y));
    if (2 * (long long)npts + (long long)npts * c->D > RT_NOALLOC_SCRATCH_DOUBLES)
        return rt_fail(c, RT_ERR_ARG, "rt_sample_outgoing_noalloc: 2*npts + npts*nDir exceeds the " +
                                          std::to_string(RT_NOALLOC_SCRATCH_DOUBLES) + " doubles of its staging array");
    if (!c->d_scratch) return rt_fail(c, RT_ERR_STATE, "no-allocation staging array not set up");
    double *px = c->d_scratch, *py = px + npts, *dout = py + npts;
    RT_HIP(c, hipMemcpyAsync(px, x, npts * sizeof(double), hipMemcpyHostToDevice, c->stream));
    RT_HIP(c, hipMemcpyAsync(py, y, npts * sizeof(double), hipMemcpyHostToDevice, c->stream));
    TRY(launch_sample(c, c->d_I[c->cur], c->scaled ? 1 : 0, npts, px, py, dout));
    RT_HIP(c, hipMemcpyAsync(out, dout, (size_t)npts * c->D * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_adjoint(rt_ctx *c, int32_t npts, const double *x, const double *y, const double *dL_dout, double *grad_eta,
               double *grad_kappa0)
{
    if (!c) return rt_fail(nullptr, RT_ERR_ARG, "null context");
    NEED(c, x); NEED(c, y); NEED(c, dL_dout); NEED(c, grad_eta); NEED(c, grad_kappa0);
    if (!c->ck_valid) return rt_fail(c, RT_ERR_STATE, "rt_adjoint needs a preceding rt_forward_checkpointed run");
    TRY(check_adjoint_ready(c));
    TRY(check_points(c, npts, x, y));
    if (!all_finite(dL_dout, (size_t)npts * c->D)) return rt_fail(c, RT_ERR_NONFINITE, "dL_dout contains NaN/Inf");
    const size_t N = c->N, ND = N * c->D, bytes = ND * tsize(c);
    const int every = c->ck_every;
    TRY(ensure_segments(c, every));
    TRY(dalloc(c, &c->d_lam[0], bytes));
    TRY(dalloc(c, &c->d_lam[1], bytes));
    TRY(dalloc(c, &c->d_grad[0], N * sizeof(double)));
    TRY(dalloc(c, &c->d_grad[1], N * sizeof(double)));
    TRY(upload_points(c, npts, x, y, dL_dout));
    /* the recomputation and the seed use the checkpoints' state convention;
     * the current state (d_I) is not touched, and its flag comes back on
     * every return */
    struct ScaleGuard {
        rt_ctx *c;
        bool saved;
        ~ScaleGuard() { c->scaled = saved; }
    } guard{c, c->scaled};
    c->scaled = c->ck_scaled;
    RT_HIP(c, hipMemsetAsync(c->d_grad[0], 0, N * sizeof(double), c->stream));
    RT_HIP(c, hipMemsetAsync(c->d_grad[1], 0, N * sizeof(double), c->stream));
    int lc = 0;
    RT_HIP(c, hipMemsetAsync(c->d_lam[lc], 0, bytes, c->stream));
    /* dL/d(final state) from the sampled outputs */
    TRY(launch_sample_adjoint(c, c->d_lam[lc], c->ck_scaled ? 1 : 0, npts, c->d_pts, c->d_pts + npts,
                              c->d_pts + 2 * npts));
    /* segment buffer b (0, 1): the state after step i and its records */
    auto seg_at = [&](int b, int i) { return seg_state(c, b, i); };
    auto rec_at = [&](int b, int i) { return seg_record(c, b, i); };
    /* buffer 0 already holds the last segment, recorded by the forward run */
    const bool have_last = c->seg_valid;
    c->seg_valid = false; /* the recomputation below overwrites it */
    /* halo and excised cells of every segment state were filled by
     * rt_forward_checkpointed (a valid checkpoint means an unchanged state) */
    /* the adjoint step clears the active cells of the buffer it reads, which
     * then takes the next step's scatter; what accumulates in fixed cells is
     * never read (fixed cells do not depend on active ones) */
    RT_HIP(c, hipMemsetAsync(c->d_lam[1 - lc], 0, bytes, c->stream));
    /* recompute segment s into buffer b on stream2, keeping every state
     * after a step */
    auto recompute = [&](int s, int b) -> int {
        const int steps = std::min(every, c->ck_steps - s * every);
        const void *prev = (const char *)c->d_ckpt + (size_t)s * bytes;
        hipStream_t keep = c->stream;
        c->stream = c->stream2;
        int r = RT_OK;
        for (int i = 0; i < steps && r == RT_OK; i++) {
            r = launch_fused_step_rec(c, prev, seg_at(b, i), rec_at(b, i));
            prev = seg_at(b, i);
        }
        c->stream = keep;
        return r;
    };
    /* stream2 starts after the set-up above */
    RT_HIP(c, hipEventRecord(c->ev_go, c->stream));
    RT_HIP(c, hipStreamWaitEvent(c->stream2, c->ev_go, 0));
    const int last = c->ck_count - 1;
    if (last >= 0) {
        if (!have_last) TRY(recompute(last, 0));
        RT_HIP(c, hipEventRecord(c->ev_seg[0], c->stream2));
    }
    for (int s = last, b = 0; s >= 0; s--, b ^= 1) {
        if (s >= 1) {
            /* buffer 1-b last served segment s+1's adjoint steps */
            if (s < last) RT_HIP(c, hipStreamWaitEvent(c->stream2, c->ev_adj[1 - b], 0));
            TRY(recompute(s - 1, 1 - b));
            RT_HIP(c, hipEventRecord(c->ev_seg[1 - b], c->stream2));
        }
        RT_HIP(c, hipStreamWaitEvent(c->stream, c->ev_seg[b], 0));
        /* adjoint steps, last to first */
        const int steps = std::min(every, c->ck_steps - s * every);
        for (int i = steps - 1; i >= 0; i--) {
            TRY(launch_adjoint_step(c, c->d_lam[lc], c->d_lam[1 - lc], seg_at(b, i), rec_at(b, i)));
            lc = 1 - lc;
        }
        RT_HIP(c, hipEventRecord(c->ev_adj[b], c->stream));
    }
    RT_HIP(c, hipMemcpyAsync(grad_eta, c->d_grad[0], N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    RT_HIP(c, hipMemcpyAsync(grad_kappa0, c->d_grad[1], N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_build_flux_tables(int32_t nOrder, int32_t capacity, double *relativeFlux, double *sigma, double *normalization,
                         int32_t *count)
{
    if (nOrder < 4 || nOrder > 4096 || capacity < 0 || !count || (capacity > 0 && (!relativeFlux || !sigma || !normalization)))
        return rt_fail(nullptr, RT_ERR_ARG, "rt_build_flux_tables: bad arguments");
    Stencil s(nOrder, 0, true);
    const int n = s.fluxToSigmaTable.size;
    *count = n;
    for (int k = 0; k < n && k < capacity; k++) {
        relativeFlux[k] = s.fluxToSigmaTable.inputs[k];
        sigma[k] = s.fluxToSigmaTable.outputs[k];
        normalization[k] = s.fluxToNormalizationTable.outputs[k];
    }
    return RT_OK;
}

int rt_build_stencil(int32_t nOrder, double *w, double *phi, double *cx, double *cy, int32_t *Nc, double *Wc)
{
    return rt_build_stencil_ghost(nOrder, 0, w, phi, cx, cy, Nc, Wc);
}

int rt_build_stencil_ghost(int32_t nOrder, int32_t nGhost, double *w, double *phi, double *cx, double *cy, int32_t *Nc,
                           double *Wc)
{
    if (nOrder < 1 || nOrder > 4096 || nGhost < 0 || nGhost > 4096 || !w || !phi || !cx || !cy)
        return rt_fail(nullptr, RT_ERR_ARG, "rt_build_stencil: bad arguments");
    if ((Nc || Wc) && nOrder < 4) return rt_fail(nullptr, RT_ERR_ARG, "rt_build_stencil: interpolation tables need nOrder >= 4");
    std::string why;
    if (!Stencil::GhostLayoutValid(nOrder, nGhost, &why)) return rt_fail(nullptr, RT_ERR_ARG, "rt_build_stencil: " + why);
    Stencil s(nOrder, nGhost, Nc || Wc);
    for (int d = 0; d < nOrder + nGhost; d++) {
        w[d] = s.W(d);
        phi[d] = s.Phi(d);
        cx[d] = s.Cx(d);
        cy[d] = s.Cy(d);
    }
    if (Nc && Wc) {
        const InterpolationGrid &g = s.interpolationGrid;
        for (size_t i = 0; i < g.nGrid; i++)
            for (int m = 0; m < 4; m++) {
                Nc[4 * i + m] = (int32_t)g.neighbourIndexesCubic[i][m];
                Wc[4 * i + m] = g.neighbourWeightsCubic[i][m];
            }
    }
    return RT_OK;
}

int rt_build_metric(int32_t kind, double m, double a, int32_t nx, int32_t ny, double dx, double dy, double startx, double starty,
                    double *metric, double *alpha, double *t11, double *t12, double *t21, double *t22)
{
    if (nx < 3 || ny < 3 || !(dx > 0) || !(dy > 0)) return rt_fail(nullptr, RT_ERR_ARG, "rt_build_metric: bad grid");
    /* rebuild the reference Grid from the stored-grid description */
    Grid grid(nx - 2, ny - 2, Coord(startx + dx, starty + dy), Coord(startx + (nx - 2) * dx, starty + (ny - 2) * dy));
    grid.dx = dx;
    grid.dy = dy;
    grid.startx = startx;
    grid.starty = starty;
    grid.endx = startx + (nx - 1) * dx;
    grid.endy = starty + (ny - 1) * dy;
    Metric *M = nullptr;
    if (kind == 0) M = new Minkowski(grid, m, a);
    else if (kind == 1) M = new SchwarzSchild(grid, m, a);
    else if (kind == 2) M = new KerrSchild(grid, m, a);
    else return rt_fail(nullptr, RT_ERR_ARG, "rt_build_metric: unknown spacetime kind");
    if (metric) M->PackGeodesicBlock(metric);
    const size_t N = grid.nxy;
    for (size_t ij = 0; ij < N; ij++) {
        if (alpha) alpha[ij] = M->alpha[ij];
        if (t11) t11[ij] = M->tetrad11_ul[ij];
        if (t12) t12[ij] = M->tetrad12_ul[ij];
        if (t21) t21[ij] = M->tetrad21_ul[ij];
        if (t22) t22[ij] = M->tetrad22_ul[ij];
    }
    delete M;
    return RT_OK;
}

} /* extern "C" */
