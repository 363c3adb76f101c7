// Host-only stand-in for the C ABI calls the C++ classes make (include/rt2d.h),
// for memory checks of those classes on a machine without a GPU.  Each entry
// point validates its arguments as the library does, reads every element of
// its input buffers and writes every element of its output buffers, at the
// sizes the header documents (N = nx*ny stored cells, nDir, nCoeff).  It
// computes nothing: results are zeros.  Built with AddressSanitizer and
// UBSan together with the classes and a user program (Makefile target
// `asan`), it makes every too-small caller buffer and every heap misuse in
// the host code, teardown included, a reported error.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include "rt2d.h"

struct rt_ctx {
    rt_desc d;
    size_t N, D, F;
    std::string err;
    volatile double sink = 0;
};

namespace {
thread_local std::string g_err;
void take(rt_ctx *c, const double *p, size_t n)
{
    double s = 0;
    for (size_t k = 0; k < n; k++) s += p[k];
    c->sink = c->sink + s;
}
void give(double *p, size_t n)
{
    for (size_t k = 0; k < n; k++) p[k] = 0.0;
}
int bad(rt_ctx *c, const char *msg)
{
    if (c) c->err = msg;
    else g_err = msg;
    return RT_ERR_ARG;
}
} // namespace

extern "C" {

int rt_abi_version(void) { return RT2D_ABI_VERSION; }
const char *rt_last_error(const rt_ctx *c) { return c ? c->err.c_str() : g_err.c_str(); }

int rt_create(const rt_desc *desc, rt_ctx **out)
{
    if (!desc || !out) return bad(nullptr, "rt_create: null argument");
    if (desc->nx < 3 || desc->ny < 3 || desc->nDir < 4 || desc->nCoeff < 1 || desc->nCoeff > 8)
        return bad(nullptr, "rt_create: bad sizes");
    rt_ctx *c = new rt_ctx();
    c->d = *desc;
    c->N = (size_t)desc->nx * desc->ny;
    c->D = (size_t)desc->nDir;
    c->F = (size_t)desc->nCoeff;
    *out = c;
    return RT_OK;
}
void rt_destroy(rt_ctx *c) { delete c; }
int rt_synchronize(rt_ctx *c) { return c ? RT_OK : bad(nullptr, "null context"); }

int rt_set_stencil(rt_ctx *c, const double *w, const double *phi, const double *cx, const double *cy, int32_t nGrid,
                   const int32_t *Nc, const double *Wc)
{
    if (!c || !w || !phi || !cx || !cy || nGrid < 0) return bad(c, "rt_set_stencil: bad argument");
    for (const double *p : {w, phi, cx, cy}) take(c, p, c->D);
    if (nGrid > 0) {
        if (!Nc || !Wc) return bad(c, "rt_set_stencil: tables missing");
        for (int32_t k = 0; k < 4 * nGrid; k++) {
            if (Nc[k] < 0 || (size_t)Nc[k] >= c->D) return bad(c, "rt_set_stencil: neighbour index out of range");
        }
        take(c, Wc, 4 * (size_t)nGrid);
    }
    return RT_OK;
}
int rt_set_streaming_stencil(rt_ctx *c, const double *w, const double *phi, const double *cx, const double *cy)
{
    if (!c || !w || !phi || !cx || !cy) return bad(c, "rt_set_streaming_stencil: null argument");
    for (const double *p : {w, phi, cx, cy}) take(c, p, c->F);
    return RT_OK;
}
int rt_set_geometry(rt_ctx *c, const double *alpha, const double *t11, const double *t12, const double *t21,
                    const double *t22)
{
    if (!c || !alpha || !t11 || !t12 || !t21 || !t22) return bad(c, "rt_set_geometry: null argument");
    for (const double *p : {alpha, t11, t12, t21, t22}) take(c, p, c->N);
    return RT_OK;
}
int rt_set_lab_frame(rt_ctx *c, const double *const tetrad[9], const double *g11, const double *g12,
                     const double *g22)
{
    if (!c || !tetrad || !g11 || !g12 || !g22) return bad(c, "rt_set_lab_frame: null argument");
    for (int k = 0; k < 9; k++) {
        if (!tetrad[k]) return bad(c, "rt_set_lab_frame: null tetrad array");
        take(c, tetrad[k], c->N);
    }
    for (const double *p : {g11, g12, g22}) take(c, p, c->N);
    return RT_OK;
}
int rt_build_fourier_tables(rt_ctx *c, const double *metric, double tol)
{
    if (!c || !metric || !(tol > 0)) return bad(c, "rt_build_fourier_tables: bad argument");
    take(c, metric, c->N * RT_NMETRIC);
    return RT_OK;
}
int rt_rebuild_fourier_tables(rt_ctx *c, double tol) { return c && tol > 0 ? RT_OK : bad(c, "bad argument"); }
int rt_get_fourier_tables(rt_ctx *c, double *cS, double *cX, double *cY, double *cCx, double *cCy)
{
    if (!c || !cS || !cX || !cY || !cCx || !cCy) return bad(c, "rt_get_fourier_tables: null argument");
    for (double *p : {cS, cX, cY, cCx, cCy}) give(p, c->N * c->F);
    return RT_OK;
}
int rt_set_fields(rt_ctx *c, const double *eta, const double *k0, const double *k1, const double *kA, const double *ux,
                  const double *uy)
{
    if (!c) return bad(c, "null context");
    for (const double *p : {eta, k0, k1, kA, ux, uy})
        if (p) take(c, p, c->N);
    return RT_OK;
}
int rt_upload_I(rt_ctx *c, const double *I)
{
    if (!c || !I) return bad(c, "rt_upload_I: null argument");
    take(c, I, c->N * c->D);
    return RT_OK;
}
int rt_download_I(rt_ctx *c, double *I)
{
    if (!c || !I) return bad(c, "rt_download_I: null argument");
    give(I, c->N * c->D);
    return RT_OK;
}
int rt_get_moments(rt_ctx *c, double *E, double *Fx, double *Fy, double *Pxx, double *Pxy, double *Pyy)
{
    if (!c) return bad(c, "null context");
    for (double *p : {E, Fx, Fy, Pxx, Pxy, Pyy})
        if (p) give(p, c->N);
    return RT_OK;
}
int rt_set_moments(rt_ctx *c, const double *E, const double *Fx, const double *Fy, const double *Pxx,
                   const double *Pxy, const double *Pyy)
{
    if (!c) return bad(c, "null context");
    for (const double *p : {E, Fx, Fy, Pxx, Pxy, Pyy})
        if (p) take(c, p, c->N);
    return RT_OK;
}
int rt_get_moments_lf(rt_ctx *c, double *E, double *Fx, double *Fy, double *Pxx, double *Pxy, double *Pyy,
                      double *F)
{
    if (!c) return bad(c, "null context");
    for (double *p : {E, Fx, Fy, Pxx, Pxy, Pyy, F})
        if (p) give(p, c->N);
    return RT_OK;
}
int rt_set_sources(rt_ctx *c, const uint8_t *isSource, const double *I)
{
    if (!c) return bad(c, "null context");
    if (!isSource) return RT_OK; /* clears the source cells */
    if (!I) return bad(c, "rt_set_sources: null intensities");
    size_t n = 0;
    for (size_t k = 0; k < c->N; k++) n += isSource[k] ? 1 : 0;
    take(c, I, c->N * c->D);
    c->sink = c->sink + (double)n;
    return RT_OK;
}
int rt_set_source_rotation(rt_ctx *c, const double *angle)
{
    if (!c || !angle) return bad(c, "rt_set_source_rotation: null argument");
    take(c, angle, c->N);
    return RT_OK;
}
int rt_set_rotation(rt_ctx *c, const double *angle, const double *angleNew)
{
    if (!c || !angle) return bad(c, "rt_set_rotation: null argument");
    take(c, angle, c->N);
    if (angleNew) take(c, angleNew, c->N);
    return RT_OK;
}
int rt_get_rotation(rt_ctx *c, double *angle, double *angleNew)
{
    if (!c) return bad(c, "null context");
    if (angle) give(angle, c->N);
    if (angleNew) give(angleNew, c->N);
    return RT_OK;
}
int rt_iteration_stats(rt_ctx *c, int32_t *mx, double *sum)
{
    if (!c) return bad(c, "null context");
    if (mx) *mx = 0;
    if (sum) *sum = 0;
    return RT_OK;
}
int rt_reset_iteration_stats(rt_ctx *c) { return c ? RT_OK : bad(c, "null context"); }
int rt_get_iteration_counts(rt_ctx *c, int32_t *count)
{
    if (!c || !count) return bad(c, "rt_get_iteration_counts: null argument");
    for (size_t k = 0; k < c->N; k++) count[k] = 0;
    return RT_OK;
}
int rt_step(rt_ctx *c, int32_t nSteps, int32_t) { return c && nSteps >= 0 ? RT_OK : bad(c, "rt_step: bad argument"); }
int rt_stream(rt_ctx *c) { return c ? RT_OK : bad(c, "null context"); }
int rt_compute_moments(rt_ctx *c) { return c ? RT_OK : bad(c, "null context"); }
int rt_collide(rt_ctx *c) { return c ? RT_OK : bad(c, "null context"); }
int rt_update_rotation(rt_ctx *c) { return c ? RT_OK : bad(c, "null context"); }

} // extern "C"
