/*
 * Internal declarations shared by the gfx950 kernels and the C-ABI layer.
 */
#ifndef RT_INTERNAL_H
#define RT_INTERNAL_H
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include "../../include/rt2d.h"
#include "rt_math.h"

#define RT_MAX_FUSED_DIR 64
#define RT_FUSED_STRIP 64 /* cells per wavefront row segment */
/* Lambda-cycle statistics: RT_STAT_SLOTS (sum, max) pairs RT_STAT_STRIDE
 * words (one 128-byte line) apart.  Each wavefront adds its partial to the
 * pair its key selects, so the atomics of one launch spread over many
 * lines instead of queueing on one; the host reduces the pairs. */
#define RT_STAT_SLOTS 64
#define RT_STAT_STRIDE 16
__device__ inline unsigned long long *rt_stat_slot(unsigned long long *stats, unsigned key)
{
    return stats + (size_t)(key & (RT_STAT_SLOTS - 1)) * RT_STAT_STRIDE;
}

#define RT_FUSED_WAVES 4
/* LDS intensity rows of one grid row: directions -2..D+3 (mod D), so that
 * every 5-tap angular window [s0, s0+4] with s0 in [-2, D-1] (and the window
 * of fine angle nGrid, which equals angle 0) is contiguous; the window of
 * direction s0 starts at LDS row s0 + RT_RING_PAD */
#define RT_RING_PAD 2
#define RT_RING_ROWS(D) ((D) + 6)

/* Per-direction constants of the fused kernel, uploaded once per stencil.
 * Indexing by a wave-uniform direction lets the compiler use scalar loads. */
struct DirTables {
    double basis[RT_MAX_FUSED_DIR][8];  /* Fourier basis B_k(phi_d), k < nCoeff (<= 8) */
    float basisf[RT_MAX_FUSED_DIR][8];
    double w6[RT_MAX_FUSED_DIR][6];     /* w_d * {1, nx, ny, nx^2, nx ny, ny^2} */
    double cx[RT_MAX_FUSED_DIR], cy[RT_MAX_FUSED_DIR];
    float cxf[RT_MAX_FUSED_DIR], cyf[RT_MAX_FUSED_DIR];
    int flat_di[RT_MAX_FUSED_DIR], flat_dj[RT_MAX_FUSED_DIR]; /* floor of the flat departure offset */
    double flat_fx[RT_MAX_FUSED_DIR], flat_fy[RT_MAX_FUSED_DIR];
    float flat_fxf[RT_MAX_FUSED_DIR], flat_fyf[RT_MAX_FUSED_DIR];
    double S[3][6];      /* sum_d w X, sum_d w nx X, sum_d w ny X for X in the moment list */
    int woff[10];        /* window start offset per fine-angle residue */
    double wA[10][5], wD[10][5];   /* combined 5-tap angular weights: C = A + t*D */
    float wAf[10][5], wDf[10][5];
    /* FP32 row per residue for 16-byte LDS reads: wA[5], wD[5], woff (int bits), 0 */
    float wrowf[11][12];
    double ang_u;        /* nGrid / (2 pi) */
    float ang_uf;
    int quarter;         /* nGrid / 4 */
    int nGrid;
};

/* Per-cell inputs of the fused kernel, SoA [RT_CELL_FIELDS][N] in the kernel
 * precision. */
enum { CF_ETA = 0, CF_K0, CF_K1, CF_KA, CF_UX, CF_UY, CF_ADT, CF_IA3, CF_ACT, RT_CELL_FIELDS };
/* Cell mask bits. */
enum { MASK_ACTIVE = 1, MASK_NEAR_BH = 2 };

struct rt_ctx {
    rt_desc desc;
    int N = 0, D = 0, nCoeff = 0;
    int ncu = 0; /* compute units of the device */
    bool fp32 = false;
    hipStream_t stream = nullptr;
    std::string err;

    /* stencil */
    bool has_stencil = false, has_sstencil = false;
    double *d_w = nullptr, *d_phi = nullptr, *d_cx = nullptr, *d_cy = nullptr;
    int nGrid = 0;
    int32_t *d_Nc = nullptr;
    double *d_Wc = nullptr;
    double *d_basis = nullptr;    /* [D][nCoeff] Fourier basis at intensity angles (FP64) */
    double sw[16], sphi[16], scx[16], scy[16]; /* streaming stencil (host copy) */
    DirTables *d_dirt = nullptr;
    bool fused_tables_ok = false; /* periodic angular tables verified */

    /* geometry */
    bool has_geometry = false, has_tables = false;
    double *d_alpha = nullptr, *d_t11 = nullptr, *d_t12 = nullptr, *d_t21 = nullptr, *d_t22 = nullptr;
    double *d_coef = nullptr;     /* SoA [5 series][nCoeff][N], FP64 */
    double *d_metric = nullptr;   /* [N][RT_NMETRIC] block of the last table build, kept for rebuilds */
    /* the same coefficients in the fused kernel's form and precision: S - 1,
     * X and Y relative to the cell centre in cells, velocity as is */
    void *d_coefT = nullptr;
    /* FP32 fixed curved streaming: angular window of every (cell, direction)
     * from the FP64 series, [D][N] (k_angle_rec) */
    uint32_t *d_arec = nullptr;
    uint8_t *d_mask = nullptr;
    double max_disp = 0;          /* max |departure - cell| in cells (curved tables) */
    /* bit d set: the departure point of direction d (< 64) is excised */
    unsigned long long *d_excl = nullptr;
    unsigned long long *d_red = nullptr; /* one-word reduction target */

    /* fields */
    bool has_fields = false;
    double *d_f64 = nullptr;      /* [6][N] eta, kappa0, kappa1, kappaA, ux, uy */
    void *d_cell = nullptr;       /* [RT_CELL_FIELDS][N] in kernel precision */
    bool any_u = false;

    /* state */
    bool has_I = false;
    void *d_I[2] = {nullptr, nullptr};
    int cur = 0;
    /* FP64 fixed streaming: streamed intensities in true units [D][N], the
     * hand-over from the fused stream to the exact collide */
    double *d_J = nullptr;
    /* curved fused stepping keeps the state as alpha^-3 I in d_I[cur] */
    bool scaled = false;
    /* the fixed (halo and excised) cells of both state buffers are equal,
     * as the fused step needs; anything that writes them clears this */
    bool fixed_synced = false;
    /* CF_ACT of d_cell lacks the zero-start marks (launch_mark_zero_start):
     * set when the cell rows or the direction tables are rebuilt */
    bool z0_dirty = true;
    double *d_mom = nullptr;      /* [6][N] */
    /* lab frame: tetrad [9][N] and gamma_ll [3][N]; LF moments [7][N] */
    double *d_frame = nullptr;
    double *d_lf = nullptr;
    bool has_frame = false;
    /* source cells re-imposed before every step: cell indices and their
     * intensities SoA [nDir][nsrc], FP64 */
    int32_t *d_src_idx = nullptr;
    double *d_src_I = nullptr;
    int nsrc = 0, src_cap = 0;
    /* adaptive streaming: rotationAngle is d_rot[rcur], rotationAngleNew
     * d_rot[1 - rcur] (they swap with every stream, as in the reference);
     * d_rtrig [4][N] holds sin, cos of the new angle and sin, cos of the
     * inverse of the current angle of every cell for the next stream.
     * d_src_rot [nsrc]: angle re-imposed on each source cell (NaN: none). */
    double *d_rot[2] = {nullptr, nullptr};
    int rcur = 0;
    double *d_rtrig = nullptr;
    /* FP32-mixed adaptive contexts: FP32 copies of d_rtrig and of the cubic
     * weights, and (curved) d_coefT holding the relative-form series */
    float *d_rtrigf = nullptr;
    float *d_Wcf = nullptr;
    bool coefT_valid = false;
    /* weighted power sums of the unrotated intensity stencil, for the
     * adaptive collide's moment form: sum w, w cx, w cy, w cx^2, w cx cy,
     * w cy^2, w cx^3, w cx^2 cy, w cx cy^2, w cy^3 */
    double stencil_sums[10] = {};
    double *d_src_rot = nullptr;
    bool src_rot_set = false;
    int32_t *d_count = nullptr;   /* [N] per-method collide cycle counts */
    unsigned long long *d_stats = nullptr; /* RT_STAT_SLOTS (sum, max) pairs */
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    /* adjoint: checkpoints of the last rt_forward_checkpointed run (states
     * before steps 0, every, 2 every, ...), the recomputed segment, per-step
     * records, adjoint state buffers and the parameter gradients */
    bool ck_valid = false;
    /* state convention (alpha^-3 I or not) of the checkpointed run: the
     * checkpoints keep it even if a later call unscales the current state */
    bool ck_scaled = false;
    int ck_steps = 0, ck_every = 0, ck_count = 0, ck_cap = 0, seg_cap = 0;
    void *d_ckpt = nullptr;
    void *d_seg = nullptr;
    double *d_rec_m = nullptr;
    uint8_t *d_rec_k = nullptr;
    unsigned long long *d_rec_mask = nullptr; /* [every][8][N]: clamp bits per recording wavefront */
    /* the adjoint recomputes segment s-1 on stream2 while segment s's adjoint
     * steps run on `stream`: d_seg and the records hold two segments */
    hipStream_t stream2 = nullptr;
    /* segment buffer 0 holds the last segment of the latest
     * rt_forward_checkpointed run, recorded as it ran */
    bool seg_valid = false;
    hipEvent_t ev_seg[2] = {nullptr, nullptr}, ev_adj[2] = {nullptr, nullptr}, ev_go = nullptr;
    int rec_waves = RT_FUSED_WAVES;             /* wavefronts of the last recording launch (4 or 8) */
    void *d_lam[2] = {nullptr, nullptr};
    double *d_grad[2] = {nullptr, nullptr};
    double *d_pts = nullptr; /* sample points and their gradients */
    int pts_cap = 0;
    /* this device's address of the no-allocation staging array
     * (rt_sample_outgoing_noalloc) */
    double *d_scratch = nullptr;
};

int rt_fail(rt_ctx *c, int code, const std::string &msg);
#define TRY(x)                                                                                     \
    do {                                                                                           \
        int _r = (x);                                                                              \
        if (_r != RT_OK) return _r;                                                                \
    } while (0)
int rt_hip_check(rt_ctx *c, hipError_t e, const char *what);
#define RT_HIP(c, call)                                                                    \
    do {                                                                                   \
        hipError_t _e = (call);                                                            \
        if (_e != hipSuccess) return rt_hip_check((c), _e, #call);                         \
    } while (0)

/* kernel launchers (kernels_*.hip) */
int launch_transpose_in(rt_ctx *c, const double *aos, void *soa);
int launch_transpose_out(rt_ctx *c, const void *soa, double *aos);
int launch_copy_fixed_cells(rt_ctx *c, const void *src, void *dst);
/* Per-step records the adjoint needs from a forward step. */
struct RecordPtrs {
    double *m;                /* [6][N] moments of the streamed intensities */
    uint8_t *k;               /* [N] lambda cycles */
    unsigned long long *mask; /* [4][N] corner clamp bits (curved) */
};
int launch_fused_step_rec(rt_ctx *c, const void *in, void *out, const RecordPtrs &rec);
/* state <- I alpha^-3 (inverse = false) or I <- state alpha^3 (inverse = true)
 * for every cell with a finite positive lapse */
int launch_scale_state(rt_ctx *c, void *buf, bool inverse);
int launch_prepare_cells(rt_ctx *c);
int launch_mark_zero_start(rt_ctx *c);
int launch_ref_stream(rt_ctx *c, const void *in, void *out);
/* adaptive streaming (kernels_adaptive.hip): rotation update (or only the
 * sines and cosines), the rotated stream, source-cell angles */
int launch_rotation(rt_ctx *c, bool update);
int launch_stream_adaptive(rt_ctx *c, const void *in, void *out);
int launch_restore_source_rot(rt_ctx *c);
bool adaptive_f32_ready(const rt_ctx *c);
int launch_collide_adaptive(rt_ctx *c, const void *in, void *out);
inline bool rt_is_adaptive(const rt_desc &d) { return d.streaming == RT_FLAT_ADAPTIVE || d.streaming == RT_CURVED_ADAPTIVE; }
inline bool rt_is_curved(const rt_desc &d) { return d.streaming == RT_CURVED_FIXED || d.streaming == RT_CURVED_ADAPTIVE; }
int launch_ref_moments(rt_ctx *c, const void *in);
int launch_ref_collide(rt_ctx *c, const void *in, void *out);
int launch_fused_step(rt_ctx *c, const void *in, void *out, bool want_moments);
/* One step of the product path: FP32-mixed runs the fused step; FP64 runs
 * the fused stream into d_J and the exact collide (launch_collide_exact). */
int launch_product_step(rt_ctx *c, const void *in, void *out, bool want_moments);
/* ComputeMomentsIF + Collide in the reference's exact arithmetic on the
 * true-unit streamed intensities J; rec (adjoint recording) receives the
 * moments of J and the cycle counts. */
int launch_collide_exact(rt_ctx *c, const void *J, void *out, bool want_moments, const RecordPtrs *rec);
bool fused_supported(const rt_ctx *c);
/* curved state as alpha^-3 I, fixed cells equal in both buffers */
extern "C" int prepare_fused(rt_ctx *c);
int launch_geodesic_tables(rt_ctx *c, const double *d_metric, double tol);
int launch_departure_bound(rt_ctx *c, double *max_disp);
int launch_fill(rt_ctx *c, double *p, size_t n, double v);
int launch_coef_rel(rt_ctx *c);
int launch_moments_lf(rt_ctx *c);
int launch_restore_sources(rt_ctx *c, void *state, bool scaled);
int launch_adjoint_step(rt_ctx *c, const void *lam_in, void *lam_out, const void *Inext, const RecordPtrs &rec);
int launch_sample(rt_ctx *c, const void *I, int scaled, int npts, const double *px, const double *py, double *out);
int launch_sample_adjoint(rt_ctx *c, void *lam, int scaled, int npts, const double *px, const double *py,
                          const double *g);
#endif
