/*
 * rt2d.h - C ABI of the MI355X (gfx950) stream-collide library for
 * 2dRadiation-style general-relativistic lattice-Boltzmann radiation transport.
 *
 * The reference (Tom-Olsen/2dRadiation) is a C++/OpenMP code whose hot path is
 * the per-step loop of Radiation::RunSimulation (src/Radiation.cpp:784-832):
 * StreamFlatFixed / StreamCurvedFixed, then ComputeMomentsIF, then Collide, with
 * the geodesic Fourier tables built once by UpdateFourierCoefficients.  It has
 * no FFI of its own; these entry points sit *under* its C++ methods (the C++
 * classes in 2dradiation_amd/cpp call them) and are what a ctypes/cgo/JNI
 * binding would call.  Plain pointers and sizes only.
 *
 * Conventions
 *   - Every entry point returns RT_OK (0) or a negative RT_ERR_* code; the
 *     message is available from rt_last_error(ctx) (ctx may be NULL for
 *     errors raised before a context exists).  The reference reports errors by
 *     ExitOnError (src/Utility.hh:16-20); the C++ layer maps non-zero codes to
 *     that behaviour.
 *   - Host arrays use the reference layouts: per-cell arrays [j*nx+i] over the
 *     stored grid including the 1-cell halo (src/Grid.cpp:44-47); intensities
 *     AoS I[ij*nDir+d] (src/Radiation.cpp:89-96); Fourier coefficients
 *     [ij*nCoeff+k] (src/Radiation.cpp:98-101).  Inputs are copied; the caller
 *     keeps ownership.  On the device intensities are SoA [d][ij].
 *   - One context is used by one host thread at a time; its work is ordered on
 *     its own HIP stream.  Separate contexts are independent.
 */
#ifndef RT2D_H
#define RT2D_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RT2D_ABI_VERSION 1

enum rt_status {
    RT_OK = 0,
    RT_ERR_ARG = -1,         /* invalid argument (sizes, null pointers, ranges) */
    RT_ERR_STATE = -2,       /* call out of order (e.g. step before fields are set) */
    RT_ERR_HIP = -3,         /* HIP runtime error */
    RT_ERR_OOM = -4,         /* device allocation failed */
    RT_ERR_UNSUPPORTED = -5, /* valid request this build does not implement */
    RT_ERR_NONFINITE = -6    /* NaN/Inf found in an input field */
};

enum rt_precision { RT_FP64 = 0, RT_FP32_MIXED = 1 };

/* Values mirror the reference's StreamingType (src/Config.hh:6-13). */
enum rt_streaming {
    RT_FLAT_FIXED = 0,
    RT_FLAT_ADAPTIVE = 1,
    RT_CURVED_FIXED = 2,
    RT_CURVED_ADAPTIVE = 3,
    RT_GEODESIC_FIXED = 4   /* not implemented: RT_ERR_UNSUPPORTED */
};

typedef struct rt_ctx rt_ctx;

/* Problem description.  Grid fields are those of the reference Grid
 * (src/Grid.cpp:4-37): nx, ny count stored cells including the halo. */
typedef struct rt_desc {
    int32_t nx, ny;
    double dx, dy, dt;
    double startx, starty;
    int32_t nDir;        /* intensity stencil size (Stencil::nDir) */
    int32_t nCoeff;      /* Fourier coefficients per series (streaming Stencil::nCoefficients) */
    int32_t precision;   /* rt_precision */
    int32_t streaming;   /* rt_streaming */
    double excision_r2;  /* Metric::InsideBH(xy) is x*x+y*y <= excision_r2; negative: never */
    int32_t device;      /* HIP device ordinal */
} rt_desc;

/* Number of metric arrays rt_build_fourier_tables takes, and their order in
 * the per-cell block metric[ij*RT_NMETRIC + q] (names as in src/Metric.h). */
#define RT_NMETRIC 29
enum rt_metric_index {
    RT_M_ALPHA = 0, RT_M_BETA1_U, RT_M_BETA2_U,
    RT_M_D1_ALPHA, RT_M_D2_ALPHA,
    RT_M_D1_BETA1_U, RT_M_D1_BETA2_U, RT_M_D2_BETA1_U, RT_M_D2_BETA2_U,
    RT_M_K11, RT_M_K12, RT_M_K22,
    RT_M_GAMMA11_UU, RT_M_GAMMA12_UU, RT_M_GAMMA22_UU,
    RT_M_D1_GAMMA11_LL, RT_M_D1_GAMMA12_LL, RT_M_D1_GAMMA22_LL,
    RT_M_D2_GAMMA11_LL, RT_M_D2_GAMMA12_LL, RT_M_D2_GAMMA22_LL,
    RT_M_TETRAD11, RT_M_TETRAD12, RT_M_TETRAD21, RT_M_TETRAD22,
    RT_M_TETRADINV11, RT_M_TETRADINV12, RT_M_TETRADINV21, RT_M_TETRADINV22
};

/* ---- lifetime / diagnostics ------------------------------------------ */
/* Replaces: construction of Radiation (src/Radiation.cpp:3-83); allocates
 * the device state for one problem. */
int rt_create(const rt_desc *desc, rt_ctx **out);
/* Replaces: Radiation::~Radiation (src/Radiation.cpp:84-87). */
void rt_destroy(rt_ctx *ctx);
/* Error text of the last failing call on ctx (or of rt_create if ctx is NULL). */
const char *rt_last_error(const rt_ctx *ctx);
/* Waits for all work queued on the context's stream. */
int rt_synchronize(rt_ctx *ctx);
int rt_abi_version(void);

/* ---- set-up ------------------------------------------------------------ */
/* Intensity stencil and its cubic angular interpolation tables.
 * Replaces: Stencil::{W,Phi,Cx,Cy} and InterpolationGrid::neighbourIndexesCubic /
 * neighbourWeightsCubic (src/Stencil.h:10-91, src/Stencil.cpp:4-162).
 * Nc, Wc have nGrid*4 entries. */
int rt_set_stencil(rt_ctx *ctx, const double *w, const double *phi, const double *cx,
                   const double *cy, int32_t nGrid, const int32_t *Nc, const double *Wc);
/* Streaming stencil used for the Fourier fit of the geodesic tables
 * (Stencil(nF,0,false), exe/main.cpp:547): nCoeff entries each. */
int rt_set_streaming_stencil(rt_ctx *ctx, const double *w, const double *phi,
                             const double *cx, const double *cy);
/* Cell-centre lapse and spatial tetrad (Metric::alpha, tetrad{11,12,21,22}_ul,
 * src/Metric.h:38, 60-62), each [N]. */
int rt_set_geometry(rt_ctx *ctx, const double *alpha, const double *t11, const double *t12,
                    const double *t21, const double *t22);
/* Geodesic Fourier coefficients computed elsewhere, each [N*nCoeff]
 * (Radiation::coefficients{S,X,Y,Cx,Cy}, src/Radiation.h:75-79). */
int rt_set_fourier_tables(rt_ctx *ctx, const double *cS, const double *cX, const double *cY,
                          const double *cCx, const double *cCy);
/* Builds the same tables on the GPU in FP64: per interior cell and streaming
 * direction an adaptive RK45 backtrace over dt, then the Fourier fit.
 * Replaces: Radiation::UpdateFourierCoefficients (src/Radiation.cpp:191-237)
 * with RK45_GeodesicEquation<-1> (src/GeodesicEquationSolver.cpp:78-188).
 * metric is the per-cell block [N*RT_NMETRIC]; rk45_tol is the velocity
 * truncation tolerance (the reference hard-codes 1e-4). */
int rt_build_fourier_tables(rt_ctx *ctx, const double *metric, double rk45_tol);
/* The same RK45 build from the metric block of the last
 * rt_build_fourier_tables, kept on the device: the per-step table update of
 * Config::updateFourierHarmonics (src/Radiation.cpp:790-791) without a host
 * transfer. */
int rt_rebuild_fourier_tables(rt_ctx *ctx, double rk45_tol);
/* Copies the current tables back to the host ([N*nCoeff] each). */
int rt_get_fourier_tables(rt_ctx *ctx, double *cS, double *cX, double *cY, double *cCx,
                          double *cCy);
/* Fluid fields (Radiation::eta, kappa0, kappa1, kappaA, ux, uy;
 * src/Radiation.h:63-68), each [N].  Rejects NaN/Inf. */
int rt_set_fields(rt_ctx *ctx, const double *eta, const double *kappa0, const double *kappa1,
                  const double *kappaA, const double *ux, const double *uy);

/* ---- state --------------------------------------------------------------- */
/* Host AoS FP64 intensities [N*nDir] in/out (Radiation::I). */
int rt_upload_I(rt_ctx *ctx, const double *I);
int rt_download_I(rt_ctx *ctx, double *I);
/* Inertial-frame moments E,Fx,Fy,Pxx,Pxy,Pyy (Radiation::E..Pyy), each [N]. */
int rt_get_moments(rt_ctx *ctx, double *E, double *Fx, double *Fy, double *Pxx, double *Pxy,
                   double *Pyy);
/* Source cells (Config::keepSourceNodesActive: LoadInitialData at the start
 * of every step, src/Radiation.cpp:790-791 and :124-178): rt_step puts the
 * given intensities (host AoS [N*nDir], read at cells with isSource[ij] != 0)
 * back into those cells before each step.  isSource == NULL removes them.
 * The adjoint and batches do not take source cells (RT_ERR_UNSUPPORTED). */
int rt_set_sources(rt_ctx *ctx, const uint8_t *isSource, const double *I);
/* Lab-frame moments (Radiation::ComputeMomentsLF, src/Radiation.cpp:270-301):
 * the inertial-frame energy-momentum tensor [[E,Fx,Fy],[Fx,Pxx,Pxy],
 * [Fy,Pxy,Pyy]] of the moments last computed (rt_compute_moments, or rt_step
 * with want_moments) transformed with the tetrad, T_LF[a][b] =
 * sum_AB T_IF[A][B] e[a][A] e[b][B] (TransformIFtoLF, src/SpecialMath.cpp:85-94),
 * and F = sqrt(|gamma_ij F^i F^j|) (Norm2, src/SpecialMath.cpp:20-27).
 * Cells inside the excision radius get 0.  rt_set_lab_frame takes the full
 * tetrad e[a][A] (Metric::tetradAB_ul, 9 arrays [N] in the order 00, 01, 02,
 * 10, ..., 22) and the spatial metric gamma_ll (11, 12, 22), each [N];
 * rt_get_moments_lf writes 7 arrays [N]. */
int rt_set_lab_frame(rt_ctx *ctx, const double *const tetrad[9], const double *gamma11,
                     const double *gamma12, const double *gamma22);
int rt_get_moments_lf(rt_ctx *ctx, double *E_LF, double *Fx_LF, double *Fy_LF, double *Pxx_LF,
                      double *Pxy_LF, double *Pyy_LF, double *F_LF);
/* Lambda-iteration statistics accumulated since the last reset: the maximum
 * per-cell cycle count and the sum of all per-cell counts
 * (Radiation::maxItterationCount / averageItterationCount numerator,
 * src/Radiation.cpp:726-733). */
int rt_iteration_stats(rt_ctx *ctx, int32_t *maxCount, double *sumCount);
int rt_reset_iteration_stats(rt_ctx *ctx);
/* Per-cell cycle counts of the last reference-order rt_collide
 * (Radiation::itterationCount, src/Radiation.cpp:715-724); count gets N
 * entries, 0 for halo and excised cells.  The fused rt_step keeps only the
 * aggregate statistics above. */
int rt_get_iteration_counts(rt_ctx *ctx, int32_t *count);

/* ---- hot path ------------------------------------------------------------ */
/* nSteps full steps.  FlatFixed and CurvedFixed: stream + moments + collide
 * fused in one kernel where rt_uses_fused_path says so, else the
 * reference-order kernels; FlatAdaptive and CurvedAdaptive: rotation update,
 * adaptive stream, rotated moments and collide (reference-order kernels in
 * FP64 contexts; FP32-mixed contexts stream in FP32 and collide in moment
 * form, the λ iteration on the moments of each cell's rotated stencil).
 * Replaces: the loop body of Radiation::RunSimulation (src/Radiation.cpp:784-832).
 * If want_moments is non-zero the moments of the last step are kept for
 * rt_get_moments (the reference-order kernels always keep them). */
int rt_step(rt_ctx *ctx, int32_t nSteps, int32_t want_moments);
/* Reference-order single methods (FP64 arithmetic, same operation order as
 * the reference), each followed by the reference's buffer swap (for adaptive
 * streaming rt_stream also swaps the two angle buffers).
 * Replace: Radiation::StreamFlatFixed / StreamFlatAdaptive / StreamCurvedFixed /
 * StreamCurvedAdaptive (:380-414, :415-453, :454-512, :513-575),
 * ComputeMomentsIF (:239-269) and Collide (:641-736). */
int rt_stream(rt_ctx *ctx);
int rt_compute_moments(rt_ctx *ctx);
int rt_collide(rt_ctx *ctx);

/* ---- adaptive streaming (FlatAdaptive, CurvedAdaptive) -------------------
 * Every cell carries a rotation of the stencil, rotationAngle, and the angle
 * of the next stream, rotationAngleNew (src/Radiation.h:42-43); both swap with
 * every stream.  rt_step runs UpdateRotationMatrizes, the adaptive stream,
 * ComputeMomentsIF and Collide in the rotated frames per step
 * (src/Radiation.cpp:801-819).  Angles must lie in [-pi/2, 5pi/2], where the
 * reference's MySin/MyCos are defined; a new context holds 0 in both. */
/* Replaces: writes to Radiation::rotationAngle / rotationAngleNew
 * (src/Radiation.cpp:55, 171, 177).  angleNew NULL: 0 everywhere. */
int rt_set_rotation(rt_ctx *ctx, const double *angle, const double *angleNew);
int rt_get_rotation(rt_ctx *ctx, double *angle, double *angleNew);
/* Replaces: Radiation::UpdateRotationMatrizes (src/Radiation.cpp:361-378),
 * from the moments the context holds (rt_set_moments, or the last step). */
int rt_update_rotation(rt_ctx *ctx);
/* Rotation angle LoadInitialData re-imposes on every source cell of
 * rt_set_sources before each step (src/Radiation.cpp:171, 177): angle[N]
 * read at the source cells, NaN for a cell whose angle is left alone (the
 * reference skips cells with a vanishing initial flux).  Call after
 * rt_set_sources; a new rt_set_sources clears it. */
int rt_set_source_rotation(rt_ctx *ctx, const double *angle);
/* Replaces: writes to Radiation::E, Fx, Fy, Pxx, Pxy, Pyy, which
 * UpdateRotationMatrizes reads (AverageF, :345-359).  NULL entries are left. */
int rt_set_moments(rt_ctx *ctx, const double *E, const double *Fx, const double *Fy,
                   const double *Pxx, const double *Pxy, const double *Pyy);
/* Which implementation rt_step uses: 1 fused kernel, 0 reference-order
 * kernels (for stencils or time steps the fused kernel does not cover). */
int rt_uses_fused_path(rt_ctx *ctx, int32_t *fused);

/* ---- gradients ------------------------------------------------------------
 * No counterpart in the reference, which has no adjoint.  Differentiates the
 * fused step (fluid velocity zero) with respect to the per-cell emissivity eta
 * and absorption kappa0, for a loss on bilinear samples of the final state.
 * Requires rt_uses_fused_path == 1. */
/* Runs nSteps steps like rt_step and keeps the state before steps 0, every,
 * 2*every, ... on the device for rt_adjoint (ceil(nSteps/every) states), with
 * the last segment recorded.  The segment buffers, the adjoint state and the
 * gradients rt_adjoint uses are allocated here; rt_adjoint itself allocates
 * only its point buffer, if no earlier rt_sample_outgoing sized it. */
int rt_forward_checkpointed(rt_ctx *ctx, int32_t nSteps, int32_t every);
/* Bilinear samples of the current intensities at npts points (x, y): out
 * [npts*nDir] gets I_d at the point for the directions pointing outward
 * (n_d . (x, y) > 0) and 0 for the others (an observer ring around the
 * origin). */
int rt_sample_outgoing(rt_ctx *ctx, int32_t npts, const double *x, const double *y, double *out);
/* rt_sample_outgoing without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * points and samples are staged in one fixed device array per GPU of
 * RT_NOALLOC_SCRATCH_DOUBLES doubles (1 MiB, held until the process ends),
 * so 2*npts + npts*nDir must not exceed it (RT_ERR_ARG otherwise).  Every
 * context on a GPU shares that array: calls that overlap on one GPU race
 * (calls in sequence, the usual case, do not: each returns after its
 * samples are back on the host).  Same outputs as rt_sample_outgoing. */
#define RT_NOALLOC_SCRATCH_DOUBLES (1 << 17)
int rt_sample_outgoing_noalloc(rt_ctx *ctx, int32_t npts, const double *x, const double *y, double *out);
/* Gradient of a loss L(samples) of the state after the last
 * rt_forward_checkpointed run: dL_dout [npts*nDir] is dL/d(out) of
 * rt_sample_outgoing at the same points.  Recomputes each checkpoint segment
 * and runs the adjoint steps backwards; writes dL/d eta and dL/d kappa0 ([N]
 * each, zero for halo and excised cells).  The state is left unchanged. */
int rt_adjoint(rt_ctx *ctx, int32_t npts, const double *x, const double *y, const double *dL_dout,
               double *grad_eta, double *grad_kappa0);

/* ---- batches of independent problems ----------------------------------------
 * No counterpart in the reference (one problem per process there).  Many
 * contexts, each fully set up (uneven grids, nDir, spacetimes allowed), are
 * advanced together: per step the FP32 members in one persistent launch that
 * drains a (problem, tile) queue (or one launch per (nDir, streaming) class,
 * rt_batch_set_queue), the FP64 members as rt_step steps them, two steps
 * captured in a HIP graph and replayed.  Members must be covered by the
 * fused step (rt_uses_fused_path == 1) and live on one device; while a batch
 * exists, each member can still be used on its own. */
typedef struct rt_batch rt_batch;
int rt_batch_create(rt_ctx *const *ctxs, int32_t n, rt_batch **out);
int rt_batch_destroy(rt_batch *b);
const char *rt_batch_last_error(const rt_batch *b);
/* Advances every member by nSteps steps; returns when the work is done. */
int rt_batch_step(rt_batch *b, int32_t nSteps);
/* As rt_batch_step, timed with HIP events on the batch's stream. */
int rt_batch_time_steps(rt_batch *b, int32_t nSteps, float *ms);
/* The FP32 members step in one persistent launch that takes (problem, tile)
 * items from a queue, largest tiles first (the default), or in one launch per
 * (nDir, streaming) class.  rt_batch_autotune times 2*trial_steps steps per
 * candidate (the members' states are restored afterwards): rows per tile of
 * the queue, and rows per workgroup of each class, and keeps the queue when
 * it is faster than the classes' launches together; chosen (optional, one
 * entry per class) gets the rows per tile or workgroup, 0 = the default rule.
 * No counterpart in the reference (a batching layer of this library). */
int rt_batch_autotune(rt_batch *b, int32_t trial_steps, int32_t *chosen);
/* *queue = 1 when the FP32 members step in the queue launch; set_queue
 * chooses (1) or leaves (0) it. */
int rt_batch_uses_queue(const rt_batch *b, int32_t *queue);
int rt_batch_set_queue(rt_batch *b, int32_t on);
int rt_batch_groups(const rt_batch *b, int32_t *ngroups);

/* ---- timing helper ------------------------------------------------------- */
/* Runs rt_step(ctx, nSteps, 0) bracketed by HIP events recorded on the
 * context's own stream and returns the elapsed milliseconds. */
int rt_time_steps(rt_ctx *ctx, int32_t nSteps, float *ms);

/* ---- set-up helpers for callers without the C++ classes ------------------ */
/* Stencil(nOrder, 0, isIntensityStencil) (src/Stencil.cpp:177-297):
 * w, phi, cx, cy get nOrder entries; if Nc/Wc are non-null they get the
 * InterpolationGrid(10*nOrder) cubic tables (nGrid = 10*nOrder, *4 each). */
/* The flux-to-sigma tables of Stencil(nOrder) (Stencil::PopulateLookUpTable,
 * src/Stencil.cpp:378-455): relative flux |F|/E of a von Mises distribution,
 * its concentration sigma and log-normalisation, increasing in the flux.
 * *count receives the number of entries; at most `capacity` are written. */
int rt_build_flux_tables(int32_t nOrder, int32_t capacity, double *relativeFlux, double *sigma,
                         double *normalization, int32_t *count);
int rt_build_stencil(int32_t nOrder, double *w, double *phi, double *cx, double *cy,
                     int32_t *Nc, double *Wc);
/* Stencil(nOrder, nGhost, true) with nGhost weight-0 ghost directions near the
 * stencil's forward direction (src/Stencil.cpp:177-297), sorted by angle
 * (SortDirections, :299-332): w, phi, cx, cy get nOrder+nGhost entries and
 * Nc/Wc the InterpolationGrid(10*nOrder) tables.  nGhost > 0 needs an even
 * nOrder; layouts whose ghost counts do not add up (the reference then
 * overruns its ghost array) are rejected. */
int rt_build_stencil_ghost(int32_t nOrder, int32_t nGhost, double *w, double *phi, double *cx,
                           double *cy, int32_t *Nc, double *Wc);
/* Metric arrays of Minkowski (kind 0), SchwarzSchild (1) or KerrSchild (2)
 * on a reference Grid(nx_, ny_, start, end) (src/Metric.cpp:126-336):
 * fills the per-cell block [N*RT_NMETRIC] (metric may be NULL) and, if non-null,
 * alpha and the spatial tetrad t11,t12,t21,t22 ([N] each). */
int rt_build_metric(int32_t kind, double m, double a, int32_t nx, int32_t ny, double dx,
                    double dy, double startx, double starty, double *metric, double *alpha,
                    double *t11, double *t12, double *t21, double *t22);

#ifdef __cplusplus
}
#endif
#endif
