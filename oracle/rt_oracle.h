/*
 * CPU oracle for the stream-collide path of 2dRadiation.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference
 * algorithm (Tom-Olsen/2dRadiation, src/Radiation.cpp, src/Stencil.cpp,
 * src/Metric.cpp, src/Spacetimes.cpp, src/GeodesicEquationSolver.cpp,
 * src/FourierHarmonics.cpp, src/Utility.hh).  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it, and only as the checker or the
 * CPU baseline.  The product library never links or calls it.
 *
 * Layouts follow the reference: intensities AoS I[ij*nDir + d], cell index
 * ij = i + j*nx (i fastest), Fourier coefficients [ij*nCoeff + k].
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Hastings-1955 polynomial trig (src/Utility.hh:133-360) ---- */
double or_mysin(double x);
double or_mycos(double x);
double or_myatan(double z);
double or_myatan2(double y, double x);
double or_phi(double x, double y); /* Tensor2::Phi, src/DataTypes.hh:166-170 */

/* ---- Grid (src/Grid.cpp:4-37) ---- */
typedef struct {
    int32_t nx, ny;            /* stored cells, including the 1-cell halo */
    double dx, dy, dt;
    double startx, starty, endx, endy;
} or_grid;
void or_grid_make(int32_t nx_, int32_t ny_, double sx, double sy, double ex, double ey,
                  double cfl, or_grid *g);

/* ---- Stencil (src/Stencil.cpp:177-297) and InterpolationGrid (:4-162) ---- */
void or_stencil(int32_t nOrder, double *w, double *phi, double *cx, double *cy);
void or_interp_grid(int32_t nDir, const double *phi, int32_t nGrid, int32_t *Nc, double *Wc);

/* ---- Fourier harmonics (src/FourierHarmonics.cpp:5-62) ---- */
double or_fourier_basis(int32_t k, double phi);

/* ---- Metric arrays (src/Metric.cpp:126-336, src/Spacetimes.cpp) ----
 * kind: 0 Minkowski, 1 SchwarzSchild, 2 KerrSchild.  out is [OR_NM][N]. */
enum {
    OR_G00L, OR_G01L, OR_G02L, OR_G11L, OR_G12L, OR_G22L,
    OR_G00U, OR_G01U, OR_G02U, OR_G11U, OR_G12U, OR_G22U,
    OR_ALPHA, OR_B1U, OR_B2U, OR_B1L, OR_B2L,
    OR_GA11L, OR_GA12L, OR_GA22L, OR_GA11U, OR_GA12U, OR_GA22U,
    OR_DA1, OR_DA2,
    OR_D1B1U, OR_D1B2U, OR_D2B1U, OR_D2B2U,
    OR_D1GA11L, OR_D1GA12L, OR_D1GA22L, OR_D2GA11L, OR_D2GA12L, OR_D2GA22L,
    OR_K11, OR_K12, OR_K22,
    OR_T00, OR_T01, OR_T02, OR_T10, OR_T11, OR_T12, OR_T20, OR_T21, OR_T22,
    OR_TI00, OR_TI01, OR_TI02, OR_TI10, OR_TI11, OR_TI12, OR_TI20, OR_TI21, OR_TI22,
    OR_NM
};
void or_metric_function(int32_t kind, double m, double a, double x, double y, double *g9);
void or_invert3(const double *m9, double *inv9);
void or_metric_build(const or_grid *g, int32_t kind, double m, double a, double *out);

/* ---- Geodesic tables (src/Radiation.cpp:191-237, GeodesicEquationSolver.cpp:78-188) ----
 * bh_r2 < 0 means nothing is excised.  Coefficient arrays are [N*nF]. */
double or_rk45(const or_grid *g, const double *metric, double bh_r2, double tol,
               double dt, double *x, double *y, double *vx, double *vy);
void or_geodesic_tables(const or_grid *g, const double *metric, int32_t nF, double bh_r2,
                        double tol, double *cS, double *cX, double *cY, double *cCx, double *cCy);

/* ---- Simulation (src/Radiation.cpp:239-343, 380-414, 454-512, 641-736, 738-832) ---- */
typedef struct {
    or_grid grid;
    int32_t nDir;
    const double *w, *phi, *cx, *cy;   /* [nDir] */
    int32_t nGrid;
    const int32_t *Nc;                 /* [nGrid*4] */
    const double *Wc;                  /* [nGrid*4] */
    int32_t streaming;                 /* 0 FlatFixed, 1 FlatAdaptive, 2 CurvedFixed, 3 CurvedAdaptive */
    int32_t nCoeff;
    const double *cS, *cX, *cY, *cCx, *cCy; /* [N*nCoeff], curved only */
    const double *alpha;               /* [N] */
    const double *t11, *t12, *t21, *t22; /* [N] spatial tetrad (ul) */
    double bh_r2;                      /* excision threshold on r^2, <0: none */
    const double *eta, *kappa0, *kappa1, *kappaA, *ux, *uy; /* [N] */
    /* rotationAngle / rotationAngleNew [N] (src/Radiation.h:42-43); rot NULL
     * means 0 everywhere, as for fixed streaming */
    const double *rot, *rotNew;
} or_sim;

/* Runs nSteps full steps (stream, moments, collide) in place on I (AoS).
 * moments (if non-null) receives E,Fx,Fy,Pxx,Pxy,Pyy [6][N] as the reference
 * holds them after the last step.  itcount (if non-null, [nSteps*N]) gets the
 * per-cell lambda-iteration counts of every step. */
int or_simulate(const or_sim *s, int32_t nSteps, double *I, double *moments,
                int32_t *maxIt, double *sumIt, int32_t *itcount);

/* Single reference methods, for fine-grained checks. Inew/I are AoS. */
void or_stream_flat(const or_sim *s, const double *I, double *Inew);
void or_stream_curved(const or_sim *s, const double *I, double *Inew);
void or_moments(const or_sim *s, const double *I, double *mom /*[6][N]*/);
void or_moments_lf(const or_grid *g, const double *mom /*[6][N]*/, const double *tetrad /*[9][N]*/,
                   const double *gamma /*[3][N]*/, double r2, double *out /*[7][N]*/);
/* Collide in place on the moment arrays; writes Inew for interior active cells. */
void or_collide(const or_sim *s, const double *I, double *Inew, double *mom, int32_t *count);
/* Collide of one active cell ij (src/Radiation.cpp:650-724): Icell/Inew_cell
 * are its nDir intensities, m its six moments (in: of Icell; out: after the
 * last cycle).  Returns the cycle count. */
int or_collide_cell(const or_sim *s, long ij, const double *Icell, double *Inew_cell, double m[6]);
/* Pieces of a cycle of Collide for fluid velocity u = (u1, u2) in the
 * inertial frame: the fluid-frame (E, Fx, Fy) of the moments m
 * (src/Radiation.cpp:664-672) and, for a direction n, (nFF1, nFF2, A)
 * (:689-698). */
void or_fluid_frame(double u1, double u2, const double m[6], double out[3]);
void or_fluid_dir(double u1, double u2, double n1, double n2, double out[3]);

/* ---- Adaptive streaming (rt_adaptive.c) ---- */
double or_myacos(double x);                                /* src/Utility.hh:453-471 */
void or_rotate(double angle, double x, double y, double *ox, double *oy); /* DataTypes.hh:630-635 */
/* Stencil(nOrder, nGhost) (src/Stencil.cpp:177-332), sorted by angle; nOrder+nGhost
 * entries each.  Returns 0, or -1 (bad sizes) / -2 (ghost counts do not add up). */
int or_stencil_ghost(int32_t nOrder, int32_t nGhost, double *w, double *phi, double *cx, double *cy);
/* UpdateRotationMatrizes (src/Radiation.cpp:345-378) */
void or_update_rotation(const or_sim *s, const double *mom, const double *rot, double *rotNew);
/* StreamFlatAdaptive / StreamCurvedAdaptive (:415-453, :513-575), using s->rot and s->rotNew */
void or_stream_flat_adaptive(const or_sim *s, const double *I, double *Inew);
void or_stream_curved_adaptive(const or_sim *s, const double *I, double *Inew);
/* nSteps adaptive steps (s->streaming 1 or 3).  mom [6][N], rot, rotNew [N]
 * are the reference's E..Pyy, rotationAngle, rotationAngleNew before the
 * loop and are updated in place. */
int or_simulate_adaptive(const or_sim *s, int32_t nSteps, double *I, double *mom, double *rot, double *rotNew,
                         int32_t *maxIt, double *sumIt, int32_t *itcount);

int32_t or_num_threads(void);
void or_set_threads(int32_t n);

#ifdef __cplusplus
}
#endif
#endif
