/*
 * FP64 CPU adjoint of the stream-collide step: the oracle for the GPU
 * gradients of an observer-ring loss with respect to the per-cell emissivity
 * eta and absorption kappa0.
 *
 * TEST INFRASTRUCTURE ONLY (see rt_oracle.h).  The reference has no adjoint;
 * what is differentiated is the reference's forward step for fluid velocity
 * u = 0:
 *   J = Stream(I)                 StreamFlatFixed / StreamCurvedFixed
 *                                 (src/Radiation.cpp:380-414, 454-512), including
 *                                 the clamp >= 0 of IntensityAt (:342)
 *   Inew = Collide(J)             Collide (src/Radiation.cpp:641-736) with u = 0,
 *                                 where every lambda cycle is affine in the moments:
 *     v_0 = (E, Fx, Fy)(J),  v_k = q m(J) + a q (eta + k0 v_{k-1}[0]) S0
 *                                 + cc (v_{k-1}[1] Sx + v_{k-1}[2] Sy),
 *     Inew_d = q J_d + a q (eta + k0 E_{K-1}) + cc (Fx_{K-1} nx_d + Fy_{K-1} ny_d),
 *     q = 1 / (1 + a (kA + k0)),  cc = 3 k1 a q,  a = alpha dt,
 *   with K per cell the cycle count of the reference's own per-direction
 *   cycle on the same J (or_collide_cell).  Halo and excised
 *   cells keep their values (they are constants and carry no gradient).
 * The cycle count K is held fixed in the derivative (the rule is piecewise
 * constant).  Gradients: reverse mode through the moment recursion for J,
 * forward mode for the two per-cell parameters.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "rt_oracle.h"

#define ADJ_PI 3.14159265358979323846
#define OR_ADJ_MAXDIR 4096

/* One cell-direction gather written as a list of (state index, weight) taps
 * and per-corner clamps: out = scale * sum_c bw[c] * r3[c] * max(0, pre[c]),
 * pre[c] = sum_t wt[c][t] * I[idx[c][t]]. */
typedef struct {
    int ncorner, ntap;
    long idx[4][8];
    double wt[4][8];
    double bw[4], r3[4], pre[4];
    double scale;
    int zero; /* excised departure: out = 0 */
} adj_gather;

static double fourier_value(const double *coef, const double *B, int nF)
{
    double r = 0;
    for (int k = 0; k < nF; k++) r += coef[k] * B[k];
    return r;
}

static void describe_gather(const or_sim *s, const double *basis, const double *I, int i, int j, int d,
                            adj_gather *G)
{
    const or_grid *g = &s->grid;
    const int D = s->nDir, nx = g->nx;
    const long ij = i + (long)j * nx;
    memset(G, 0, sizeof *G);
    G->scale = 1.0;
    double iT, jT, vx = 0, vy = 0;
    if (s->streaming == 0) {
        double xT = g->startx + i * g->dx, yT = g->starty + j * g->dy;
        xT -= s->cx[d] * g->dt;
        yT -= s->cy[d] * g->dt;
        iT = (xT - g->startx) / g->dx;
        jT = (yT - g->starty) / g->dy;
    } else {
        const int nF = s->nCoeff;
        const double *B = basis + (long)d * nF;
        double sh = fourier_value(s->cS + ij * nF, B, nF);
        double xT = fourier_value(s->cX + ij * nF, B, nF);
        double yT = fourier_value(s->cY + ij * nF, B, nF);
        vx = fourier_value(s->cCx + ij * nF, B, nF);
        vy = fourier_value(s->cCy + ij * nF, B, nF);
        if (s->bh_r2 >= 0 && xT * xT + yT * yT <= s->bh_r2) {
            G->zero = 1;
            return;
        }
        iT = (xT - g->startx) / g->dx;
        jT = (yT - g->starty) / g->dy;
        G->scale = sh * sh * sh;
    }
    const long i0 = (long)floor(iT), j0 = (long)floor(jT);
    const double fx = iT - i0, fy = jT - j0;
    const long ci[4] = {i0, i0, i0 + 1, i0 + 1}, cj[4] = {j0, j0 + 1, j0, j0 + 1};
    const double bw[4] = {(1 - fx) * (1 - fy), (1 - fx) * fy, fx * (1 - fy), fx * fy};
    G->ncorner = 4;
    for (int c = 0; c < 4; c++) {
        const long k = ci[c] + cj[c] * nx;
        G->bw[c] = bw[c];
        if (s->streaming == 0) {
            G->ntap = 1;
            G->idx[c][0] = k * D + d;
            G->wt[c][0] = 1.0;
            G->r3[c] = 1.0;
            G->pre[c] = I[k * D + d];
        } else {
            const double r = s->alpha[ij] / s->alpha[k];
            G->r3[c] = r * (r * r);
            const double nG = (double)s->nGrid;
            double dd = nG * or_phi(vx, vy) / (2.0 * ADJ_PI);
            if (dd < 1e-8) dd = 1e-8;
            if (dd > nG - 1e-8) dd = nG - 1e-8;
            const long d0 = (long)floor(dd) % s->nGrid, d1 = (d0 + 1) % s->nGrid;
            const double t = dd - d0;
            G->ntap = 8;
            double pre = 0;
            for (int m = 0; m < 4; m++) {
                G->idx[c][m] = k * D + s->Nc[4 * d0 + m];
                G->wt[c][m] = (1.0 - t) * s->Wc[4 * d0 + m];
                G->idx[c][4 + m] = k * D + s->Nc[4 * d1 + m];
                G->wt[c][4 + m] = t * s->Wc[4 * d1 + m];
            }
            for (int m = 0; m < 8; m++) pre += G->wt[c][m] * I[G->idx[c][m]];
            G->pre[c] = pre;
        }
    }
}

static int cell_active(const or_sim *s, int i, int j)
{
    const or_grid *g = &s->grid;
    if (i < 1 || j < 1 || i > g->nx - 2 || j > g->ny - 2) return 0;
    const double x = g->startx + i * g->dx, y = g->starty + j * g->dy;
    return !(s->bh_r2 >= 0 && x * x + y * y <= s->bh_r2);
}

/* Stencil sums S0[X] = sum_d w X_d, Sx[X] = sum_d w nx X_d, Sy[X] = sum_d w ny X_d
 * for X in {1, nx, ny, nx^2, nx ny, ny^2}. */
static void stencil_sums(const or_sim *s, double S[3][6])
{
    memset(S, 0, sizeof(double) * 18);
    for (int d = 0; d < s->nDir; d++) {
        const double n1 = s->cx[d], n2 = s->cy[d], w = s->w[d];
        const double X[6] = {1.0, n1, n2, n1 * n1, n1 * n2, n2 * n2};
        for (int q = 0; q < 6; q++) {
            S[0][q] += w * X[q];
            S[1][q] += w * n1 * X[q];
            S[2][q] += w * n2 * X[q];
        }
    }
}

typedef struct {
    double q, cc, a, eta, k0, k1;
    double mJ[6];
    int K;
    double v[3];     /* (E, Fx, Fy) entering the last cycle */
    double te[3];    /* d v / d eta */
    double tk[3];    /* d v / d kappa0 */
} adj_cell;

/* Moment-space collide of one cell: K, the final moments entering the last
 * cycle and their tangents with respect to eta and kappa0. */
static void collide_cell(const or_sim *s, const double S[3][6], long ij, const double *J, adj_cell *c)
{
    const int D = s->nDir;
    const double a = s->alpha[ij] * s->grid.dt;
    c->a = a;
    c->eta = s->eta[ij];
    c->k0 = s->kappa0[ij];
    c->k1 = s->kappa1[ij];
    c->q = 1.0 / (1.0 + a * (s->kappaA[ij] + c->k0));
    c->cc = 3.0 * c->k1 * a * c->q;
    const double dq = -a * c->q * c->q, dcc = 3.0 * c->k1 * a * dq;
    double e = 0, fx = 0, fy = 0, pxx = 0, pxy = 0, pyy = 0;
    for (int d = 0; d < D; d++) { /* ComputeMomentsIF order, src/Radiation.cpp:250-262 */
        const double n1 = s->cx[d], n2 = s->cy[d], w = s->w[d] * J[ij * D + d];
        e += w;
        fx += w * n1;
        fy += w * n2;
        pxx += w * n1 * n1;
        pxy += w * n1 * n2;
        pyy += w * n2 * n2;
    }
    const double mJ[6] = {e, fx, fy, pxx, pxy, pyy};
    memcpy(c->mJ, mJ, sizeof mJ);
    /* the cycle count is the reference's: its per-direction cycle decides
     * it (a moment that is rounding noise, exactly 0 in some cells, stops
     * the loop through 0/0), and the recursion below runs that many cycles */
    double mref[6], Iscr[OR_ADJ_MAXDIR];
    memcpy(mref, mJ, sizeof mref);
    const int Kref = or_collide_cell(s, ij, J + ij * D, Iscr, mref);
    double prev[6], te_prev[3] = {0, 0, 0}, tk_prev[3] = {0, 0, 0};
    memcpy(prev, mJ, sizeof prev);
    double v[3] = {mJ[0], mJ[1], mJ[2]}, te[3] = {0, 0, 0}, tk[3] = {0, 0, 0};
    int K = 0;
    while (K < Kref) {
        K++;
        /* the moments entering this cycle and their tangents */
        v[0] = prev[0]; v[1] = prev[1]; v[2] = prev[2];
        memcpy(te, te_prev, sizeof te);
        memcpy(tk, tk_prev, sizeof tk);
        const double b = a * c->q * (c->eta + c->k0 * prev[0]);
        double nw[6];
        for (int X = 0; X < 6; X++) nw[X] = c->q * mJ[X] + b * S[0][X] + c->cc * (prev[1] * S[1][X] + prev[2] * S[2][X]);
        /* tangents of the new (E, Fx, Fy) */
        double ne[3], nk[3];
        for (int X = 0; X < 3; X++) {
            const double db_de = a * c->q * (1.0 + c->k0 * te[0]);
            const double db_dk = a * (dq * (c->eta + c->k0 * v[0]) + c->q * (v[0] + c->k0 * tk[0]));
            ne[X] = db_de * S[0][X] + c->cc * (te[1] * S[1][X] + te[2] * S[2][X]);
            nk[X] = dq * mJ[X] + db_dk * S[0][X] + dcc * (v[1] * S[1][X] + v[2] * S[2][X]) +
                    c->cc * (tk[1] * S[1][X] + tk[2] * S[2][X]);
        }
        memcpy(te_prev, ne, sizeof ne);
        memcpy(tk_prev, nk, sizeof nk);
        memcpy(prev, nw, sizeof prev);
    }
    c->K = K;
    memcpy(c->v, v, sizeof v);
    memcpy(c->te, te, sizeof te);
    memcpy(c->tk, tk, sizeof tk);
}

static void stream_all(const or_sim *s, const double *basis, const double *I, double *J)
{
    const or_grid *g = &s->grid;
    const int D = s->nDir;
    for (int j = 1; j < g->ny - 1; j++)
        for (int i = 1; i < g->nx - 1; i++) {
            const long ij = i + (long)j * g->nx;
            for (int d = 0; d < D; d++) {
                if (!cell_active(s, i, j)) {
                    J[ij * D + d] = 0;
                    continue;
                }
                adj_gather G;
                describe_gather(s, basis, I, i, j, d, &G);
                double out = 0;
                if (!G.zero)
                    for (int c = 0; c < G.ncorner; c++) out += G.bw[c] * G.r3[c] * (G.pre[c] > 0 || s->streaming == 0 ? G.pre[c] : 0.0);
                J[ij * D + d] = G.scale * out;
            }
        }
}

static double *make_basis(const or_sim *s)
{
    if (s->streaming != 2) return NULL;
    double *B = (double *)malloc(sizeof(double) * s->nDir * s->nCoeff);
    for (int d = 0; d < s->nDir; d++)
        for (int k = 0; k < s->nCoeff; k++) B[(long)d * s->nCoeff + k] = or_fourier_basis(k, s->phi[d]);
    return B;
}

static int has_velocity(const or_sim *s)
{
    const long N = (long)s->grid.nx * s->grid.ny;
    for (long k = 0; k < N; k++)
        if (s->ux[k] != 0.0 || s->uy[k] != 0.0) return 1;
    return 0;
}

/* One forward step I -> Inew (AoS), moment-space collide. */
static void forward_step(const or_sim *s, const double *basis, const double S[3][6], const double *I, double *J,
                         double *Inew)
{
    const or_grid *g = &s->grid;
    const int D = s->nDir;
    const long N = (long)g->nx * g->ny;
    stream_all(s, basis, I, J);
    memcpy(Inew, I, sizeof(double) * N * D);
    for (int j = 1; j < g->ny - 1; j++)
        for (int i = 1; i < g->nx - 1; i++) {
            if (!cell_active(s, i, j)) continue;
            const long ij = i + (long)j * g->nx;
            adj_cell c;
            collide_cell(s, S, ij, J, &c);
            const double b = c.a * c.q * (c.eta + c.k0 * c.v[0]);
            for (int d = 0; d < D; d++)
                Inew[ij * D + d] = c.q * J[ij * D + d] + b + c.cc * (c.v[1] * s->cx[d] + c.v[2] * s->cy[d]);
        }
}

/* nSteps forward steps from I0; states gets (nSteps+1) AoS states. */
int or_adj_forward(const or_sim *s, int32_t nSteps, const double *I0, double *states)
{
    if (has_velocity(s)) return -1;
    const long ND = (long)s->grid.nx * s->grid.ny * s->nDir;
    double S[3][6];
    stencil_sums(s, S);
    double *basis = make_basis(s);
    double *J = (double *)malloc(sizeof(double) * ND);
    memcpy(states, I0, sizeof(double) * ND);
    for (int n = 0; n < nSteps; n++) forward_step(s, basis, S, states + n * ND, J, states + (n + 1) * ND);
    free(J);
    free(basis);
    return 0;
}

/* Observer ring: bilinear samples of I_d at npts points for the directions
 * pointing outward (n_d . (x, y) > 0), 0 for the others.  obs is [npts][nDir].
 * Returns L = 1/2 sum (obs - target)^2; if dLdI is non-null it receives dL/dI
 * (AoS, overwritten). */
double or_observer_loss(const or_sim *s, const double *I, int32_t npts, const double *px, const double *py,
                        const double *target, double *obs, double *dLdI)
{
    const or_grid *g = &s->grid;
    const int D = s->nDir;
    const long ND = (long)g->nx * g->ny * D;
    if (dLdI) memset(dLdI, 0, sizeof(double) * ND);
    double L = 0;
    for (int p = 0; p < npts; p++) {
        const double iT = (px[p] - g->startx) / g->dx, jT = (py[p] - g->starty) / g->dy;
        const long i0 = (long)floor(iT), j0 = (long)floor(jT);
        const double fx = iT - i0, fy = jT - j0;
        const long k[4] = {i0 + j0 * g->nx, i0 + (j0 + 1) * g->nx, i0 + 1 + j0 * g->nx, i0 + 1 + (j0 + 1) * g->nx};
        const double bw[4] = {(1 - fx) * (1 - fy), (1 - fx) * fy, fx * (1 - fy), fx * fy};
        for (int d = 0; d < D; d++) {
            double o = 0;
            const int out = s->cx[d] * px[p] + s->cy[d] * py[p] > 0;
            if (out)
                for (int c = 0; c < 4; c++) o += bw[c] * I[k[c] * D + d];
            if (obs) obs[(long)p * D + d] = o;
            const double r = out ? o - target[(long)p * D + d] : 0.0;
            L += 0.5 * r * r;
            if (dLdI && out)
                for (int c = 0; c < 4; c++) dLdI[k[c] * D + d] += r * bw[c];
        }
    }
    return L;
}

/* Backward sweep over stored states (nSteps+1 AoS states from
 * or_adj_forward).  lamT = dL/dI at the last state (AoS); grad_eta and
 * grad_kappa0 ([N]) are overwritten with dL/d eta and dL/d kappa0. */
int or_adjoint_run(const or_sim *s, int32_t nSteps, const double *states, const double *lamT, double *grad_eta,
                   double *grad_kappa0)
{
    if (has_velocity(s)) return -1;
    const or_grid *g = &s->grid;
    const int D = s->nDir;
    const long N = (long)g->nx * g->ny, ND = N * D;
    double S[3][6];
    stencil_sums(s, S);
    double *basis = make_basis(s);
    double *J = (double *)malloc(sizeof(double) * ND);
    double *lam = (double *)malloc(sizeof(double) * ND);
    double *lamPrev = (double *)malloc(sizeof(double) * ND);
    memcpy(lam, lamT, sizeof(double) * ND);
    memset(grad_eta, 0, sizeof(double) * N);
    memset(grad_kappa0, 0, sizeof(double) * N);
    for (int n = nSteps - 1; n >= 0; n--) {
        const double *I = states + n * ND;
        stream_all(s, basis, I, J);
        memset(lamPrev, 0, sizeof(double) * ND);
        for (int j = 1; j < g->ny - 1; j++)
            for (int i = 1; i < g->nx - 1; i++) {
                if (!cell_active(s, i, j)) continue;
                const long ij = i + (long)j * g->nx;
                adj_cell c;
                collide_cell(s, S, ij, J, &c);
                double L0 = 0, Lx = 0, Ly = 0, LJ = 0;
                for (int d = 0; d < D; d++) {
                    const double l = lam[ij * D + d];
                    L0 += l;
                    Lx += l * s->cx[d];
                    Ly += l * s->cy[d];
                    LJ += l * J[ij * D + d];
                }
                const double q = c.q, a = c.a, k0 = c.k0, cc = c.cc;
                const double dq = -a * q * q, dcc = 3.0 * c.k1 * a * dq;
                grad_eta[ij] += a * q * (1.0 + k0 * c.te[0]) * L0 + cc * (c.te[1] * Lx + c.te[2] * Ly);
                grad_kappa0[ij] += dq * LJ + a * (dq * (c.eta + k0 * c.v[0]) + q * c.v[0] + q * k0 * c.tk[0]) * L0 +
                                   dcc * (c.v[1] * Lx + c.v[2] * Ly) + cc * (c.tk[1] * Lx + c.tk[2] * Ly);
                /* reverse through v_k = c0 + A v_{k-1}, A = [a q k0 S0 | cc Sx | cc Sy] (3x3) */
                double A[3][3];
                for (int X = 0; X < 3; X++) {
                    A[X][0] = a * q * k0 * S[0][X];
                    A[X][1] = cc * S[1][X];
                    A[X][2] = cc * S[2][X];
                }
                double u[3] = {a * q * k0 * L0, cc * Lx, cc * Ly}, Cbar[3] = {0, 0, 0};
                for (int k = c.K - 1; k >= 1; k--) {
                    double nu[3];
                    for (int X = 0; X < 3; X++) Cbar[X] += u[X];
                    for (int Y = 0; Y < 3; Y++) nu[Y] = A[0][Y] * u[0] + A[1][Y] * u[1] + A[2][Y] * u[2];
                    memcpy(u, nu, sizeof u);
                }
                const double gb[3] = {q * Cbar[0] + u[0], q * Cbar[1] + u[1], q * Cbar[2] + u[2]};
                /* J-bar, then the transpose of the gather */
                for (int d = 0; d < D; d++) {
                    const double jb = q * lam[ij * D + d] + s->w[d] * (gb[0] + gb[1] * s->cx[d] + gb[2] * s->cy[d]);
                    if (jb == 0.0) continue;
                    adj_gather G;
                    describe_gather(s, basis, I, i, j, d, &G);
                    if (G.zero) continue;
                    for (int cn = 0; cn < G.ncorner; cn++) {
                        if (s->streaming == 2 && !(G.pre[cn] > 0)) continue;
                        const double f = jb * G.scale * G.bw[cn] * G.r3[cn];
                        for (int t = 0; t < G.ntap; t++) lamPrev[G.idx[cn][t]] += f * G.wt[cn][t];
                    }
                }
            }
        /* fixed cells are constants: their adjoint is dropped */
        for (int j = 0; j < g->ny; j++)
            for (int i = 0; i < g->nx; i++)
                if (!cell_active(s, i, j))
                    for (int d = 0; d < D; d++) lamPrev[(i + (long)j * g->nx) * D + d] = 0;
        double *t = lam;
        lam = lamPrev;
        lamPrev = t;
    }
    free(J);
    free(lam);
    free(lamPrev);
    free(basis);
    return 0;
}
