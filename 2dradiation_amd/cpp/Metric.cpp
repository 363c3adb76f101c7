#include "Metric.h"
#include "Parallel.h"
#include "../../include/rt2d.h"

Metric::Metric(Grid &grid_, double m_, double a_) : grid(grid_), m(m_), a(a_)
{
    DoubleBuffer *all[] = {
        &g00_ll, &g01_ll, &g02_ll, &g11_ll, &g12_ll, &g22_ll, &g00_uu, &g01_uu, &g02_uu, &g11_uu,
        &g12_uu, &g22_uu, &d1_g00_lll, &d1_g01_lll, &d1_g02_lll, &d1_g11_lll, &d1_g12_lll,
        &d1_g22_lll, &d2_g00_lll, &d2_g01_lll, &d2_g02_lll, &d2_g11_lll, &d2_g12_lll, &d2_g22_lll,
        &d1_g00_luu, &d1_g01_luu, &d1_g02_luu, &d1_g11_luu, &d1_g12_luu, &d1_g22_luu, &d2_g00_luu,
        &d2_g01_luu, &d2_g02_luu, &d2_g11_luu, &d2_g12_luu, &d2_g22_luu, &alpha, &beta1_u,
        &beta2_u, &beta1_l, &beta2_l, &gamma11_ll, &gamma12_ll, &gamma22_ll, &gamma11_uu,
        &gamma12_uu, &gamma22_uu, &d1_alpha_l, &d2_alpha_l, &d1_beta1_lu, &d1_beta2_lu,
        &d2_beta1_lu, &d2_beta2_lu, &d1_beta1_ll, &d1_beta2_ll, &d2_beta1_ll, &d2_beta2_ll,
        &d1_gamma11_lll, &d1_gamma12_lll, &d1_gamma22_lll, &d2_gamma11_lll, &d2_gamma12_lll,
        &d2_gamma22_lll, &d1_gamma11_luu, &d1_gamma12_luu, &d1_gamma22_luu, &d2_gamma11_luu,
        &d2_gamma12_luu, &d2_gamma22_luu, &K11_ll, &K12_ll, &K22_ll, &tetrad00_ul, &tetrad01_ul,
        &tetrad02_ul, &tetrad10_ul, &tetrad11_ul, &tetrad12_ul, &tetrad20_ul, &tetrad21_ul,
        &tetrad22_ul, &tetrad00_lu, &tetrad01_lu, &tetrad02_lu, &tetrad10_lu, &tetrad11_lu,
        &tetrad12_lu, &tetrad20_lu, &tetrad21_lu, &tetrad22_lu};
    for (DoubleBuffer *b : all) b->assign(grid.nxy, 0.0);
}

void Metric::Initialize()
{
    InitializeMetricOnGrid();
    InitializeMetricDerivativesOnGrid();
    InitializeAdmComponentsOnGrid();
    InitializeBoostedTetradOnGrid();
}

// Reference src/Metric.cpp:126-150.
void Metric::InitializeMetricOnGrid()
{
    ParallelFor(grid.nxy, [this](size_t ij) {
        Tensor3x3 gl = MetricFunction(grid.xy(ij));
        Tensor3x3 gu = gl.Invert();
        g00_ll[ij] = gl[{0, 0}]; g01_ll[ij] = gl[{0, 1}]; g02_ll[ij] = gl[{0, 2}];
        g11_ll[ij] = gl[{1, 1}]; g12_ll[ij] = gl[{1, 2}]; g22_ll[ij] = gl[{2, 2}];
        g00_uu[ij] = gu[{0, 0}]; g01_uu[ij] = gu[{0, 1}]; g02_uu[ij] = gu[{0, 2}];
        g11_uu[ij] = gu[{1, 1}]; g12_uu[ij] = gu[{1, 2}]; g22_uu[ij] = gu[{2, 2}];
    });
}

// Fourth-order central differences with step 1e-8 of g_ll and of its inverse
// (reference src/Metric.cpp:189-274).
void Metric::InitializeMetricDerivativesOnGrid()
{
    const double dk = 1e-8;
    auto deriv = [this, dk](const Coord &xy, int k, bool inverse) {
        double ex = (k == 1), ey = (k == 2);
        Tensor3x3 g[4];
        const double off[4] = {-2, -1, 1, 2};
        for (int q = 0; q < 4; q++) {
            g[q] = MetricFunction(Coord(xy[1] + off[q] * dk * ex, xy[2] + off[q] * dk * ey));
            if (inverse) g[q] = g[q].Invert();
        }
        Tensor3x3 r;
        for (int q = 0; q < 9; q++)
            r.data[q] = (1.0 / 12.0 * g[0].data[q] - 1 / 12.0 * g[3].data[q] -
                         2.0 / 3.0 * g[1].data[q] + 2.0 / 3.0 * g[2].data[q]) / dk;
        return r;
    };
    ParallelFor(grid.nxy, [&](size_t ij) {
        Coord xy = grid.xy(ij);
        Tensor3x3 xl = deriv(xy, 1, false), yl = deriv(xy, 2, false);
        Tensor3x3 xu = deriv(xy, 1, true), yu = deriv(xy, 2, true);
        d1_g00_lll[ij] = xl[{0, 0}]; d1_g01_lll[ij] = xl[{0, 1}]; d1_g02_lll[ij] = xl[{0, 2}];
        d1_g11_lll[ij] = xl[{1, 1}]; d1_g12_lll[ij] = xl[{1, 2}]; d1_g22_lll[ij] = xl[{2, 2}];
        d2_g00_lll[ij] = yl[{0, 0}]; d2_g01_lll[ij] = yl[{0, 1}]; d2_g02_lll[ij] = yl[{0, 2}];
        d2_g11_lll[ij] = yl[{1, 1}]; d2_g12_lll[ij] = yl[{1, 2}]; d2_g22_lll[ij] = yl[{2, 2}];
        d1_g00_luu[ij] = xu[{0, 0}]; d1_g01_luu[ij] = xu[{0, 1}]; d1_g02_luu[ij] = xu[{0, 2}];
        d1_g11_luu[ij] = xu[{1, 1}]; d1_g12_luu[ij] = xu[{1, 2}]; d1_g22_luu[ij] = xu[{2, 2}];
        d2_g00_luu[ij] = yu[{0, 0}]; d2_g01_luu[ij] = yu[{0, 1}]; d2_g02_luu[ij] = yu[{0, 2}];
        d2_g11_luu[ij] = yu[{1, 1}]; d2_g12_luu[ij] = yu[{1, 2}]; d2_g22_luu[ij] = yu[{2, 2}];
    });
}

// 2+1 split: lapse, shift, spatial metric, their derivatives and the
// extrinsic curvature (reference src/Metric.cpp:276-336).
void Metric::InitializeAdmComponentsOnGrid()
{
    ParallelFor(grid.nxy, [this](size_t ij) {
        double al = 1.0 / std::sqrt(-g00_uu[ij]);
        alpha[ij] = al;
        beta1_u[ij] = al * al * g01_uu[ij];
        beta2_u[ij] = al * al * g02_uu[ij];
        beta1_l[ij] = g01_ll[ij];
        beta2_l[ij] = g02_ll[ij];
        gamma11_ll[ij] = g11_ll[ij];
        gamma12_ll[ij] = g12_ll[ij];
        gamma22_ll[ij] = g22_ll[ij];
        gamma11_uu[ij] = g11_uu[ij] + beta1_u[ij] * beta1_u[ij] / (al * al);
        gamma12_uu[ij] = g12_uu[ij] + beta1_u[ij] * beta2_u[ij] / (al * al);
        gamma22_uu[ij] = g22_uu[ij] + beta2_u[ij] * beta2_u[ij] / (al * al);
        d1_alpha_l[ij] = d1_g00_luu[ij] / (2.0 * std::pow(-g00_uu[ij], 3.0 / 2.0));
        d2_alpha_l[ij] = d2_g00_luu[ij] / (2.0 * std::pow(-g00_uu[ij], 3.0 / 2.0));
        d1_beta1_lu[ij] = al * al * d1_g01_luu[ij] + 2.0 * g01_uu[ij] * al * d1_alpha_l[ij];
        d2_beta1_lu[ij] = al * al * d2_g01_luu[ij] + 2.0 * g01_uu[ij] * al * d2_alpha_l[ij];
        d1_beta2_lu[ij] = al * al * d1_g02_luu[ij] + 2.0 * g02_uu[ij] * al * d1_alpha_l[ij];
        d2_beta2_lu[ij] = al * al * d2_g02_luu[ij] + 2.0 * g02_uu[ij] * al * d2_alpha_l[ij];
        d1_beta1_ll[ij] = d1_g01_lll[ij];
        d1_beta2_ll[ij] = d1_g02_lll[ij];
        d2_beta1_ll[ij] = d2_g01_lll[ij];
        d2_beta2_ll[ij] = d2_g02_lll[ij];
        d1_gamma11_lll[ij] = d1_g11_lll[ij];
        d2_gamma11_lll[ij] = d2_g11_lll[ij];
        d1_gamma12_lll[ij] = d1_g12_lll[ij];
        d2_gamma12_lll[ij] = d2_g12_lll[ij];
        d1_gamma22_lll[ij] = d1_g22_lll[ij];
        d2_gamma22_lll[ij] = d2_g22_lll[ij];
        double a3 = std::pow(al, 3);
        auto dgu = [&](double dg, double dbA, double bA, double bB, double dbB, double da) {
            return dg + (dbA * bB + bA * dbB) / (al * al) - 2.0 * bA * bB * da / a3;
        };
        d1_gamma11_luu[ij] = dgu(d1_g11_luu[ij], d1_beta1_lu[ij], beta1_u[ij], beta1_u[ij], d1_beta1_lu[ij], d1_alpha_l[ij]);
        d1_gamma12_luu[ij] = dgu(d1_g12_luu[ij], d1_beta1_lu[ij], beta1_u[ij], beta2_u[ij], d1_beta2_lu[ij], d1_alpha_l[ij]);
        d1_gamma22_luu[ij] = dgu(d1_g22_luu[ij], d1_beta2_lu[ij], beta2_u[ij], beta2_u[ij], d1_beta2_lu[ij], d1_alpha_l[ij]);
        d2_gamma11_luu[ij] = dgu(d2_g11_luu[ij], d2_beta1_lu[ij], beta1_u[ij], beta1_u[ij], d2_beta1_lu[ij], d2_alpha_l[ij]);
        d2_gamma12_luu[ij] = dgu(d2_g12_luu[ij], d2_beta1_lu[ij], beta1_u[ij], beta2_u[ij], d2_beta2_lu[ij], d2_alpha_l[ij]);
        d2_gamma22_luu[ij] = dgu(d2_g22_luu[ij], d2_beta2_lu[ij], beta2_u[ij], beta2_u[ij], d2_beta2_lu[ij], d2_alpha_l[ij]);
        // K_ab = (2 gamma_ac d_b beta^c + d_c gamma_ab beta^c) / (2 alpha), static metric.
        const double gam[2][2] = {{gamma11_ll[ij], gamma12_ll[ij]}, {gamma12_ll[ij], gamma22_ll[ij]}};
        const double dbu[2][2] = {{d1_beta1_lu[ij], d1_beta2_lu[ij]}, {d2_beta1_lu[ij], d2_beta2_lu[ij]}};
        const double dg[2][2][2] = {{{d1_gamma11_lll[ij], d1_gamma12_lll[ij]}, {d1_gamma12_lll[ij], d1_gamma22_lll[ij]}},
                                    {{d2_gamma11_lll[ij], d2_gamma12_lll[ij]}, {d2_gamma12_lll[ij], d2_gamma22_lll[ij]}}};
        const double bu[2] = {beta1_u[ij], beta2_u[ij]};
        double K[2][2];
        for (int p = 0; p < 2; p++)
            for (int q = 0; q < 2; q++) {
                double k = -0.0;
                for (int c = 0; c < 2; c++) k += 2.0 * gam[p][c] * dbu[q][c] + dg[c][p][q] * bu[c];
                K[p][q] = k / (2.0 * al);
            }
        K11_ll[ij] = K[0][0];
        K12_ll[ij] = K[0][1];
        K22_ll[ij] = K[1][1];
    });
}

// Tetrad from the Eulerian observer and a Gram-Schmidt-like spatial basis
// (reference src/Metric.cpp:152-187), and its inverse.
void Metric::InitializeBoostedTetradOnGrid()
{
    ParallelFor(grid.nxy, [this](size_t ij) {
        double al = alpha[ij];
        double aa = 1.0 / std::sqrt(gamma11_uu[ij]);
        double bb = aa * aa * gamma12_uu[ij];
        double m11 = 1 / aa, m21 = bb / aa, m22 = 1;
        double g2 = m22 * gamma22_ll[ij] * m22;
        Tensor3x3 t(1.0 / al, 0, 0, -beta1_u[ij] / al, m11, 0, -beta2_u[ij] / al, m21, 1 / std::sqrt(g2));
        Tensor3x3 ti = t.Invert();
        tetrad00_ul[ij] = t.data[0]; tetrad01_ul[ij] = t.data[1]; tetrad02_ul[ij] = t.data[2];
        tetrad10_ul[ij] = t.data[3]; tetrad11_ul[ij] = t.data[4]; tetrad12_ul[ij] = t.data[5];
        tetrad20_ul[ij] = t.data[6]; tetrad21_ul[ij] = t.data[7]; tetrad22_ul[ij] = t.data[8];
        tetrad00_lu[ij] = ti.data[0]; tetrad01_lu[ij] = ti.data[1]; tetrad02_lu[ij] = ti.data[2];
        tetrad10_lu[ij] = ti.data[3]; tetrad11_lu[ij] = ti.data[4]; tetrad12_lu[ij] = ti.data[5];
        tetrad20_lu[ij] = ti.data[6]; tetrad21_lu[ij] = ti.data[7]; tetrad22_lu[ij] = ti.data[8];
    });
}

Tensor3x3 Metric::GetTetrad(size_t ij) const
{
    return Tensor3x3(tetrad00_ul[ij], tetrad01_ul[ij], tetrad02_ul[ij], tetrad10_ul[ij],
                     tetrad11_ul[ij], tetrad12_ul[ij], tetrad20_ul[ij], tetrad21_ul[ij], tetrad22_ul[ij]);
}
Tensor3x3 Metric::GetTetradInverse(size_t ij) const
{
    return Tensor3x3(tetrad00_lu[ij], tetrad01_lu[ij], tetrad02_lu[ij], tetrad10_lu[ij],
                     tetrad11_lu[ij], tetrad12_lu[ij], tetrad20_lu[ij], tetrad21_lu[ij], tetrad22_lu[ij]);
}

void Metric::PackGeodesicBlock(double *out) const
{
    const DoubleBuffer *src[RT_NMETRIC] = {
        &alpha, &beta1_u, &beta2_u, &d1_alpha_l, &d2_alpha_l, &d1_beta1_lu, &d1_beta2_lu,
        &d2_beta1_lu, &d2_beta2_lu, &K11_ll, &K12_ll, &K22_ll, &gamma11_uu, &gamma12_uu,
        &gamma22_uu, &d1_gamma11_lll, &d1_gamma12_lll, &d1_gamma22_lll, &d2_gamma11_lll,
        &d2_gamma12_lll, &d2_gamma22_lll, &tetrad11_ul, &tetrad12_ul, &tetrad21_ul, &tetrad22_ul,
        &tetrad11_lu, &tetrad12_lu, &tetrad21_lu, &tetrad22_lu};
    ParallelFor(grid.nxy, [&](size_t ij) {
        for (int q = 0; q < RT_NMETRIC; q++) out[ij * RT_NMETRIC + q] = (*src[q])[ij];
    });
}

// ---- values at a point -------------------------------------------------
double Metric::InterpolateArrayTo_ij(const DoubleBuffer &array, const Coord &ij) const
{
    const size_t i0 = (size_t)std::floor(ij[1]), j0 = (size_t)std::floor(ij[2]);
    const size_t i1 = i0 + 1, j1 = j0 + 1;
    return rtm::bilinear<double>(ij[1] - i0, ij[2] - j0, array[grid.Index(i0, j0)], array[grid.Index(i0, j1)],
                                 array[grid.Index(i1, j0)], array[grid.Index(i1, j1)]);
}
double Metric::GetAlpha(const Coord &xy) const
{
    return grid.OutsideDomain(xy) ? 1.0 : InterpolateArrayTo_ij(alpha, grid.ij(xy));
}
Tensor2 Metric::GetBeta_u(const Coord &xy) const
{
    if (grid.OutsideDomain(xy)) return Tensor2(0.0);
    const Coord ij = grid.ij(xy);
    return Tensor2(InterpolateArrayTo_ij(beta1_u, ij), InterpolateArrayTo_ij(beta2_u, ij));
}
Tensor3x3 Metric::GetTetrad(const Coord &xy) const
{
    if (grid.OutsideDomain(xy)) return Tensor3x3(1, 0, 0, 0, 1, 0, 0, 0, 1);
    const Coord ij = grid.ij(xy);
    const DoubleBuffer *e[9] = {&tetrad00_ul, &tetrad01_ul, &tetrad02_ul, &tetrad10_ul, &tetrad11_ul,
                                &tetrad12_ul, &tetrad20_ul, &tetrad21_ul, &tetrad22_ul};
    Tensor3x3 t;
    for (int k = 0; k < 9; k++) t.data[k] = InterpolateArrayTo_ij(*e[k], ij);
    return t;
}
Tensor3x3 Metric::GetTetradInverse(const Coord &xy) const
{
    if (grid.OutsideDomain(xy)) return Tensor3x3(1, 0, 0, 0, 1, 0, 0, 0, 1);
    const Coord ij = grid.ij(xy);
    const DoubleBuffer *e[9] = {&tetrad00_lu, &tetrad01_lu, &tetrad02_lu, &tetrad10_lu, &tetrad11_lu,
                                &tetrad12_lu, &tetrad20_lu, &tetrad21_lu, &tetrad22_lu};
    Tensor3x3 t;
    for (int k = 0; k < 9; k++) t.data[k] = InterpolateArrayTo_ij(*e[k], ij);
    return t;
}
Tensor3x3 Metric::GetMetric_ll(const Coord &xy) const
{
    if (grid.OutsideDomain(xy)) return Tensor3x3(-1, 0, 0, 0, 1, 0, 0, 0, 1);
    const Coord ij = grid.ij(xy);
    const double g00 = InterpolateArrayTo_ij(g00_ll, ij), g11 = InterpolateArrayTo_ij(g11_ll, ij),
                 g22 = InterpolateArrayTo_ij(g22_ll, ij), g01 = InterpolateArrayTo_ij(g01_ll, ij),
                 g02 = InterpolateArrayTo_ij(g02_ll, ij), g12 = InterpolateArrayTo_ij(g12_ll, ij);
    return Tensor3x3(g00, g01, g02, g01, g11, g12, g02, g12, g22);
}
Tensor3 Metric::uEulObs(size_t ij) const
{
    const double a = GetAlpha(ij);
    const Tensor2 b = GetBeta_u(ij);
    return Tensor3(1.0 / a, -b[1] / a, -b[2] / a);
}
Tensor3 Metric::uEulObs(const Coord &xy) const
{
    if (grid.OutsideDomain(xy)) return Tensor3(1, 0, 0);
    const double a = GetAlpha(xy);
    const Tensor2 b = GetBeta_u(xy);
    return Tensor3(1.0 / a, -b[1] / a, -b[2] / a);
}

// ---- remaining getters (src/Metric.cpp:406-800) ----
namespace {
using Arr = const DoubleBuffer *;
// symmetric 3x3 from the arrays 00, 01, 02, 11, 12, 22
Tensor3x3 sym3(const double v[6]) { return Tensor3x3(v[0], v[1], v[2], v[1], v[3], v[4], v[2], v[4], v[5]); }
} // namespace

Tensor3x3 Metric::GetMetric_uu(size_t ij) const
{
    const double v[6] = {g00_uu[ij], g01_uu[ij], g02_uu[ij], g11_uu[ij], g12_uu[ij], g22_uu[ij]};
    return sym3(v);
}
Tensor3x3 Metric::GetMetric_uu(const Coord &xy) const
{
    if (grid.OutsideDomain(xy)) return Tensor3x3(-1, 0, 0, 0, 1, 0, 0, 0, 1);
    const Coord ij = grid.ij(xy);
    const Arr a[6] = {&g00_uu, &g01_uu, &g02_uu, &g11_uu, &g12_uu, &g22_uu};
    double v[6];
    for (int k = 0; k < 6; k++) v[k] = InterpolateArrayTo_ij(*a[k], ij);
    return sym3(v);
}

/* d_a g_bc with a = t (zero), x, y and the symmetric pair bc from the six
 * arrays of each spatial derivative */
static Tensor3x3x3 deriv_sym3(const double (&v)[2][6])
{
    Tensor3x3x3 t(0.0);
    for (int a = 1; a < 3; a++) {
        const Tensor3x3 m = sym3(v[a - 1]);
        for (int b = 0; b < 3; b++)
            for (int c = 0; c < 3; c++) t[{a, b, c}] = m[{b, c}];
    }
    return t;
}
Tensor3x3x3 Metric::GetDerivMetric_lll(size_t ij) const
{
    const double v[2][6] = {{d1_g00_lll[ij], d1_g01_lll[ij], d1_g02_lll[ij], d1_g11_lll[ij], d1_g12_lll[ij], d1_g22_lll[ij]},
                            {d2_g00_lll[ij], d2_g01_lll[ij], d2_g02_lll[ij], d2_g11_lll[ij], d2_g12_lll[ij], d2_g22_lll[ij]}};
    return deriv_sym3(v);
}
Tensor3x3x3 Metric::GetDerivMetric_lll(const Coord &xy) const
{
    if (grid.OutsideDomain(xy)) return Tensor3x3x3(0.0);
    const Coord ij = grid.ij(xy);
    const Arr a[2][6] = {{&d1_g00_lll, &d1_g01_lll, &d1_g02_lll, &d1_g11_lll, &d1_g12_lll, &d1_g22_lll},
                         {&d2_g00_lll, &d2_g01_lll, &d2_g02_lll, &d2_g11_lll, &d2_g12_lll, &d2_g22_lll}};
    double v[2][6];
    for (int d = 0; d < 2; d++)
        for (int k = 0; k < 6; k++) v[d][k] = InterpolateArrayTo_ij(*a[d][k], ij);
    return deriv_sym3(v);
}
Tensor3x3x3 Metric::GetDerivMetric_luu(size_t ij) const
{
    const double v[2][6] = {{d1_g00_luu[ij], d1_g01_luu[ij], d1_g02_luu[ij], d1_g11_luu[ij], d1_g12_luu[ij], d1_g22_luu[ij]},
                            {d2_g00_luu[ij], d2_g01_luu[ij], d2_g02_luu[ij], d2_g11_luu[ij], d2_g12_luu[ij], d2_g22_luu[ij]}};
    return deriv_sym3(v);
}
Tensor3x3x3 Metric::GetDerivMetric_luu(const Coord &xy) const
{
    if (grid.OutsideDomain(xy)) return Tensor3x3x3(0.0);
    const Coord ij = grid.ij(xy);
    const Arr a[2][6] = {{&d1_g00_luu, &d1_g01_luu, &d1_g02_luu, &d1_g11_luu, &d1_g12_luu, &d1_g22_luu},
                         {&d2_g00_luu, &d2_g01_luu, &d2_g02_luu, &d2_g11_luu, &d2_g12_luu, &d2_g22_luu}};
    double v[2][6];
    for (int d = 0; d < 2; d++)
        for (int k = 0; k < 6; k++) v[d][k] = InterpolateArrayTo_ij(*a[d][k], ij);
    return deriv_sym3(v);
}

Tensor2 Metric::GetBeta_l(const Coord &xy) const
{
    if (grid.OutsideDomain(xy)) return Tensor2(0.0);
    const Coord ij = grid.ij(xy);
    return Tensor2(InterpolateArrayTo_ij(beta1_l, ij), InterpolateArrayTo_ij(beta2_l, ij));
}

/* 2x2 from four arrays (xx, xy, yx, yy) at cell ij or at a point; outside
 * the domain the given flat value */
static Tensor2x2 mat2_at(const Metric &m, const Arr (&a)[4], const Coord &xy, const Tensor2x2 &outside)
{
    if (m.grid.OutsideDomain(xy)) return outside;
    const Coord ij = m.grid.ij(xy);
    return Tensor2x2(m.InterpolateArrayTo_ij(*a[0], ij), m.InterpolateArrayTo_ij(*a[1], ij),
                     m.InterpolateArrayTo_ij(*a[2], ij), m.InterpolateArrayTo_ij(*a[3], ij));
}
static Tensor2x2 mat2_ij(const Arr (&a)[4], size_t ij) { return Tensor2x2((*a[0])[ij], (*a[1])[ij], (*a[2])[ij], (*a[3])[ij]); }

Tensor2x2 Metric::GetGamma_ll(size_t ij) const
{
    const Arr a[4] = {&gamma11_ll, &gamma12_ll, &gamma12_ll, &gamma22_ll};
    return mat2_ij(a, ij);
}
Tensor2x2 Metric::GetGamma_ll(const Coord &xy) const
{
    const Arr a[4] = {&gamma11_ll, &gamma12_ll, &gamma12_ll, &gamma22_ll};
    return mat2_at(*this, a, xy, Tensor2x2(1, 0, 0, 1));
}
Tensor2x2 Metric::GetGamma_uu(size_t ij) const
{
    const Arr a[4] = {&gamma11_uu, &gamma12_uu, &gamma12_uu, &gamma22_uu};
    return mat2_ij(a, ij);
}
Tensor2x2 Metric::GetGamma_uu(const Coord &xy) const
{
    const Arr a[4] = {&gamma11_uu, &gamma12_uu, &gamma12_uu, &gamma22_uu};
    return mat2_at(*this, a, xy, Tensor2x2(1, 0, 0, 1));
}
Tensor2 Metric::GetDerivAlpha_l(const Coord &xy) const
{
    if (grid.OutsideDomain(xy)) return Tensor2(0.0);
    const Coord ij = grid.ij(xy);
    return Tensor2(InterpolateArrayTo_ij(d1_alpha_l, ij), InterpolateArrayTo_ij(d2_alpha_l, ij));
}
Tensor2x2 Metric::GetDerivBeta_lu(size_t ij) const
{
    const Arr a[4] = {&d1_beta1_lu, &d1_beta2_lu, &d2_beta1_lu, &d2_beta2_lu};
    return mat2_ij(a, ij);
}
Tensor2x2 Metric::GetDerivBeta_lu(const Coord &xy) const
{
    const Arr a[4] = {&d1_beta1_lu, &d1_beta2_lu, &d2_beta1_lu, &d2_beta2_lu};
    return mat2_at(*this, a, xy, Tensor2x2(0.0));
}
Tensor2x2 Metric::GetDerivBeta_ll(size_t ij) const
{
    const Arr a[4] = {&d1_beta1_ll, &d1_beta2_ll, &d2_beta1_ll, &d2_beta2_ll};
    return mat2_ij(a, ij);
}
Tensor2x2 Metric::GetDerivBeta_ll(const Coord &xy) const
{
    const Arr a[4] = {&d1_beta1_ll, &d1_beta2_ll, &d2_beta1_ll, &d2_beta2_ll};
    return mat2_at(*this, a, xy, Tensor2x2(0.0));
}
Tensor2x2 Metric::GetExtrCurv_ll(size_t ij) const
{
    const Arr a[4] = {&K11_ll, &K12_ll, &K12_ll, &K22_ll};
    return mat2_ij(a, ij);
}
Tensor2x2 Metric::GetExtrCurv_ll(const Coord &xy) const
{
    const Arr a[4] = {&K11_ll, &K12_ll, &K12_ll, &K22_ll};
    return mat2_at(*this, a, xy, Tensor2x2(0.0));
}

/* d_a gamma_bc for a, b, c in {1, 2} */
static Tensor2x2x2 deriv_sym2(const double (&v)[2][3])
{
    return Tensor2x2x2(v[0][0], v[0][1], v[0][1], v[0][2], v[1][0], v[1][1], v[1][1], v[1][2]);
}
Tensor2x2x2 Metric::GetDerivGamma_lll(size_t ij) const
{
    const double v[2][3] = {{d1_gamma11_lll[ij], d1_gamma12_lll[ij], d1_gamma22_lll[ij]},
                            {d2_gamma11_lll[ij], d2_gamma12_lll[ij], d2_gamma22_lll[ij]}};
    return deriv_sym2(v);
}
Tensor2x2x2 Metric::GetDerivGamma_lll(const Coord &xy) const
{
    if (grid.OutsideDomain(xy)) return Tensor2x2x2(0.0);
    const Coord ij = grid.ij(xy);
    const Arr a[2][3] = {{&d1_gamma11_lll, &d1_gamma12_lll, &d1_gamma22_lll},
                         {&d2_gamma11_lll, &d2_gamma12_lll, &d2_gamma22_lll}};
    double v[2][3];
    for (int d = 0; d < 2; d++)
        for (int k = 0; k < 3; k++) v[d][k] = InterpolateArrayTo_ij(*a[d][k], ij);
    return deriv_sym2(v);
}
Tensor2x2x2 Metric::GetDerivGamma_luu(size_t ij) const
{
    const double v[2][3] = {{d1_gamma11_luu[ij], d1_gamma12_luu[ij], d1_gamma22_luu[ij]},
                            {d2_gamma11_luu[ij], d2_gamma12_luu[ij], d2_gamma22_luu[ij]}};
    return deriv_sym2(v);
}
Tensor2x2x2 Metric::GetDerivGamma_luu(const Coord &xy) const
{
    if (grid.OutsideDomain(xy)) return Tensor2x2x2(0.0);
    const Coord ij = grid.ij(xy);
    const Arr a[2][3] = {{&d1_gamma11_luu, &d1_gamma12_luu, &d1_gamma22_luu},
                         {&d2_gamma11_luu, &d2_gamma12_luu, &d2_gamma22_luu}};
    double v[2][3];
    for (int d = 0; d < 2; d++)
        for (int k = 0; k < 3; k++) v[d][k] = InterpolateArrayTo_ij(*a[d][k], ij);
    return deriv_sym2(v);
}
