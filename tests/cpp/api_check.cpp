// Host-side checks of the reference-API helpers user programs call to build
// initial data (no GPU needed): metric values at points, frame transforms,
// null normalisation, the Eulerian observer's velocity, Lorentz boosts and
// the vector arithmetic.  Each check is fixed by an identity, not by the
// formula it tests.  Prints one line per failure and "api_check OK" at the end.
#include <cmath>
#include <cstdio>
#include "Radiation.h"

static int failures = 0;
static void check(bool ok, const char *what, double value)
{
    if (!ok) {
        std::printf("FAIL %s (%.17g)\n", what, value);
        failures++;
    }
}

// g_{mu nu} a^mu b^nu
static double dot4(const Tensor3x3 &g, const Tensor3 &a, const Tensor3 &b)
{
    double s = 0;
    for (int m = 0; m < 3; m++)
        for (int n = 0; n < 3; n++) s += g[{m, n}] * a[m] * b[n];
    return s;
}

int main()
{
    Grid grid(40, 30, Coord(-4, -3), Coord(4, 3));
    grid.SetCFL(0.5);
    SchwarzSchild bh(grid, 1.0, 0.0);
    KerrSchild ks(grid, 1.0, 0.5);

    // metric at a cell centre equals the cell value; flat values outside
    const size_t ij = grid.Index(30, 20);
    const Coord xy = grid.xy(30, 20);
    check(std::abs(bh.GetAlpha(xy) - bh.GetAlpha(ij)) < 1e-14, "GetAlpha(xy) at a cell centre", bh.GetAlpha(xy));
    check(bh.GetAlpha(Coord(40, 40)) == 1.0, "GetAlpha outside the domain", bh.GetAlpha(Coord(40, 40)));
    check(bh.GetMetric_ll(Coord(-40, 0))[{0, 0}] == -1.0, "GetMetric_ll outside the domain", 0);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            const double d = ks.GetTetrad(xy)[{a, b}] - ks.GetTetrad(ij)[{a, b}];
            check(std::abs(d) < 1e-14, "GetTetrad(xy) at a cell centre", d);
        }
    // halfway between two cells: the mean of the two (bilinear in x)
    const Coord mid(0.5 * (grid.x((size_t)30) + grid.x((size_t)31)), grid.y((size_t)20));
    const double am = 0.5 * (bh.alpha[grid.Index(30, 20)] + bh.alpha[grid.Index(31, 20)]);
    check(std::abs(bh.GetAlpha(mid) - am) < 1e-14, "GetAlpha between cells", bh.GetAlpha(mid) - am);

    // tetrad and its inverse undo each other; e^T g e = diag(-1, 1, 1)
    const Tensor3x3 e = ks.GetTetrad(ij), ei = ks.GetTetradInverse(ij), g = ks.GetMetric_ll(ij);
    const Tensor3 v(0.3, -0.7, 0.2);
    const Tensor3 back = TransformLFtoIF(TransformIFtoLF(v, e), ei);
    for (int k = 0; k < 3; k++) check(std::abs(back[k] - v[k]) < 1e-12, "IF->LF->IF round trip", back[k] - v[k]);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            const Tensor3 ea(e[{0, a}], e[{1, a}], e[{2, a}]), eb(e[{0, b}], e[{1, b}], e[{2, b}]);
            const double eta = (a == b) ? (a == 0 ? -1.0 : 1.0) : 0.0;
            check(std::abs(dot4(g, ea, eb) - eta) < 1e-10, "tetrad orthonormality", dot4(g, ea, eb) - eta);
        }

    // NullNormalize gives a null vector with the same time component and
    // spatial direction; an observer sees it at unit speed
    const Tensor3 u = NullNormalize(Tensor3(1.0, 0.4, -0.9), g);
    check(std::abs(dot4(g, u, u)) < 1e-12, "NullNormalize: null", dot4(g, u, u));
    check(u[0] == 1.0 && std::abs(u[1] / u[2] - 0.4 / -0.9) < 1e-12, "NullNormalize: direction", u[1] / u[2]);
    const Tensor2 vIF = Vec2ObservedByEulObs<LF, IF>(u, ij, ks);
    check(std::abs(vIF.EuklNorm() - 1.0) < 1e-10, "null vector seen at unit speed", vIF.EuklNorm() - 1.0);
    const Tensor2 vLF = Vec2ObservedByEulObs<LF, LF>(u, xy, ks);
    const Tensor2 vIF2 = TransformLFtoIF(vLF, ks.GetTetradInverse(xy));
    check(std::abs(vIF2[1] - vIF[1]) < 1e-12 && std::abs(vIF2[2] - vIF[2]) < 1e-12, "LF->LF then tetrad = LF->IF", 0);

    // Lorentz boosts keep the Minkowski metric
    const Tensor3x3 L = BoostMatrix(Tensor2(0.5, -0.3));
    const Tensor3x3 eta(-1, 0, 0, 0, 1, 0, 0, 0, 1);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double s = 0;
            for (int m = 0; m < 3; m++)
                for (int n = 0; n < 3; n++) s += L[{m, a}] * eta[{m, n}] * L[{n, b}];
            check(std::abs(s - eta[{a, b}]) < 1e-12, "boost preserves eta", s - eta[{a, b}]);
        }
    const Tensor3x3 I3 = BoostMatrix(Tensor2(0.0, 0.0));
    check(I3[{0, 0}] == 1 && I3[{1, 1}] == 1 && I3[{0, 1}] == 0, "zero boost is the identity", 0);

    // vector arithmetic and helpers
    const Coord c = (Coord(4, 4) + Coord(8, 8)) / 2;
    check(c[1] == 6 && c[2] == 6, "Coord arithmetic", c[1]);
    check((Coord(1, 2) - Coord(4, 6)).EuklNorm() == 5.0, "Coord difference norm", 0);
    const Tensor2 n = Tensor2(3, 4).EuklNormalized();
    check(std::abs(n[1] - 0.6) < 1e-15 && std::abs(n[2] - 0.8) < 1e-15, "EuklNormalized", n[1]);

    // Logger: step count and the frame writer's share of the time
    Stencil s(16), st(5, 0, false);
    Logger log(s, st, bh);
    log.SetValues("check", 10.3 * grid.dt);
    check(log.timeSteps == 11 && std::abs(log.simTime - 11 * grid.dt) < 1e-15, "Logger::SetValues", log.timeSteps);
    log.AddTimeMeasurement("Total Time", 2.0);
    log.AddTimeMeasurement(Logger::kFrameWriterScope, 0.5);
    check(log.ComputationTime() == 1.5, "Logger::ComputationTime", log.ComputationTime());

    // the remaining metric getters, by identities: g^uu g_ll = 1 and
    // gamma^uu gamma_ll = 1 at a cell and between cells; flat values outside;
    // derivative arrays against central differences of the metric function;
    // no extrinsic curvature for the static Schwarzschild slicing
    const Coord pt(0.3 * grid.x((size_t)33) + 0.7 * grid.x((size_t)34), grid.y((size_t)24)); // r = 3.4
    for (const Metric *mp : {(const Metric *)&bh, (const Metric *)&ks}) {
        const Metric &M = *mp;
        for (int pass = 0; pass < 2; pass++) {
            const Tensor3x3 gu = pass ? M.GetMetric_uu(pt) : M.GetMetric_uu(ij), gl = pass ? M.GetMetric_ll(pt) : M.GetMetric_ll(ij);
            const Tensor2x2 hu = pass ? M.GetGamma_uu(pt) : M.GetGamma_uu(ij), hl = pass ? M.GetGamma_ll(pt) : M.GetGamma_ll(ij);
            const double tol = pass ? 5e-3 : 1e-12; // between cells: bilinear values of each array
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    double sg = 0;
                    for (int m = 0; m < 3; m++) sg += gu[{a, m}] * gl[{m, b}];
                    check(std::abs(sg - (a == b)) < tol, "g^uu g_ll = 1", sg);
                }
            const Tensor2x2 h = hu * hl;
            check(std::abs(h[{1, 1}] - 1) < tol && std::abs(h[{1, 2}]) < tol && std::abs(h[{2, 2}] - 1) < tol,
                  "gamma^uu gamma_ll = 1", h[{1, 2}]);
            const Tensor2x2 hi = hl.Invert();
            check(std::abs(hi[{1, 1}] - hu[{1, 1}]) < 10 * tol, "Tensor2x2::Invert", hi[{1, 1}] - hu[{1, 1}]);
        }
        const Coord far(40, -40);
        check(M.GetMetric_uu(far)[{0, 0}] == -1 && M.GetGamma_ll(far)[{1, 1}] == 1 && M.GetGamma_uu(far)[{2, 2}] == 1,
              "flat metric outside the domain", 0);
        check(M.GetDerivAlpha_l(far)[1] == 0 && M.GetExtrCurv_ll(far)[{1, 2}] == 0 && M.GetDerivBeta_lu(far)[{2, 1}] == 0 &&
                  M.GetDerivGamma_lll(far)[{2, 1, 2}] == 0 && M.GetDerivMetric_lll(far)[{1, 0, 0}] == 0,
              "zero derivatives outside the domain", 0);
        check(M.GetMinkowskiMetric_ll(ij)[{0, 0}] == -1 && M.GetMinkowskiGamma_uu(pt)[{1, 2}] == 0, "Minkowski getters", 0);
        // d_x g_00 and d_y g_12 at the cell against central differences of g(x, y)
        const double h = 1e-5;
        Metric &Mm = const_cast<Metric &>(M);
        const Tensor3x3 gxp = Mm.MetricFunction(Coord(xy[1] + h, xy[2])), gxm = Mm.MetricFunction(Coord(xy[1] - h, xy[2]));
        const Tensor3x3 gyp = Mm.MetricFunction(Coord(xy[1], xy[2] + h)), gym = Mm.MetricFunction(Coord(xy[1], xy[2] - h));
        const Tensor3x3x3 dg = M.GetDerivMetric_lll(ij);
        const double fx = (gxp[{0, 0}] - gxm[{0, 0}]) / (2 * h), fy = (gyp[{1, 2}] - gym[{1, 2}]) / (2 * h);
        check(std::abs(dg[{1, 0, 0}] - fx) < 1e-6 && std::abs(dg[{2, 1, 2}] - fy) < 1e-6 && dg[{0, 1, 1}] == 0,
              "GetDerivMetric_lll against central differences", dg[{1, 0, 0}] - fx);
        const Tensor2x2x2 dga = M.GetDerivGamma_lll(ij);
        check(std::abs(dga[{1, 1, 1}] - (gxp[{1, 1}] - gxm[{1, 1}]) / (2 * h)) < 1e-6 &&
                  std::abs(dga[{2, 1, 2}] - fy) < 1e-6, "GetDerivGamma_lll against central differences", dga[{1, 1, 1}]);
        check(M.GetDerivAlpha_l(ij)[1] == M.d1_alpha_l[ij] && M.GetBeta_l(ij)[2] == M.beta2_l[ij] &&
                  M.GetDerivBeta_ll(ij)[{2, 1}] == M.d2_beta1_ll[ij] && M.GetDerivGamma_luu(ij)[{1, 2, 2}] == M.d1_gamma22_luu[ij] &&
                  M.GetDerivMetric_luu(ij)[{2, 0, 1}] == M.d2_g01_luu[ij] && M.GetExtrCurv_ll(ij)[{2, 2}] == M.K22_ll[ij],
              "cell getters read their arrays", 0);
    }
    // Schwarzschild coordinates: lapse sqrt(1 - 2/r), d_x alpha = x / (r^3 alpha); static slicing K = 0
    {
        const double r = std::sqrt(xy[1] * xy[1] + xy[2] * xy[2]), al = std::sqrt(1 - 2.0 / r);
        check(std::abs(bh.GetDerivAlpha_l(ij)[1] - xy[1] / (r * r * r * al)) < 1e-6, "d_x alpha (Schwarzschild)",
              bh.GetDerivAlpha_l(ij)[1]);
        const Tensor2x2 K = bh.GetExtrCurv_ll(ij);
        check(std::abs(K[{1, 1}]) < 1e-6 && std::abs(K[{1, 2}]) < 1e-6, "no extrinsic curvature (Schwarzschild)", K[{1, 1}]);
    }
    check(grid.OutsideDomain(-0.5, 3.0) && !grid.OutsideDomain(3.0, 4.0) && grid.OutsideDomain(2.0, grid.ny - 0.5),
          "OutsideDomain(i, j)", 0);

    // linear angular tables: weights sum to 1 and reproduce the angle itself
    // (every fine angle sits at pi in the rotated frame the weights use)
    Stencil s20(20, 4);
    const InterpolationGrid &ig = s20.interpolationGrid;
    for (size_t gi = 0; gi < ig.nGrid; gi++) {
        const double phi = 2.0 * rtm::kPi * gi / ig.nGrid;
        double sw = 0, sa = 0;
        for (int k = 0; k < 2; k++) {
            const double a = std::fmod(s20.Phi(ig.neighbourIndexesLinear[gi][k]) + rtm::kPi - phi + 2 * rtm::kPi, 2 * rtm::kPi);
            sw += ig.neighbourWeightsLinear[gi][k];
            sa += ig.neighbourWeightsLinear[gi][k] * a;
        }
        check(std::abs(sw - 1) < 1e-12 && std::abs(sa - rtm::kPi) < 1e-12, "linear angular weights", sa - rtm::kPi);
    }
    grid.Print();
    ig.Print();
    s20.WriteToCsv();

    if (failures == 0) std::printf("api_check OK\n");
    return failures == 0 ? 0 : 1;
}
