// Host-side helpers of the Radiation class (GetTempCoordinate,
// GetTemp2VelocityIF, GetFrequencyShift, IntensityAt, AverageF,
// InitialDataLFtoIF, NormalizationOverwrite) checked by identities on the
// host buffers.  Built against the host-only stand-in of the C ABI
// (rt_host_stub.cpp): no GPU.  Prints "radiation_helpers OK" at the end.
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

int main()
{
    Grid grid(30, 30, Coord(-1, -1), Coord(1, 1));
    grid.SetCFL(0.5);
    Minkowski metric(grid, 1.0, 0.0);
    Stencil stencil(16);
    Stencil streamingStencil(5, 0, false);
    Config config = {.name = "helpers", .simTime = 0, .writeData = false, .printResults = false,
                     .streamingType = StreamingType::FlatFixed, .initialDataType = InitialDataType::Intensities};
    Radiation radiation(metric, stencil, streamingStencil, config);

    // the tables the constructor starts from map every cell to itself with
    // no frequency shift, at every angle
    const size_t ij = grid.Index(11, 17);
    for (double a : {0.1, 1.7, 3.3, 5.9}) {
        const Coord p = radiation.GetTempCoordinate(ij, a);
        check(std::abs(p[1] - grid.x((size_t)11)) < 1e-12 && std::abs(p[2] - grid.y((size_t)17)) < 1e-12,
              "GetTempCoordinate of the identity tables", p[1] - grid.x((size_t)11));
        check(std::abs(radiation.GetFrequencyShift(ij, a) - 1.0) < 1e-12, "GetFrequencyShift of the identity tables",
              radiation.GetFrequencyShift(ij, a));
        const Tensor2 v = radiation.GetTemp2VelocityIF(ij, a);
        check(std::isfinite(v[1]) && std::isfinite(v[2]), "GetTemp2VelocityIF finite", v[1]);
    }

    // IntensityAt: node values at the stencil's own directions, a smooth
    // profile between them, and the clamp at 0 on the negative lobe of a spike
    for (size_t d = 0; d < stencil.nDir; d++) radiation.I[radiation.Index(ij, d)] = 1.0 + 0.5 * std::cos(stencil.Phi(d));
    /* the angle goes through the reference's Hastings atan2 (order 9, about
     * 1e-5 rad from the true angle), which moves the fine index by ~3e-4 */
    for (size_t d = 0; d < stencil.nDir; d++) {
        const double want = 1.0 + 0.5 * std::cos(stencil.Phi(d));
        const double got = radiation.IntensityAt(ij, stencil.C(d));
        check(std::abs(got - want) < 2e-5, "IntensityAt at a stencil direction", got - want);
    }
    for (double a : {0.05, 2.2, 4.0}) {
        const double got = radiation.IntensityAt(ij, Tensor2(std::cos(a), std::sin(a)));
        check(std::abs(got - (1.0 + 0.5 * std::cos(a))) < 5e-3, "IntensityAt between directions", got);
    }
    radiation.rotationAngle[ij] = 1.0; // the velocity is read in the cell's turned frame
    {
        const double a = 2.5;
        const double got = radiation.IntensityAt(ij, Tensor2(std::cos(a), std::sin(a)));
        check(std::abs(got - (1.0 + 0.5 * std::cos(a - 1.0))) < 5e-3, "IntensityAt in a turned cell", got);
    }
    radiation.rotationAngle[ij] = 0.0;
    for (size_t d = 0; d < stencil.nDir; d++) radiation.I[radiation.Index(ij, d)] = d == 5 ? 1.0 : 0.0;
    double lowest = 1.0, highest = 0.0;
    for (int k = 0; k < 400; k++) {
        const double a = 2.0 * rtm::kPi * (k + 0.5) / 400;
        const double got = radiation.IntensityAt(ij, Tensor2(std::cos(a), std::sin(a)));
        lowest = std::min(lowest, got);
        highest = std::max(highest, got);
    }
    check(lowest == 0.0 && highest > 0.9, "IntensityAt clamps the cubic lobes at 0", lowest);

    // AverageF: the 3x3 mean
    for (size_t k = 0; k < grid.nxy; k++) {
        radiation.Fx[k] = (double)grid.i(k);
        radiation.Fy[k] = 2.0;
    }
    const Tensor2 f = radiation.AverageF(7, 9);
    check(f[1] == 7.0 && f[2] == 2.0, "AverageF", f[1]);

    // initial-data helpers: flat space keeps E; the normalisation of sigma = 0 is log(sum w) = 0
    radiation.initialE_LF[ij] = 2.0;
    radiation.initialFx_LF[ij] = 0.5;
    const Tensor3 m = radiation.InitialDataLFtoIF(ij);
    check(std::abs(m[0] - 2.0) < 1e-14 && m[1] == 0.5, "InitialDataLFtoIF in flat space", m[0]);
    radiation.sigmaOverwrite = 0.0;
    check(std::abs(radiation.NormalizationOverwrite()) < 1e-14, "NormalizationOverwrite at sigma 0",
          radiation.NormalizationOverwrite());

    if (failures == 0) std::printf("radiation_helpers OK\n");
    return failures == 0 ? 0 : 1;
}
