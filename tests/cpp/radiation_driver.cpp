// A user program written against the reference's C++ API (compare
// exe/main.cpp:12-71 and exe/benchmark.cpp:12-78 of the reference): it
// builds Grid, Stencil, Metric and Radiation, fills the fluid fields and calls
// RunSimulation.  Used by the GPU tests to check that reference-style code
// compiles unchanged and reproduces the oracle.
//
//   radiation_driver flat   out.bin   Minkowski 100x100 on [-1,1]^2, 20 directions,
//                                     CFL 0.5, 100 steps, Gaussian eta (sigma 0.1)
//   radiation_driver curved out.bin   Schwarzschild 40x40 on [-4,4]^2, 16 directions,
//                                     10 steps, eta = 1 on 3 < r < 3.5
//   radiation_driver frames out.bin   Minkowski 40x40 on [-1,1]^2, 20 directions, 12 steps,
//                                     Gaussian eta, writeData with writePeriod 5 dt: CSV
//                                     frames and Log.txt under ../output/driver_frames/
//   radiation_driver vonmises out.bin Minkowski 20x20 on [-1,1]^2, 16 directions, no steps:
//                                     initial data given as lab-frame moments at three cells
//   radiation_driver adaptive out.bin Minkowski 30x30 on [-1,1]^2, Stencil(16,4), FlatAdaptive,
//                                     20 steps, kappa0 0.5, kappa1 0.2, three source cells
//                                     (keepSourceNodesActive) with intensities 1 + 0.5 cx_d
//                                     turned to the angles 0.3, 2.0, 4.5
//   radiation_driver adaptive_curved out.bin  Schwarzschild 30x30 on [-4,4]^2, Stencil(16,4),
//                                     CurvedAdaptive, 8 steps, eta = 1 on 3 < r < 3.5
// out.bin: I (nxy*nDir doubles) then E, Fx, Fy (nxy doubles each); the
// adaptive modes append rotationAngle (nxy doubles).
#include <cmath>
#include <cstdio>
#include <cstring>
#include "Radiation.h"
#include "Spacetimes.h"

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s flat|curved out.bin\n", argv[0]);
        return 2;
    }
    const bool adaptiveCurved = std::strcmp(argv[1], "adaptive_curved") == 0;
    const bool adaptive = adaptiveCurved || std::strcmp(argv[1], "adaptive") == 0;
    const bool curved = adaptiveCurved || std::strcmp(argv[1], "curved") == 0;
    const bool frames = std::strcmp(argv[1], "frames") == 0;
    const bool vonmises = std::strcmp(argv[1], "vonmises") == 0;
    const size_t n = adaptive ? 30 : (vonmises ? 20 : ((curved || frames) ? 40 : 100));
    const double L = curved ? 4.0 : 1.0;
    Grid grid(n, n, Coord(-L, -L), Coord(L, L));
    grid.SetCFL(0.5);
    Minkowski flat(grid, 1.0, 0.0);
    SchwarzSchild bh(grid, 1.0, 0.0);
    Metric &metric = curved ? (Metric &)bh : (Metric &)flat;
    Stencil stencil((curved || vonmises || adaptive) ? 16 : 20, adaptive ? 4 : 0);
    Stencil streamingStencil(5, 0, false);

    const int steps = adaptiveCurved ? 8 : (adaptive ? 20 : (vonmises ? 0 : (curved ? 10 : (frames ? 12 : 100))));
    Config config = {
        .name = curved ? "driver_curved" : (frames ? "driver_frames" : "driver_flat"),
        .simTime = steps * grid.dt,
        .writePeriod = frames ? 5 * grid.dt : 1e9,
        .keepSourceNodesActive = adaptive && !curved,
        .writeData = frames,
        .printSetup = false,
        .printProgress = false,
        .printResults = true,
        .saveInitialData = frames,
        .streamingType = adaptive ? (curved ? StreamingType::CurvedAdaptive : StreamingType::FlatAdaptive)
                                  : (curved ? StreamingType::CurvedFixed : StreamingType::FlatFixed),
        .initialDataType = vonmises ? InitialDataType::Moments : InitialDataType::Intensities,
    };
    Radiation radiation(metric, stencil, streamingStencil, config);
    if (vonmises) {
        /* (E, Fx, Fy) in the lab frame at three cells: relative flux 0.5, an
         * isotropic cell, and |F| > E (clipped to |F| = E) */
        const size_t cells[3] = {grid.Index(5, 5), grid.Index(10, 12), grid.Index(15, 7)};
        const double mom[3][3] = {{2.0, 0.6, 0.8}, {1.5, 0.0, 0.0}, {1.0, -3.0, 4.0}};
        for (int c = 0; c < 3; c++) {
            radiation.isInitialGridPoint[cells[c]] = true;
            radiation.initialE_LF[cells[c]] = mom[c][0];
            radiation.initialFx_LF[cells[c]] = mom[c][1];
            radiation.initialFy_LF[cells[c]] = mom[c][2];
        }
    }
    if (adaptive && !curved) {
        const size_t cells[3] = {grid.Index(8, 8), grid.Index(20, 12), grid.Index(14, 22)};
        const double angle[3] = {0.3, 2.0, 4.5};
        for (int c = 0; c < 3; c++) {
            radiation.isInitialGridPoint[cells[c]] = true;
            radiation.initialFluxAngle_IF[cells[c]] = angle[c];
            for (size_t d = 0; d < stencil.nDir; d++)
                radiation.initialI[radiation.Index(cells[c], d)] = 1.0 + 0.5 * stencil.Cx(d);
        }
    }
    for (size_t j = 0; j < grid.ny; j++)
        for (size_t i = 0; i < grid.nx; i++) {
            const size_t ij = grid.Index(i, j);
            const double x = grid.x(i), y = grid.y(j), r = std::sqrt(x * x + y * y);
            if (adaptive && !curved) {
                radiation.kappa0[ij] = 0.5;
                radiation.kappa1[ij] = 0.2;
                continue;
            }
            radiation.eta[ij] = curved ? ((r > 3 && r < 3.5) ? 1.0 : 0.0) : std::exp(-(x * x + y * y) / (2 * 0.1 * 0.1));
        }
    radiation.RunSimulation();

    std::FILE *f = std::fopen(argv[2], "wb");
    if (!f) return 1;
    std::fwrite(radiation.I.data(), sizeof(double), radiation.I.size(), f);
    std::fwrite(radiation.E.data(), sizeof(double), grid.nxy, f);
    std::fwrite(radiation.Fx.data(), sizeof(double), grid.nxy, f);
    std::fwrite(radiation.Fy.data(), sizeof(double), grid.nxy, f);
    if (adaptive) std::fwrite(radiation.rotationAngle.data(), sizeof(double), grid.nxy, f);
    std::fclose(f);
    std::printf("maxItterationCount %d\n", radiation.maxItterationCount);
    return 0;
}
