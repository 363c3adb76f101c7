// Radiation driver over the C ABI (reference: src/Radiation.cpp).
#include "Radiation.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <iostream>
#include "rt2d.h"

void Radiation::Check(int status) const
{
    if (status != RT_OK) ExitOnError(std::string("rt2d: ") + rt_last_error(ctx));
}

void Radiation::RequireSupportedStreaming() const
{
    if (config.streamingType != StreamingType::FlatFixed && config.streamingType != StreamingType::CurvedFixed)
        ExitOnError("streaming type " + StreamingName(config.streamingType) +
                    " is not part of this library (FlatFixed and CurvedFixed are)");
}

// Reference constructor: src/Radiation.cpp:3-82.  Buffers get the same sizes;
// the Fourier tables start as the identity map (S = 1, X = x, Y = y, and the
// velocity series from the intensity stencil's first directions, as in the
// reference) projected with Fourier::GetCoefficients
// (src/FourierHarmonics.cpp:33-45).
Radiation::Radiation(Metric &metric_, Stencil &stencil_, Stencil &streamingStencil_, Config config_)
    : grid(metric_.grid), metric(metric_), stencil(stencil_), streamingStencil(streamingStencil_), config(config_)
{
    RequireSupportedStreaming();
    const size_t nxy = grid.nxy, D = stencil.nDir, nF = streamingStencil.nCoefficients;
    isInitialGridPoint = new bool[nxy]();
    for (DoubleBuffer *b : {&initialE_LF, &initialFx_LF, &initialFy_LF, &initialPxx_LF, &initialPxy_LF, &initialPyy_LF,
                            &initialFluxAngle_IF, &rotationAngle, &rotationAngleNew, &E, &Fx, &Fy, &Pxx, &Pxy, &Pyy,
                            &E_LF, &Fx_LF, &Fy_LF, &Pxx_LF, &Pxy_LF, &Pyy_LF, &F_LF, &kappa0, &kappa1, &kappaA, &eta,
                            &ux, &uy})
        b->assign(nxy, 0.0);
    initialI.assign(nxy * D, 0.0);
    I.assign(nxy * D, 0.0);
    Inew.assign(nxy * D, 0.0);
    for (DoubleBuffer *b : {&coefficientsS, &coefficientsX, &coefficientsY, &coefficientsCx, &coefficientsCy})
        b->assign(nxy * nF, 0.0);
    itterationCount.assign(nxy, 0);

    const size_t nS = streamingStencil.nDir;
    std::vector<double> B(nS * nF);
    for (size_t d = 0; d < nS; d++)
        for (size_t k = 0; k < nF; k++) B[d * nF + k] = rtm::fourier_basis((int)k, streamingStencil.Phi(d));
    for (size_t ij = 0; ij < nxy; ij++) {
        const Coord xy = grid.xy(ij);
        for (size_t d = 0; d < nS; d++) {
            const double w = streamingStencil.W(d);
            const bool has = d < stencil.nDir;
            const double data[5] = {1.0, xy[1], xy[2], has ? stencil.Cx(d) : 0.0, has ? stencil.Cy(d) : 0.0};
            DoubleBuffer *dst[5] = {&coefficientsS, &coefficientsX, &coefficientsY, &coefficientsCx, &coefficientsCy};
            for (int s = 0; s < 5; s++) {
                const double c = data[s] * w;
                for (size_t k = 0; k < nF; k++) (*dst[s])[HarmonicIndex(k, ij)] += c * B[d * nF + k];
            }
        }
    }

    rt_desc desc{};
    desc.nx = (int32_t)grid.nx;
    desc.ny = (int32_t)grid.ny;
    desc.dx = grid.dx;
    desc.dy = grid.dy;
    desc.dt = grid.dt;
    desc.startx = grid.startx;
    desc.starty = grid.starty;
    desc.nDir = (int32_t)D;
    desc.nCoeff = (int32_t)nF;
    desc.precision = config.precision == Precision::FP32Mixed ? RT_FP32_MIXED : RT_FP64;
    desc.streaming = config.streamingType == StreamingType::CurvedFixed ? RT_CURVED_FIXED : RT_FLAT_FIXED;
    desc.excision_r2 = config.excisionRadius >= 0 ? config.excisionRadius * config.excisionRadius : metric.ExcisionR2();
    desc.device = config.device;
    if (rt_create(&desc, &ctx) != RT_OK) ExitOnError(std::string("rt2d: ") + rt_last_error(nullptr));

    const InterpolationGrid &ig = stencil.interpolationGrid;
    std::vector<int32_t> Nc(4 * ig.nGrid);
    std::vector<double> Wc(4 * ig.nGrid);
    for (size_t g = 0; g < ig.nGrid; g++)
        for (int m = 0; m < 4; m++) {
            Nc[4 * g + m] = (int32_t)ig.neighbourIndexesCubic[g][m];
            Wc[4 * g + m] = ig.neighbourWeightsCubic[g][m];
        }
    Check(rt_set_stencil(ctx, stencil.WData(), stencil.PhiData(), stencil.CxData(), stencil.CyData(),
                         (int32_t)ig.nGrid, Nc.data(), Wc.data()));
    Check(rt_set_streaming_stencil(ctx, streamingStencil.WData(), streamingStencil.PhiData(), streamingStencil.CxData(),
                                   streamingStencil.CyData()));
    Check(rt_set_geometry(ctx, metric.alpha.data(), metric.tetrad11_ul.data(), metric.tetrad12_ul.data(),
                          metric.tetrad21_ul.data(), metric.tetrad22_ul.data()));
}

Radiation::~Radiation()
{
    rt_destroy(ctx);
    delete[] isInitialGridPoint;
}

void Radiation::UploadFields()
{
    Check(rt_set_fields(ctx, eta.data(), kappa0.data(), kappa1.data(), kappaA.data(), ux.data(), uy.data()));
}
void Radiation::UploadI() { Check(rt_upload_I(ctx, I.data())); }
void Radiation::DownloadI() { Check(rt_download_I(ctx, I.data())); }
void Radiation::DownloadMoments()
{
    Check(rt_get_moments(ctx, E.data(), Fx.data(), Fy.data(), Pxx.data(), Pxy.data(), Pyy.data()));
}

// Reference: src/Radiation.cpp:124-189.  Initial intensities are copied for
// the marked cells; initial moments need the reference's flux-to-sigma
// lookup tables, which this library does not build.
void Radiation::LoadInitialData()
{
    bool any = false;
    for (size_t ij = 0; ij < grid.nxy; ij++) any = any || isInitialGridPoint[ij];
    if (!any) return;
    if (config.initialDataType == InitialDataType::Moments)
        ExitOnError("initial data given as moments is not part of this library; use InitialDataType::Intensities");
    for (size_t ij = 0; ij < grid.nxy; ij++) {
        if (!isInitialGridPoint[ij]) continue;
        rotationAngle[ij] = 0;
        for (size_t d = 0; d < stencil.nDir; d++) I[Index(ij, d)] = initialI[Index(ij, d)];
    }
}

// Reference: src/Radiation.cpp:191-237 (GPU RK45 backtrace, FP64).
void Radiation::UpdateFourierCoefficients()
{
    if (config.streamingType != StreamingType::CurvedFixed) return;
    std::vector<double> block(grid.nxy * RT_NMETRIC);
    metric.PackGeodesicBlock(block.data());
    Check(rt_build_fourier_tables(ctx, block.data(), config.rk45Tolerance));
    Check(rt_get_fourier_tables(ctx, coefficientsS.data(), coefficientsX.data(), coefficientsY.data(),
                                coefficientsCx.data(), coefficientsCy.data()));
    tablesOnDevice = true;
}

// Reference: src/Radiation.cpp:239-269.
void Radiation::ComputeMomentsIF()
{
    UploadFields();
    UploadI();
    Check(rt_compute_moments(ctx));
    DownloadMoments();
}

void Radiation::ComputeMomentsLF() { ExitOnError("lab-frame moments are not part of this library"); }

// Reference: src/Radiation.cpp:380-414; the streamed intensities end in I
// and the previous state in Inew, as after the reference's swap.
void Radiation::StreamFlatFixed()
{
    if (config.streamingType != StreamingType::FlatFixed)
        ExitOnError("StreamFlatFixed needs a Radiation built with StreamingType::FlatFixed");
    UploadFields();
    UploadI();
    Check(rt_stream(ctx));
    std::swap(I, Inew);
    DownloadI();
}

// Reference: src/Radiation.cpp:454-512.
void Radiation::StreamCurvedFixed()
{
    if (config.streamingType != StreamingType::CurvedFixed)
        ExitOnError("StreamCurvedFixed needs a Radiation built with StreamingType::CurvedFixed");
    if (!tablesOnDevice) UpdateFourierCoefficients();
    UploadFields();
    UploadI();
    Check(rt_stream(ctx));
    std::swap(I, Inew);
    DownloadI();
}

void Radiation::StreamFlatAdaptive() { ExitOnError("FlatAdaptive streaming is not part of this library"); }
void Radiation::StreamCurvedAdaptive() { ExitOnError("CurvedAdaptive streaming is not part of this library"); }
void Radiation::StreamGeodesicFixed() { ExitOnError("GeodesicFixed streaming is not part of this library"); }

// Reference: src/Radiation.cpp:641-736.  The lambda iteration starts from the
// moments of I (the reference's ComputeMomentsIF call just before Collide);
// afterwards I holds the collided state and E..Pyy the final moments.
void Radiation::Collide()
{
    if (config.streamingType == StreamingType::CurvedFixed && !tablesOnDevice) UpdateFourierCoefficients();
    UploadFields();
    UploadI();
    Check(rt_compute_moments(ctx));
    Check(rt_reset_iteration_stats(ctx));
    Check(rt_collide(ctx));
    std::swap(I, Inew);
    DownloadI();
    DownloadMoments();
    Check(rt_get_iteration_counts(ctx, itterationCount.data()));
    int32_t mx = 0;
    double sum = 0;
    Check(rt_iteration_stats(ctx, &mx, &sum));
    maxItterationCount = std::max(maxItterationCount, (int)mx);
    averageItterationCount += sum;
}

// Reference: src/Radiation.cpp:738-885.  The whole step loop runs on the GPU
// (rt_step: stream, moments and collide fused per step).  Frame output is
// not part of this library: with writeData set, the final IF moments are
// left in E..Pyy instead.
void Radiation::RunSimulation()
{
    LoadInitialData();
    UpdateFourierCoefficients();

    const int timeSteps = (int)std::ceil(config.simTime / grid.dt);
    config.simTime = timeSteps * grid.dt;
    if (config.printSetup) {
        std::cout << " nx           = " << grid.nx << "\n";
        std::cout << " ny           = " << grid.ny << "\n";
        std::cout << " nDir         = " << stencil.nDir << "\n";
        std::cout << " nFourier     = " << streamingStencil.nDir << "\n";
        std::cout << " simTime      = " << config.simTime << "\n";
        std::cout << " dx           = " << grid.dx << "\n";
        std::cout << " dy           = " << grid.dy << "\n";
        std::cout << " dt           = " << grid.dt << "\n";
        std::cout << " timeSteps    = " << timeSteps << std::endl;
    }
    if (config.writeData)
        std::cerr << "note: CSV frame output is not part of this library; final inertial-frame moments are in "
                     "E, Fx, Fy, Pxx, Pxy, Pyy"
                  << std::endl;

    UploadFields();
    UploadI();
    Check(rt_reset_iteration_stats(ctx));
    const auto t0 = std::chrono::steady_clock::now();
    if (!config.updateFourierHarmonics && !config.keepSourceNodesActive) {
        Check(rt_step(ctx, timeSteps, 1));
    } else {
        // per-step hooks of the reference loop (src/Radiation.cpp:790-793)
        for (int n = 0; n < timeSteps; n++) {
            if (config.updateFourierHarmonics) UpdateFourierCoefficients();
            if (config.keepSourceNodesActive) {
                DownloadI();
                LoadInitialData();
                UploadI();
            }
            Check(rt_step(ctx, 1, n == timeSteps - 1 ? 1 : 0));
        }
    }
    Check(rt_synchronize(ctx));
    const auto t1 = std::chrono::steady_clock::now();
    const double seconds = std::chrono::duration<double>(t1 - t0).count();
    DownloadI();
    if (timeSteps > 0) DownloadMoments();

    int32_t mx = 0;
    double sum = 0;
    Check(rt_iteration_stats(ctx, &mx, &sum));
    maxItterationCount = mx;
    averageItterationCount = timeSteps > 0 ? sum / ((double)timeSteps * (grid.nx - 2) * (grid.ny - 2)) : 0.0;
    if (config.printResults) {
        std::cout << "total time: " << seconds << "s" << std::endl;
        std::cout << "Max Lambda Itteration Count     = " << maxItterationCount << std::endl;
        std::cout << "Average Lambda Itteration Count = " << averageItterationCount << std::endl;
        std::cout << "Benchmark: " << grid.nxy * timeSteps / (1e6 * seconds) << "MLUPS" << std::endl;
    }
}
