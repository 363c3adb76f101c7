// Radiation driver over the C ABI (reference: src/Radiation.cpp).
#include "Radiation.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include "rt2d.h"

void Radiation::Check(int status) const
{
    if (status != RT_OK) ExitOnError(std::string("rt2d: ") + rt_last_error(ctx));
}

void Radiation::RequireSupportedStreaming() const
{
    if (config.streamingType == StreamingType::GeodesicFixed)
        ExitOnError("streaming type GeodesicFixed (a performance test of the reference) is not part of this library");
}
bool Radiation::IsAdaptive() const
{
    return config.streamingType == StreamingType::FlatAdaptive || config.streamingType == StreamingType::CurvedAdaptive;
}
bool Radiation::IsCurved() const
{
    return config.streamingType == StreamingType::CurvedFixed || config.streamingType == StreamingType::CurvedAdaptive;
}
void Radiation::UploadRotation()
{
    if (IsAdaptive()) Check(rt_set_rotation(ctx, rotationAngle.data(), rotationAngleNew.data()));
}
void Radiation::DownloadRotation()
{
    if (IsAdaptive()) Check(rt_get_rotation(ctx, rotationAngle.data(), rotationAngleNew.data()));
}

// Reference constructor: src/Radiation.cpp:3-82.  Buffers get the same sizes;
// the Fourier tables start as the identity map (S = 1, X = x, Y = y, and the
// velocity series from the intensity stencil's first directions, as in the
// reference) projected with Fourier::GetCoefficients
// (src/FourierHarmonics.cpp:33-45).
Radiation::Radiation(Metric &metric_, Stencil &stencil_, Stencil &streamingStencil_, Config config_)
    : grid(metric_.grid), metric(metric_), stencil(stencil_), streamingStencil(streamingStencil_), config(config_),
      logger(stencil_, streamingStencil_, metric_)
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
    desc.streaming = (int32_t)config.streamingType; /* rt_streaming mirrors StreamingType */
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
    const double *tetrad[9] = {metric.tetrad00_ul.data(), metric.tetrad01_ul.data(), metric.tetrad02_ul.data(),
                               metric.tetrad10_ul.data(), metric.tetrad11_ul.data(), metric.tetrad12_ul.data(),
                               metric.tetrad20_ul.data(), metric.tetrad21_ul.data(), metric.tetrad22_ul.data()};
    Check(rt_set_lab_frame(ctx, tetrad, metric.gamma11_ll.data(), metric.gamma12_ll.data(),
                           metric.gamma22_ll.data()));
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

// Reference: src/Radiation.cpp:303-320 (Fourier::GetValue,
// src/FourierHarmonics.cpp:56-62, of the cell's series at the angle).
static double SeriesValue(const DoubleBuffer &coefficients, size_t first, size_t nF, double angle)
{
    double v = 0;
    for (size_t k = 0; k < nF; k++) v += coefficients[first + k] * rtm::fourier_basis((int)k, angle);
    return v;
}
Coord Radiation::GetTempCoordinate(size_t ij, double angle) const
{
    const size_t nF = streamingStencil.nCoefficients, f0 = HarmonicIndex(0, ij);
    return Coord(SeriesValue(coefficientsX, f0, nF, angle), SeriesValue(coefficientsY, f0, nF, angle));
}
Tensor2 Radiation::GetTemp2VelocityIF(size_t ij, double angle) const
{
    const size_t nF = streamingStencil.nCoefficients, f0 = HarmonicIndex(0, ij);
    return Tensor2(SeriesValue(coefficientsCx, f0, nF, angle), SeriesValue(coefficientsCy, f0, nF, angle));
}
double Radiation::GetFrequencyShift(size_t ij, double angle) const
{
    return SeriesValue(coefficientsS, HarmonicIndex(0, ij), streamingStencil.nCoefficients, angle);
}

// Reference: src/Radiation.cpp:322-343: the velocity turned into the cell's
// frame, the fine angle, two 4-point cubic sums, their lerp, clamped at 0.
double Radiation::IntensityAt(size_t ij, Tensor2 vTempIF) const
{
    vTempIF = RotationMatrix(rotationAngle[ij]).Inverse() * vTempIF;
    const InterpolationGrid &g = stencil.interpolationGrid;
    const double d = g.d(vTempIF.Phi());
    const size_t d0 = (size_t)((int)std::floor(d) % (int)g.nGrid), d1 = (d0 + 1) % g.nGrid;
    double v0 = 0, v1 = 0;
    for (size_t k = 0; k < 4; k++) {
        v0 += g.neighbourWeightsCubic[d0][k] * I[Index(ij, g.neighbourIndexesCubic[d0][k])];
        v1 += g.neighbourWeightsCubic[d1][k] * I[Index(ij, g.neighbourIndexesCubic[d1][k])];
    }
    const double t = d - d0;
    return std::max(0.0, v0 * (1.0 - t) + v1 * t);
}

// Reference: src/Radiation.cpp:345-359.
Tensor2 Radiation::AverageF(size_t i, size_t j) const
{
    Tensor2 f(0.0);
    for (int b = -1; b <= 1; b++)
        for (int a = -1; a <= 1; a++) {
            const size_t k = grid.Index(i + a, j + b);
            f[1] += Fx[k];
            f[2] += Fy[k];
        }
    f[1] /= 9.0;
    f[2] /= 9.0;
    return f;
}

// Reference: src/Radiation.cpp:103-123 (InitialDataLFtoIF).  The lab-frame
// tensor is transformed with the inverse tetrad; as in the reference only
// its energy density is taken from the result, the flux from the lab-frame
// tensor, and |F| > E is clipped to |F| = E.
Tensor3 Radiation::InitialDataLFtoIF(size_t ij) const
{
    const double T[3][3] = {{initialE_LF[ij], initialFx_LF[ij], initialFy_LF[ij]},
                            {initialFx_LF[ij], initialPxx_LF[ij], initialPxy_LF[ij]},
                            {initialFy_LF[ij], initialPxy_LF[ij], initialPyy_LF[ij]}};
    const Tensor3x3 e = metric.GetTetrad(ij).Invert();
    double E_IF = 0;
    for (int A = 0; A < 3; A++)
        for (int B = 0; B < 3; B++) E_IF += T[A][B] * e[{0, A}] * e[{0, B}];
    double Fx_IF = T[0][1], Fy_IF = T[0][2];
    const double F_IF = Tensor2(Fx_IF, Fy_IF).EuklNorm();
    if (F_IF > E_IF) {
        Fx_IF *= E_IF / F_IF;
        Fy_IF *= E_IF / F_IF;
    }
    return Tensor3(E_IF, Fx_IF, Fy_IF);
}

// Reference: src/Radiation.cpp:179-185 (NormalizationOverwrite).
double Radiation::NormalizationOverwrite() const
{
    double n = 0;
    for (size_t d = 0; d < stencil.nDir; d++) n += stencil.W(d) * std::exp(sigmaOverwrite * Tensor2::Dot(Tensor2(1, 0), stencil.C(d)));
    return std::log(n);
}

// Reference: src/Radiation.cpp:124-178.  For the marked cells: intensities as
// given (InitialDataType::Intensities), or (Moments) a von Mises distribution
// I_d = E exp(sigma n.C_d - normalization) around the inertial-frame flux
// direction n, with sigma and the normalization from the stencil's
// flux-to-sigma tables at |F|/E (or sigmaOverwrite).  Adaptive streaming
// puts the distribution at angle 0 of the cell's stencil and turns the
// stencil to the flux direction (rotationAngle); fixed streaming sets the
// angle to 0.  Cells without flux keep their angle.  sourceAngle records, per
// cell, the angle written here (NaN: none), which keepSourceNodesActive
// re-imposes with the intensities.
void Radiation::LoadInitialData()
{
    const bool adaptive = IsAdaptive();
    const double normalizationOverwrite = (0 <= sigmaOverwrite) ? NormalizationOverwrite() : 0.0;
    sourceAngle.assign(grid.nxy, std::nan(""));
    for (size_t ij = 0; ij < grid.nxy; ij++) {
        if (!isInitialGridPoint[ij]) continue;
        if (config.initialDataType == InitialDataType::Intensities) {
            rotationAngle[ij] = adaptive ? initialFluxAngle_IF[ij] : 0.0;
            sourceAngle[ij] = rotationAngle[ij];
            for (size_t d = 0; d < stencil.nDir; d++) I[Index(ij, d)] = initialI[Index(ij, d)];
            continue;
        }
        const Tensor3 m = InitialDataLFtoIF(ij);
        const double E_IF = m[0];
        const Tensor2 F(m[1], m[2]);
        const double Fn = F.EuklNorm();
        const double relativeF = Fn / E_IF;
        Tensor2 dir = adaptive ? Tensor2(1, 0) : Tensor2(m[1] / Fn, m[2] / Fn);
        if (Fn < MIN_FLUX_NORM) dir = Tensor2(1, 0);
        if (!(Fn < MIN_FLUX_NORM)) {
            rotationAngle[ij] = adaptive ? std::fmod(MyAtan2(m[2], m[1]) + 2.0 * rtm::kPi, 2.0 * rtm::kPi) : 0.0;
            sourceAngle[ij] = rotationAngle[ij];
        }
        double sigma = stencil.fluxToSigmaTable.Evaluate(relativeF);
        double normalization = stencil.fluxToNormalizationTable.Evaluate(relativeF);
        if (0 <= sigmaOverwrite) {
            sigma = sigmaOverwrite;
            normalization = normalizationOverwrite;
        }
        for (size_t d = 0; d < stencil.nDir; d++)
            I[Index(ij, d)] = E_IF * std::exp(sigma * Tensor2::Dot(dir, stencil.C(d)) - normalization);
    }
}

// Reference: src/Radiation.cpp:191-237 (GPU RK45 backtrace, FP64).
void Radiation::UpdateFourierCoefficients()
{
    if (!IsCurved()) return;
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
    UploadRotation();
    Check(rt_compute_moments(ctx));
    DownloadMoments();
}

// Reference: src/Radiation.cpp:270-301.  The lab-frame transform runs on the
// GPU on the inertial-frame moments of the current I, which also end in
// E..Pyy.
void Radiation::ComputeMomentsLF()
{
    UploadI();
    UploadRotation();
    Check(rt_compute_moments(ctx));
    DownloadMoments();
    DownloadMomentsLF();
}
void Radiation::DownloadMomentsLF()
{
    Check(rt_get_moments_lf(ctx, E_LF.data(), Fx_LF.data(), Fy_LF.data(), Pxx_LF.data(), Pxy_LF.data(),
                            Pyy_LF.data(), F_LF.data()));
}

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

// Reference: src/Radiation.cpp:361-378.  The new angles of the interior
// cells come from the 3x3 average of Fx, Fy; the result is rotationAngleNew.
void Radiation::UpdateRotationMatrizes()
{
    if (!IsAdaptive()) ExitOnError("UpdateRotationMatrizes needs a Radiation built with adaptive streaming");
    Check(rt_set_moments(ctx, E.data(), Fx.data(), Fy.data(), Pxx.data(), Pxy.data(), Pyy.data()));
    UploadRotation();
    Check(rt_update_rotation(ctx));
    DownloadRotation();
}

// Reference: src/Radiation.cpp:415-453 and 513-575: stencils turned by
// rotationAngleNew, neighbours read in their rotationAngle frames; I and the
// two angle buffers are swapped afterwards, as in the reference.
void Radiation::StreamFlatAdaptive()
{
    if (config.streamingType != StreamingType::FlatAdaptive)
        ExitOnError("StreamFlatAdaptive needs a Radiation built with StreamingType::FlatAdaptive");
    UploadFields();
    UploadI();
    UploadRotation();
    Check(rt_stream(ctx));
    std::swap(I, Inew);
    DownloadI();
    DownloadRotation();
}
void Radiation::StreamCurvedAdaptive()
{
    if (config.streamingType != StreamingType::CurvedAdaptive)
        ExitOnError("StreamCurvedAdaptive needs a Radiation built with StreamingType::CurvedAdaptive");
    if (!tablesOnDevice) UpdateFourierCoefficients();
    UploadFields();
    UploadI();
    UploadRotation();
    Check(rt_stream(ctx));
    std::swap(I, Inew);
    DownloadI();
    DownloadRotation();
}
void Radiation::StreamGeodesicFixed() { ExitOnError("GeodesicFixed streaming is not part of this library"); }

// Reference: src/Radiation.cpp:641-736.  The lambda iteration starts from the
// moments of I (the reference's ComputeMomentsIF call just before Collide);
// afterwards I holds the collided state and E..Pyy the final moments.
void Radiation::Collide()
{
    if (IsCurved() && !tablesOnDevice) UpdateFourierCoefficients();
    UploadFields();
    UploadI();
    UploadRotation();
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

// Reference: src/Radiation.cpp:738-885.  The steps between two frames run on
// the GPU as one rt_step call (stream, moments and collide fused per step).
// A frame is written after the step at which the time since the last frame
// reaches writePeriod - 1e-8, as in the reference, plus the initial frame
// (saveInitialData) and a final one; each holds the lab-frame E, Fx, Fy and
// |F| (ComputeMomentsIF, ComputeMomentsLF, Grid::WriteFrametoCsv) under
// OUTPUTDIR/<name>/Moments/.  `logger` gets the run's values and times: the
// whole loop as "Total Time" and the CSV writes under the reference's frame
// writer scope, so its computation time excludes frame output, like the
// reference's; Log.txt is written with writeData.
void Radiation::RunSimulation()
{
    LoadInitialData();
    UpdateFourierCoefficients();

    logger.SetValues(config.name, config.simTime);
    config.simTime = logger.simTime;
    const int timeSteps = logger.timeSteps;
    const std::string directoryPath = logger.directoryPath;
    if (config.printSetup) {
        std::cout << " nx           = " << grid.nx << "\n";
        std::cout << " ny           = " << grid.ny << "\n";
        std::cout << " nDir         = " << stencil.nDir << "\n";
        std::cout << " nFourier     = " << streamingStencil.nDir << "\n";
        std::cout << " simTime      = " << config.simTime << "\n";
        std::cout << " writePeriod  = " << config.writePeriod << "\n";
        std::cout << " dx           = " << grid.dx << "\n";
        std::cout << " dy           = " << grid.dy << "\n";
        std::cout << " dt           = " << grid.dt << "\n";
        std::cout << " timeSteps    = " << timeSteps << "\n";
        std::cout << " filesToWrite = " << std::floor(config.simTime / config.writePeriod) << std::endl;
    }

    UploadFields();
    UploadI();
    if (IsAdaptive()) {
        /* UpdateRotationMatrizes of the first step reads the moments held now */
        UploadRotation();
        Check(rt_set_moments(ctx, E.data(), Fx.data(), Fy.data(), Pxx.data(), Pxy.data(), Pyy.data()));
    }
    if (config.keepSourceNodesActive) {
        /* the source cells' initial intensities (and, for adaptive streaming,
         * their angles) are re-imposed on the GPU before every step */
        std::vector<uint8_t> src(grid.nxy);
        bool any = false;
        for (size_t ij = 0; ij < grid.nxy; ij++) {
            src[ij] = isInitialGridPoint[ij] ? 1 : 0;
            any = any || isInitialGridPoint[ij];
        }
        Check(rt_set_sources(ctx, src.data(), I.data()));
        if (IsAdaptive() && any) Check(rt_set_source_rotation(ctx, sourceAngle.data()));
    }
    Check(rt_reset_iteration_stats(ctx));
    using clock = std::chrono::steady_clock;
    const auto loopStart = clock::now();
    double currentTime = config.t0, timeSinceLastFrame = 0, computeSeconds = 0, writeSeconds = 0;
    auto frame = [&]() {
        Check(rt_compute_moments(ctx));
        DownloadMoments();
        DownloadMomentsLF();
        const auto w0 = clock::now();
        grid.WriteFrametoCsv((float)currentTime, E_LF, Fx_LF, Fy_LF, F_LF, directoryPath + "/Moments/");
        writeSeconds += std::chrono::duration<double>(clock::now() - w0).count();
        timeSinceLastFrame = 0;
    };
    if (config.writeData && config.saveInitialData) frame();
    const bool perStep = config.updateFourierHarmonics;
    for (int n = 0; n < timeSteps;) {
        /* the steps up to the next frame, or to the end */
        int k = 0;
        bool writeNow = false;
        while (n + k < timeSteps && !writeNow) {
            if (config.printProgress) std::cout << "\n" << n + k << "," << Format(currentTime, 4) << "," << std::flush;
            k++;
            currentTime += grid.dt;
            timeSinceLastFrame += grid.dt;
            writeNow = config.writeData && timeSinceLastFrame >= config.writePeriod - 1e-8;
        }
        const int last = (n + k == timeSteps) ? 1 : 0;
        const auto t0 = clock::now();
        if (!perStep) {
            Check(rt_step(ctx, k, last));
        } else {
            // per-step table rebuild of the reference loop (src/Radiation.cpp:790-791),
            // from the metric block kept on the device
            for (int s = 0; s < k; s++) {
                if (IsCurved()) Check(rt_rebuild_fourier_tables(ctx, config.rk45Tolerance));
                Check(rt_step(ctx, 1, last && s == k - 1 ? 1 : 0));
            }
        }
        Check(rt_synchronize(ctx));
        computeSeconds += std::chrono::duration<double>(clock::now() - t0).count();
        n += k;
        if (writeNow) frame();
    }
    if (timeSteps > 0 && !(config.writeData && timeSinceLastFrame == 0)) DownloadMoments();
    if (config.writeData && timeSinceLastFrame != 0) frame();
    DownloadI();
    DownloadRotation();
    const double totalSeconds = std::chrono::duration<double>(clock::now() - loopStart).count();

    int32_t mx = 0;
    double sum = 0;
    Check(rt_iteration_stats(ctx, &mx, &sum));
    maxItterationCount = mx;
    averageItterationCount = timeSteps > 0 ? sum / ((double)timeSteps * (grid.nx - 2) * (grid.ny - 2)) : 0.0;
    logger.maxItterationCount = maxItterationCount;
    logger.averageItterationCount = averageItterationCount;
    logger.AddTimeMeasurement("Total Time", totalSeconds);
    logger.AddTimeMeasurement(Logger::kFrameWriterScope, writeSeconds);
    logger.AddTimeMeasurement("GPU steps (rt_step)", computeSeconds);
    std::cout << "total time: " << totalSeconds << "s" << std::endl;
    if (config.printResults) {
        std::cout << "Max Lambda Itteration Count     = " << maxItterationCount << std::endl;
        std::cout << "Average Lambda Itteration Count = " << averageItterationCount << std::endl;
        std::cout << "Computation Time: " << logger.ComputationTime() << "s" << std::endl;
        std::cout << "Benchmark: " << grid.nxy * timeSteps / (1e6 * logger.ComputationTime()) << "MLUPS" << std::endl;
    }
    if (config.writeData) logger.Write();
    if (config.keepSourceNodesActive) Check(rt_set_sources(ctx, nullptr, nullptr));
}
