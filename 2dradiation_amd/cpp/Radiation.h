// Radiation transport driver with the reference's public interface
// (reference: src/Radiation.h, src/Radiation.cpp).  Public members and
// methods keep the reference's names so user programs compile unchanged; the
// work runs on the GPU through the C ABI in include/rt2d.h.
//
// State model: the host buffers (I, E..Pyy, the fluid fields, ...) are the
// user-visible state, as in the reference.  The single-step methods
// (StreamFlatFixed, ComputeMomentsIF, Collide, ...) upload what they read and
// download what they write, so they can be mixed with host code freely.
// RunSimulation uploads once, keeps the state on the device for the whole
// loop and downloads I and the moments at the end.
//
// Frame output (writeData) follows the reference: lab-frame moments computed
// on the GPU, CSV frames under OUTPUTDIR/<name>/Moments/ and Log.txt.
//
// Streaming types: FlatFixed, CurvedFixed (fused kernel where it applies) and
// FlatAdaptive, CurvedAdaptive (per-cell rotated stencils; rotationAngle and
// rotationAngleNew travel to and from the GPU with the state).  Not part of
// this library (exits with a message, like the reference's ExitOnError):
// GeodesicFixed, the reference's performance test.
#ifndef RT2D_RADIATION_H
#define RT2D_RADIATION_H
#include "Config.hh"
#include "Logger.hh"
#include "Metric.h"
#include "SpecialMath.h"
#include "Spacetimes.h"
#include "Stencil.h"

struct rt_ctx;

class Radiation {
  private:
    static constexpr int HALO = 1;
    static constexpr double MIN_FLUX_NORM = 1e-16;
    static constexpr double MIN_ENERGY_DENSITY = 1e-16;
    static constexpr double LAMBDA_ITTERATION_TOLERENCE = 1e-4;
    static constexpr int MAX_LAMBDA_ITERATIONS = 100;

    rt_ctx *ctx = nullptr;
    bool tablesOnDevice = false;
    void Check(int status) const;
    void UploadFields();
    void UploadI();
    void DownloadI();
    void DownloadMoments();
    void DownloadMomentsLF();
    void RequireSupportedStreaming() const;
    bool IsAdaptive() const;
    bool IsCurved() const;
    void UploadRotation();
    void DownloadRotation();
    std::vector<double> sourceAngle; // angle LoadInitialData wrote per cell, NaN: none

  public:
    Grid &grid;
    Metric &metric;
    Stencil &stencil;
    Stencil &streamingStencil;
    Config config;
    Logger logger;
    double sigmaOverwrite = -1;

    bool *isInitialGridPoint;
    DoubleBuffer initialI;
    DoubleBuffer initialFluxAngle_IF;
    DoubleBuffer initialE_LF, initialFx_LF, initialFy_LF;
    DoubleBuffer initialPxx_LF, initialPxy_LF, initialPyy_LF;

    DoubleBuffer rotationAngle, rotationAngleNew;

    DoubleBuffer E, Fx, Fy, Pxx, Pxy, Pyy;
    DoubleBuffer E_LF, Fx_LF, Fy_LF, Pxx_LF, Pxy_LF, Pyy_LF, F_LF;

    DoubleBuffer kappa0, kappa1, kappaA, eta, ux, uy;

    DoubleBuffer I, Inew;

    DoubleBuffer coefficientsS, coefficientsX, coefficientsY, coefficientsCx, coefficientsCy;

    IntBuffer itterationCount;
    int maxItterationCount = 0;
    double averageItterationCount = 0;

    Radiation() = delete;
    Radiation(Metric &metric, Stencil &stencil, Stencil &streamingStencil, Config config);
    Radiation(const Radiation &) = delete;
    Radiation &operator=(const Radiation &) = delete;
    ~Radiation();

    size_t Index(size_t ij, size_t d) const { return d + ij * stencil.nDir; }
    size_t Index(size_t i, size_t j, size_t d) const { return d + grid.Index(i, j) * stencil.nDir; }
    size_t HarmonicIndex(size_t f, size_t ij) const { return f + ij * streamingStencil.nCoefficients; }

    // Host-side helpers of the reference (src/Radiation.cpp:303-359), on the
    // host buffers: the departure point, departure velocity and frequency
    // shift series of cell ij at an angle (from coefficients{X,Y,Cx,Cy,S},
    // which RunSimulation and UpdateFourierCoefficients fill), the angular
    // interpolation of I at a velocity in the cell's frame, and the 3x3 mean
    // of Fx, Fy around an interior cell.
    Coord GetTempCoordinate(size_t ij, double angle) const;
    Tensor2 GetTemp2VelocityIF(size_t ij, double angle) const;
    double GetFrequencyShift(size_t ij, double angle) const;
    double IntensityAt(size_t ij, Tensor2 vTempIF) const;
    Tensor2 AverageF(size_t i, size_t j) const;

    Tensor3 InitialDataLFtoIF(size_t ij) const;
    double NormalizationOverwrite() const;

    void LoadInitialData();
    void UpdateFourierCoefficients();
    void ComputeMomentsIF();
    void ComputeMomentsLF();
    void UpdateRotationMatrizes();

    void StreamFlatFixed();
    void StreamFlatAdaptive();
    void StreamCurvedFixed();
    void StreamCurvedAdaptive();
    void StreamGeodesicFixed();

    void Collide();

    void RunSimulation();

    // The device context, for callers that want the C ABI directly (e.g. to
    // time steps); owned by this object.
    rt_ctx *Context() const { return ctx; }
};
#endif
