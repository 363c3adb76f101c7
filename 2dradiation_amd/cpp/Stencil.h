// Velocity stencil, its cubic angular interpolation grid and the
// flux-to-sigma lookup tables of the von Mises initial data (reference:
// src/Stencil.h, src/Stencil.cpp:4-162, 177-297, 378-455).  Intensity
// stencils may carry nGhost weight-0 ghost directions for adaptive streaming.
#ifndef RT2D_STENCIL_H
#define RT2D_STENCIL_H
#include <array>
#include "DataTypes.hh"

struct Stencil;

struct InterpolationGrid {
    size_t nGrid = 0;
    // per fine angle: the two stencil directions around it and their linear
    // Lagrange weights, and the four around it with cubic weights
    std::vector<std::array<size_t, 2>> neighbourIndexesLinear;
    std::vector<std::array<double, 2>> neighbourWeightsLinear;
    std::vector<std::array<size_t, 4>> neighbourIndexesCubic;
    std::vector<std::array<double, 4>> neighbourWeightsCubic;

    InterpolationGrid() = default;
    InterpolationGrid(size_t nGrid, const Stencil &stencil);
    double d(double phi) const;
    double Phi(size_t d) const { return 2.0 * rtm::kPi * d / nGrid; }
    double Cx(size_t d) const { return MyCos(Phi(d)); }
    double Cy(size_t d) const { return MySin(Phi(d)); }
    Tensor2 C(size_t d) const { return Tensor2(Cx(d), Cy(d)); }
    // InterpolationGrid::Print (reference src/Stencil.cpp:164-173)
    void Print() const;
};

struct Stencil {
    std::string name;
    size_t nDir, nGhost, nOrder, nCoefficients;
    InterpolationGrid interpolationGrid;
    // relative flux |F|/E of a von Mises distribution -> its concentration
    // sigma and log-normalisation (intensity stencils only)
    LookUpTable<double> fluxToSigmaTable;
    LookUpTable<double> fluxToNormalizationTable;
    double sigmaMax = 0;
    double relativeFluxMax = 0;
    static constexpr double deltaPhi = rtm::kPi / 8.0 + 1e-8;
    static constexpr double maxInterpolationError = 0.01;

  private:
    DoubleBuffer w, phi, cx, cy;

  public:
    Stencil(size_t nOrder, int nGhost = 0, bool isIntensityStencil = true);
    // Whether Stencil(nOrder, nGhost) has a well-defined ghost layout (even
    // nOrder, ghost counts per bucket adding up to nGhost); why gets the reason.
    static bool GhostLayoutValid(size_t nOrder, int nGhost, std::string *why = nullptr);
    double W(size_t d) const { return w[d]; }
    double Phi(size_t d) const { return phi[d]; }
    double Cx(size_t d) const { return cx[d]; }
    double Cy(size_t d) const { return cy[d]; }
    Tensor2 C(size_t d) const { return Tensor2(cx[d], cy[d]); }
    const double *WData() const { return w.data(); }
    const double *PhiData() const { return phi.data(); }
    const double *CxData() const { return cx.data(); }
    const double *CyData() const { return cy.data(); }
    // Stencil::Print / PrintAll (reference src/Stencil.cpp:355-369)
    void Print() const;
    void PrintAll() const;
    // Stencil::WriteToCsv (reference src/Stencil.cpp:370-376):
    // OUTPUTDIR<name>.csv with "#d,w,phi,cx,cy,cz" and one line per direction
    void WriteToCsv() const;
    void PopulateLookUpTable();
};
#endif
