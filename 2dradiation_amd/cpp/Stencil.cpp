#include "Stencil.h"
#include <algorithm>

// InterpolationGrid (reference src/Stencil.cpp:4-38, 109-162): for every fine
// angle 2*pi*i/nGrid, the four stencil directions around the nearest one and
// their cubic Lagrange weights, with angles rotated so the fine angle sits at pi.
InterpolationGrid::InterpolationGrid(size_t nGrid_, const Stencil &s) : nGrid(nGrid_)
{
    const double twoPi = 2.0 * rtm::kPi;
    const long n = (long)s.nDir;
    neighbourIndexesCubic.resize(nGrid);
    neighbourWeightsCubic.resize(nGrid);
    for (size_t i = 0; i < nGrid; i++) {
        double phi = 2.0 * rtm::kPi * i / nGrid;
        long index = -1;
        double minDist = 10;
        for (long d = 0; d < n; d++) {
            double delta = std::abs(phi - s.Phi(d));
            double dist = std::min(delta, twoPi - delta);
            if (dist < minDist) {
                minDist = dist;
                index = d;
            }
        }
        double phiMin = std::fmod(s.Phi(index) + rtm::kPi - phi, twoPi);
        long first = (phiMin < rtm::kPi) ? index - 1 : index - 2;
        std::array<size_t, 4> nb;
        for (int k = 0; k < 4; k++) nb[k] = (size_t)(((first + k) % n + n) % n);
        std::array<double, 4> wt;
        for (int a = 0; a < 4; a++) {
            double weight = 1;
            double phi_a = std::fmod(s.Phi(nb[a]) + rtm::kPi - phi + twoPi, twoPi);
            for (int b = 0; b < 4; b++) {
                double phi_b = std::fmod(s.Phi(nb[b]) + rtm::kPi - phi + twoPi, twoPi);
                if (a != b) weight *= (rtm::kPi - phi_b) / (phi_a - phi_b);
            }
            wt[a] = weight;
        }
        neighbourIndexesCubic[i] = nb;
        neighbourWeightsCubic[i] = wt;
    }
}
double InterpolationGrid::d(double phi) const
{
    return std::clamp(nGrid * phi / (2.0 * rtm::kPi), 1e-8, nGrid - 1e-8);
}

// Stencil (reference src/Stencil.cpp:177-297) without ghost directions:
// uniform weights 1/n at angles 2*pi*(d+1/2)/n, x-cosines through MyCos.
Stencil::Stencil(size_t nOrder_, int nGhost_, bool isIntensityStencil)
{
    if (nGhost_ != 0)
        ExitOnError("Stencil with nGhost>0 (adaptive streaming) is not supported by this library.");
    if (nOrder_ < 4 && isIntensityStencil) ExitOnError("Intensity stencil needs at least 4 directions.");
    name = "Stencil" + std::to_string(nOrder_) + "." + std::to_string(nGhost_);
    nDir = nOrder_;
    nGhost = 0;
    nOrder = nOrder_;
    nCoefficients = nOrder_;
    w.resize(nDir);
    phi.resize(nDir);
    cx.resize(nDir);
    cy.resize(nDir);
    for (size_t d = 0; d < nDir; d++) {
        w[d] = 1.0 / nDir;
        phi[d] = 2.0 * rtm::kPi * (d + 0.5) / nDir;
        cx[d] = MyCos(phi[d]);
        cy[d] = MySin(phi[d]);
    }
    if (isIntensityStencil) {
        interpolationGrid = InterpolationGrid(10 * nOrder, *this);
        PopulateLookUpTable();
        sigmaMax = fluxToSigmaTable.outputs[fluxToSigmaTable.size - 1];
        relativeFluxMax = fluxToSigmaTable.inputs[fluxToSigmaTable.size - 1];
    }
}

// Stencil::PopulateLookUpTable (reference src/Stencil.cpp:378-455): walk the
// von Mises concentration sigma of a unit-energy distribution along x up in
// steps of 10^-refinement.  A sigma is accepted, and (relative flux, sigma)
// and (relative flux, log-normalisation) enter the tables, while the mean
// relative error of the cubic angular interpolation at the fine angles within
// deltaPhi/2 of the flux direction stays <= maxInterpolationError; otherwise
// the step is undone and refined, until the refinement reaches 8.
void Stencil::PopulateLookUpTable()
{
    const Tensor2 F(1, 0);
    std::vector<double> I(nDir);
    double sigma = 0;
    int refinement = 1;
    while (true) {
        double normalization = 0;
        for (size_t d = 0; d < nDir; d++) normalization += W(d) * std::exp(sigma * Tensor2::Dot(F, C(d)));
        normalization = std::log(normalization);

        double Fx = 0, Fy = 0;
        for (size_t d = 0; d < nDir; d++) {
            I[d] = std::exp(sigma * Tensor2::Dot(F, C(d)) - normalization);
            const double c = W(d) * I[d];
            Fx += c * Cx(d);
            Fy += c * Cy(d);
        }
        const double relativeF = Tensor2(Fx, Fy).EuklNorm();

        double error = 0;
        int count = 0;
        for (size_t g = 0; g < interpolationGrid.nGrid; g++) {
            if (MyAcos(Tensor2::Dot(F, interpolationGrid.C(g))) > deltaPhi / 2.0) continue;
            const double exact = std::exp(sigma * Tensor2::Dot(F, interpolationGrid.C(g)) - normalization);
            double cubic = 0;
            for (int p = 0; p < 4; p++)
                cubic += interpolationGrid.neighbourWeightsCubic[g][p] * I[interpolationGrid.neighbourIndexesCubic[g][p]];
            error += std::abs((exact - cubic) / exact);
            count++;
        }
        error /= count;

        if (error > maxInterpolationError) {
            sigma -= std::pow(10, -refinement);
            refinement++;
            sigma += std::pow(10, -refinement);
            if (refinement >= 8) break;
        } else {
            fluxToSigmaTable.Add(relativeF, sigma);
            fluxToNormalizationTable.Add(relativeF, normalization);
            sigma += std::pow(10, -refinement);
        }
    }
}
