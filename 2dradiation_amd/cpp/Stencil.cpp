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
    if (isIntensityStencil) interpolationGrid = InterpolationGrid(10 * nOrder, *this);
}
