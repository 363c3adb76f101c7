#include "Stencil.h"
#include <algorithm>

// InterpolationGrid (reference src/Stencil.cpp:4-38, 57-162): for every fine
// angle 2*pi*i/nGrid, the two (linear) and four (cubic) stencil directions
// around the nearest one and their Lagrange weights, with angles rotated so
// the fine angle sits at pi.
InterpolationGrid::InterpolationGrid(size_t nGrid_, const Stencil &s) : nGrid(nGrid_)
{
    const double twoPi = 2.0 * rtm::kPi;
    const long n = (long)s.nDir;
    neighbourIndexesLinear.resize(nGrid);
    neighbourWeightsLinear.resize(nGrid);
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
        /* linear: the nearest direction and its neighbour on the other side
         * of the fine angle */
        const long firstL = (phiMin < rtm::kPi) ? index : index - 1;
        std::array<size_t, 2> nl;
        for (int k = 0; k < 2; k++) nl[k] = (size_t)(((firstL + k) % n + n) % n);
        std::array<double, 2> wl;
        for (int a = 0; a < 2; a++) {
            const double phi_a = std::fmod(s.Phi(nl[a]) + rtm::kPi - phi + twoPi, twoPi);
            const double phi_b = std::fmod(s.Phi(nl[1 - a]) + rtm::kPi - phi + twoPi, twoPi);
            wl[a] = (rtm::kPi - phi_b) / (phi_a - phi_b);
        }
        neighbourIndexesLinear[i] = nl;
        neighbourWeightsLinear[i] = wl;
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

namespace {
// Ghost directions of Stencil(nOrder, nGhost) (reference src/Stencil.cpp:202-266).
// Buckets are the gaps between neighbouring Fourier directions that both lie
// within deltaPhi of pi (or, if there are none, the gap around pi); nGhost
// samples of the arccos distribution over the buckets' span decide how many
// ghosts each bucket gets; they are spread evenly inside their bucket and the
// whole set is turned by pi, so the ghosts sit around the forward direction
// phi = 0 of the rotated stencil.  Returns false if the counts miss nGhost.
bool GhostAngles(size_t nOrder, int nGhost, std::vector<double> &out)
{
    const double twoPi = 2.0 * rtm::kPi;
    auto circ = [&](double a, double b) {
        const double delta = std::abs(a - b);
        return std::min(delta, twoPi - delta);
    };
    std::vector<double> lo, hi;
    for (size_t d = 0; d < nOrder; d++) {
        const double p0 = 2.0 * rtm::kPi * (d + 0.5) / nOrder;
        const double p1 = 2.0 * rtm::kPi * ((d + 1) % nOrder + 0.5) / nOrder;
        if (circ(p0, rtm::kPi) <= Stencil::deltaPhi && circ(p1, rtm::kPi) <= Stencil::deltaPhi) {
            lo.push_back(p0);
            hi.push_back(p1);
        }
    }
    if (lo.empty()) {
        const double first = 2.0 * rtm::kPi * 0.5 / nOrder;
        lo.push_back(rtm::kPi - first);
        hi.push_back(rtm::kPi + first);
    }
    const double from = lo.front(), to = hi.back();
    std::vector<int> count(lo.size(), 0);
    for (int g = 0; g < nGhost; g++) {
        const double q = MyAcos(1.0 - 2.0 * ((g + 0.5) / nGhost)) / rtm::kPi;
        const double angle = from + (to - from) * q;
        for (size_t b = 0; b < lo.size(); b++)
            if (circ(angle, (lo[b] + hi[b]) / 2.0) < (hi[b] - lo[b]) / 2.0) count[b]++;
    }
    out.clear();
    for (size_t b = 0; b < lo.size(); b++)
        for (int k = 0; k < count[b]; k++) {
            const double a = lo[b] + (hi[b] - lo[b]) * (k + 0.5) / count[b] + rtm::kPi;
            out.push_back(std::fmod(a + twoPi, twoPi));
        }
    return (int)out.size() == nGhost;
}
} // namespace

bool Stencil::GhostLayoutValid(size_t nOrder, int nGhost, std::string *why)
{
    if (nGhost < 0) {
        if (why) *why = "nGhost must be >= 0";
        return false;
    }
    if (nGhost == 0) return true;
    if (nOrder % 2 != 0) {
        if (why) *why = "Stencil odd nOrder not allowed when using nGhost>0.";
        return false;
    }
    std::vector<double> g;
    if (!GhostAngles(nOrder, nGhost, g)) {
        if (why) *why = "ghost directions of Stencil(" + std::to_string(nOrder) + "," + std::to_string(nGhost) +
                        ") do not fill their buckets";
        return false;
    }
    return true;
}

// Stencil (reference src/Stencil.cpp:177-297, 299-332): nOrder Fourier
// directions with weights 1/nOrder at angles 2*pi*(d+1/2)/nOrder (x-cosines
// through MyCos), nGhost weight-0 ghost directions, all sorted by angle.
Stencil::Stencil(size_t nOrder_, int nGhost_, bool isIntensityStencil)
{
    if (!isIntensityStencil && nGhost_ > 0) ExitOnError("Stencil with nGhost>0 is only allowed for intensity stencils.");
    if (nOrder_ < 4 && isIntensityStencil) ExitOnError("Intensity stencil needs at least 4 directions.");
    std::string why;
    if (!GhostLayoutValid(nOrder_, nGhost_, &why)) ExitOnError(why);
    name = "Stencil" + std::to_string(nOrder_) + "." + std::to_string(nGhost_);
    nGhost = (size_t)nGhost_;
    nOrder = nOrder_;
    nDir = nOrder + nGhost;
    nCoefficients = nOrder_;
    std::vector<std::array<double, 4>> dirs; // (phi, w, cx, cy)
    for (size_t d = 0; d < nOrder; d++) {
        const double p = 2.0 * rtm::kPi * (d + 0.5) / nOrder;
        dirs.push_back({p, 1.0 / nOrder, MyCos(p), MySin(p)});
    }
    std::vector<double> ghosts;
    if (nGhost > 0) GhostAngles(nOrder, nGhost_, ghosts);
    for (double p : ghosts) dirs.push_back({p, 0.0, MyCos(p), MySin(p)});
    std::stable_sort(dirs.begin(), dirs.end(), [](const auto &a, const auto &b) { return a[0] < b[0]; });
    w.resize(nDir);
    phi.resize(nDir);
    cx.resize(nDir);
    cy.resize(nDir);
    for (size_t d = 0; d < nDir; d++) {
        phi[d] = dirs[d][0];
        w[d] = dirs[d][1];
        cx[d] = dirs[d][2];
        cy[d] = dirs[d][3];
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

void InterpolationGrid::Print() const
{
    std::cout << "        d,\tneighbours,\tweights,\n";
    for (size_t d = 0; d < nGrid; d++) {
        const auto &n = neighbourIndexesCubic[d];
        const auto &w = neighbourWeightsCubic[d];
        std::cout << Format(d) << ",\t(" << n[0] << "," << n[1] << "," << n[2] << "," << n[3] << ")";
        std::cout << ",\t(" << Format(w[0]) << "," << Format(w[1]) << "," << Format(w[2]) << "," << Format(w[3])
                  << "),\n";
    }
    std::cout << std::endl;
}

void Stencil::WriteToCsv() const
{
    std::ofstream out(OUTPUTDIR + name + ".csv");
    out << "#d,w,phi,cx,cy,cz\n";
    for (size_t d = 0; d < nDir; d++)
        out << Format(d) << "," << Format(W(d)) << "," << Format(Phi(d)) << "," << Format(Cx(d)) << "," << Format(Cy(d))
            << ",0\n";
}

void Stencil::Print() const
{
    std::cout << name << ":\n";
    std::cout << "           nDir: " << nDir << "\n";
    std::cout << "         nGhost: " << nGhost << "\n";
    std::cout << "       sigmaMax: " << sigmaMax << "\n";
    std::cout << "relativeFluxMax: " << relativeFluxMax << "\n";
}
void Stencil::PrintAll() const
{
    std::cout << "        d,\t        w,\t      phi,\t       cx,\t       cy,\n";
    for (size_t d = 0; d < nDir; d++)
        std::cout << Format(d) << ",\t" << Format(W(d)) << ",\t" << Format(Phi(d)) << ",\t" << Format(Cx(d)) << ",\t"
                  << Format(Cy(d)) << ",\n";
    std::cout << std::endl;
}
