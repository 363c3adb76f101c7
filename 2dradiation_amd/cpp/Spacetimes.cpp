#include "Spacetimes.h"

Minkowski::Minkowski(Grid &grid_, double m_, double a_) : Metric(grid_, m_, a_) { Initialize(); }
Tensor3x3 Minkowski::MetricFunction(const Coord &) { return Tensor3x3(-1, 0, 0, 0, 1, 0, 0, 0, 1); }

// Schwarzschild in Schwarzschild coordinates, Cartesian form
// (reference src/Spacetimes.cpp:30-67).
SchwarzSchild::SchwarzSchild(Grid &grid_, double m_, double a_) : Metric(grid_, m_, a_) { Initialize(); }
Tensor3x3 SchwarzSchild::MetricFunction(const Coord &xy)
{
    double rs = 2.0 * m;
    double x = xy[1], y = xy[2];
    double r2 = x * x + y * y;
    double r = std::sqrt(r2);
    if (!(r > 2.0 * m)) return Tensor3x3(-1, 0, 0, 0, 1, 0, 0, 0, 1);
    Tensor3x3 g(0.0);
    g[{0, 0}] = -1.0 + rs / r;
    g[{1, 1}] = x * x / (r2 - rs * r) + y * y / r2;
    g[{2, 2}] = x * x / r2 + y * y / (r2 - rs * r);
    g[{2, 1}] = g[{1, 2}] = rs * x * y / (r2 * (r - rs));
    return g;
}

// Kerr-Schild form g = eta + 2 H l l (reference src/Spacetimes.cpp:70-111).
KerrSchild::KerrSchild(Grid &grid_, double m_, double a_) : Metric(grid_, m_, a_) { Initialize(); }
Tensor3x3 KerrSchild::MetricFunction(const Coord &xy)
{
    double x = xy[1], y = xy[2];
    double a2 = a * a;
    double R2 = x * x + y * y;
    if (!(R2 > a2)) return Tensor3x3(-1, 0, 0, 0, 1, 0, 0, 0, 1);
    double r2 = R2 - a2;
    double r = std::sqrt(r2);
    double rho2 = r2 + a2;
    double H = m / r;
    const double l[3] = {1, (r * x + a * y) / rho2, (r * y - a * x) / rho2};
    Tensor3x3 g;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            g[{i, j}] = (i == 0 && j == 0) ? -1.0 + 2.0 * H * l[i] * l[j]
                                           : (double)(i == j) + 2.0 * H * l[i] * l[j];
    return g;
}
