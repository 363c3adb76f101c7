// Spacetimes (reference: src/Spacetimes.h, src/Spacetimes.cpp:1-111).
#ifndef RT2D_SPACETIMES_H
#define RT2D_SPACETIMES_H
#include "Metric.h"

class Minkowski : public Metric {
  public:
    Minkowski(Grid &grid_, double m_ = 0, double a_ = 0);
    Minkowski(const Minkowski &) = delete;
    bool InsideBH(const Coord &) override { return false; }
    double ExcisionR2() override { return -1.0; }
    std::string Name() override { return "Minkowski"; }
    Tensor3x3 MetricFunction(const Coord &xy) override;
};

class SchwarzSchild : public Metric {
  public:
    SchwarzSchild(Grid &grid_, double m_ = 1, double a_ = 0);
    SchwarzSchild(const SchwarzSchild &) = delete;
    bool InsideBH(const Coord &xy) override { return xy.EuklNormSquared() <= ExcisionR2(); }
    double ExcisionR2() override { return 2.1 * 2.1 * m * m; }
    std::string Name() override { return "SchwarzSchild"; }
    Tensor3x3 MetricFunction(const Coord &xy) override;
};

class KerrSchild : public Metric {
  public:
    KerrSchild(Grid &grid_, double m_ = 1, double a_ = 0);
    KerrSchild(const KerrSchild &) = delete;
    bool InsideBH(const Coord &xy) override { return xy.EuklNormSquared() <= ExcisionR2(); }
    double ExcisionR2() override { return 2.1 * 2.1 * m * m; }
    std::string Name() override { return "KerrSchild"; }
    Tensor3x3 MetricFunction(const Coord &xy) override;
};
#endif
