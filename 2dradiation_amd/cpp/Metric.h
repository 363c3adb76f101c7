// Metric data on the grid (reference: src/Metric.h, src/Metric.cpp:126-336).
// Arrays keep the reference names.  The time-derivative arrays d0_* are
// identically zero for the static spacetimes here and are not stored.
#ifndef RT2D_METRIC_H
#define RT2D_METRIC_H
#include "Grid.h"

class Metric {
  public:
    Grid &grid;
    double m = 1.0;
    double a = 0.0;

    DoubleBuffer g00_ll, g01_ll, g02_ll, g11_ll, g12_ll, g22_ll;
    DoubleBuffer g00_uu, g01_uu, g02_uu, g11_uu, g12_uu, g22_uu;
    DoubleBuffer d1_g00_lll, d1_g01_lll, d1_g02_lll, d1_g11_lll, d1_g12_lll, d1_g22_lll;
    DoubleBuffer d2_g00_lll, d2_g01_lll, d2_g02_lll, d2_g11_lll, d2_g12_lll, d2_g22_lll;
    DoubleBuffer d1_g00_luu, d1_g01_luu, d1_g02_luu, d1_g11_luu, d1_g12_luu, d1_g22_luu;
    DoubleBuffer d2_g00_luu, d2_g01_luu, d2_g02_luu, d2_g11_luu, d2_g12_luu, d2_g22_luu;
    DoubleBuffer alpha, beta1_u, beta2_u, beta1_l, beta2_l;
    DoubleBuffer gamma11_ll, gamma12_ll, gamma22_ll, gamma11_uu, gamma12_uu, gamma22_uu;
    DoubleBuffer d1_alpha_l, d2_alpha_l;
    DoubleBuffer d1_beta1_lu, d1_beta2_lu, d2_beta1_lu, d2_beta2_lu;
    DoubleBuffer d1_beta1_ll, d1_beta2_ll, d2_beta1_ll, d2_beta2_ll;
    DoubleBuffer d1_gamma11_lll, d1_gamma12_lll, d1_gamma22_lll;
    DoubleBuffer d2_gamma11_lll, d2_gamma12_lll, d2_gamma22_lll;
    DoubleBuffer d1_gamma11_luu, d1_gamma12_luu, d1_gamma22_luu;
    DoubleBuffer d2_gamma11_luu, d2_gamma12_luu, d2_gamma22_luu;
    DoubleBuffer K11_ll, K12_ll, K22_ll;
    DoubleBuffer tetrad00_ul, tetrad01_ul, tetrad02_ul, tetrad10_ul, tetrad11_ul, tetrad12_ul,
        tetrad20_ul, tetrad21_ul, tetrad22_ul;
    DoubleBuffer tetrad00_lu, tetrad01_lu, tetrad02_lu, tetrad10_lu, tetrad11_lu, tetrad12_lu,
        tetrad20_lu, tetrad21_lu, tetrad22_lu;

    Metric(Grid &grid_, double m_, double a_);
    virtual ~Metric() = default;
    virtual std::string Name() = 0;
    virtual Tensor3x3 MetricFunction(const Coord &xy) = 0;
    virtual bool InsideBH(const Coord &xy) = 0;
    // x*x + y*y threshold of InsideBH, negative if nothing is excised.
    virtual double ExcisionR2() = 0;

    double GetAlpha(size_t ij) const { return alpha[ij]; }
    Tensor3x3 GetTetrad(size_t ij) const;
    Tensor3x3 GetTetradInverse(size_t ij) const;
    Tensor2 GetBeta_u(size_t ij) const { return Tensor2(beta1_u[ij], beta2_u[ij]); }
    Tensor3x3 GetMetric_ll(size_t ij) const
    {
        return Tensor3x3(g00_ll[ij], g01_ll[ij], g02_ll[ij], g01_ll[ij], g11_ll[ij], g12_ll[ij], g02_ll[ij], g12_ll[ij],
                         g22_ll[ij]);
    }
    // Values at a point (reference src/Metric.cpp:338-359, 383-405, 541-608):
    // bilinear in the cell arrays inside the domain, flat space outside.
    double InterpolateArrayTo_ij(const DoubleBuffer &array, const Coord &ij) const;
    double GetAlpha(const Coord &xy) const;
    Tensor2 GetBeta_u(const Coord &xy) const;
    Tensor3x3 GetTetrad(const Coord &xy) const;
    Tensor3x3 GetTetradInverse(const Coord &xy) const;
    Tensor3x3 GetMetric_ll(const Coord &xy) const;
    // four-velocity of the Eulerian observer (src/Metric.cpp:368-382)
    Tensor3 uEulObs(size_t ij) const;
    Tensor3 uEulObs(const Coord &xy) const;

    // The remaining getters of the reference (src/Metric.h:86-137,
    // src/Metric.cpp:406-800): the value of cell ij, or at a point the
    // bilinear value inside the domain and flat space outside it (zero
    // derivatives and extrinsic curvature).  Time derivatives are zero for
    // the static spacetimes here.
    Tensor3x3 GetMetric_uu(size_t ij) const;
    Tensor3x3 GetMetric_uu(const Coord &xy) const;
    Tensor3x3 GetMinkowskiMetric_ll(size_t) const { return Tensor3x3(-1, 0, 0, 0, 1, 0, 0, 0, 1); }
    Tensor3x3 GetMinkowskiMetric_ll(const Coord &) const { return Tensor3x3(-1, 0, 0, 0, 1, 0, 0, 0, 1); }
    Tensor3x3 GetMinkowskiMetric_uu(size_t) const { return Tensor3x3(-1, 0, 0, 0, 1, 0, 0, 0, 1); }
    Tensor3x3 GetMinkowskiMetric_uu(const Coord &) const { return Tensor3x3(-1, 0, 0, 0, 1, 0, 0, 0, 1); }
    Tensor3x3x3 GetDerivMetric_lll(size_t ij) const;
    Tensor3x3x3 GetDerivMetric_lll(const Coord &xy) const;
    Tensor3x3x3 GetDerivMetric_luu(size_t ij) const;
    Tensor3x3x3 GetDerivMetric_luu(const Coord &xy) const;
    Tensor2 GetBeta_l(size_t ij) const { return Tensor2(beta1_l[ij], beta2_l[ij]); }
    Tensor2 GetBeta_l(const Coord &xy) const;
    Tensor2x2 GetGamma_ll(size_t ij) const;
    Tensor2x2 GetGamma_ll(const Coord &xy) const;
    Tensor2x2 GetGamma_uu(size_t ij) const;
    Tensor2x2 GetGamma_uu(const Coord &xy) const;
    Tensor2x2 GetMinkowskiGamma_ll(size_t) const { return Tensor2x2(1, 0, 0, 1); }
    Tensor2x2 GetMinkowskiGamma_ll(const Coord &) const { return Tensor2x2(1, 0, 0, 1); }
    Tensor2x2 GetMinkowskiGamma_uu(size_t) const { return Tensor2x2(1, 0, 0, 1); }
    Tensor2x2 GetMinkowskiGamma_uu(const Coord &) const { return Tensor2x2(1, 0, 0, 1); }
    Tensor2 GetDerivAlpha_l(size_t ij) const { return Tensor2(d1_alpha_l[ij], d2_alpha_l[ij]); }
    Tensor2 GetDerivAlpha_l(const Coord &xy) const;
    Tensor2x2 GetDerivBeta_lu(size_t ij) const;
    Tensor2x2 GetDerivBeta_lu(const Coord &xy) const;
    Tensor2x2 GetDerivBeta_ll(size_t ij) const;
    Tensor2x2 GetDerivBeta_ll(const Coord &xy) const;
    Tensor2x2x2 GetDerivGamma_lll(size_t ij) const;
    Tensor2x2x2 GetDerivGamma_lll(const Coord &xy) const;
    Tensor2x2x2 GetDerivGamma_luu(size_t ij) const;
    Tensor2x2x2 GetDerivGamma_luu(const Coord &xy) const;
    Tensor2x2 GetExtrCurv_ll(size_t ij) const;
    Tensor2x2 GetExtrCurv_ll(const Coord &xy) const;
    // Per-cell block [nxy * RT_NMETRIC] consumed by rt_build_fourier_tables.
    void PackGeodesicBlock(double *out) const;

  protected:
    void InitializeMetricOnGrid();
    void InitializeMetricDerivativesOnGrid();
    void InitializeAdmComponentsOnGrid();
    void InitializeBoostedTetradOnGrid();
    void Initialize();
};
#endif
