/*
 * Scalar math shared by the host C++ classes and the gfx950 kernels.
 *
 * Hastings (1955) polynomial approximations of order 9, as used by the
 * reference for every angle on the hot path (src/Utility.hh:133-360,
 * order fixed by APPROXIMATION_ORDER in src/ControlFlow.hh:15):
 *   atan(z)  = z*(a1 + z^2*(a3 + z^2*(a5 + z^2*(a7 + z^2*a9))))
 *   sin(x)   = x'*(b1 + x'^2*(...)), x' = 2x/pi after folding x into [-pi/2, pi/2]
 *   cos(x)   = sin(x +- pi/2) evaluated through float (the reference's MyCos
 *              takes and returns float, so stencil x-cosines are float-rounded)
 *   atan2    = octant reduction with the float comparison |x|>|y|.
 */
#ifndef RT_MATH_H
#define RT_MATH_H
#include <math.h>

#if defined(__HIPCC__)
#define RT_HD __host__ __device__
#else
#define RT_HD
#endif

namespace rtm {

constexpr double kPi = 3.14159265358979323846;
constexpr double kPi2 = 1.57079632679489661923;
constexpr double kSqrt2 = 1.4142135623730951;

/* Hastings atan, order 9 (src/Utility.hh:179-188). */
template <typename T>
RT_HD inline T hastings_atan(T z)
{
    const T a1 = T(0.9998660), a3 = T(-0.3302995), a5 = T(0.1801410), a7 = T(-0.0851330),
            a9 = T(0.0208351);
    T zz = z * z;
    return z * (a1 + zz * (a3 + zz * (a5 + zz * (a7 + zz * a9))));
}

/* MyAtan2 (src/Utility.hh:220-254), double arithmetic. */
RT_HD inline double my_atan2(double y, double x)
{
    if (x != 0.0) {
        if (fabsf((float)x) > fabsf((float)y)) {
            const double z = y / x;
            if (x > 0.0) return hastings_atan(z);
            else if (y >= 0.0) return hastings_atan(z) + kPi;
            else return hastings_atan(z) - kPi;
        } else {
            const double z = x / y;
            if (y > 0.0) return -hastings_atan(z) + kPi2;
            else return -hastings_atan(z) - kPi2;
        }
    } else {
        if (y > 0.0) return kPi2;
        else if (y < 0.0) return -kPi2;
    }
    return 0.0;
}

/* Tensor2::Phi (src/DataTypes.hh:166-170): angle in [0, 2pi). */
RT_HD inline double my_phi(double x, double y)
{
    double p = my_atan2(y, x);
    return (p < 0) ? p + 2.0 * kPi : p;
}

/* MySin (src/Utility.hh:272-323) for x in [-pi/2, 5pi/2]; NaN outside
 * (the reference aborts there). */
RT_HD inline double my_sin(double x)
{
    if (-kPi2 <= x && x <= kPi2) {
    } else if (kPi2 <= x && x <= 3.0 * kPi2) x = -(x - kPi);
    else if (3.0 * kPi2 <= x && x <= 5.0 * kPi2) x = x - 2.0 * kPi;
    else return NAN;
    x *= 2.0 / kPi;
    const double a1 = 1.57079631847, a3 = -0.64596371106, a5 = 0.07968967928,
                 a7 = -0.00467376557, a9 = 0.00015148419;
    double xx = x * x;
    return x * (a1 + xx * (a3 + xx * (a5 + xx * (a7 + xx * a9))));
}

/* MyCos (src/Utility.hh:340-359): float argument folding, float result. */
RT_HD inline double my_cos(double xd)
{
    const float pi = (float)kPi, pi_2 = (float)kPi2;
    float x = (float)xd;
    if (x >= pi) {
        float t = x - pi_2;
        return (double)(float)(-my_sin((double)t));
    }
    float t = x + pi_2;
    return (double)(float)my_sin((double)t);
}

/* Fourier::Basis (src/FourierHarmonics.cpp:5-15): 1, sqrt2 cos(phi),
 * sqrt2 sin(phi), sqrt2 cos(2phi), sqrt2 sin(2phi), ... */
RT_HD inline double fourier_basis(int k, double phi)
{
    if (k == 0) return 1.0;
    else if (k % 2 != 0) return kSqrt2 * cos(phi * (k + 1.0) / 2.0);
    else return kSqrt2 * sin(phi * k / 2.0);
}

/* BilinearInterpolation (src/Interpolation.cpp:27-37), operation order kept. */
template <typename T>
RT_HD inline T bilinear(T x, T y, T f00, T f01, T f10, T f11)
{
    T f0 = f00 * (T(1) - x) + f10 * x;
    T f1 = f01 * (T(1) - x) + f11 * x;
    return f0 * (T(1) - y) + f1 * y;
}

} // namespace rtm
#endif
