// Frame transforms used to set up initial data (reference: src/SpecialMath.h,
// src/SpecialMath.cpp:38-70, 123-187; frame tags IF/LF from
// src/ControlFlow.hh:5-9).  IF is the inertial (tetrad) frame of the local
// Eulerian observer, LF the lab (coordinate) frame.
#ifndef RT2D_SPECIALMATH_H
#define RT2D_SPECIALMATH_H
#include <type_traits>
#include "Metric.h"

class IF {};
class LF {};

// Spatial part of a vector through the tetrad e^a_A (IF -> LF) or its
// inverse (LF -> IF).
inline Tensor2 TransformIFtoLF(const Tensor2 &v, const Tensor3x3 &tetrad)
{
    return Tensor2(tetrad[{1, 1}] * v[1] + tetrad[{1, 2}] * v[2], tetrad[{2, 1}] * v[1] + tetrad[{2, 2}] * v[2]);
}
inline Tensor2 TransformLFtoIF(const Tensor2 &v, const Tensor3x3 &tetradInverse)
{
    return TransformIFtoLF(v, tetradInverse);
}
inline Tensor3 TransformIFtoLF(const Tensor3 &v, const Tensor3x3 &tetrad)
{
    Tensor3 r;
    for (int a = 0; a < 3; a++) r[a] = tetrad[{a, 0}] * v[0] + tetrad[{a, 1}] * v[1] + tetrad[{a, 2}] * v[2];
    return r;
}
inline Tensor3 TransformLFtoIF(const Tensor3 &v, const Tensor3x3 &tetradInverse)
{
    return TransformIFtoLF(v, tetradInverse);
}
// Rank-2 tensors: T'^{ab} = e^a_A e^b_B T^{AB}.
inline Tensor3x3 TransformIFtoLF(const Tensor3x3 &T, const Tensor3x3 &tetrad)
{
    Tensor3x3 r(0.0);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            for (int A = 0; A < 3; A++)
                for (int B = 0; B < 3; B++) r[{a, b}] += T[{A, B}] * tetrad[{a, A}] * tetrad[{b, B}];
    return r;
}
inline Tensor3x3 TransformLFtoIF(const Tensor3x3 &T, const Tensor3x3 &tetradInverse)
{
    return TransformIFtoLF(T, tetradInverse);
}

// Scales the spatial part of (u0, u1, u2) by the positive root x of
// g_{mu nu} w^mu w^nu = 0 with w = (u0, x u1, x u2): a null vector with the
// same time component and spatial direction.
inline Tensor3 NullNormalize(const Tensor3 &u, const Tensor3x3 &g_ll)
{
    double A = 0, B = 0;
    for (int i = 1; i < 3; i++)
        for (int j = 1; j < 3; j++) A += g_ll[{i, j}] * u[i] * u[j];
    for (int i = 1; i < 3; i++) B += 2.0 * g_ll[{0, i}] * u[0] * u[i];
    const double Cc = g_ll[{0, 0}] * u[0] * u[0];
    double x = (-B + std::sqrt(B * B - 4.0 * A * Cc)) / (2.0 * A);
    if (A == 0) x = -Cc / A;
    if (A == 0 && B == 0) x = 1.0;
    return Tensor3(u[0], x * u[1], x * u[2]);
}

// 3-velocity the Eulerian observer measures for the four-velocity u given in
// frame In, expressed in frame Out (at a cell index or a point): from IF,
// v = (u^1, u^2) / alpha; from LF, v = (u^i + beta^i) / alpha; the result is
// carried into Out by the tetrad (IF -> LF) or its inverse (LF -> IF).
template <class FrameIn, class FrameOut, class Where>
Tensor2 Vec2ObservedByEulObs(const Tensor3 &u, const Where &at, Metric &metric)
{
    const double alpha = metric.GetAlpha(at);
    if constexpr (std::is_same<FrameIn, IF>::value) {
        const Tensor2 v(u[1] / alpha, u[2] / alpha);
        if constexpr (std::is_same<FrameOut, IF>::value) return v;
        else return TransformIFtoLF(v, metric.GetTetrad(at));
    } else {
        const Tensor2 beta = metric.GetBeta_u(at);
        const Tensor2 v((u[1] + beta[1]) / alpha, (u[2] + beta[2]) / alpha);
        if constexpr (std::is_same<FrameOut, LF>::value) return v;
        else return TransformLFtoIF(v, metric.GetTetradInverse(at));
    }
}
#endif
