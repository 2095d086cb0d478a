// Interpolated B-H curve, device side (InterpolatedNonlinearBHCurve,
// src/similie/physics/magnetostatics/nonlinear_magnetostatics.py:14-181):
// h(b) is linear between samples, bracketed on q = b², constant h₀ below
// the first sample and extrapolated along the last segment past the end.
#pragma once

#include "operators.hpp"

namespace sil {

__device__ inline int bracket_q(const BhCurve& c, double q)
{
    if (q <= c.q[0]) return 0;
    for (int i = 0; i + 1 < c.n; ++i)
        if (q <= c.q[i + 1]) return i;
    return c.n - 2;
}

__device__ inline double h_from_b(const BhCurve& c, double b)
{
    if (b <= c.b[0]) return c.h[0];
    if (b >= c.b[c.n - 1]) {
        const int i = c.n - 2;
        return c.h[i] + c.dh_db[i] * (b - c.b[i]);
    }
    const int i = bracket_q(c, b * b);
    return c.h[i] + c.dh_db[i] * (b - c.b[i]);
}

__device__ inline double dh_db_from_b(const BhCurve& c, double b)
{
    if (b <= c.b[0]) return c.dh_db[0];
    if (b >= c.b[c.n - 1]) return c.dh_db[c.n - 2];
    return c.dh_db[bracket_q(c, b * b)];
}

}  // namespace sil
