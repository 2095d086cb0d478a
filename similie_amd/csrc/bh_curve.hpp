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

// f, f′, f″ of f(q) = h(√q)/√q (h_over_b_from_q and its q-derivatives)
__device__ inline void h_over_b_q(const BhCurve& c, double q, double& f, double& f1, double& f2)
{
    if (q == 0.0) {
        f = dh_db_from_b(c, 0.0);
        f1 = 0.0;
        f2 = 0.0;
        return;
    }
    const double b = sqrt(q);
    const double h = h_from_b(c, b), dh = dh_db_from_b(c, b);
    f = h / b;
    const double d1 = (dh * b - h) / (b * b);             // d(h/b)/db
    const double d2 = 2.0 * (h - dh * b) / (b * b * b);   // d²(h/b)/db²
    f1 = d1 / (2.0 * b);
    f2 = d2 / (4.0 * q) - d1 / (4.0 * q * b);
}

}  // namespace sil
