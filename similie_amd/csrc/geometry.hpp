// Per-node geometry of the discrete Hodge star, device side.
//
// The reference builds ⋆ per node from the position and metric tensors
// through reconstruction → continuous star → circumcentric reduction
// (src/similie/exterior/hodge_star.hpp:224-358,
// reduction_and_reconstruction.hpp:535-647); for every non-degenerate node
// that chain equals the volume ratio sign(I‖J)·V_J/V_I of the legacy form
// (reduction_and_reconstruction.hpp:403-432), V_S = sqrt|det Gram_S| of the
// metric-weighted edge vectors (volume.hpp:55-155).  This header evaluates
// that ratio directly, with the reference's degeneracy rules: a node whose
// reconstruction Jacobian of any |I|-subset is below 1e-14 in magnitude, or
// whose metric has a zero Gauss–Jordan pivot, gets zero coefficients.
#pragma once

#include "common.hpp"

namespace sil {

__device__ inline int64_t flat_index(const DevGeom& g, const int64_t* idx)
{
    int64_t lin = 0;
    for (int a = 0; a < g.ndim; ++a) lin = lin * g.ext[a] + idx[a];
    return lin;
}

// forward edge, backward at the last node, zero on a single-node axis
// (volume.hpp:55-82) — rectilinear form
__device__ inline double axis_spacing(const DevGeom& g, int a, int64_t i)
{
    const double* c = g.axis[a];
    int64_t n = g.ext[a];
    if (i + 1 < n) return c[i + 1] - c[i];
    if (i > 0) return c[i] - c[i - 1];
    return 0.0;
}

struct NodeGeo {
    int n;
    bool simple;     // rectilinear with identity / diagonal metric
    double s[4];     // simple: sqrt(g_aa)·h_a
    double h[4];     // simple: raw spacings
    double e[4][4];  // general: edge vectors
    double m[16];    // general: metric
    bool metric_ok;
};

__device__ inline bool gauss_jordan_ok(int n, const double* a_in)
{
    // zero-pivot detection of misc::math::invert (small_matrix.hpp:116-180)
    double w[16];
    for (int i = 0; i < n * n; ++i) w[i] = a_in[i];
    for (int i = 0; i < n; ++i) {
        int p = i;
        double pa = fabs(w[i * n + i]);
        for (int r = i + 1; r < n; ++r)
            if (fabs(w[r * n + i]) > pa) {
                p = r;
                pa = fabs(w[r * n + i]);
            }
        if (pa == 0.0) return false;
        if (p != i)
            for (int c = 0; c < n; ++c) {
                double t = w[i * n + c];
                w[i * n + c] = w[p * n + c];
                w[p * n + c] = t;
            }
        double d = w[i * n + i];
        for (int c = 0; c < n; ++c) w[i * n + c] /= d;
        for (int r = 0; r < n; ++r) {
            if (r == i) continue;
            double f = w[r * n + i];
            for (int c = 0; c < n; ++c) w[r * n + c] -= f * w[i * n + c];
        }
    }
    return true;
}

__device__ inline double lu_det(int n, const double* a_in)
{
    if (n == 0) return 1.0;
    double a[16];
    for (int i = 0; i < n * n; ++i) a[i] = a_in[i];
    double det = 1.0;
    int sign = 1;
    for (int i = 0; i < n; ++i) {
        int p = i;
        double pa = fabs(a[i * n + i]);
        for (int r = i + 1; r < n; ++r)
            if (fabs(a[r * n + i]) > pa) {
                p = r;
                pa = fabs(a[r * n + i]);
            }
        if (pa == 0.0) return 0.0;
        if (p != i) {
            sign = -sign;
            for (int c = 0; c < n; ++c) {
                double t = a[i * n + c];
                a[i * n + c] = a[p * n + c];
                a[p * n + c] = t;
            }
        }
        double d = a[i * n + i];
        det *= d;
        for (int r = i + 1; r < n; ++r) {
            double f = a[r * n + i] / d;
            for (int c = i + 1; c < n; ++c) a[r * n + c] -= f * a[i * n + c];
        }
    }
    return sign * det;
}

__device__ inline double cofactor_det(int k, const double* a)
{
    // k ≤ 4, expansion along the first row (reduction_and_reconstruction.hpp:163-187)
    if (k == 0) return 1.0;
    if (k == 1) return a[0];
    if (k == 2) return a[0] * a[3] - a[1] * a[2];
    if (k == 3)
        return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
               a[2] * (a[3] * a[7] - a[4] * a[6]);
    double det = 0.0;
    for (int col = 0; col < 4; ++col) {
        double mnr[9];
        int q = 0;
        for (int r = 1; r < 4; ++r)
            for (int c = 0; c < 4; ++c)
                if (c != col) mnr[q++] = a[r * 4 + c];
        double sub = mnr[0] * (mnr[4] * mnr[8] - mnr[5] * mnr[7]) - mnr[1] * (mnr[3] * mnr[8] - mnr[5] * mnr[6]) +
                     mnr[2] * (mnr[3] * mnr[7] - mnr[4] * mnr[6]);
        det += ((col % 2 == 0) ? 1.0 : -1.0) * a[col] * sub;
    }
    return det;
}

__device__ inline int sym_index(int n, int i, int j)
{
    if (i > j) {
        int t = i;
        i = j;
        j = t;
    }
    int off = 0;
    for (int r = 0; r < i; ++r) off += n - r;
    return off + (j - i);
}

__device__ inline void load_node_geo(const DevGeom& g, const int64_t* idx, NodeGeo& ng)
{
    int n = g.ndim;
    ng.n = n;
    int64_t lin = flat_index(g, idx);
    ng.simple = g.rectilinear && g.metric_kind != SIL_METRIC_SYMMETRIC;
    if (ng.simple) {
        ng.metric_ok = true;
        for (int a = 0; a < n; ++a) {
            double h = axis_spacing(g, a, idx[a]);
            double gm = (g.metric_kind == SIL_METRIC_DIAGONAL) ? g.metric[lin * n + a] : 1.0;
            if (gm == 0.0) ng.metric_ok = false;
            ng.h[a] = h;
            ng.s[a] = sqrt(fabs(gm)) * fabs(h);
        }
        return;
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double v;
            if (g.metric_kind == SIL_METRIC_IDENTITY)
                v = (i == j) ? 1.0 : 0.0;
            else if (g.metric_kind == SIL_METRIC_DIAGONAL)
                v = (i == j) ? g.metric[lin * n + i] : 0.0;
            else
                v = g.metric[lin * (n * (n + 1) / 2) + sym_index(n, i, j)];
            ng.m[i * n + j] = v;
        }
    ng.metric_ok = gauss_jordan_ok(n, ng.m);
    for (int a = 0; a < n; ++a) {
        int64_t f[4], b[4];
        for (int q = 0; q < n; ++q) f[q] = b[q] = idx[q];
        f[a] += 1;
        b[a] -= 1;
        double c[4], o[4];
        bool hf = f[a] < g.ext[a], hb = b[a] >= 0;
        if (g.rectilinear) {
            for (int q = 0; q < n; ++q) {
                c[q] = 0.0;
                o[q] = 0.0;
            }
            double h = axis_spacing(g, a, idx[a]);
            for (int q = 0; q < n; ++q) ng.e[a][q] = (q == a) ? h : 0.0;
            continue;
        }
        const double* P = g.nodal;
        for (int q = 0; q < n; ++q) c[q] = P[lin * n + q];
        if (hf) {
            int64_t lf = flat_index(g, f);
            for (int q = 0; q < n; ++q) o[q] = P[lf * n + q];
            for (int q = 0; q < n; ++q) ng.e[a][q] = o[q] - c[q];
        } else if (hb) {
            int64_t lb = flat_index(g, b);
            for (int q = 0; q < n; ++q) o[q] = P[lb * n + q];
            for (int q = 0; q < n; ++q) ng.e[a][q] = c[q] - o[q];
        } else {
            for (int q = 0; q < n; ++q) ng.e[a][q] = 0.0;
        }
    }
}

__device__ inline double node_volume(const NodeGeo& ng, unsigned mask)
{
    int n = ng.n;
    if (ng.simple) {
        double v = 1.0;
        for (int a = 0; a < n; ++a)
            if (mask & (1u << a)) v *= ng.s[a];
        return v;
    }
    int ids[4], k = 0;
    for (int a = 0; a < n; ++a)
        if (mask & (1u << a)) ids[k++] = a;
    if (k == 0) return 1.0;
    double gram[16];
    for (int i = 0; i < k; ++i)
        for (int j = i; j < k; ++j) {
            double v = 0.0;
            for (int p = 0; p < n; ++p)
                for (int q = 0; q < n; ++q) v += ng.e[ids[i]][p] * ng.m[p * n + q] * ng.e[ids[j]][q];
            gram[i * k + j] = v;
            gram[j * k + i] = v;
        }
    return sqrt(fabs(lu_det(k, gram)));
}

// any |I|-subset reconstruction Jacobian below 1e-14, or a singular metric
__device__ inline bool node_degenerate(const NodeGeo& ng, int k)
{
    if (!ng.metric_ok) return true;
    if (k == 0) return false;
    int n = ng.n;
    int cnt = c_combos.count[n][k];
    for (int c = 0; c < cnt; ++c) {
        unsigned m = c_combos.mask[n][k][c];
        int ids[4], q = 0;
        for (int a = 0; a < n; ++a)
            if (m & (1u << a)) ids[q++] = a;
        double jd;
        if (ng.simple) {
            jd = 1.0;
            for (int i = 0; i < k; ++i) jd *= ng.h[ids[i]];
        } else {
            double jac[16];
            for (int col = 0; col < k; ++col)
                for (int row = 0; row < k; ++row) jac[row * k + col] = ng.e[ids[col]][ids[row]];
            jd = cofactor_det(k, jac);
        }
        if (fabs(jd) < 1e-14) return true;
    }
    return false;
}

// coefficient of ⋆ from the axes set `src` onto its complement
__device__ inline double node_hodge(const NodeGeo& ng, unsigned src)
{
    int n = ng.n;
    unsigned all = (1u << n) - 1u;
    unsigned tgt = all & ~src;
    if (node_degenerate(ng, popcount4(src))) return 0.0;
    double vi = node_volume(ng, src);
    if (vi == 0.0) return 0.0;
    return concat_sign(src, tgt, n) * node_volume(ng, tgt) / vi;
}

}  // namespace sil
