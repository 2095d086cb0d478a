/*
 * CPU ORACLE — test infrastructure only (see dec_oracle.c).
 *
 * Reference-structured CPU baseline of the staged 3D 1-form Hodge Laplacian,
 * timed by bench.py on the host cores: the same data flow as the reference's
 * StagedLaplacian for an intermediate-rank cochain
 * (src/similie/exterior/laplacian.hpp:369-647): four per-node ⋆ tables
 * (27 + 27 + 9 + 27 = 90 doubles per node, filled once as
 * fill_discrete_hodge_star does, hodge_star.hpp:453-494) and six passes over
 * the grid per application — dω·⋆ into a dual buffer, dᵀ·⋆ into the output,
 * ⋆ω into a second dual buffer, dᵀ·⋆ into the codifferential buffer, d of it
 * into a buffer, and the final addition — each parallelised with OpenMP over
 * the slowest axis.  Table contraction follows tensor_prod over full natural
 * indices (tensor_prod.hpp:96-221).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

double orc_hodge_coefficient(const void* g, const long* idx, unsigned src, unsigned tgt);

typedef struct {
    int ndim;
    long ext[4];
    int pos_kind;
    const double* axis[4];
    const double* nodal;
    int metric_kind;
    const double* metric;
} orc_geom_view; /* layout of orc_geom in dec_oracle.c */

typedef struct {
    long nx, ny, nz;
    double* h1; /* [node][9][3]  ⋆ of dω (2-form → dual 1-form) */
    double* h2; /* [node][9][3]  dual ⋆ (dual 2-form → 1-form) */
    double* h3; /* [node][3][3]  ⋆ of ω (1-form → dual 2-form) */
    double* h4; /* [node][27]    dual ⋆ (dual 3-form → 0-form) */
    double *dual2, *dual1, *codiff, *cob;
} orc_staged;

static const unsigned M2[3] = {3u, 5u, 6u}; /* XY XZ YZ */
static const unsigned M1[3] = {1u, 2u, 4u}; /* X Y Z */

static int perm_parity(const int* t, int k)
{
    int odd = 0;
    for (int i = 0; i < k; ++i)
        for (int j = i + 1; j < k; ++j)
            if (t[i] > t[j]) odd ^= 1;
    return odd;
}

static int idx2(unsigned m)
{
    for (int i = 0; i < 3; ++i)
        if (M2[i] == m) return i;
    return -1;
}

int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void* orc_staged_lap3d1_create(const long* ext, const double* const* axis, const double* metric_diag)
{
    orc_staged* s = (orc_staged*)calloc(1, sizeof(orc_staged));
    s->nx = ext[0];
    s->ny = ext[1];
    s->nz = ext[2];
    long nn = s->nx * s->ny * s->nz;
    s->h1 = (double*)malloc(sizeof(double) * nn * 27);
    s->h2 = (double*)malloc(sizeof(double) * nn * 27);
    s->h3 = (double*)malloc(sizeof(double) * nn * 9);
    s->h4 = (double*)malloc(sizeof(double) * nn * 27);
    s->dual2 = (double*)malloc(sizeof(double) * nn * 3);
    s->dual1 = (double*)malloc(sizeof(double) * nn * 3);
    s->codiff = (double*)malloc(sizeof(double) * nn);
    s->cob = (double*)malloc(sizeof(double) * nn * 3);
    orc_geom_view g;
    memset(&g, 0, sizeof g);
    g.ndim = 3;
    for (int a = 0; a < 3; ++a) {
        g.ext[a] = ext[a];
        g.axis[a] = axis[a];
    }
    g.metric_kind = metric_diag ? 1 : 0;
    g.metric = metric_diag;
#pragma omp parallel for schedule(static)
    for (long lin = 0; lin < nn; ++lin) {
        long idx[3] = {lin / (s->ny * s->nz), (lin / s->nz) % s->ny, lin % s->nz};
        /* full natural source tuples with the 1/k! · sign(perm) of the filled tensor */
        for (int f = 0; f < 9; ++f) {
            int t[2] = {f / 3, f % 3};
            unsigned src = (1u << t[0]) | (1u << t[1]);
            int uniq = t[0] != t[1];
            double sg = perm_parity(t, 2) ? -0.5 : 0.5;
            for (int q = 0; q < 3; ++q) {
                double v1 = 0.0;
                if (uniq && ((7u & ~src) == M1[q])) v1 = sg * orc_hodge_coefficient(&g, idx, src, M1[q]);
                s->h1[(lin * 9 + f) * 3 + q] = v1;
                s->h2[(lin * 9 + f) * 3 + q] = v1; /* dual 2-form → 1-form: same rule and node */
            }
        }
        for (int f = 0; f < 3; ++f)
            for (int q = 0; q < 3; ++q)
                s->h3[(lin * 3 + f) * 3 + q] =
                    ((7u & ~M1[f]) == M2[q]) ? orc_hodge_coefficient(&g, idx, M1[f], M2[q]) : 0.0;
        double c3 = orc_hodge_coefficient(&g, idx, 7u, 0u);
        for (int f = 0; f < 27; ++f) {
            int t[3] = {f / 9, (f / 3) % 3, f % 3};
            int uniq = t[0] != t[1] && t[0] != t[2] && t[1] != t[2];
            s->h4[lin * 27 + f] = uniq ? (perm_parity(t, 3) ? -1.0 : 1.0) / 6.0 * c3 : 0.0;
        }
    }
    return s;
}

void orc_staged_lap3d1_destroy(void* p)
{
    orc_staged* s = (orc_staged*)p;
    if (!s) return;
    free(s->h1);
    free(s->h2);
    free(s->h3);
    free(s->h4);
    free(s->dual2);
    free(s->dual1);
    free(s->codiff);
    free(s->cob);
    free(s);
}

/* natural (a,b) access of an antisymmetric 2-form stored XY XZ YZ */
static inline double nat2(const double* w, int a, int b)
{
    if (a == b) return 0.0;
    int m = idx2((1u << a) | (1u << b));
    return a < b ? w[m] : -w[m];
}

/* natural (a,b,c) access of the single-component 3-form */
static inline double nat3(double v, int a, int b, int c)
{
    if (a == b || a == c || b == c) return 0.0;
    int t[3] = {a, b, c};
    return perm_parity(t, 3) ? -v : v;
}

void orc_staged_lap3d1_run(void* p, const double* w, double* out, const long* oext)
{
    orc_staged* s = (orc_staged*)p;
    const long nx = s->nx, ny = s->ny, nz = s->nz;
    const long ox = oext[0], oy = oext[1], oz = oext[2];
#define L3(i, j, k) (((i) * ny + (j)) * nz + (k))
    /* pass 1: dω (clamped) → ⋆ → dual2 */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nx; ++i)
        for (long j = 0; j < ny; ++j)
            for (long k = 0; k < nz; ++k) {
                long n = L3(i, j, k);
                long ip = i + 1 < nx ? i + 1 : i, jp = j + 1 < ny ? j + 1 : j, kp = k + 1 < nz ? k + 1 : k;
                double F[3];
                F[0] = (w[L3(ip, j, k) * 3 + 1] - w[n * 3 + 1]) - (w[L3(i, jp, k) * 3 + 0] - w[n * 3 + 0]);
                F[1] = (w[L3(ip, j, k) * 3 + 2] - w[n * 3 + 2]) - (w[L3(i, j, kp) * 3 + 0] - w[n * 3 + 0]);
                F[2] = (w[L3(i, jp, k) * 3 + 2] - w[n * 3 + 2]) - (w[L3(i, j, kp) * 3 + 1] - w[n * 3 + 1]);
                for (int q = 0; q < 3; ++q) {
                    double acc = 0.0;
                    for (int f = 0; f < 9; ++f) acc += s->h1[(n * 9 + f) * 3 + q] * nat2(F, f / 3, f % 3);
                    s->dual2[n * 3 + q] = acc;
                }
            }
    /* pass 2: dᵀ (zero outside) → dual ⋆ → out (δd) */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < ox; ++i)
        for (long j = 0; j < oy; ++j)
            for (long k = 0; k < oz; ++k) {
                long n = L3(i, j, k);
                const double* e = s->dual2;
                double ex[3] = {0, 0, 0}, ey[3] = {0, 0, 0}, ez[3] = {0, 0, 0};
                if (i > 0) memcpy(ex, e + L3(i - 1, j, k) * 3, sizeof ex);
                if (j > 0) memcpy(ey, e + L3(i, j - 1, k) * 3, sizeof ey);
                if (k > 0) memcpy(ez, e + L3(i, j, k - 1) * 3, sizeof ez);
                const double* e0 = e + n * 3;
                double G[3];
                G[0] = (e0[1] - ex[1]) - (e0[0] - ey[0]);
                G[1] = (e0[2] - ex[2]) - (e0[0] - ez[0]);
                G[2] = (e0[2] - ey[2]) - (e0[1] - ez[1]);
                long o = ((i * oy + j) * oz + k) * 3;
                for (int q = 0; q < 3; ++q) {
                    double acc = 0.0;
                    for (int f = 0; f < 9; ++f) acc += s->h2[(n * 9 + f) * 3 + q] * nat2(G, f / 3, f % 3);
                    out[o + q] = acc;
                }
            }
    /* pass 3: ⋆ω → dual1 (dual 2-form, XY XZ YZ) */
#pragma omp parallel for schedule(static)
    for (long n = 0; n < nx * ny * nz; ++n)
        for (int q = 0; q < 3; ++q) {
            double acc = 0.0;
            for (int f = 0; f < 3; ++f) acc += s->h3[(n * 3 + f) * 3 + q] * w[n * 3 + f];
            s->dual1[n * 3 + q] = acc;
        }
    /* pass 4: dᵀ → dual 3-form → dual ⋆, sign −1 → codiff (full domain) */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nx; ++i)
        for (long j = 0; j < ny; ++j)
            for (long k = 0; k < nz; ++k) {
                long n = L3(i, j, k);
                const double* e = s->dual1;
                double bx = i > 0 ? e[L3(i - 1, j, k) * 3 + 2] : 0.0;
                double by = j > 0 ? e[L3(i, j - 1, k) * 3 + 1] : 0.0;
                double bz = k > 0 ? e[L3(i, j, k - 1) * 3 + 0] : 0.0;
                double H = (e[n * 3 + 2] - bx) - (e[n * 3 + 1] - by) + (e[n * 3 + 0] - bz);
                double acc = 0.0;
                for (int f = 0; f < 27; ++f) acc += s->h4[n * 27 + f] * nat3(H, f / 9, (f / 3) % 3, f % 3);
                s->codiff[n] = -acc;
            }
    /* pass 5: d(δω), clamped → cob on the output domain */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < ox; ++i)
        for (long j = 0; j < oy; ++j)
            for (long k = 0; k < oz; ++k) {
                long n = L3(i, j, k);
                long ip = i + 1 < nx ? i + 1 : i, jp = j + 1 < ny ? j + 1 : j, kp = k + 1 < nz ? k + 1 : k;
                long o = ((i * oy + j) * oz + k) * 3;
                s->cob[o + 0] = s->codiff[L3(ip, j, k)] - s->codiff[n];
                s->cob[o + 1] = s->codiff[L3(i, jp, k)] - s->codiff[n];
                s->cob[o + 2] = s->codiff[L3(i, j, kp)] - s->codiff[n];
            }
    /* pass 6: Δ = δd + dδ */
#pragma omp parallel for schedule(static)
    for (long t = 0; t < ox * oy * oz * 3; ++t) out[t] += s->cob[t];
#undef L3
}
