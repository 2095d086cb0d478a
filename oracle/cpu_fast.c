/*
 * CPU ORACLE — test infrastructure only (see dec_oracle.c).
 *
 * CPU baselines timed by bench.py beside the GPU kernels, built on the host
 * that runs them with -O3 -march=native -fopenmp (Makefile target
 * `native`); the parity oracle keeps its own literal build.
 *
 *  - orc_fused_lap3d1: the 3D Hodge Laplacian of a 1-form on a rectilinear
 *    grid with a diagonal metric in one sweep per thread over its own range
 *    of x planes (the same mathematics as the staged passes of
 *    src/similie/exterior/laplacian.hpp:369-647, with the stars as volume
 *    ratios, reduction_and_reconstruction.hpp:403-432): per plane the seven
 *    star coefficients of every node are computed once and the intermediate
 *    forms ⋆ω, ⋆dω and δω are kept in rolling plane buffers.
 *  - orc_fast_poisson2d_cg / orc_fast_ms3d_cg: FP64 conjugate gradients
 *    (x0 = 0, stop on the recurrence residual ‖r‖2 ≤ tol·‖b‖2 as the
 *    reference's Ginkgo call does, minimize_strong_formulation_residual.hpp:
 *    1063-1130), Jacobi or none, every vector operation one OpenMP pass,
 *    with the operators of solver_oracle.c in gather form.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* fused 3D 1-form Laplacian                                           */
/* ------------------------------------------------------------------ */

typedef struct {
    double a1, a2, a3, b1, b2, b3, c;
} coefs7;

/* stars at one node from s_a² = |g_aa| h_a²; degenerate nodes (an edge
 * below 1e-14, a face below 1e-14, a zero metric entry) get zeros */
static inline coefs7 node_coefs(double gx, double gy, double gz, double hx, double hy, double hz)
{
    coefs7 k;
    const int mok = gx != 0.0 && gy != 0.0 && gz != 0.0;
    const double ax = fabs(hx), ay = fabs(hy), az = fabs(hz);
    const int ok1 = mok && ax >= 1e-14 && ay >= 1e-14 && az >= 1e-14;
    const int ok2 = mok && ax * ay >= 1e-14 && ax * az >= 1e-14 && ay * az >= 1e-14;
    const int ok3 = mok && ax * ay * az >= 1e-14;
    const double X = fabs(gx) * ax * ax, Y = fabs(gy) * ay * ay, Z = fabs(gz) * az * az;
    const double v = X * Y * Z;
    const double c = v > 0.0 ? 1.0 / sqrt(v) : 0.0;
    const double c1 = ok1 ? c : 0.0, c2 = ok2 ? c : 0.0;
    k.a1 = X * c2;
    k.a2 = Y * c2;
    k.a3 = Z * c2;
    k.b1 = Y * Z * c1;
    k.b2 = X * Z * c1;
    k.b3 = X * Y * c1;
    k.c = ok3 ? c : 0.0;
    return k;
}

/* forward edge lengths, the backward one at the last node */
static void spacing(const double* c, long n, double* h)
{
    for (long i = 0; i + 1 < n; ++i) h[i] = c[i + 1] - c[i];
    h[n - 1] = n > 1 ? c[n - 1] - c[n - 2] : 0.0;
}

typedef struct {
    coefs7* K;  /* [plane slot][y][z] */
    double* ep; /* ⋆ω: [slot][y][z][3] */
    double* et; /* ⋆dω: [slot][y][z][3] */
    double* dw; /* δω: [slot][y][z] */
} planes;

void orc_fused_lap3d1(const long* ext, const double* const* axis, const double* g, const double* w, double* out,
                      const long* oext)
{
    const long nx = ext[0], ny = ext[1], nz = ext[2];
    const long ox = oext[0], oy = oext[1], oz = oext[2];
    const long pl = ny * nz;
    double* h[3];
    for (int a = 0; a < 3; ++a) {
        h[a] = (double*)malloc(sizeof(double) * ext[a]);
        spacing(axis[a], ext[a], h[a]);
    }
#pragma omp parallel
    {
        int nt = 1, id = 0;
#ifdef _OPENMP
        nt = omp_get_num_threads();
        id = omp_get_thread_num();
#endif
        const long xa = ox * id / nt, xb = ox * (id + 1) / nt;
        planes P;
        P.K = (coefs7*)malloc(sizeof(coefs7) * 2 * pl);
        P.ep = (double*)malloc(sizeof(double) * 2 * pl * 3);
        P.et = (double*)malloc(sizeof(double) * 2 * pl * 3);
        P.dw = (double*)malloc(sizeof(double) * 2 * pl);
#define WAT(q, j, k, c) w[(((q) * ny + (j)) * nz + (k)) * 3 + (c)]
        /* slot of plane q: q & 1.  K and ⋆ω of plane q (zero outside the grid) */
#define LOAD_PLANE(qq)                                                                                          \
    do {                                                                                                        \
        const int ss = (int)((qq) & 1);                                                                         \
        for (long j = 0; j < ny; ++j)                                                                           \
            for (long k = 0; k < nz; ++k) {                                                                     \
                const long m = j * nz + k;                                                                      \
                double* e = &P.ep[(ss * pl + m) * 3];                                                           \
                if ((qq) < 0 || (qq) >= nx) {                                                                   \
                    memset(&P.K[ss * pl + m], 0, sizeof(coefs7));                                               \
                    e[0] = e[1] = e[2] = 0.0;                                                                   \
                    continue;                                                                                   \
                }                                                                                               \
                const double* gg = g + ((((qq)) * ny + j) * nz + k) * 3;                                        \
                const coefs7 K = node_coefs(gg[0], gg[1], gg[2], h[0][(qq)], h[1][j], h[2][k]);                 \
                P.K[ss * pl + m] = K;                                                                           \
                e[0] = K.b1 * WAT((qq), j, k, 0);                                                               \
                e[1] = -K.b2 * WAT((qq), j, k, 1);                                                              \
                e[2] = K.b3 * WAT((qq), j, k, 2);                                                               \
            }                                                                                                   \
    } while (0)
/* ⋆dω of plane qq (forward differences, clamped) into slot(qq); zero outside */
#define STAR_DW(qq)                                                                                             \
    do {                                                                                                        \
        const int ss = (int)((qq) & 1);                                                                         \
        for (long j = 0; j < ny; ++j)                                                                           \
            for (long k = 0; k < nz; ++k) {                                                                     \
                const long m = j * nz + k;                                                                      \
                double* et = &P.et[(ss * pl + m) * 3];                                                          \
                if ((qq) < 0) {                                                                                 \
                    et[0] = et[1] = et[2] = 0.0;                                                                \
                    continue;                                                                                   \
                }                                                                                               \
                const coefs7 K = P.K[ss * pl + m];                                                              \
                const long qp = (qq) + 1 < nx ? (qq) + 1 : (qq), jp = j + 1 < ny ? j + 1 : j,                   \
                           kp = k + 1 < nz ? k + 1 : k;                                                         \
                const double w0 = WAT((qq), j, k, 0), w1 = WAT((qq), j, k, 1), w2 = WAT((qq), j, k, 2);         \
                const double Fxy = (WAT(qp, j, k, 1) - w1) - (WAT((qq), jp, k, 0) - w0);                        \
                const double Fxz = (WAT(qp, j, k, 2) - w2) - (WAT((qq), j, kp, 0) - w0);                        \
                const double Fyz = (WAT((qq), jp, k, 2) - w2) - (WAT((qq), j, kp, 1) - w1);                     \
                et[0] = K.a1 * Fyz;                                                                             \
                et[1] = -K.a2 * Fxz;                                                                            \
                et[2] = K.a3 * Fxy;                                                                             \
            }                                                                                                   \
    } while (0)
/* δω of plane qq into slot(qq): backward differences of ⋆ω, zero outside;
 * needs ⋆ω of qq and qq − 1 */
#define DELTA(qq)                                                                                               \
    do {                                                                                                        \
        const int ss = (int)((qq) & 1), sp = 1 - ss;                                                            \
        for (long j = 0; j < ny; ++j)                                                                           \
            for (long k = 0; k < nz; ++k) {                                                                     \
                const long m = j * nz + k;                                                                      \
                const double* e = &P.ep[(ss * pl + m) * 3];                                                     \
                const double ex0 = P.ep[(sp * pl + m) * 3 + 0];                                                 \
                const double ey1 = j > 0 ? P.ep[(ss * pl + m - nz) * 3 + 1] : 0.0;                              \
                const double ez2 = k > 0 ? P.ep[(ss * pl + m - 1) * 3 + 2] : 0.0;                              \
                const double H = (e[0] - ex0) - (e[1] - ey1) + (e[2] - ez2);                                    \
                P.dw[ss * pl + m] = -P.K[ss * pl + m].c * H;                                                    \
            }                                                                                                   \
    } while (0)
        if (xa < xb) {
            /* prologue: K, ⋆ω of planes xa−1 and xa; ⋆dω(xa−1); δω(xa) */
            LOAD_PLANE(xa - 1);
            LOAD_PLANE(xa);
            STAR_DW(xa - 1);
            DELTA(xa);
        }
        for (long q = xa; q < xb; ++q) {
            /* top of step q: K, ⋆ω of planes q, q−1; ⋆dω(q−1); δω(q) */
            const int s = (int)(q & 1), s1 = 1 - s;
            LOAD_PLANE(q + 1); /* replaces plane q − 1 */
            STAR_DW(q);
            DELTA(q + 1);
            const int hx1 = q + 1 < nx;
            for (long j = 0; j < oy; ++j)
                for (long k = 0; k < oz; ++k) {
                    const long m = j * nz + k;
                    const coefs7 K = P.K[s * pl + m];
                    const double* e = &P.et[(s * pl + m) * 3];
                    const double* ex = &P.et[(s1 * pl + m) * 3];
                    const double ey0 = j > 0 ? P.et[(s * pl + m - nz) * 3 + 0] : 0.0;
                    const double ey2 = j > 0 ? P.et[(s * pl + m - nz) * 3 + 2] : 0.0;
                    const double ez0 = k > 0 ? P.et[(s * pl + m - 1) * 3 + 0] : 0.0;
                    const double ez1 = k > 0 ? P.et[(s * pl + m - 1) * 3 + 1] : 0.0;
                    const double Gxy = (e[1] - ex[1]) - (e[0] - ey0);
                    const double Gxz = (e[2] - ex[2]) - (e[0] - ez0);
                    const double Gyz = (e[2] - ey2) - (e[1] - ez1);
                    const double dc = P.dw[s * pl + m];
                    const double ddx = hx1 ? P.dw[s1 * pl + m] - dc : 0.0;
                    const double ddy = j + 1 < ny ? P.dw[s * pl + m + nz] - dc : 0.0;
                    const double ddz = k + 1 < nz ? P.dw[s * pl + m + 1] - dc : 0.0;
                    double* o = out + ((q * oy + j) * oz + k) * 3;
                    o[0] = K.a1 * Gyz + ddx;
                    o[1] = -K.a2 * Gxz + ddy;
                    o[2] = K.a3 * Gxy + ddz;
                }
        }
#undef LOAD_PLANE
#undef STAR_DW
#undef DELTA
#undef WAT
        free(P.K);
        free(P.ep);
        free(P.et);
        free(P.dw);
    }
    for (int a = 0; a < 3; ++a) free(h[a]);
}

/* ------------------------------------------------------------------ */
/* conjugate gradients                                                 */
/* ------------------------------------------------------------------ */

typedef void (*fast_matvec)(const void* ctx, const double* in, double* out);

/* x0 = 0, Jacobi (inv_diag) or none; returns the iterations, res[0] = the
 * final recurrence ‖r‖2, res[1] = ‖b‖2.  max_it < 0: exactly −max_it
 * iterations whatever the residual (timing samples). */
static int fast_cg(long n, fast_matvec A, const void* ctx, const double* b, const double* inv_diag, double tol,
                   int max_it, double* x, double* res)
{
    double* r = (double*)malloc(sizeof(double) * n);
    double* p = (double*)malloc(sizeof(double) * n);
    double* q = (double*)malloc(sizeof(double) * n);
    double bb = 0.0, rz = 0.0;
#pragma omp parallel for reduction(+ : bb, rz) schedule(static)
    for (long i = 0; i < n; ++i) {
        x[i] = 0.0;
        r[i] = b[i];
        const double z = inv_diag ? b[i] * inv_diag[i] : b[i];
        p[i] = z;
        bb += b[i] * b[i];
        rz += b[i] * z;
    }
    const int fixed = max_it < 0;
    const int limit = fixed ? -max_it : max_it;
    double rr = bb;
    int it = 0;
    while (it < limit && (fixed || (bb > 0.0 && rr > tol * tol * bb))) {
        A(ctx, p, q);
        double pq = 0.0;
#pragma omp parallel for reduction(+ : pq) schedule(static)
        for (long i = 0; i < n; ++i) pq += p[i] * q[i];
        const double alpha = pq != 0.0 ? rz / pq : 0.0;
        double rz_new = 0.0, rr_new = 0.0;
#pragma omp parallel for reduction(+ : rz_new, rr_new) schedule(static)
        for (long i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            const double ri = r[i] - alpha * q[i];
            r[i] = ri;
            rr_new += ri * ri;
            rz_new += ri * (inv_diag ? ri * inv_diag[i] : ri);
        }
        const double beta = rz != 0.0 ? rz_new / rz : 0.0;
        rz = rz_new;
        rr = rr_new;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) p[i] = (inv_diag ? r[i] * inv_diag[i] : r[i]) + beta * p[i];
        ++it;
    }
    res[0] = sqrt(rr);
    res[1] = sqrt(bb);
    free(r);
    free(p);
    free(q);
    return it;
}

/* Dirichlet-zero Poisson on an nx × ny grid (x slowest), the operator of
 * orc_poisson2d_apply with its edge weights precomputed */
typedef struct {
    long nx, ny;
    double *wxp, *wxm, *wyp, *wym;
} p2_fast;

static void p2_fast_mv(const void* c, const double* u, double* out)
{
    const p2_fast* k = (const p2_fast*)c;
    const long nx = k->nx, ny = k->ny;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nx; ++i)
        for (long j = 0; j < ny; ++j) {
            const long n = i * ny + j;
            if (i == 0 || j == 0 || i == nx - 1 || j == ny - 1) {
                out[n] = 0.0;
                continue;
            }
            const double un = u[n];
            const double ue = (i + 1 < nx - 1) ? u[n + ny] : 0.0, uw = (i - 1 > 0) ? u[n - ny] : 0.0;
            const double uN = (j + 1 < ny - 1) ? u[n + 1] : 0.0, us = (j - 1 > 0) ? u[n - 1] : 0.0;
            out[n] = k->wxp[n] * (un - ue) + k->wxm[n] * (un - uw) + k->wyp[n] * (un - uN) + k->wym[n] * (un - us);
        }
}

int orc_fast_poisson2d_cg(long nx, long ny, const double* hx, const double* hy, const double* kappa, const double* b,
                          double tol, int max_it, int jacobi, double* x, double* res)
{
    const long n = nx * ny;
    p2_fast c = {nx, ny, (double*)calloc(n, 8), (double*)calloc(n, 8), (double*)calloc(n, 8), (double*)calloc(n, 8)};
    double* inv = jacobi ? (double*)calloc(n, 8) : NULL;
#pragma omp parallel for schedule(static)
    for (long i = 1; i < nx - 1; ++i)
        for (long j = 1; j < ny - 1; ++j) {
            const long m = i * ny + j;
            const double kxp = kappa ? 0.5 * (kappa[m] + kappa[m + ny]) : 1.0;
            const double kxm = kappa ? 0.5 * (kappa[m] + kappa[m - ny]) : 1.0;
            const double kyp = kappa ? 0.5 * (kappa[m] + kappa[m + 1]) : 1.0;
            const double kym = kappa ? 0.5 * (kappa[m] + kappa[m - 1]) : 1.0;
            c.wxp[m] = kxp * hy[j] / hx[i];
            c.wxm[m] = kxm * hy[j] / hx[i - 1];
            c.wyp[m] = kyp * hx[i] / hy[j];
            c.wym[m] = kym * hx[i] / hy[j - 1];
            if (inv) inv[m] = 1.0 / (c.wxp[m] + c.wxm[m] + c.wyp[m] + c.wym[m]);
        }
    const int it = fast_cg(n, p2_fast_mv, &c, b, inv, tol, max_it, x, res);
    free(c.wxp);
    free(c.wxm);
    free(c.wyp);
    free(c.wym);
    free(inv);
    return it;
}

/* magnetostatics curl-curl operator (solver_oracle.c orc_ms_apply, y sign
 * +1) in gather form: pass 1 H = w^B ν B at interior nodes, pass 2 the
 * transposed curl of H plus the gauge penalty; boundary rows identity */
typedef struct {
    long nx, ny, nz;
    double *lx, *ly, *lz; /* local spacings */
    const double* nu;     /* per node */
    double penalty;
    double* H; /* [node][3] */
} ms_fast;

static double lsp(const double* c, long n, long i)
{
    if (n < 2) return 1.0;
    if (i == 0) return c[1] - c[0];
    if (i + 1 == n) return c[n - 1] - c[n - 2];
    return 0.5 * (c[i + 1] - c[i - 1]);
}

static void ms_fast_mv(const void* c, const double* x, double* out)
{
    const ms_fast* k = (const ms_fast*)c;
    const long nx = k->nx, ny = k->ny, nz = k->nz, sy = nx, sz = nx * ny;
#define INT(i, j, kk) ((i) > 0 && (j) > 0 && (kk) > 0 && (i) + 1 < nx && (j) + 1 < ny && (kk) + 1 < nz)
#define A(i, j, kk, comp) (INT(i, j, kk) ? x[3 * ((i) + nx * ((j) + ny * (kk))) + (comp)] : 0.0)
#pragma omp parallel for schedule(static)
    for (long kk = 0; kk < nz; ++kk)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i) {
                double* h = k->H + 3 * (i + nx * (j + ny * kk));
                if (!INT(i, j, kk)) {
                    h[0] = h[1] = h[2] = 0.0;
                    continue;
                }
                const double a0 = A(i, j, kk, 0), a1 = A(i, j, kk, 1), a2 = A(i, j, kk, 2);
                const double bx = A(i, j + 1, kk, 2) - a2 - A(i, j, kk + 1, 1) + a1;
                const double by = -A(i + 1, j, kk, 2) + a2 + A(i, j, kk + 1, 0) - a0;
                const double bz = A(i + 1, j, kk, 1) - a1 - A(i, j + 1, kk, 0) + a0;
                const double nu = k->nu[i + nx * (j + ny * kk)];
                const double lx = k->lx[i], ly = k->ly[j], lz = k->lz[kk];
                h[0] = lx / (ly * lz) * nu * bx;
                h[1] = ly / (lx * lz) * nu * by;
                h[2] = lz / (lx * ly) * nu * bz;
            }
#pragma omp parallel for schedule(static)
    for (long kk = 0; kk < nz; ++kk)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i) {
                const long nd = i + nx * (j + ny * kk);
                if (!INT(i, j, kk)) {
                    out[3 * nd] = x[3 * nd];
                    out[3 * nd + 1] = x[3 * nd + 1];
                    out[3 * nd + 2] = x[3 * nd + 2];
                    continue;
                }
                const double* h = k->H + 3 * nd;
                const double* hmx = k->H + 3 * (nd - 1);
                const double* hmy = k->H + 3 * (nd - sy);
                const double* hmz = k->H + 3 * (nd - sz);
                const double lx = k->lx[i], ly = k->ly[j], lz = k->lz[kk];
                out[3 * nd + 0] = hmz[1] - h[1] - hmy[2] + h[2] + k->penalty * (ly * lz / lx) * x[3 * nd + 0];
                out[3 * nd + 1] = -hmz[0] + h[0] + hmx[2] - h[2] + k->penalty * (lx * lz / ly) * x[3 * nd + 1];
                out[3 * nd + 2] = hmy[0] - h[0] - hmx[1] + h[1] + k->penalty * (lx * ly / lz) * x[3 * nd + 2];
            }
#undef A
#undef INT
}

/* nodal μ → ν; Jacobi from the operator's diagonal when jacobi != 0
 * (the diagonal of orc_ms_diagonal: for an interior DOF the four H terms
 * that contain it, plus the penalty) */
int orc_fast_ms3d_cg(long nx, long ny, long nz, const double* xc, const double* yc, const double* zc,
                     const double* mu, double penalty, const double* b, double tol, int max_it, int jacobi,
                     double* x, double* res)
{
    const long nn = nx * ny * nz, n = 3 * nn;
    ms_fast c;
    c.nx = nx;
    c.ny = ny;
    c.nz = nz;
    c.lx = (double*)malloc(8 * nx);
    c.ly = (double*)malloc(8 * ny);
    c.lz = (double*)malloc(8 * nz);
    for (long i = 0; i < nx; ++i) c.lx[i] = lsp(xc, nx, i);
    for (long j = 0; j < ny; ++j) c.ly[j] = lsp(yc, ny, j);
    for (long kk = 0; kk < nz; ++kk) c.lz[kk] = lsp(zc, nz, kk);
    double* nu = (double*)malloc(8 * nn);
#pragma omp parallel for schedule(static)
    for (long t = 0; t < nn; ++t) nu[t] = 1.0 / mu[t];
    c.nu = nu;
    c.penalty = penalty;
    c.H = (double*)malloc(8 * n);
    double* inv = NULL;
    if (jacobi) {
        inv = (double*)malloc(8 * n);
#define INT(i, j, kk) ((i) > 0 && (j) > 0 && (kk) > 0 && (i) + 1 < nx && (j) + 1 < ny && (kk) + 1 < nz)
#pragma omp parallel for schedule(static)
        for (long kk = 0; kk < nz; ++kk)
            for (long j = 0; j < ny; ++j)
                for (long i = 0; i < nx; ++i) {
                    const long nd = i + nx * (j + ny * kk);
                    if (!INT(i, j, kk)) {
                        inv[3 * nd] = inv[3 * nd + 1] = inv[3 * nd + 2] = 1.0;
                        continue;
                    }
                    /* W_m(s) = w^B_m(s) ν(s) at interior s, else 0 */
#define W(m, ii, jj, kq)                                                                                        \
    (INT(ii, jj, kq) ? ((m) == 0   ? c.lx[ii] / (c.ly[jj] * c.lz[kq])                                          \
                        : (m) == 1 ? c.ly[jj] / (c.lx[ii] * c.lz[kq])                                          \
                                   : c.lz[kq] / (c.lx[ii] * c.ly[jj])) *                                        \
                           nu[(ii) + nx * ((jj) + ny * (kq))]                                                   \
                     : 0.0)
                    const double lx = c.lx[i], ly = c.ly[j], lz = c.lz[kk];
                    const double d0 = W(1, i, j, kk - 1) + W(1, i, j, kk) + W(2, i, j - 1, kk) + W(2, i, j, kk) +
                                      penalty * ly * lz / lx;
                    const double d1 = W(0, i, j, kk - 1) + W(0, i, j, kk) + W(2, i - 1, j, kk) + W(2, i, j, kk) +
                                      penalty * lx * lz / ly;
                    const double d2 = W(0, i, j - 1, kk) + W(0, i, j, kk) + W(1, i - 1, j, kk) + W(1, i, j, kk) +
                                      penalty * lx * ly / lz;
                    inv[3 * nd] = 1.0 / d0;
                    inv[3 * nd + 1] = 1.0 / d1;
                    inv[3 * nd + 2] = 1.0 / d2;
                }
#undef W
#undef INT
    }
    const int it = fast_cg(n, ms_fast_mv, &c, b, inv, tol, max_it, x, res);
    free(c.lx);
    free(c.ly);
    free(c.lz);
    free(nu);
    free(c.H);
    free(inv);
    return it;
}
