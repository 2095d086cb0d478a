/*
 * CPU ORACLE — test infrastructure only (see dec_oracle.c).
 *
 * Restates the 3D magnetostatics operator of the reference's ONELAB
 * interface, its right-hand side and nodal permeability rules, its assembled
 * CSR form, the SPD sparse approximate inverse preconditioner of Ginkgo
 * v1.11.0 (docker/similie_env/Dockerfile:70; a dependency the reference does
 * not vendor, so the published FSAI algorithm is restated), the 2D Dirichlet
 * Poisson operators the solvers run on, and textbook conjugate gradients.
 * Used only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* magnetostatics 3D operator                                          */
/* ------------------------------------------------------------------ */

/* local_spacing (onelab_interface/magnetostatics_onelab.hpp:856-870) */
static double local_spacing(const double* c, long n, long i)
{
    if (n < 2) return 1.0;
    if (i == 0) return c[1] - c[0];
    if (i + 1 == n) return c[n - 1] - c[n - 2];
    return 0.5 * (c[i + 1] - c[i - 1]);
}

typedef struct {
    long nx, ny, nz;
    const double *x, *y, *z;
} ms_grid;

static long dof(const ms_grid* g, long i, long j, long k, int c) { return 3 * (i + g->nx * (j + g->ny * k)) + c; }

static int is_boundary(const ms_grid* g, long i, long j, long k)
{
    return i == 0 || j == 0 || k == 0 || i + 1 == g->nx || j + 1 == g->ny || k + 1 == g->nz;
}

/* The 3D operator's switches (magnetostatics_onelab.hpp:796-828, read there
 * from SIMILIE_MAGNETIC_Y_RESPONSE_SIGN_3D, SIMILIE_DIVERGENCE_GAUGE_3D and
 * SIMILIE_VECTOR_POTENTIAL_GAUGE_SIGN_3D): the sign of the y moment's
 * response, and the divergence gauge factor·δᵀ W δ (factor = penalty ·
 * gauge sign) in place of the penalty term P·w^A·x.  Oracle state shared by
 * every 3D function below; set with orc_ms_set_switches. */
static struct {
    double y_sign;
    int div_gauge;
    double gauge_sign;
} g_sw = {1.0, 0, 1.0};

void orc_ms_set_switches(double y_sign, int div_gauge, double gauge_sign)
{
    g_sw.y_sign = y_sign < 0.0 ? -1.0 : 1.0;
    g_sw.div_gauge = div_gauge != 0;
    g_sw.gauge_sign = gauge_sign < 0.0 ? -1.0 : 1.0;
}

/* edge of the Hodge star's cell along an axis at node i: the forward edge,
 * the backward one at the last node (dec_oracle.c's ⋆ restatement) */
static double cell_edge(const double* c, long n, long i)
{
    if (n < 2) return 1.0;
    return i + 1 < n ? c[i + 1] - c[i] : c[i] - c[i - 1];
}

/* row n = (i,j,k) of the divergence gauge's codifferential
 * (magnetostatics_onelab.hpp:1893-1941: Codifferential::value of
 * codifferential.hpp:191-310 at an identity metric, interior potential DOFs
 * only).  On a rectilinear grid it reduces to the backward difference
 *   δA(n) = Σ_c [A_c(n−e_c) / (e_c(n) e_c(n−e_c)) − A_c(n) / e_c(n)²]
 * (e = cell_edge; pinned against dec_oracle.c's codifferential).  Entries in
 * the stencil's order: (n−e_x, x), (n−e_y, y), (n−e_z, z), (n, x), (n, y),
 * (n, z); returns their number. */
static int ms_codiff_row(const ms_grid* g, long i, long j, long k, long* cols, double* coefs)
{
    const long idx[3] = {i, j, k}, ext[3] = {g->nx, g->ny, g->nz};
    const double* ax[3] = {g->x, g->y, g->z};
    double e[3];
    for (int d = 0; d < 3; ++d) e[d] = cell_edge(ax[d], ext[d], idx[d]);
    int cnt = 0;
    for (int c = 0; c < 3; ++c) {
        long q[3] = {i, j, k};
        q[c] -= 1;
        if (q[c] < 0 || is_boundary(g, q[0], q[1], q[2])) continue;
        cols[cnt] = 3 * (q[0] + g->nx * (q[1] + g->ny * q[2])) + c;
        coefs[cnt] = 1.0 / (e[c] * cell_edge(ax[c], ext[c], q[c]));
        ++cnt;
    }
    if (!is_boundary(g, i, j, k))
        for (int c = 0; c < 3; ++c) {
            cols[cnt] = 3 * (i + g->nx * (j + g->ny * k)) + c;
            coefs[cnt] = -1.0 / (e[c] * e[c]);
            ++cnt;
        }
    return cnt;
}

/* the rows of the codifferential holding DOF (m, c) of an interior node m:
 * n = m (coefficient −1/e_c(m)²) and n = m + e_c (1/(e_c(m+e_c) e_c(m))),
 * ascending */
static int ms_codiff_slots(const ms_grid* g, long i, long j, long k, int c, long* rows, double* coefs)
{
    const long idx[3] = {i, j, k}, ext[3] = {g->nx, g->ny, g->nz};
    const double* ax[3] = {g->x, g->y, g->z};
    const double ec = cell_edge(ax[c], ext[c], idx[c]);
    long q[3] = {i, j, k};
    q[c] += 1;
    rows[0] = i + g->nx * (j + g->ny * k);
    coefs[0] = -1.0 / (ec * ec);
    rows[1] = q[0] + g->nx * (q[1] + g->ny * q[2]);
    coefs[1] = 1.0 / (cell_edge(ax[c], ext[c], q[c]) * ec);
    return 2;
}

/* δA at every node through ms_codiff_row (for the pinning tests) */
void orc_ms_codifferential(long nx, long ny, long nz, const double* xc, const double* yc, const double* zc,
                           const double* in, double* out)
{
    ms_grid g = {nx, ny, nz, xc, yc, zc};
    for (long n = 0; n < nx * ny * nz; ++n) {
        long cols[6];
        double cf[6];
        int cnt = ms_codiff_row(&g, n % nx, (n / nx) % ny, n / (nx * ny), cols, cf);
        double v = 0.0;
        for (int t = 0; t < cnt; ++t) v += cf[t] * in[cols[t]];
        out[n] = v;
    }
}

/* codifferential weight of node n: ℓx ℓy ℓz (:1896-1898) */
static double ms_codiff_weight(const ms_grid* g, long n)
{
    long i = n % g->nx, j = (n / g->nx) % g->ny, k = n / (g->nx * g->ny);
    return local_spacing(g->x, g->nx, i) * local_spacing(g->y, g->ny, j) * local_spacing(g->z, g->nz, k);
}

/* the gauge term of every interior row, added to out (:2359-2392): P·w^A·in,
 * or with the divergence gauge Σ_slots factor·w(n)·coef·(δin)(n) */
static void ms_add_gauge(const ms_grid* g, double penalty, const double* in, double* out)
{
    long nx = g->nx, ny = g->ny, nz = g->nz, nn = nx * ny * nz;
    double* dv = NULL;
    if (g_sw.div_gauge) {
        dv = (double*)malloc(sizeof(double) * (nn > 0 ? nn : 1));
        for (long n = 0; n < nn; ++n) {
            long cols[6];
            double cf[6];
            int cnt = ms_codiff_row(g, n % nx, (n / nx) % ny, n / (nx * ny), cols, cf);
            double v = 0.0;
            for (int t = 0; t < cnt; ++t) v += cf[t] * in[cols[t]];
            dv[n] = v;
        }
    }
    const double factor = penalty * g_sw.gauge_sign;
    for (long k = 1; k + 1 < nz; ++k)
        for (long j = 1; j + 1 < ny; ++j)
            for (long i = 1; i + 1 < nx; ++i) {
                double dx = local_spacing(g->x, nx, i), dy = local_spacing(g->y, ny, j), dz = local_spacing(g->z, nz, k);
                double wa[3] = {dy * dz / dx, dx * dz / dy, dx * dy / dz};
                for (int c = 0; c < 3; ++c) {
                    long r = dof(g, i, j, k, c);
                    if (!g_sw.div_gauge) {
                        out[r] += penalty * wa[c] * in[r];
                        continue;
                    }
                    long rows[2];
                    double cf[2];
                    int ns = ms_codiff_slots(g, i, j, k, c, rows, cf);
                    double gr = 0.0;
                    for (int t = 0; t < ns; ++t) gr += factor * ms_codiff_weight(g, rows[t]) * cf[t] * dv[rows[t]];
                    out[r] += gr;
                }
            }
    free(dv);
}

/* the gauge entries of interior row (i,j,k,c) for the assembly (:3204-3240),
 * in slot order then the codifferential row's stencil order */
static void ms_gauge_entries(const ms_grid* g, long i, long j, long k, int c, double penalty, int* count, int* lc,
                             double* lv);

/* The moment (magnetic induction) stencils of forward_value
 * (src/similie/physics/magnetostatics/magnetostatics_quantities.hpp:173-196):
 * B_x = (dA)_yz, B_y = -(dA)_xz, B_z = (dA)_xy, each a forward difference of
 * the vector potential around the node, with the sign layout of
 * Coboundary::run.  Entries: (di,dj,dk, potential component, coefficient). */
typedef struct {
    int di, dj, dk, comp;
    double coeff;
} ms_tap;

static const ms_tap MOMENT[3][4] = {
    /* B_x = A_z(s+e_y) - A_z(s) - A_y(s+e_z) + A_y(s) */
    {{0, 1, 0, 2, 1.0}, {0, 0, 0, 2, -1.0}, {0, 0, 1, 1, -1.0}, {0, 0, 0, 1, 1.0}},
    /* B_y = -(A_z(s+e_x) - A_z(s) - A_x(s+e_z) + A_x(s)) */
    {{1, 0, 0, 2, -1.0}, {0, 0, 0, 2, 1.0}, {0, 0, 1, 0, 1.0}, {0, 0, 0, 0, -1.0}},
    /* B_z = A_y(s+e_x) - A_y(s) - A_x(s+e_y) + A_x(s) */
    {{1, 0, 0, 1, 1.0}, {0, 0, 0, 1, -1.0}, {0, 1, 0, 0, -1.0}, {0, 0, 0, 0, 1.0}},
};

/* apply_with_precomputed_stencils (magnetostatics_onelab.hpp:2241-2396) in
 * scatter form: for every interior sampled node s, B(s) is gathered from the
 * interior potential DOFs, weighted by the induction Hodge factor
 * (:873-893) and ∂H/∂B = B/μ of the linear Hamiltonian
 * (physics/magnetostatics/linear_magnetostatics.py:19), and scattered back
 * through the transposed stencil; interior rows add the vector-potential
 * gauge penalty (:2359-2367), boundary rows are the identity. */
void orc_ms_apply(long nx, long ny, long nz, const double* xc, const double* yc, const double* zc,
                  const double* mu, double penalty, double y_sign, const double* in, double* out)
{
    ms_grid g = {nx, ny, nz, xc, yc, zc};
    long ndof = 3 * nx * ny * nz;
    for (long r = 0; r < ndof; ++r) out[r] = 0.0;
    for (long k = 1; k + 1 < nz; ++k)
        for (long j = 1; j + 1 < ny; ++j)
            for (long i = 1; i + 1 < nx; ++i) {
                double dx = local_spacing(xc, nx, i), dy = local_spacing(yc, ny, j), dz = local_spacing(zc, nz, k);
                double wb[3] = {dx / (dy * dz), dy / (dx * dz), dz / (dx * dy)};
                double nu = 1.0 / mu[i + nx * (j + ny * k)];
                double B[3];
                for (int m = 0; m < 3; ++m) {
                    double b = 0.0;
                    for (int t = 0; t < 4; ++t) {
                        const ms_tap* tp = &MOMENT[m][t];
                        long pi = i + tp->di, pj = j + tp->dj, pk = k + tp->dk;
                        if (is_boundary(&g, pi, pj, pk)) continue;
                        b += tp->coeff * in[dof(&g, pi, pj, pk, tp->comp)];
                    }
                    B[m] = b;
                }
                for (int m = 0; m < 3; ++m) {
                    double h = wb[m] * nu * B[m] * (m == 1 ? y_sign * g_sw.y_sign : 1.0);
                    for (int t = 0; t < 4; ++t) {
                        const ms_tap* tp = &MOMENT[m][t];
                        long pi = i + tp->di, pj = j + tp->dj, pk = k + tp->dk;
                        if (is_boundary(&g, pi, pj, pk)) continue;
                        out[dof(&g, pi, pj, pk, tp->comp)] += tp->coeff * h;
                    }
                }
            }
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i)
                if (is_boundary(&g, i, j, k))
                    for (int c = 0; c < 3; ++c) out[dof(&g, i, j, k, c)] = in[dof(&g, i, j, k, c)];
    ms_add_gauge(&g, penalty, in, out);
}

/* right-hand side: ⅛·|dual face|·J over the ≤8 cells around an interior
 * node (magnetostatics_onelab.hpp:4919-4975).  cell_j is [cell][3] with
 * cells indexed ci + (nx-1)(cj + (ny-1) ck). */
void orc_ms_rhs(long nx, long ny, long nz, const double* xc, const double* yc, const double* zc,
                const double* cell_j, double* rhs)
{
    ms_grid g = {nx, ny, nz, xc, yc, zc};
    long cx = nx - 1, cy = ny - 1, cz = nz - 1;
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i)
                for (int c = 0; c < 3; ++c) {
                    long r = dof(&g, i, j, k, c);
                    rhs[r] = 0.0;
                    if (is_boundary(&g, i, j, k)) continue;
                    for (int dk = -1; dk <= 0; ++dk)
                        for (int dj = -1; dj <= 0; ++dj)
                            for (int di = -1; di <= 0; ++di) {
                                long ci = i + di, cj = j + dj, ck = k + dk;
                                if (ci < 0 || cj < 0 || ck < 0 || ci >= cx || cj >= cy || ck >= cz) continue;
                                double wx = xc[ci + 1] - xc[ci], wy = yc[cj + 1] - yc[cj], wz = zc[ck + 1] - zc[ck];
                                double face = wx * wy;
                                if (c == 0) face = wy * wz;
                                else if (c == 1) face = wx * wz;
                                rhs[r] += 0.125 * face * cell_j[3 * (ci + cx * (cj + cy * ck)) + c];
                            }
                }
}

/* nodal permeability = mean over adjacent cells (magnetostatics_onelab.hpp
 * :5029-5066) */
void orc_ms_nodal_mu(long nx, long ny, long nz, const double* cell_mu, double mu0, double* mu)
{
    long cx = nx - 1, cy = ny - 1, cz = nz - 1;
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i) {
                double acc = 0.0;
                long cnt = 0;
                for (int dk = -1; dk <= 0; ++dk)
                    for (int dj = -1; dj <= 0; ++dj)
                        for (int di = -1; di <= 0; ++di) {
                            long ci = i + di, cj = j + dj, ck = k + dk;
                            if (ci < 0 || cj < 0 || ck < 0 || ci >= cx || cj >= cy || ck >= cz) continue;
                            acc += cell_mu[ci + cx * (cj + cy * ck)];
                            ++cnt;
                        }
                mu[i + nx * (j + ny * k)] = cnt == 0 ? mu0 : acc / (double)cnt;
            }
}

/* Jacobi diagonal of the operator above */
void orc_ms_diagonal(long nx, long ny, long nz, const double* xc, const double* yc, const double* zc,
                     const double* mu, double penalty, double* diag)
{
    long ndof = 3 * nx * ny * nz;
    double* e = (double*)calloc(ndof, sizeof(double));
    double* o = (double*)malloc(sizeof(double) * ndof);
    /* columns are probed in 27 colour classes so that probes never overlap */
    for (int ci = 0; ci < 3; ++ci)
        for (int cj = 0; cj < 3; ++cj)
            for (int ck = 0; ck < 3; ++ck)
                for (int c = 0; c < 3; ++c) {
                    for (long r = 0; r < ndof; ++r) e[r] = 0.0;
                    for (long k = ck; k < nz; k += 3)
                        for (long j = cj; j < ny; j += 3)
                            for (long i = ci; i < nx; i += 3) e[3 * (i + nx * (j + ny * k)) + c] = 1.0;
                    orc_ms_apply(nx, ny, nz, xc, yc, zc, mu, penalty, 1.0, e, o);
                    for (long k = ck; k < nz; k += 3)
                        for (long j = cj; j < ny; j += 3)
                            for (long i = ci; i < nx; i += 3) {
                                long r = 3 * (i + nx * (j + ny * k)) + c;
                                diag[r] = o[r];
                            }
                }
    free(e);
    free(o);
}

/* ------------------------------------------------------------------ */
/* assembled operator (CSR) and the SPD sparse approximate inverse      */
/* ------------------------------------------------------------------ */

#define ORC_ROW_MAX 128 /* MATRIX_ROW_MAX_SIZE, magnetostatics_onelab.hpp:3145 */

typedef struct {
    long moment_row;
    double coeff;
} orc_tslot;

static void add_entry(int* count, int* lc, double* lv, long col, double value)
{
    if (value == 0.0) return;
    for (int s = 0; s < *count; ++s)
        if (lc[s] == (int)col) {
            lv[s] += value;
            return;
        }
    lc[*count] = (int)col;
    lv[*count] = value;
    ++*count;
}

static void ms_gauge_entries(const ms_grid* g, long i, long j, long k, int c, double penalty, int* count, int* lc,
                             double* lv)
{
    if (!g_sw.div_gauge) {
        double dx = local_spacing(g->x, g->nx, i), dy = local_spacing(g->y, g->ny, j), dz = local_spacing(g->z, g->nz, k);
        double wa = c == 0 ? dy * dz / dx : (c == 1 ? dx * dz / dy : dx * dy / dz);
        add_entry(count, lc, lv, dof(g, i, j, k, c), penalty * wa);
        return;
    }
    const double factor = penalty * g_sw.gauge_sign;
    long rows[2];
    double rcf[2];
    int ns = ms_codiff_slots(g, i, j, k, c, rows, rcf);
    for (int t = 0; t < ns; ++t) {
        long n = rows[t];
        long cols[6];
        double cf[6];
        int cnt = ms_codiff_row(g, n % g->nx, (n / g->nx) % g->ny, n / (g->nx * g->ny), cols, cf);
        double w = ms_codiff_weight(g, n);
        for (int q = 0; q < cnt; ++q) add_entry(count, lc, lv, cols[q], factor * w * rcf[t] * cf[q]);
    }
}

/* assemble_matrix_data (onelab_interface/magnetostatics_onelab.hpp:3115-3536)
 * for the linear equations (∂H/∂B = δ_ab/μ): the transposed moment slots of
 * every potential DOF are collected in sampled-node order (the precompute of
 * :1950-1981, restated serially); an interior row starts with the gauge term
 * P·w^A_c on its diagonal and adds, for every slot (moment row 3s+m,
 * coefficient c_r), c_r·(w^B_m(s)·ν(s))·c_t for every interior tap t of the
 * moment stencil of m at s; zero contributions are skipped and repeated
 * columns summed (add_entry, :3176-3190).  Boundary rows hold 1 on the
 * diagonal.  Rows come out with ascending columns.  cols/vals must hold
 * 17·3·nx·ny·nz entries; returns the number of stored entries, or -1 when a
 * row exceeds ORC_ROW_MAX. */
long orc_ms_assemble_csr(long nx, long ny, long nz, const double* xc, const double* yc, const double* zc,
                         const double* mu, double penalty, double y_sign, long* row_ptr, int* cols, double* vals)
{
    ms_grid g = {nx, ny, nz, xc, yc, zc};
    long ndof = 3 * nx * ny * nz;
    /* a component is a tap of two moment stencils, twice each: ≤ 4 slots per DOF */
    int* tcount = (int*)calloc(ndof, sizeof(int));
    orc_tslot* tslot = (orc_tslot*)malloc(sizeof(orc_tslot) * 4 * (ndof > 0 ? ndof : 1));
    for (long k = 1; k + 1 < nz; ++k)
        for (long j = 1; j + 1 < ny; ++j)
            for (long i = 1; i + 1 < nx; ++i)
                for (int m = 0; m < 3; ++m)
                    for (int t = 0; t < 4; ++t) {
                        const ms_tap* tp = &MOMENT[m][t];
                        long pi = i + tp->di, pj = j + tp->dj, pk = k + tp->dk;
                        if (is_boundary(&g, pi, pj, pk)) continue;
                        long r = dof(&g, pi, pj, pk, tp->comp);
                        orc_tslot* sl = &tslot[4 * r + tcount[r]++];
                        sl->moment_row = 3 * (i + nx * (j + ny * k)) + m;
                        sl->coeff = tp->coeff;
                    }
    long nnz = 0;
    int lc[ORC_ROW_MAX];
    double lv[ORC_ROW_MAX];
    for (long r = 0; r < ndof; ++r) {
        long node = r / 3;
        int c = (int)(r % 3);
        long i = node % nx, j = (node / nx) % ny, k = node / (nx * ny);
        row_ptr[r] = nnz;
        if (is_boundary(&g, i, j, k)) {
            cols[nnz] = (int)r;
            vals[nnz] = 1.0;
            ++nnz;
            continue;
        }
        int count = 0;
        ms_gauge_entries(&g, i, j, k, c, penalty, &count, lc, lv);
        for (int sidx = 0; sidx < tcount[r]; ++sidx) {
            long mrow = tslot[4 * r + sidx].moment_row;
            double rc = tslot[4 * r + sidx].coeff;
            long s = mrow / 3;
            int m = (int)(mrow % 3);
            long si = s % nx, sj = (s / nx) % ny, sk = s / (nx * ny);
            double sx = local_spacing(xc, nx, si), sy = local_spacing(yc, ny, sj), sz = local_spacing(zc, nz, sk);
            double ys = y_sign * g_sw.y_sign;
            double hf = m == 0 ? sx / (sy * sz) : (m == 1 ? ys * (sy / (sx * sz)) : sz / (sx * sy));
            double jc = hf * (1.0 / mu[s]);
            if (jc == 0.0) continue;
            for (int t = 0; t < 4; ++t) {
                const ms_tap* tp = &MOMENT[m][t];
                long pi = si + tp->di, pj = sj + tp->dj, pk = sk + tp->dk;
                if (is_boundary(&g, pi, pj, pk)) continue;
                if (count >= ORC_ROW_MAX) {
                    free(tcount);
                    free(tslot);
                    return -1;
                }
                add_entry(&count, lc, lv, dof(&g, pi, pj, pk, tp->comp), rc * jc * tp->coeff);
            }
        }
        /* ascending columns (insertion sort; rows are short) */
        for (int a = 1; a < count; ++a) {
            int kc = lc[a];
            double kv = lv[a];
            int b = a - 1;
            while (b >= 0 && lc[b] > kc) {
                lc[b + 1] = lc[b];
                lv[b + 1] = lv[b];
                --b;
            }
            lc[b + 1] = kc;
            lv[b + 1] = kv;
        }
        for (int a = 0; a < count; ++a) {
            cols[nnz] = lc[a];
            vals[nnz] = lv[a];
            ++nnz;
        }
    }
    row_ptr[ndof] = nnz;
    free(tcount);
    free(tslot);
    return nnz;
}

/* SpdIsai of Ginkgo v1.11.0 (gko::preconditioner::SpdIsai, sparsity power 1;
 * selected at src/similie/solvers/minimize_strong_formulation_residual.hpp
 * :963-967), i.e. the factorized sparse approximate inverse of Kolotilina &
 * Yeremin (1993): for row i let J = the stored columns j ≤ i of row i
 * (ascending, i last) and S = A[J, J] (entry (a, b), b ≤ a, read from row
 * J[a]); solve S y = e_|J| by Cholesky and set G[i, J] = y / √y_|J|, so that
 * M = Gᵀ G ≈ A⁻¹.  Columns of A must be ascending.  g_row_ptr has n+1
 * entries, g_cols / g_vals room for nnz(A).  Returns 0, -(i+1) when row i
 * lacks its diagonal, i+1 when its local system is not positive definite,
 * or -(n+1) - i when J holds more than 64 columns. */
long orc_spd_isai(long n, const long* row_ptr, const int* cols, const double* vals, long* g_row_ptr, int* g_cols,
                  double* g_vals)
{
    enum { MAXJ = 64 };
    double S[MAXJ][MAXJ], y[MAXJ];
    int J[MAXJ];
    long out = 0;
    for (long i = 0; i < n; ++i) {
        g_row_ptr[i] = out;
        int m = 0;
        for (long e = row_ptr[i]; e < row_ptr[i + 1] && cols[e] <= i; ++e) {
            if (m == MAXJ) return -(n + 1) - i;
            J[m++] = cols[e];
        }
        if (m == 0 || J[m - 1] != i) return -(i + 1);
        for (int a = 0; a < m; ++a) {
            for (int b = 0; b <= a; ++b) S[a][b] = 0.0;
            for (long e = row_ptr[J[a]]; e < row_ptr[J[a] + 1]; ++e)
                for (int b = 0; b <= a; ++b)
                    if (cols[e] == J[b]) S[a][b] = vals[e];
        }
        /* Cholesky S = L Lᵀ, L stored in the lower triangle of S */
        for (int a = 0; a < m; ++a) {
            for (int b = 0; b < a; ++b) {
                double v = S[a][b];
                for (int c = 0; c < b; ++c) v -= S[a][c] * S[b][c];
                S[a][b] = v / S[b][b];
            }
            double d = S[a][a];
            for (int c = 0; c < a; ++c) d -= S[a][c] * S[a][c];
            if (!(d > 0.0)) return i + 1;
            S[a][a] = sqrt(d);
        }
        /* L w = e_m, then Lᵀ y = w */
        for (int a = 0; a < m; ++a) {
            double v = a == m - 1 ? 1.0 : 0.0;
            for (int c = 0; c < a; ++c) v -= S[a][c] * y[c];
            y[a] = v / S[a][a];
        }
        for (int a = m - 1; a >= 0; --a) {
            double v = y[a];
            for (int c = a + 1; c < m; ++c) v -= S[c][a] * y[c];
            y[a] = v / S[a][a];
        }
        double scale = sqrt(y[m - 1]);
        for (int a = 0; a < m; ++a) {
            g_cols[out] = J[a];
            g_vals[out] = y[a] / scale;
            ++out;
        }
    }
    g_row_ptr[n] = out;
    return 0;
}

/* ------------------------------------------------------------------ */
/* 2D Dirichlet Poisson operators                                      */
/* ------------------------------------------------------------------ */

/* ⋆0·Δ of the DEC 0-form Laplacian on a rectilinear 2D grid, restricted to
 * interior nodes with zero Dirichlet values (= d0ᵀ ⋆1 d0): edge weights
 * w_x(m) = h_y(m)/h_x(m), w_y(m) = h_x(m)/h_y(m) from the per-node ⋆ rule.
 * u and out are full nx×ny arrays (x slowest); boundary entries of out are 0. */
void orc_poisson2d_apply(long nx, long ny, const double* hx, const double* hy, const double* kappa,
                         const double* u, double* out)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nx; ++i)
        for (long j = 0; j < ny; ++j) {
            long n = i * ny + j;
            if (i == 0 || j == 0 || i == nx - 1 || j == ny - 1) {
                out[n] = 0.0;
                continue;
            }
            double un = u[n];
            double acc = 0.0;
            /* edge weights: w_x at (i,j) couples (i,j)-(i+1,j) */
            double kx_p = kappa ? 0.5 * (kappa[n] + kappa[n + ny]) : 1.0;
            double kx_m = kappa ? 0.5 * (kappa[n] + kappa[n - ny]) : 1.0;
            double ky_p = kappa ? 0.5 * (kappa[n] + kappa[n + 1]) : 1.0;
            double ky_m = kappa ? 0.5 * (kappa[n] + kappa[n - 1]) : 1.0;
            double wxp = kx_p * hy[j] / hx[i], wxm = kx_m * hy[j] / hx[i - 1];
            double wyp = ky_p * hx[i] / hy[j], wym = ky_m * hx[i] / hy[j - 1];
            double ue = (i + 1 < nx - 1) ? u[n + ny] : 0.0;
            double uw = (i - 1 > 0) ? u[n - ny] : 0.0;
            double un_ = (j + 1 < ny - 1) ? u[n + 1] : 0.0;
            double us = (j - 1 > 0) ? u[n - 1] : 0.0;
            acc = wxp * (un - ue) + wxm * (un - uw) + wyp * (un - un_) + wym * (un - us);
            out[n] = acc;
        }
}

/* ------------------------------------------------------------------ */
/* textbook conjugate gradients (Hestenes–Stiefel)                     */
/* ------------------------------------------------------------------ */

typedef void (*orc_matvec_fn)(void* ctx, const double* in, double* out);

/* fixed blocks summed in order: the result does not depend on the thread
 * count or on the order in which threads finish (an OpenMP reduction
 * combines its partial sums in an unspecified order) */
#define ORC_DOT_BLOCK 4096
static double dot(long n, const double* a, const double* b)
{
    long nb = (n + ORC_DOT_BLOCK - 1) / ORC_DOT_BLOCK;
    double stack_part[64];
    double* part = nb <= 64 ? stack_part : (double*)malloc(sizeof(double) * nb);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < nb; ++q) {
        long e = (q + 1) * ORC_DOT_BLOCK < n ? (q + 1) * ORC_DOT_BLOCK : n;
        double s = 0.0;
        for (long i = q * ORC_DOT_BLOCK; i < e; ++i) s += a[i] * b[i];
        part[q] = s;
    }
    double s = 0.0;
    for (long q = 0; q < nb; ++q) s += part[q];
    if (part != stack_part) free(part);
    return s;
}

/* x0 = 0; stop when ‖r_k‖2 ≤ tol·‖b‖2 on the recurrence residual or after
 * max_it iterations (the stopping rule of the reference's Ginkgo CG call,
 * src/similie/solvers/minimize_strong_formulation_residual.hpp:1063-1130).
 * diag (optional) applies Jacobi preconditioning.  Returns iterations;
 * res_out[0] = final recurrence ‖r‖2, res_out[1] = ‖b‖2. */
int orc_cg(long n, orc_matvec_fn A, void* ctx, const double* b, const double* diag, double tol, int max_it,
           double* x, double* res_out)
{
    double* r = (double*)malloc(sizeof(double) * n);
    double* z = (double*)malloc(sizeof(double) * n);
    double* p = (double*)malloc(sizeof(double) * n);
    double* q = (double*)malloc(sizeof(double) * n);
    for (long i = 0; i < n; ++i) {
        x[i] = 0.0;
        r[i] = b[i];
        z[i] = diag ? b[i] / diag[i] : b[i];
        p[i] = z[i];
    }
    double bnorm = sqrt(dot(n, b, b));
    double rz = dot(n, r, z);
    double rnorm = bnorm;
    int it = 0;
    if (bnorm == 0.0) {
        res_out[0] = 0.0;
        res_out[1] = 0.0;
        free(r);
        free(z);
        free(p);
        free(q);
        return 0;
    }
    while (it < max_it && rnorm > tol * bnorm) {
        A(ctx, p, q);
        double pq = dot(n, p, q);
        double alpha = rz / pq;
        for (long i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
        }
        for (long i = 0; i < n; ++i) z[i] = diag ? r[i] / diag[i] : r[i];
        double rz_new = dot(n, r, z);
        rnorm = sqrt(dot(n, r, r));
        double beta = rz_new / rz;
        rz = rz_new;
        for (long i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        ++it;
    }
    res_out[0] = rnorm;
    res_out[1] = bnorm;
    free(r);
    free(z);
    free(p);
    free(q);
    return it;
}

/* CG on the 2D Dirichlet Poisson operator above (unknowns = all nodes; the
 * boundary rows are excluded by keeping their entries of b, x at zero). */
typedef struct {
    long nx, ny;
    const double *hx, *hy, *kappa;
} p2_ctx;

static void p2_mv(void* c, const double* in, double* out)
{
    p2_ctx* k = (p2_ctx*)c;
    orc_poisson2d_apply(k->nx, k->ny, k->hx, k->hy, k->kappa, in, out);
}

int orc_poisson2d_cg(long nx, long ny, const double* hx, const double* hy, const double* kappa, const double* b,
                     double tol, int max_it, double* x, double* res_out)
{
    p2_ctx c = {nx, ny, hx, hy, kappa};
    return orc_cg(nx * ny, p2_mv, &c, b, NULL, tol, max_it, x, res_out);
}

typedef struct {
    long nx, ny, nz;
    const double *x, *y, *z, *mu;
    double penalty;
} ms_ctx;

static void ms_mv(void* c, const double* in, double* out)
{
    ms_ctx* k = (ms_ctx*)c;
    orc_ms_apply(k->nx, k->ny, k->nz, k->x, k->y, k->z, k->mu, k->penalty, 1.0, in, out);
}

int orc_ms_cg(long nx, long ny, long nz, const double* xc, const double* yc, const double* zc, const double* mu,
              double penalty, const double* b, const double* diag, double tol, int max_it, double* x, double* res_out)
{
    ms_ctx c = {nx, ny, nz, xc, yc, zc, mu, penalty};
    return orc_cg(3 * nx * ny * nz, ms_mv, &c, b, diag, tol, max_it, x, res_out);
}

/* ------------------------------------------------------------------ */
/* nonlinear magnetostatics                                            */
/* ------------------------------------------------------------------ */

#define ORC_BH_MAX 64 /* InterpolatedNonlinearBHCurve<64> (magnetostatics_onelab.hpp:4550) */

/* InterpolatedNonlinearBHCurve (src/similie/physics/magnetostatics/
 * nonlinear_magnetostatics.py:14-181): h(b) piecewise linear in b between
 * the samples, bracketed through q = b²; ν = h/b per sample with ν₀ = ν₁. */
typedef struct {
    int n;
    double b[ORC_BH_MAX], h[ORC_BH_MAX], q[ORC_BH_MAX], nu[ORC_BH_MAX];
    double dnu_dq[ORC_BH_MAX - 1], dh_db[ORC_BH_MAX - 1];
} orc_bh;

static void bh_init(orc_bh* c, const double* b, const double* h, int n)
{
    c->n = n;
    for (int i = 0; i < n; ++i) {
        c->b[i] = b[i];
        c->h[i] = h[i];
        c->q[i] = b[i] * b[i];
    }
    c->nu[0] = c->h[1] / c->b[1];
    for (int i = 1; i < n; ++i) c->nu[i] = c->h[i] / c->b[i];
    for (int i = 0; i + 1 < n; ++i) {
        c->dnu_dq[i] = (c->nu[i + 1] - c->nu[i]) / (c->q[i + 1] - c->q[i]);
        c->dh_db[i] = (c->h[i + 1] - c->h[i]) / (c->b[i + 1] - c->b[i]);
    }
}

static int bh_bracket_q(const orc_bh* c, double q)
{
    if (q <= c->q[0]) return 0;
    for (int i = 0; i + 1 < c->n; ++i)
        if (q <= c->q[i + 1]) return i;
    return c->n - 2;
}

static double bh_h_from_b(const orc_bh* c, double b)
{
    if (b <= c->b[0]) return c->h[0];
    if (b >= c->b[c->n - 1]) {
        int i = c->n - 2;
        return c->h[i] + c->dh_db[i] * (b - c->b[i]);
    }
    int i = bh_bracket_q(c, b * b);
    return c->h[i] + c->dh_db[i] * (b - c->b[i]);
}

static double bh_dh_db_from_b(const orc_bh* c, double b)
{
    if (b <= c->b[0]) return c->dh_db[0];
    if (b >= c->b[c->n - 1]) return c->dh_db[c->n - 2];
    return c->dh_db[bh_bracket_q(c, b * b)];
}

static double bh_h_over_b_from_b(const orc_bh* c, double b)
{
    if (b == 0.0) return bh_dh_db_from_b(c, 0.0);
    return bh_h_from_b(c, b) / b;
}

static double bh_d_h_over_b_db_from_b(const orc_bh* c, double b)
{
    if (b == 0.0) return 0.0;
    return (bh_dh_db_from_b(c, b) * b - bh_h_from_b(c, b)) / (b * b);
}

static double bh_d2_h_over_b_db2_from_b(const orc_bh* c, double b)
{
    if (b == 0.0) return 0.0;
    double intercept = bh_h_from_b(c, b) - bh_dh_db_from_b(c, b) * b;
    return 2.0 * intercept / (b * b * b);
}

static double bh_f(const orc_bh* c, double q) { return bh_h_over_b_from_b(c, sqrt(q)); }

static double bh_f1(const orc_bh* c, double q)
{
    if (q == 0.0) return 0.0;
    double b = sqrt(q);
    return bh_d_h_over_b_db_from_b(c, b) / (2.0 * b);
}

static double bh_f2(const orc_bh* c, double q)
{
    if (q == 0.0) return 0.0;
    double b = sqrt(q);
    return bh_d2_h_over_b_db2_from_b(c, b) / (4.0 * q) - bh_d_h_over_b_db_from_b(c, b) / (4.0 * q * b);
}

/* one B-H curve function at `arg` (test hook): 0 h_from_b, 1 dh_db_from_b,
 * 2 h_over_b_from_q, 3 d_h_over_b_dq_from_q, 4 d2_h_over_b_dq2_from_q,
 * 5 nu at sample `arg` */
double orc_bh_eval(const double* b, const double* h, int n, int which, double arg)
{
    orc_bh c;
    bh_init(&c, b, h, n);
    switch (which) {
    case 0: return bh_h_from_b(&c, arg);
    case 1: return bh_dh_db_from_b(&c, arg);
    case 2: return bh_f(&c, arg);
    case 3: return bh_f1(&c, arg);
    case 4: return bh_f2(&c, arg);
    default: return c.nu[(int)arg];
    }
}

/* MaterialMagnetostaticsHamiltonian (magnetostatics_onelab.hpp:957-1011):
 * a node whose ferromagnetic fraction exceeds 0.5 follows the generated
 * NonlinearMagnetostaticsHamiltonian — H = ½ q f(q), q = |B|², f = h/b of the
 * curve (nonlinear_magnetostatics.py:184-227), so ∂H/∂B_m = B_m (f + q f′)
 * and ∂²H/∂B_m∂B_c = δ_mc (f + q f′) + 2 B_m B_c (2 f′ + q f″) (the
 * derivatives generate_cpp_hamiltonian.py:306-397 emits); other nodes are
 * linear: B_m / μ and δ_mc / μ. */
static void nl_response(const orc_bh* c, int ferro, double mu, const double B[3], double H[3], double J[9])
{
    if (ferro) {
        double q = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
        double f = bh_f(c, q), f1 = bh_f1(c, q), f2 = bh_f2(c, q);
        double a = f + q * f1, cc = 2.0 * (2.0 * f1 + q * f2);
        for (int m = 0; m < 3; ++m) {
            if (H) H[m] = B[m] * a;
            if (J)
                for (int n = 0; n < 3; ++n) J[3 * m + n] = (m == n ? a : 0.0) + cc * B[m] * B[n];
        }
    } else {
        for (int m = 0; m < 3; ++m) {
            if (H) H[m] = B[m] / mu;
            if (J)
                for (int n = 0; n < 3; ++n) J[3 * m + n] = m == n ? 1.0 / mu : 0.0;
        }
    }
}

static void moments_at(const ms_grid* g, long i, long j, long k, const double* x, double B[3])
{
    for (int m = 0; m < 3; ++m) {
        double v = 0.0;
        for (int t = 0; t < 4; ++t) {
            const ms_tap* tp = &MOMENT[m][t];
            long pi = i + tp->di, pj = j + tp->dj, pk = k + tp->dk;
            if (is_boundary(g, pi, pj, pk)) continue;
            v += tp->coeff * x[dof(g, pi, pj, pk, tp->comp)];
        }
        B[m] = v;
    }
}

/* the operator of apply_with_precomputed_stencils (magnetostatics_onelab.hpp
 * :2241-2396) with the material Hamiltonian (state = NULL: out = F(in)) or
 * apply_jacobian (:2918-3113; out = J(state)·in); boundary rows identity,
 * interior rows add the gauge term P·w^A·in */
static void nl_operator(long nx, long ny, long nz, const double* xc, const double* yc, const double* zc,
                        const double* mu, const double* ferro, const orc_bh* c, double penalty, const double* state,
                        const double* in, double* out)
{
    ms_grid g = {nx, ny, nz, xc, yc, zc};
    long ndof = 3 * nx * ny * nz;
    for (long r = 0; r < ndof; ++r) out[r] = 0.0;
    for (long k = 1; k + 1 < nz; ++k)
        for (long j = 1; j + 1 < ny; ++j)
            for (long i = 1; i + 1 < nx; ++i) {
                long s = i + nx * (j + ny * k);
                double dx = local_spacing(xc, nx, i), dy = local_spacing(yc, ny, j), dz = local_spacing(zc, nz, k);
                double wb[3] = {dx / (dy * dz), dy / (dx * dz), dz / (dx * dy)};
                int fe = ferro[s] > 0.5;
                double Bin[3], H[3];
                moments_at(&g, i, j, k, in, Bin);
                if (state) {
                    double B0[3], J[9];
                    moments_at(&g, i, j, k, state, B0);
                    nl_response(c, fe, mu[s], B0, NULL, J);
                    for (int m = 0; m < 3; ++m) H[m] = J[3 * m] * Bin[0] + J[3 * m + 1] * Bin[1] + J[3 * m + 2] * Bin[2];
                } else {
                    nl_response(c, fe, mu[s], Bin, H, NULL);
                }
                for (int m = 0; m < 3; ++m) {
                    double hm = wb[m] * H[m] * (m == 1 ? g_sw.y_sign : 1.0);
                    for (int t = 0; t < 4; ++t) {
                        const ms_tap* tp = &MOMENT[m][t];
                        long pi = i + tp->di, pj = j + tp->dj, pk = k + tp->dk;
                        if (is_boundary(&g, pi, pj, pk)) continue;
                        out[dof(&g, pi, pj, pk, tp->comp)] += tp->coeff * hm;
                    }
                }
            }
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i)
                if (is_boundary(&g, i, j, k))
                    for (int cc = 0; cc < 3; ++cc) out[dof(&g, i, j, k, cc)] = in[dof(&g, i, j, k, cc)];
    ms_add_gauge(&g, penalty, in, out);
}

void orc_nl_apply(long nx, long ny, long nz, const double* xc, const double* yc, const double* zc, const double* mu,
                  const double* ferro, const double* bh_b, const double* bh_h, int nbh, double penalty,
                  const double* in, double* out)
{
    orc_bh c;
    bh_init(&c, bh_b, bh_h, nbh);
    nl_operator(nx, ny, nz, xc, yc, zc, mu, ferro, &c, penalty, NULL, in, out);
}

void orc_nl_jacobian_apply(long nx, long ny, long nz, const double* xc, const double* yc, const double* zc,
                           const double* mu, const double* ferro, const double* bh_b, const double* bh_h, int nbh,
                           double penalty, const double* state, const double* in, double* out)
{
    orc_bh c;
    bh_init(&c, bh_b, bh_h, nbh);
    nl_operator(nx, ny, nz, xc, yc, zc, mu, ferro, &c, penalty, state, in, out);
}

/* assemble_matrix_data(model, state) (onelab_interface/magnetostatics_onelab
 * .hpp:3115-3536) of the Jacobian J(state): as orc_ms_assemble_csr, each
 * transposed moment slot (moment row 3s+m, coefficient c_r) now contributes,
 * for every moment component m2 whose coefficient w^B_m(s)·∂²H/∂B_m∂B_m2 at
 * the state's induction is nonzero, c_r·(that coefficient)·c_t for every
 * interior tap t of the moment stencil of m2 at s (:3243-3320).  Rows come
 * out with ascending columns; cols/vals must hold 27·3·nx·ny·nz entries;
 * returns the number of stored entries, or -1 when a row exceeds
 * ORC_ROW_MAX. */
long orc_nl_jacobian_assemble_csr(long nx, long ny, long nz, const double* xc, const double* yc, const double* zc,
                                  const double* mu, const double* ferro, const double* bh_b, const double* bh_h,
                                  int nbh, double penalty, const double* state, long* row_ptr, int* cols,
                                  double* vals)
{
    orc_bh cv;
    bh_init(&cv, bh_b, bh_h, nbh);
    ms_grid g = {nx, ny, nz, xc, yc, zc};
    long ndof = 3 * nx * ny * nz, nn = nx * ny * nz;
    int* tcount = (int*)calloc(ndof, sizeof(int));
    orc_tslot* tslot = (orc_tslot*)malloc(sizeof(orc_tslot) * 4 * (ndof > 0 ? ndof : 1));
    double* Jn = (double*)calloc(9 * (nn > 0 ? nn : 1), sizeof(double)); /* ∂²H/∂B∂B per interior node */
    for (long k = 1; k + 1 < nz; ++k)
        for (long j = 1; j + 1 < ny; ++j)
            for (long i = 1; i + 1 < nx; ++i) {
                long sn = i + nx * (j + ny * k);
                double B0[3];
                moments_at(&g, i, j, k, state, B0);
                nl_response(&cv, ferro[sn] > 0.5, mu[sn], B0, NULL, Jn + 9 * sn);
                for (int m = 0; m < 3; ++m)
                    for (int t = 0; t < 4; ++t) {
                        const ms_tap* tp = &MOMENT[m][t];
                        long pi = i + tp->di, pj = j + tp->dj, pk = k + tp->dk;
                        if (is_boundary(&g, pi, pj, pk)) continue;
                        long r = dof(&g, pi, pj, pk, tp->comp);
                        orc_tslot* sl = &tslot[4 * r + tcount[r]++];
                        sl->moment_row = 3 * sn + m;
                        sl->coeff = tp->coeff;
                    }
            }
    long nnz = 0;
    int lc[ORC_ROW_MAX];
    double lv[ORC_ROW_MAX];
    for (long r = 0; r < ndof; ++r) {
        long node = r / 3;
        int c = (int)(r % 3);
        long i = node % nx, j = (node / nx) % ny, k = node / (nx * ny);
        row_ptr[r] = nnz;
        if (is_boundary(&g, i, j, k)) {
            cols[nnz] = (int)r;
            vals[nnz] = 1.0;
            ++nnz;
            continue;
        }
        int count = 0;
        ms_gauge_entries(&g, i, j, k, c, penalty, &count, lc, lv);
        for (int sidx = 0; sidx < tcount[r]; ++sidx) {
            long mrow = tslot[4 * r + sidx].moment_row;
            double rc = tslot[4 * r + sidx].coeff;
            long sn = mrow / 3;
            int m = (int)(mrow % 3);
            long si = sn % nx, sj = (sn / nx) % ny, sk = sn / (nx * ny);
            double sx = local_spacing(xc, nx, si), sy = local_spacing(yc, ny, sj), sz = local_spacing(zc, nz, sk);
            double hf = m == 0 ? sx / (sy * sz) : (m == 1 ? g_sw.y_sign * (sy / (sx * sz)) : sz / (sx * sy));
            for (int m2 = 0; m2 < 3; ++m2) {
                double jc = hf * Jn[9 * sn + 3 * m + m2];
                if (jc == 0.0) continue;
                for (int t = 0; t < 4; ++t) {
                    const ms_tap* tp = &MOMENT[m2][t];
                    long pi = si + tp->di, pj = sj + tp->dj, pk = sk + tp->dk;
                    if (is_boundary(&g, pi, pj, pk)) continue;
                    if (count >= ORC_ROW_MAX) {
                        free(tcount);
                        free(tslot);
                        free(Jn);
                        return -1;
                    }
                    add_entry(&count, lc, lv, dof(&g, pi, pj, pk, tp->comp), rc * jc * tp->coeff);
                }
            }
        }
        for (int a = 1; a < count; ++a) {
            int kc = lc[a];
            double kv = lv[a];
            int b = a - 1;
            while (b >= 0 && lc[b] > kc) {
                lc[b + 1] = lc[b];
                lv[b + 1] = lv[b];
                --b;
            }
            lc[b + 1] = kc;
            lv[b + 1] = kv;
        }
        for (int a = 0; a < count; ++a) {
            cols[nnz] = lc[a];
            vals[nnz] = lv[a];
            ++nnz;
        }
    }
    row_ptr[ndof] = nnz;
    free(tcount);
    free(tslot);
    free(Jn);
    return nnz;
}

/* Jacobi diagonal of J(state), probed column by column in 27 colour
 * classes (the same probing as orc_ms_diagonal) */
void orc_nl_jacobian_diagonal(long nx, long ny, long nz, const double* xc, const double* yc, const double* zc,
                              const double* mu, const double* ferro, const double* bh_b, const double* bh_h, int nbh,
                              double penalty, const double* state, double* diag)
{
    orc_bh c;
    bh_init(&c, bh_b, bh_h, nbh);
    long ndof = 3 * nx * ny * nz;
    double* e = (double*)calloc(ndof, sizeof(double));
    double* o = (double*)malloc(sizeof(double) * ndof);
    for (int ci = 0; ci < 3; ++ci)
        for (int cj = 0; cj < 3; ++cj)
            for (int ck = 0; ck < 3; ++ck)
                for (int cc = 0; cc < 3; ++cc) {
                    for (long r = 0; r < ndof; ++r) e[r] = 0.0;
                    for (long k = ck; k < nz; k += 3)
                        for (long j = cj; j < ny; j += 3)
                            for (long i = ci; i < nx; i += 3) e[3 * (i + nx * (j + ny * k)) + cc] = 1.0;
                    nl_operator(nx, ny, nz, xc, yc, zc, mu, ferro, &c, penalty, state, e, o);
                    for (long k = ck; k < nz; k += 3)
                        for (long j = cj; j < ny; j += 3)
                            for (long i = ci; i < nx; i += 3) {
                                long r = 3 * (i + nx * (j + ny * k)) + cc;
                                diag[r] = o[r];
                            }
                }
    free(e);
    free(o);
}

/* the nonlinear branch of minimize_strong_formulation_residual
 * (src/similie/solvers/minimize_strong_formulation_residual.hpp:1383-1569):
 * x₀ = 0, r = b − F(x); at most 30 Newton steps while ‖r‖/‖r₀‖ > tol; each
 * solves J(x) δ = r by CG to the same relative tolerance (Jacobi from the
 * Jacobian's diagonal when use_jacobi), then a backtracking line search
 * (α = 1, halved up to 8 times until ‖r(x + αδ)‖ < ‖r(x)‖, keeping the best
 * α tried).  out[0] = Newton steps, out[1] = Σ inner CG iterations,
 * out[2] = ‖r₀‖, out[3] = final ‖r‖; returns 1 when converged. */
/* one nonlinear problem for the Newton loop: F(x), J(state)·v and the
 * diagonal of J(state) */
typedef struct nl_problem {
    long n;
    void (*F)(const struct nl_problem* P, const double* x, double* fx);
    void (*J)(const struct nl_problem* P, const double* state, const double* in, double* out);
    void (*D)(const struct nl_problem* P, const double* state, double* diag);
    const void* data;
    /* optional: J(state) assembled (ascending columns; returns nnz, or < 0) */
    long (*A)(const struct nl_problem* P, const double* state, long* row_ptr, int* cols, double* vals);
} nl_problem;

/* textbook preconditioned CG as orc_cg with z = M r supplied by `M` */
typedef void (*orc_prec_fn)(void* ctx, const double* r, double* z);

static int orc_pcg(long n, orc_matvec_fn A, void* ctx, orc_prec_fn M, void* mctx, const double* b, double tol,
                   int max_it, double* x, double* res_out)
{
    double* r = (double*)malloc(sizeof(double) * n);
    double* z = (double*)malloc(sizeof(double) * n);
    double* p = (double*)malloc(sizeof(double) * n);
    double* q = (double*)malloc(sizeof(double) * n);
    for (long i = 0; i < n; ++i) {
        x[i] = 0.0;
        r[i] = b[i];
    }
    M(mctx, r, z);
    for (long i = 0; i < n; ++i) p[i] = z[i];
    double bnorm = sqrt(dot(n, b, b));
    double rz = dot(n, r, z);
    double rnorm = bnorm;
    int it = 0;
    while (bnorm > 0.0 && it < max_it && rnorm > tol * bnorm) {
        A(ctx, p, q);
        double pq = dot(n, p, q);
        double alpha = rz / pq;
        for (long i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
        }
        M(mctx, r, z);
        double rz_new = dot(n, r, z);
        rnorm = sqrt(dot(n, r, r));
        double beta = rz_new / rz;
        rz = rz_new;
        for (long i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        ++it;
    }
    res_out[0] = rnorm;
    res_out[1] = bnorm;
    free(r);
    free(z);
    free(p);
    free(q);
    return it;
}

/* z = Gᵀ (G r) with the SpdIsai factor G (CSR) */
typedef struct {
    long n;
    const long* rp;
    const int* cols;
    const double* vals;
    double* t;
} isai_ctx;

static void isai_apply(void* c, const double* r, double* z)
{
    isai_ctx* k = (isai_ctx*)c;
    for (long i = 0; i < k->n; ++i) {
        double v = 0.0;
        for (long q = k->rp[i]; q < k->rp[i + 1]; ++q) v += k->vals[q] * r[k->cols[q]];
        k->t[i] = v;
    }
    for (long i = 0; i < k->n; ++i) z[i] = 0.0;
    for (long i = 0; i < k->n; ++i)
        for (long q = k->rp[i]; q < k->rp[i + 1]; ++q) z[k->cols[q]] += k->vals[q] * k->t[i];
}

typedef struct {
    const nl_problem* P;
    const double* state;
} nl_jctx;

static void nl_problem_jmv(void* p, const double* in, double* out)
{
    nl_jctx* k = (nl_jctx*)p;
    k->P->J(k->P, k->state, in, out);
}

static int newton_core(const nl_problem* P, const double* b, double tol, int max_inner, int use_jacobi, double* x,
                       double* out)
{
    long n = P->n;
    double* r = (double*)malloc(sizeof(double) * n);
    double* fx = (double*)malloc(sizeof(double) * n);
    double* delta = (double*)malloc(sizeof(double) * n);
    double* xc_ = (double*)malloc(sizeof(double) * n);
    double* rc = (double*)malloc(sizeof(double) * n);
    double* diag = use_jacobi == 1 ? (double*)malloc(sizeof(double) * n) : NULL;
    for (long i = 0; i < n; ++i) x[i] = 0.0;
    P->F(P, x, fx);
    for (long i = 0; i < n; ++i) r[i] = b[i] - fx[i];
    double r0 = sqrt(dot(n, r, r)), rn = r0;
    int steps = 0, inner = 0, converged = 0;
    if (r0 == 0.0) converged = 1;
    for (int it = 0; it < 30 && !converged; ++it) {
        rn = sqrt(dot(n, r, r));
        if (rn / r0 <= tol) {
            converged = 1;
            break;
        }
        nl_jctx ctx = {P, x};
        double res[2];
        if (use_jacobi == 2) {
            /* SpdIsai of J(x), assembled at every Newton step
             * (minimize_strong_formulation_residual.hpp:1432-1433) */
            long* arp = (long*)malloc(sizeof(long) * (n + 1));
            int* acol = (int*)malloc(sizeof(int) * 27 * n);
            double* aval = (double*)malloc(sizeof(double) * 27 * n);
            long* grp = (long*)malloc(sizeof(long) * (n + 1));
            int* gcol = (int*)malloc(sizeof(int) * 27 * n);
            double* gval = (double*)malloc(sizeof(double) * 27 * n);
            long annz = P->A(P, x, arp, acol, aval);
            long grc = annz >= 0 ? orc_spd_isai(n, arp, acol, aval, grp, gcol, gval) : -1;
            isai_ctx ic = {n, grp, gcol, gval, (double*)malloc(sizeof(double) * n)};
            if (grc == 0)
                inner += orc_pcg(n, nl_problem_jmv, &ctx, isai_apply, &ic, r, tol, max_inner, delta, res);
            else
                for (long i = 0; i < n; ++i) delta[i] = 0.0;
            free(ic.t);
            free(arp);
            free(acol);
            free(aval);
            free(grp);
            free(gcol);
            free(gval);
            if (grc != 0) break;
        } else {
            if (diag) P->D(P, x, diag);
            inner += orc_cg(n, nl_problem_jmv, &ctx, r, diag, tol, max_inner, delta, res);
        }
        double alpha = 1.0, best_alpha = 0.0, best = INFINITY;
        for (int ls = 0; ls < 8; ++ls) {
            for (long i = 0; i < n; ++i) xc_[i] = x[i] + alpha * delta[i];
            P->F(P, xc_, fx);
            for (long i = 0; i < n; ++i) rc[i] = b[i] - fx[i];
            double cn = sqrt(dot(n, rc, rc));
            if (cn < best) {
                best_alpha = alpha;
                best = cn;
            }
            if (cn < rn) break;
            alpha *= 0.5;
        }
        if (best_alpha != alpha) {
            alpha = best_alpha;
            for (long i = 0; i < n; ++i) xc_[i] = x[i] + alpha * delta[i];
            P->F(P, xc_, fx);
            for (long i = 0; i < n; ++i) rc[i] = b[i] - fx[i];
        }
        for (long i = 0; i < n; ++i) {
            x[i] = xc_[i];
            r[i] = rc[i];
        }
        ++steps;
    }
    rn = sqrt(dot(n, r, r));
    out[0] = steps;
    out[1] = inner;
    out[2] = r0;
    out[3] = rn;
    free(r);
    free(fx);
    free(delta);
    free(xc_);
    free(rc);
    free(diag);
    return converged;
}

typedef struct {
    long nx, ny, nz;
    const double *x, *y, *z, *mu, *ferro;
    const double *bh_b, *bh_h;
    int nbh;
    const orc_bh* c;
    double penalty;
} nl3_data;

static void nl3_F(const nl_problem* P, const double* x, double* fx)
{
    const nl3_data* d = (const nl3_data*)P->data;
    nl_operator(d->nx, d->ny, d->nz, d->x, d->y, d->z, d->mu, d->ferro, d->c, d->penalty, NULL, x, fx);
}

static void nl3_J(const nl_problem* P, const double* state, const double* in, double* out)
{
    const nl3_data* d = (const nl3_data*)P->data;
    nl_operator(d->nx, d->ny, d->nz, d->x, d->y, d->z, d->mu, d->ferro, d->c, d->penalty, state, in, out);
}

static void nl3_D(const nl_problem* P, const double* state, double* diag)
{
    const nl3_data* d = (const nl3_data*)P->data;
    orc_nl_jacobian_diagonal(d->nx, d->ny, d->nz, d->x, d->y, d->z, d->mu, d->ferro, d->bh_b, d->bh_h, d->nbh,
                             d->penalty, state, diag);
}

static long nl3_A(const nl_problem* P, const double* state, long* row_ptr, int* cols, double* vals)
{
    const nl3_data* d = (const nl3_data*)P->data;
    return orc_nl_jacobian_assemble_csr(d->nx, d->ny, d->nz, d->x, d->y, d->z, d->mu, d->ferro, d->bh_b, d->bh_h,
                                        d->nbh, d->penalty, state, row_ptr, cols, vals);
}

/* use_jacobi: 0 none, 1 Jacobi, 2 SpdIsai of the assembled Jacobian at
 * every step */
int orc_nl_newton(long nx, long ny, long nz, const double* xc, const double* yc, const double* zc, const double* mu,
                  const double* ferro, const double* bh_b, const double* bh_h, int nbh, double penalty,
                  const double* b, double tol, int max_inner, int use_jacobi, double* x, double* out)
{
    orc_bh c;
    bh_init(&c, bh_b, bh_h, nbh);
    nl3_data d = {nx, ny, nz, xc, yc, zc, mu, ferro, bh_b, bh_h, nbh, &c, penalty};
    nl_problem P = {3 * nx * ny * nz, nl3_F, nl3_J, nl3_D, &d, nl3_A};
    return newton_core(&P, b, tol, max_inner, use_jacobi, x, out);
}

/* ------------------------------------------------------------------ */
/* cell post-processing of a 3D solution                               */
/* ------------------------------------------------------------------ */

void orc_codifferential(int n, const long* ext, int pos_kind, const double* const* axis, const double* nodal,
                        int metric_kind, const double* metric, int k, const double* in, const long* out_ext,
                        double* out);

/* Per cell (i, j, k), i fastest, of the (nx-1)(ny-1)(nz-1) cells:
 *   induction  B_m = moment_m / (product of the two other cell widths), the
 *              moment read from the full potential with the forward_value
 *              stencil (magnetostatics_onelab.hpp:3826-3864, 3984-4006);
 *   field      H = h(|B|)/|B|·B on nonlinear cells (0 at B = 0), B/μ_cell
 *              otherwise, unit Hodge factors (:4007-4024,
 *              nonlinear_magnetostatics.py:266-281, linear_magnetostatics.py
 *              :38-49);
 *   stress     B_0H_0 − ½B·H, B_1H_1 − ½B·H, B_2H_2 − ½B·H, B_0H_1, B_0H_2,
 *              B_1H_2 (make_cell_post_process_fields, :4029-4046);
 *   force      per xy slice of cells, −δ of the one-forms (T_0, T_3),
 *              (T_3, T_1), (T_4, T_5) on the grid of cell centres with a
 *              stored identity metric (fill_force_density_on_cell_domain
 *              and its xy-slice driver, :4048-4168, 4215-4254), δ being
 *              this oracle's codifferential (codifferential.hpp:321-498).
 * Any output may be NULL except that force needs stress. */
static void ms2_moments(long nx, long s, const double* v, double B[3]);
void orc_reconstruction(int n, const long* ext, int pos_kind, const double* const* axis, const double* nodal,
                        int k, const double* in, double* out);

/* H and the Maxwell stress of one cell from its B */
static void cell_fields(const double B[3], int nonlinear, double mu, const orc_bh* c, double H[3], double t[6])
{
    for (int m = 0; m < 3; ++m) H[m] = 0.0;
    if (nonlinear) {
        double q = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
        double bn = sqrt(q);
        if (bn != 0.0) {
            double scale = bh_h_from_b(c, bn) / bn;
            for (int m = 0; m < 3; ++m) H[m] = 1.0 * scale * B[m];
        }
    } else {
        for (int m = 0; m < 3; ++m) H[m] = 1.0 * B[m] / mu;
    }
    double half = 0.5 * (B[0] * H[0] + B[1] * H[1] + B[2] * H[2]);
    t[0] = B[0] * H[0] - half;
    t[1] = B[1] * H[1] - half;
    t[2] = B[2] * H[2] - half;
    t[3] = B[0] * H[1];
    t[4] = B[0] * H[2];
    t[5] = B[1] * H[2];
}

/* fill_force_density_on_cell_domain (:4048-4168) on one slice of mx × my
 * cells (x fastest): −δ of (T_0, T_3), (T_3, T_1), (T_4, T_5) on the grid of
 * cell centres with a stored identity metric */
static void slice_force(const double* xc, const double* yc, long mx, long my, const double* T, double* force)
{
    double* cx = (double*)malloc(sizeof(double) * mx);
    double* cy = (double*)malloc(sizeof(double) * my);
    for (long i = 0; i < mx; ++i) cx[i] = 0.5 * (xc[i] + xc[i + 1]);
    for (long j = 0; j < my; ++j) cy[j] = 0.5 * (yc[j] + yc[j + 1]);
    const double* axis[2] = {cx, cy};
    long ext[2] = {mx, my};
    long nn = mx * my;
    double* metric = (double*)malloc(sizeof(double) * 3 * nn);
    double* one = (double*)malloc(sizeof(double) * 2 * nn);
    double* out = (double*)malloc(sizeof(double) * nn);
    for (long q = 0; q < nn; ++q) {
        metric[3 * q] = 1.0;
        metric[3 * q + 1] = 0.0;
        metric[3 * q + 2] = 1.0;
    }
    static const int pick[3][2] = {{0, 3}, {3, 1}, {4, 5}};
    for (int comp = 0; comp < 3; ++comp) {
        /* the oracle's node order has the first axis slowest */
        for (long i = 0; i < mx; ++i)
            for (long j = 0; j < my; ++j) {
                const double* t = T + 6 * (i + mx * j);
                one[2 * (i * my + j)] = t[pick[comp][0]];
                one[2 * (i * my + j) + 1] = t[pick[comp][1]];
            }
        orc_codifferential(2, ext, 0, axis, NULL, 2, metric, 1, one, ext, out);
        for (long i = 0; i < mx; ++i)
            for (long j = 0; j < my; ++j) force[3 * (i + mx * j) + comp] = -out[i * my + j];
    }
    free(cx);
    free(cy);
    free(metric);
    free(one);
    free(out);
}

void orc_ms_post_process(long nx, long ny, long nz, const double* xc, const double* yc, const double* zc,
                         const double* potential, const double* cell_mu, const double* cell_nonlinear,
                         const double* bh_b, const double* bh_h, int nbh, double* induction, double* field,
                         double* stress, double* force)
{
    ms_grid g = {nx, ny, nz, xc, yc, zc};
    long mx = nx - 1, my = ny - 1, mz = nz - 1;
    long ncell = mx * my * mz;
    orc_bh c;
    if (nbh >= 2) bh_init(&c, bh_b, bh_h, nbh);
    double* T = stress ? stress : (double*)malloc(sizeof(double) * 6 * (ncell > 0 ? ncell : 1));
    for (long kk = 0; kk < mz; ++kk)
        for (long j = 0; j < my; ++j)
            for (long i = 0; i < mx; ++i) {
                long cell = i + mx * (j + my * kk);
                double w[3] = {xc[i + 1] - xc[i], yc[j + 1] - yc[j], zc[kk + 1] - zc[kk]};
                double mom[3];
                for (int m = 0; m < 3; ++m) {
                    double v = 0.0;
                    for (int t = 0; t < 4; ++t) {
                        const ms_tap* tp = &MOMENT[m][t];
                        v += tp->coeff * potential[dof(&g, i + tp->di, j + tp->dj, kk + tp->dk, tp->comp)];
                    }
                    mom[m] = v;
                }
                double B[3] = {mom[0] / (w[1] * w[2]), mom[1] / (w[0] * w[2]), mom[2] / (w[0] * w[1])};
                double H[3];
                cell_fields(B, cell_nonlinear && cell_nonlinear[cell] != 0.0, cell_mu[cell], &c, H, T + 6 * cell);
                for (int m = 0; m < 3; ++m) {
                    if (induction) induction[3 * cell + m] = B[m];
                    if (field) field[3 * cell + m] = H[m];
                }
            }
    if (force && mx > 0 && my > 0)
        for (long kk = 0; kk < mz; ++kk) slice_force(xc, yc, mx, my, T + 6 * mx * my * kk, force + 3 * mx * my * kk);
    if (!stress) free(T);
}

/* The 2D counterpart (fill_post_process_fields_on_cell_domain, :3866-3982,
 * and fill_force_density_on_quadrilateral_grid, :4170-4213): per cell s the
 * in-plane moment pair, reduced (X: −ΔxA, Y: ΔyA) and reconstructed at the
 * cell's node with the node positions (the diagonal inverse of the edge
 * lengths, this oracle's orc_reconstruction), B = (rec_Y, rec_X, 0); H,
 * stress and force as in 3D on the single slice.  potential: A_z per node. */
void orc_ms2d_post_process(long nx, long ny, const double* xc, const double* yc, const double* potential,
                           const double* cell_mu, const double* cell_nonlinear, const double* bh_b,
                           const double* bh_h, int nbh, double* induction, double* field, double* stress,
                           double* force)
{
    long mx = nx - 1, my = ny - 1, ncell = mx * my;
    orc_bh c;
    if (nbh >= 2) bh_init(&c, bh_b, bh_h, nbh);
    double* T = stress ? stress : (double*)malloc(sizeof(double) * 6 * (ncell > 0 ? ncell : 1));
    const double* axis[2] = {xc, yc};
    long ext[2] = {nx, ny};
    double* reduced = (double*)malloc(sizeof(double) * 2 * nx * ny);
    double* rec = (double*)malloc(sizeof(double) * 2 * nx * ny);
    for (long p = 0; p < 2 * nx * ny; ++p) reduced[p] = 0.0;
    for (long j = 0; j < my; ++j)
        for (long i = 0; i < mx; ++i) {
            long s = i + nx * j;
            double Bm[3];
            ms2_moments(nx, s, potential, Bm);
            /* the oracle's node order has the first axis slowest */
            reduced[2 * (i * ny + j)] = Bm[1];
            reduced[2 * (i * ny + j) + 1] = Bm[0];
        }
    orc_reconstruction(2, ext, 0, axis, NULL, 1, reduced, rec);
    for (long j = 0; j < my; ++j)
        for (long i = 0; i < mx; ++i) {
            long cell = i + mx * j;
            double B[3] = {rec[2 * (i * ny + j) + 1], rec[2 * (i * ny + j)], 0.0};
            double H[3];
            cell_fields(B, cell_nonlinear && cell_nonlinear[cell] != 0.0, cell_mu[cell], &c, H, T + 6 * cell);
            for (int m = 0; m < 3; ++m) {
                if (induction) induction[3 * cell + m] = B[m];
                if (field) field[3 * cell + m] = H[m];
            }
        }
    if (force && mx > 0 && my > 0) slice_force(xc, yc, mx, my, T, force);
    free(reduced);
    free(rec);
    if (!stress) free(T);
}

/* ------------------------------------------------------------------ */
/* 2D magnetostatics on A_z                                            */
/* ------------------------------------------------------------------ */

/* The tables of MagnetostaticsOperator2D (magnetostatics_onelab.hpp
 * :1051-1262), restated literally.  Per node, the two moment stencils of
 * MagneticVectorPotentialToMagneticInduction<X, Y>::forward_value
 * (magnetostatics_quantities.hpp:97-122) in the stencil domain's order:
 * B_x = −A(s) + A(s+e_y), B_y = +A(s) − A(s+e_x) (flat index i + nx·j, also
 * past the grid's edge, where only boundary nodes reach).  Per row, the
 * transposed slots: sampled rows ascending, interior sampled nodes only;
 * the outer slots hold the negated coefficients. */
#define MS2_SLOTS 4
typedef struct {
    long nx, ny;
    long mcol[2][2];
    double mcoef[2][2];
    int tcount[2], ocount[2];
    long tcol[2][MS2_SLOTS], ocol[2][MS2_SLOTS];
    double tcoef[2][MS2_SLOTS], ocoef[2][MS2_SLOTS];
} ms2_row_tables;

typedef struct {
    long nx, ny;
    const double* mu;
    const double* ferro; /* NULL: linear Hamiltonian */
    const orc_bh* c;
    int criterion;
} ms2_model;

static void ms2_stencil(long nx, long s, int m, long col[2], double coef[2])
{
    if (m == 0) {
        col[0] = s;
        coef[0] = -1.0;
        col[1] = s + nx;
        coef[1] = 1.0;
    } else {
        col[0] = s;
        coef[0] = 1.0;
        col[1] = s + 1;
        coef[1] = -1.0;
    }
}

static int ms2_interior(long nx, long ny, long s)
{
    long i = s % nx, j = s / nx;
    return !(i == 0 || j == 0 || i + 1 == nx || j + 1 == ny);
}

/* the transposed and outer slots of one row: the scan of the constructor
 * (:1180-1241) restricted to the sampled rows whose stencils can reach it */
static void ms2_row(long nx, long ny, long row, ms2_row_tables* t)
{
    memset(t, 0, sizeof *t);
    long cand[3] = {row - nx, row - 1, row}; /* ascending */
    for (int q = 0; q < 3; ++q) {
        long s = cand[q];
        if (s < 0 || !ms2_interior(nx, ny, s)) continue;
        for (int m = 0; m < 2; ++m) {
            long col[2];
            double coef[2];
            ms2_stencil(nx, s, m, col, coef);
            for (int k = 0; k < 2; ++k) {
                if (col[k] != row || col[k] >= nx * ny) continue;
                t->tcol[m][t->tcount[m]] = s;
                t->tcoef[m][t->tcount[m]++] = coef[k];
                t->ocol[m][t->ocount[m]] = s;
                t->ocoef[m][t->ocount[m]++] = -coef[k];
            }
        }
    }
}

static void ms2_moments(long nx, long s, const double* v, double B[3])
{
    for (int m = 0; m < 2; ++m) {
        long col[2];
        double coef[2];
        ms2_stencil(nx, s, m, col, coef);
        double acc = 0.0;
        for (int k = 0; k < 2; ++k) acc += coef[k] * v[col[k]];
        B[m] = acc;
    }
    B[2] = 0.0;
}

/* dpotential_dt<m>(moments, s) at the state of v, or with state given the
 * Jacobian row m applied to δB (jacobian_component, :666-682) */
static double ms2_response(const ms2_model* M, long s, int m, const double* state, const double* v)
{
    double B[3], H[3], J[9];
    int fe = M->ferro && M->ferro[s] > 0.5;
    if (!state) {
        ms2_moments(M->nx, s, v, B);
        if (!M->ferro) return B[m] / M->mu[s];
        nl_response(M->c, fe, M->mu[s], B, H, NULL);
        return H[m];
    }
    double dB[3];
    ms2_moments(M->nx, s, state, B);
    ms2_moments(M->nx, s, v, dB);
    if (!M->ferro) return (1.0 / M->mu[s]) * dB[m];
    nl_response(M->c, fe, M->mu[s], B, NULL, J);
    return J[3 * m] * dB[0] + J[3 * m + 1] * dB[1];
}

/* apply (:1277-1439) / apply_jacobian (:2790-2920) of the 2D operator */
static void ms2_operator(const ms2_model* M, const double* state, const double* in, double* out)
{
    long nx = M->nx, ny = M->ny;
    for (long row = 0; row < nx * ny; ++row) {
        if (!ms2_interior(nx, ny, row)) {
            out[row] = in[row];
            continue;
        }
        ms2_row_tables t;
        ms2_row(nx, ny, row, &t);
        double residual = 0.0;
        if (M->criterion == 0 || M->criterion == 2)
            for (int m = 0; m < 2; ++m)
                for (int k = 0; k < t.tcount[m]; ++k)
                    residual += t.tcoef[m][k] * ms2_response(M, t.tcol[m][k], m, state, in);
        if (M->criterion == 1 || M->criterion == 2)
            for (int m = 0; m < 2; ++m)
                for (int k = 0; k < t.ocount[m]; ++k)
                    residual -= t.ocoef[m][k] * ms2_response(M, t.ocol[m][k], m, state, in);
        out[row] = residual;
    }
}

/* F(in) (state = NULL; linear when ferro = NULL) or J(state)·in */
void orc_ms2d_apply(long nx, long ny, const double* mu, const double* ferro, const double* bh_b, const double* bh_h,
                    int nbh, int criterion, const double* state, const double* in, double* out)
{
    orc_bh c;
    if (ferro) bh_init(&c, bh_b, bh_h, nbh);
    ms2_model M = {nx, ny, mu, ferro, ferro ? &c : NULL, criterion};
    ms2_operator(&M, state, in, out);
}

/* assemble_matrix_data(model, state) of the 2D operator
 * (onelab_interface/magnetostatics_onelab.hpp:3554-3748), restated: a
 * boundary row holds 1 on its diagonal; an interior row gathers, for every
 * transposed slot (potential criterion: moment x slots, then moment y
 * slots) and then every outer slot (moments criterion, with the negated
 * coefficient), the block of its sampled node s — at the state's induction
 * B0(s), h = ∂²H/∂B∂B(s) (1/μ on the diagonal for the linear Hamiltonian) —
 * rc·h_{x,m}·c_k over the moment-x stencil and rc·h_{y,m}·c_k over the
 * moment-y stencil of s, where m is the slot's moment; a nonzero value is
 * added to its column (summed with an earlier one in the row, or
 * appended).  Columns whose sum is zero are dropped and the rest sorted
 * (matrix_data → CSR).  ferro = NULL: the linear Hamiltonian.  cols/vals
 * hold 8 entries per row. */
static long ms2_assemble(const ms2_model* M, const double* state, long* row_ptr, int* cols, double* vals)
{
    const long nx = M->nx, ny = M->ny;
    const double* mu = M->mu;
    const double* ferro = M->ferro;
    const int criterion = M->criterion;
    long nnz = 0;
    for (long row = 0; row < nx * ny; ++row) {
        row_ptr[row] = nnz;
        if (!ms2_interior(nx, ny, row)) {
            cols[nnz] = (int)row;
            vals[nnz] = 1.0;
            ++nnz;
            continue;
        }
        ms2_row_tables t;
        ms2_row(nx, ny, row, &t);
        int count = 0;
        long lc[8];
        double lv[8];
        #define MS2_ADD(col_, val_)                                  \
            do {                                                     \
                int found_ = 0;                                      \
                for (int q_ = 0; q_ < count; ++q_)                   \
                    if (lc[q_] == (col_)) {                          \
                        lv[q_] += (val_);                            \
                        found_ = 1;                                  \
                        break;                                       \
                    }                                                \
                if (!found_) {                                       \
                    lc[count] = (col_);                              \
                    lv[count] = (val_);                              \
                    ++count;                                         \
                }                                                    \
            } while (0)
        for (int part = 0; part < 2; ++part) {
            if (part == 0 && !(criterion == 0 || criterion == 2)) continue;
            if (part == 1 && !(criterion == 1 || criterion == 2)) continue;
            for (int m = 0; m < 2; ++m) {
                int cnt = part == 0 ? t.tcount[m] : t.ocount[m];
                for (int k = 0; k < cnt; ++k) {
                    long sn = part == 0 ? t.tcol[m][k] : t.ocol[m][k];
                    double rc = part == 0 ? t.tcoef[m][k] : -t.ocoef[m][k];
                    double B[3] = {0.0, 0.0, 0.0}, J[9];
                    if (state) ms2_moments(nx, sn, state, B);
                    if (ferro) {
                        nl_response(M->c, ferro[sn] > 0.5, mu[sn], B, NULL, J);
                    } else {
                        for (int q = 0; q < 9; ++q) J[q] = 0.0;
                        J[0] = J[4] = 1.0 / mu[sn];
                    }
                    /* h_{x,m} (moment-x stencil) and h_{y,m} (moment-y stencil) */
                    double hx = m == 0 ? J[0] : J[3], hy = m == 0 ? J[1] : J[4];
                    long col[2];
                    double coef[2];
                    ms2_stencil(nx, sn, 0, col, coef);
                    for (int q = 0; q < 2; ++q) {
                        double v = rc * (hx * coef[q]);
                        if (v != 0.0) MS2_ADD(col[q], v);
                    }
                    ms2_stencil(nx, sn, 1, col, coef);
                    for (int q = 0; q < 2; ++q) {
                        double v = rc * (hy * coef[q]);
                        if (v != 0.0) MS2_ADD(col[q], v);
                    }
                }
            }
        }
        #undef MS2_ADD
        /* drop zero sums, ascending columns */
        int kept = 0;
        for (int q = 0; q < count; ++q)
            if (lv[q] != 0.0) {
                lc[kept] = lc[q];
                lv[kept] = lv[q];
                ++kept;
            }
        for (int a = 1; a < kept; ++a) {
            long kc = lc[a];
            double kv = lv[a];
            int b = a - 1;
            while (b >= 0 && lc[b] > kc) {
                lc[b + 1] = lc[b];
                lv[b + 1] = lv[b];
                --b;
            }
            lc[b + 1] = kc;
            lv[b + 1] = kv;
        }
        for (int q = 0; q < kept; ++q) {
            cols[nnz] = (int)lc[q];
            vals[nnz] = lv[q];
            ++nnz;
        }
    }
    row_ptr[nx * ny] = nnz;
    return nnz;
}

long orc_ms2d_assemble_csr(long nx, long ny, const double* mu, const double* ferro, const double* bh_b,
                           const double* bh_h, int nbh, int criterion, const double* state, long* row_ptr, int* cols,
                           double* vals)
{
    orc_bh cv;
    if (ferro) bh_init(&cv, bh_b, bh_h, nbh);
    ms2_model M = {nx, ny, mu, ferro, ferro ? &cv : NULL, criterion};
    return ms2_assemble(&M, state, row_ptr, cols, vals);
}

/* the diagonal of ms2_operator(M, state, ·), probed with unit vectors in
 * 3×3 colour classes (a row reaches columns at most one node away) */
static void ms2_probe_diagonal(const ms2_model* M, const double* state, double* diag)
{
    long n = M->nx * M->ny;
    double* e = (double*)calloc(n, sizeof(double));
    double* o = (double*)malloc(sizeof(double) * n);
    for (int ci = 0; ci < 3; ++ci)
        for (int cj = 0; cj < 3; ++cj) {
            for (long r = 0; r < n; ++r) e[r] = 0.0;
            for (long j = cj; j < M->ny; j += 3)
                for (long i = ci; i < M->nx; i += 3) e[i + M->nx * j] = 1.0;
            ms2_operator(M, state, e, o);
            for (long j = cj; j < M->ny; j += 3)
                for (long i = ci; i < M->nx; i += 3) diag[i + M->nx * j] = o[i + M->nx * j];
        }
    free(e);
    free(o);
}

/* the diagonal of the linear operator (ferro = NULL) or of J(state) (a zero
 * state when state = NULL) */
void orc_ms2d_diagonal(long nx, long ny, const double* mu, const double* ferro, const double* bh_b,
                       const double* bh_h, int nbh, int criterion, const double* state, double* diag)
{
    orc_bh c;
    if (ferro) bh_init(&c, bh_b, bh_h, nbh);
    ms2_model M = {nx, ny, mu, ferro, ferro ? &c : NULL, criterion};
    double* zero = NULL;
    if (ferro && !state) state = zero = (double*)calloc(nx * ny, sizeof(double));
    ms2_probe_diagonal(&M, ferro ? state : NULL, diag);
    free(zero);
}

/* ¼·area·J_z of the adjacent cells at interior nodes (:4413-4445) */
void orc_ms2d_rhs(long nx, long ny, const double* xc, const double* yc, const double* jz, int criterion, double* rhs)
{
    long mx = nx - 1, my = ny - 1;
    for (long j = 0; j < ny; ++j)
        for (long i = 0; i < nx; ++i) {
            long node = i + nx * j;
            if (i == 0 || j == 0 || i + 1 == nx || j + 1 == ny || criterion == 0) {
                rhs[node] = 0.0;
                continue;
            }
            double acc = 0.0;
            for (int dj = -1; dj <= 0; ++dj)
                for (int di = -1; di <= 0; ++di) {
                    long ci = i + di, cj = j + dj;
                    if (ci < 0 || cj < 0 || ci >= mx || cj >= my) continue;
                    double area = (xc[ci + 1] - xc[ci]) * (yc[cj + 1] - yc[cj]);
                    acc += 0.25 * area * jz[ci + mx * cj];
                }
            rhs[node] = acc;
        }
}

/* mean of the adjacent cells (:4467-4500) */
void orc_ms2d_nodal_mu(long nx, long ny, const double* cell_mu, double mu0, double* mu)
{
    long mx = nx - 1, my = ny - 1;
    for (long j = 0; j < ny; ++j)
        for (long i = 0; i < nx; ++i) {
            double acc = 0.0;
            long count = 0;
            for (int dj = -1; dj <= 0; ++dj)
                for (int di = -1; di <= 0; ++di) {
                    long ci = i + di, cj = j + dj;
                    if (ci < 0 || cj < 0 || ci >= mx || cj >= my) continue;
                    acc += cell_mu[ci + mx * cj];
                    ++count;
                }
            mu[i + nx * j] = count == 0 ? mu0 : acc / (double)count;
        }
}

static void ms2_F(const nl_problem* P, const double* x, double* fx) { ms2_operator((const ms2_model*)P->data, NULL, x, fx); }

static void ms2_J(const nl_problem* P, const double* state, const double* in, double* out)
{
    ms2_operator((const ms2_model*)P->data, state, in, out);
}

static void ms2_D(const nl_problem* P, const double* state, double* diag)
{
    ms2_probe_diagonal((const ms2_model*)P->data, state, diag);
}

static long ms2_A(const nl_problem* P, const double* state, long* row_ptr, int* cols, double* vals)
{
    return ms2_assemble((const ms2_model*)P->data, state, row_ptr, cols, vals);
}

/* Newton on the 2D nonlinear operator, as orc_nl_newton (use_jacobi 2:
 * SpdIsai of the Jacobian assembled at every step) */
int orc_ms2d_newton(long nx, long ny, const double* mu, const double* ferro, const double* bh_b, const double* bh_h,
                    int nbh, int criterion, const double* b, double tol, int max_inner, int use_jacobi, double* x,
                    double* out)
{
    orc_bh c;
    bh_init(&c, bh_b, bh_h, nbh);
    ms2_model M = {nx, ny, mu, ferro, &c, criterion};
    nl_problem P = {nx * ny, ms2_F, ms2_J, ms2_D, &M, ms2_A};
    return newton_core(&P, b, tol, max_inner, use_jacobi, x, out);
}
