/*
 * CPU ORACLE — test infrastructure only.
 *
 * A plain-C restatement of SimiLie's discrete-exterior-calculus algorithms
 * (sil::exterior) and of the 3D magnetostatics operator, used by the test
 * suite, __graft_entry__.smoke() and the CPU-baseline leg of bench.py as the
 * checker.  Nothing in the product library links, loads or calls this file.
 *
 * Layout convention (same as the reference's Tensor over
 * DiscreteDomain<DDimX, DDimY, ..., Index> with layout_right,
 * src/similie/tensor/tensor_impl.hpp:587-602): axis 0 slowest, component
 * index fastest.  Component order of a k-form is the lexicographic order of
 * the k-subsets of the axes, which is both the antisymmetric mem_id order
 * (src/similie/tensor/antisymmetric_tensor.hpp:72-98) and the tangent-basis
 * order produced by std::prev_permutation
 * (src/similie/exterior/local_chain.hpp:413-434).
 *
 * Every function cites the reference file:line it restates.  The
 * coboundary follows the reference's boundary()-orientation algorithm
 * literally (simplex faces, reorientation, chain sign); the GPU kernels use
 * closed-form stencils instead, so agreement between the two is a real check.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAXD 4
#define ORC_MAXC 16 /* >= C(4,2); sized for the combination tables */

/* ------------------------------------------------------------------ */
/* combinatorics                                                       */
/* ------------------------------------------------------------------ */

static int popcount_u(unsigned m) { return __builtin_popcount(m); }

static long binom(int n, int k)
{
    if (k < 0 || k > n) return 0;
    long r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

/* k-subsets of {0..n-1} as bitmasks in lexicographic order of their sorted
 * index lists.  This is the sequence std::prev_permutation walks when it is
 * started from [1,...,1,0,...,0] (local_chain.hpp:423-431). */
static int combos(int n, int k, unsigned* out)
{
    int cnt = 0;
    int perm[ORC_MAXD];
    for (int i = 0; i < n; ++i) perm[i] = i < k ? 1 : 0;
    for (;;) {
        unsigned m = 0;
        for (int i = 0; i < n; ++i)
            if (perm[i]) m |= 1u << i;
        out[cnt++] = m;
        /* prev_permutation on a 0/1 sequence */
        int i = n - 1;
        while (i > 0 && perm[i - 1] <= perm[i]) --i;
        if (i == 0) break;
        int j = n - 1;
        while (perm[j] >= perm[i - 1]) --j;
        int t = perm[i - 1];
        perm[i - 1] = perm[j];
        perm[j] = t;
        for (int a = i, b = n - 1; a < b; ++a, --b) {
            t = perm[a];
            perm[a] = perm[b];
            perm[b] = t;
        }
    }
    return cnt;
}

static int combo_index(int n, int k, unsigned mask)
{
    unsigned c[16];
    int cnt = combos(n, k, c);
    for (int i = 0; i < cnt; ++i)
        if (c[i] == mask) return i;
    return -1;
}

/* parity of the concatenated index list (src ‖ dst); +1 even, -1 odd
 * (hodge_star.hpp:104-125, reduction_and_reconstruction.hpp:282-304) */
static int concat_sign(unsigned src, unsigned dst, int n)
{
    int ids[2 * ORC_MAXD], m = 0;
    for (int a = 0; a < n; ++a)
        if (src & (1u << a)) ids[m++] = a;
    for (int a = 0; a < n; ++a)
        if (dst & (1u << a)) ids[m++] = a;
    int odd = 0;
    for (int i = 0; i < m; ++i)
        for (int j = i + 1; j < m; ++j)
            if (ids[i] > ids[j]) odd ^= 1;
    return odd ? -1 : 1;
}

/* ------------------------------------------------------------------ */
/* grid / array helpers                                                */
/* ------------------------------------------------------------------ */

static long nodes_of(int n, const long* ext)
{
    long t = 1;
    for (int a = 0; a < n; ++a) t *= ext[a];
    return t;
}

static void unflatten(int n, const long* ext, long lin, long* idx)
{
    for (int a = n - 1; a >= 0; --a) {
        idx[a] = lin % ext[a];
        lin /= ext[a];
    }
}

static long flatten(int n, const long* ext, const long* idx)
{
    long lin = 0;
    for (int a = 0; a < n; ++a) lin = lin * ext[a] + idx[a];
    return lin;
}

static int inside(int n, const long* ext, const long* idx)
{
    for (int a = 0; a < n; ++a)
        if (idx[a] < 0 || idx[a] >= ext[a]) return 0;
    return 1;
}

/* clamp_to_domain (src/similie/misc/clamp_to_domain.hpp:17-47) */
static void clamp_idx(int n, const long* ext, long* idx)
{
    for (int a = 0; a < n; ++a) {
        if (idx[a] < 0) idx[a] = 0;
        if (idx[a] >= ext[a]) idx[a] = ext[a] - 1;
    }
}

/* ------------------------------------------------------------------ */
/* A.1 coboundary d — boundary()-orientation restatement               */
/* ------------------------------------------------------------------ */

/* One oriented face produced by boundary() of a (k+1)-simplex anchored at
 * `node` and spanned by the axes of `vmask`.  Restates
 * generate_local_half_subchain (boundary.hpp:43-65), the Simplex
 * reorientation of -1 entries (simplex.hpp:41-68, 118-143) and the final
 * chain sign (boundary.hpp:187). */
typedef struct {
    long off[ORC_MAXD]; /* face anchor = node + off */
    unsigned vmask;     /* face spanning axes */
    int negative;
    int second_half;
    int zeroed_axis;
} orc_face;

static int boundary_faces(int n, unsigned vmask, orc_face* faces)
{
    int kp1 = popcount_u(vmask); /* simplex dimension K' */
    int axes[ORC_MAXD], na = 0;
    for (int a = 0; a < n; ++a)
        if (vmask & (1u << a)) axes[na++] = a;
    int cnt = 0;
    /* first half: anchored at node, vector v, negative = K' % 2 */
    for (int i = 0; i < kp1; ++i) {
        orc_face f;
        memset(&f, 0, sizeof f);
        f.vmask = vmask & ~(1u << axes[i]);
        f.negative = ((kp1 % 2) + i) % 2;
        f.second_half = 0;
        f.zeroed_axis = axes[i];
        faces[cnt++] = f;
    }
    /* second half: anchored at node + v with vector -v.  Zeroing entry i
     * leaves k = K'-1 entries equal to -1; reorientation moves the anchor
     * back by one on each of them, makes the vector positive and flips the
     * orientation once per flipped entry. */
    for (int i = 0; i < kp1; ++i) {
        orc_face f;
        memset(&f, 0, sizeof f);
        int neg = i % 2;
        unsigned vm = vmask & ~(1u << axes[i]);
        for (int a = 0; a < n; ++a) {
            long o = (vmask & (1u << a)) ? 1 : 0; /* node + v */
            if (vm & (1u << a)) {                  /* entry -1 */
                o -= 1;
                neg ^= 1;
            }
            f.off[a] = o;
        }
        f.vmask = vm;
        f.negative = neg;
        f.second_half = 1;
        f.zeroed_axis = axes[i];
        faces[cnt++] = f;
    }
    /* chain *= (K' odd ? 1 : -1) * (simplex.negative ? -1 : 1) */
    if (kp1 % 2 == 0)
        for (int i = 0; i < cnt; ++i) faces[i].negative ^= 1;
    return cnt;
}

/* deriv / coboundary with the clamped evaluator
 * (coboundary.hpp:353-427, 614-654; evaluators.hpp:30-46).
 * in: k-form over in_ext, out: (k+1)-form over out_ext (any extents; samples
 * are clamped into the input domain). */
void orc_coboundary(int n, int k, const long* in_ext, const double* in,
                    const long* out_ext, double* out)
{
    unsigned up[16], lo[16];
    int nup = combos(n, k + 1, up);
    int nlo = combos(n, k, lo);
    long nout = nodes_of(n, out_ext);
    for (long lin = 0; lin < nout; ++lin) {
        long idx[ORC_MAXD];
        unflatten(n, out_ext, lin, idx);
        for (int c = 0; c < nup; ++c) {
            orc_face faces[2 * ORC_MAXD];
            int nf = boundary_faces(n, up[c], faces);
            double acc = 0.0;
            for (int f = 0; f < nf; ++f) {
                long s[ORC_MAXD];
                for (int a = 0; a < n; ++a) s[a] = idx[a] + faces[f].off[a];
                clamp_idx(n, in_ext, s);
                int lc = -1;
                for (int q = 0; q < nlo; ++q)
                    if (lo[q] == faces[f].vmask) lc = q;
                double v = in[flatten(n, in_ext, s) * nlo + lc];
                acc += (faces[f].negative ? -1.0 : 1.0) * v;
            }
            out[lin * nup + c] = acc;
        }
    }
}

/* transposed_coboundary with the zero-outside evaluator
 * (coboundary.hpp:517-605, 656-696; evaluators.hpp:48-68):
 * second-half faces are sampled two steps back along the zeroed axis and the
 * integral is negated. */
void orc_transposed_coboundary(int n, int k, const long* in_ext, const double* in,
                               const long* out_ext, double* out)
{
    unsigned up[16], lo[16];
    int nup = combos(n, k + 1, up);
    int nlo = combos(n, k, lo);
    long nout = nodes_of(n, out_ext);
    for (long lin = 0; lin < nout; ++lin) {
        long idx[ORC_MAXD];
        unflatten(n, out_ext, lin, idx);
        for (int c = 0; c < nup; ++c) {
            orc_face faces[2 * ORC_MAXD];
            int nf = boundary_faces(n, up[c], faces);
            double acc = 0.0;
            for (int f = 0; f < nf; ++f) {
                long s[ORC_MAXD];
                for (int a = 0; a < n; ++a) s[a] = idx[a] + faces[f].off[a];
                if (faces[f].second_half) s[faces[f].zeroed_axis] -= 2;
                int lc = -1;
                for (int q = 0; q < nlo; ++q)
                    if (lo[q] == faces[f].vmask) lc = q;
                double v = inside(n, in_ext, s) ? in[flatten(n, in_ext, s) * nlo + lc] : 0.0;
                acc += (faces[f].negative ? -1.0 : 1.0) * v;
            }
            out[lin * nup + c] = -acc;
        }
    }
}

/* ------------------------------------------------------------------ */
/* geometry: positions, metric, k-volumes                              */
/* ------------------------------------------------------------------ */

enum { ORC_POS_RECTILINEAR = 0, ORC_POS_NODAL = 1 };
enum { ORC_METRIC_IDENTITY = 0, ORC_METRIC_DIAGONAL = 1, ORC_METRIC_SYMMETRIC = 2 };

typedef struct {
    int ndim;
    long ext[ORC_MAXD];
    int pos_kind;
    const double* axis[ORC_MAXD]; /* rectilinear: coordinate of node i on axis a */
    const double* nodal;          /* nodal: [node][ndim] */
    int metric_kind;
    const double* metric; /* diagonal: [node][ndim]; symmetric: [node][ndim(ndim+1)/2] */
} orc_geom;

static void position_of(const orc_geom* g, const long* idx, double* x)
{
    int n = g->ndim;
    if (g->pos_kind == ORC_POS_RECTILINEAR) {
        for (int a = 0; a < n; ++a) x[a] = g->axis[a][idx[a]];
    } else {
        long lin = flatten(n, g->ext, idx);
        for (int a = 0; a < n; ++a) x[a] = g->nodal[lin * n + a];
    }
}

/* edge_vector (volume.hpp:55-82): forward difference if the forward node
 * exists, else backward, else zero. */
static void edge_vector(const orc_geom* g, const long* idx, int axis, double* e)
{
    int n = g->ndim;
    long f[ORC_MAXD], b[ORC_MAXD];
    memcpy(f, idx, sizeof(long) * n);
    memcpy(b, idx, sizeof(long) * n);
    f[axis] += 1;
    b[axis] -= 1;
    double c[ORC_MAXD], o[ORC_MAXD];
    position_of(g, idx, c);
    if (inside(n, g->ext, f)) {
        position_of(g, f, o);
        for (int a = 0; a < n; ++a) e[a] = o[a] - c[a];
    } else if (inside(n, g->ext, b)) {
        position_of(g, b, o);
        for (int a = 0; a < n; ++a) e[a] = c[a] - o[a];
    } else {
        for (int a = 0; a < n; ++a) e[a] = 0.0;
    }
}

/* symmetric mem order XX, XY, XZ, YY, YZ, ZZ (symmetric_tensor.hpp:57-87) */
static int sym_index(int n, int i, int j)
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

static void local_metric(const orc_geom* g, const long* idx, double* m)
{
    int n = g->ndim;
    long lin = flatten(n, g->ext, idx);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double v;
            if (g->metric_kind == ORC_METRIC_IDENTITY)
                v = (i == j) ? 1.0 : 0.0;
            else if (g->metric_kind == ORC_METRIC_DIAGONAL)
                v = (i == j) ? g->metric[lin * n + i] : 0.0;
            else
                v = g->metric[lin * (n * (n + 1) / 2) + sym_index(n, i, j)];
            m[i * n + j] = v;
        }
}

/* LU determinant with partial pivoting (small_matrix.hpp:74-114) */
static double determinant(int n, const double* a_in)
{
    if (n == 0) return 1.0;
    double a[ORC_MAXD * ORC_MAXD];
    memcpy(a, a_in, sizeof(double) * n * n);
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
            double fct = a[r * n + i] / d;
            for (int c = i + 1; c < n; ++c) a[r * n + c] -= fct * a[i * n + c];
            a[r * n + i] = 0.0;
        }
    }
    return sign * det;
}

/* Gauss-Jordan inverse with partial pivoting; false on a zero pivot
 * (small_matrix.hpp:116-180) */
static int invert(int n, const double* m_in, double* inv)
{
    double w[ORC_MAXC * ORC_MAXC];
    memcpy(w, m_in, sizeof(double) * n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) inv[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int i = 0; i < n; ++i) {
        int p = i;
        double pa = fabs(w[i * n + i]);
        for (int r = i + 1; r < n; ++r)
            if (fabs(w[r * n + i]) > pa) {
                p = r;
                pa = fabs(w[r * n + i]);
            }
        if (pa == 0.0) {
            for (int r = 0; r < n * n; ++r) inv[r] = 0.0;
            return 0;
        }
        if (p != i)
            for (int c = 0; c < n; ++c) {
                double t = w[i * n + c];
                w[i * n + c] = w[p * n + c];
                w[p * n + c] = t;
                t = inv[i * n + c];
                inv[i * n + c] = inv[p * n + c];
                inv[p * n + c] = t;
            }
        double d = w[i * n + i];
        for (int c = 0; c < n; ++c) {
            w[i * n + c] /= d;
            inv[i * n + c] /= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == i) continue;
            double fct = w[r * n + i];
            for (int c = 0; c < n; ++c) {
                w[r * n + c] -= fct * w[i * n + c];
                inv[r * n + c] -= fct * inv[i * n + c];
            }
        }
    }
    return 1;
}

/* SimplexVolume (volume.hpp:101-155): sqrt|det Gram| of the metric-weighted
 * edge vectors spanning the axes of `mask`. */
static double simplex_volume(const orc_geom* g, const long* idx, unsigned mask)
{
    int n = g->ndim;
    int ids[ORC_MAXD], k = 0;
    for (int a = 0; a < n; ++a)
        if (mask & (1u << a)) ids[k++] = a;
    if (k == 0) return 1.0;
    double m[ORC_MAXD * ORC_MAXD];
    local_metric(g, idx, m);
    double e[ORC_MAXD][ORC_MAXD];
    for (int i = 0; i < k; ++i) edge_vector(g, idx, ids[i], e[i]);
    double gram[ORC_MAXD * ORC_MAXD];
    for (int i = 0; i < k; ++i)
        for (int j = i; j < k; ++j) {
            double v = 0.0;
            for (int p = 0; p < n; ++p)
                for (int q = 0; q < n; ++q) v += e[i][p] * m[p * n + q] * e[j][q];
            gram[i * k + j] = v;
            gram[j * k + i] = v;
        }
    return sqrt(fabs(determinant(k, gram)));
}

/* expansion determinant used by the reduction operators
 * (reduction_and_reconstruction.hpp:163-187) */
static double cofactor_det(int k, const double* a)
{
    if (k == 0) return 1.0;
    if (k == 1) return a[0];
    double det = 0.0;
    for (int col = 0; col < k; ++col) {
        double minor[ORC_MAXD * ORC_MAXD];
        for (int r = 1; r < k; ++r) {
            int mc = 0;
            for (int c = 0; c < k; ++c) {
                if (c == col) continue;
                minor[(r - 1) * (k - 1) + mc++] = a[r * k + c];
            }
        }
        det += ((col % 2 == 0) ? 1.0 : -1.0) * a[col] * cofactor_det(k - 1, minor);
    }
    return det;
}

/* det of the Jacobian block J[row][col] = edge(tgt[col])[src[row]]
 * (reduction_and_reconstruction.hpp:651-677, 346-366) */
static double edge_jacobian(const orc_geom* g, const long* idx, unsigned src, unsigned tgt)
{
    int n = g->ndim;
    int s[ORC_MAXD], t[ORC_MAXD], k = 0, kt = 0;
    for (int a = 0; a < n; ++a) {
        if (src & (1u << a)) s[k++] = a;
        if (tgt & (1u << a)) t[kt++] = a;
    }
    double jac[ORC_MAXD * ORC_MAXD];
    for (int c = 0; c < k; ++c) {
        double e[ORC_MAXD];
        edge_vector(g, idx, t[c], e);
        for (int r = 0; r < k; ++r) jac[r * k + c] = e[s[r]];
    }
    return cofactor_det(k, jac);
}

/* ------------------------------------------------------------------ */
/* A.2 discrete Hodge star                                             */
/* ------------------------------------------------------------------ */

/* The per-node coefficient of the discrete star of a primal |I|-form onto
 * its complement J: sign(I‖J)·V_J/V_I (legacy_discrete_hodge_value_from_ids,
 * reduction_and_reconstruction.hpp:403-432, with the 1/K! of the stored
 * value cancelled by the K! terms of the full-index tensor_prod,
 * tensor_prod.hpp:124-147).  Zero in the degenerate cases the reference's
 * circumcentric reduction turns into zeros (reduction_and_reconstruction.hpp
 * :535-647): a reconstruction Jacobian of any |I|-subset below 1e-14 in
 * magnitude, a singular metric, or a zero primal volume. */
static int node_degenerate(const orc_geom* g, const long* idx, int k)
{
    int n = g->ndim;
    unsigned c[16];
    int nc = combos(n, k, c);
    for (int i = 0; i < nc; ++i) {
        if (k == 0) break;
        double j = edge_jacobian(g, idx, c[i], c[i]);
        if (fabs(j) < 1e-14) return 1;
    }
    double m[ORC_MAXD * ORC_MAXD], inv[ORC_MAXD * ORC_MAXD];
    local_metric(g, idx, m);
    if (!invert(n, m, inv)) return 1;
    return 0;
}

double orc_hodge_coefficient(const orc_geom* g, const long* idx, unsigned src, unsigned tgt)
{
    int n = g->ndim;
    unsigned all = (1u << n) - 1u;
    if ((src & tgt) || ((src | tgt) != all)) return 0.0;
    if (node_degenerate(g, idx, popcount_u(src))) return 0.0;
    double vi = simplex_volume(g, idx, src);
    if (vi == 0.0) return 0.0;
    return concat_sign(src, tgt, n) * simplex_volume(g, idx, tgt) / vi;
}

/* The same coefficient computed the long way the reference's
 * DiscreteHodgeStar::run does it (hodge_star.hpp:224-273): reconstruction
 * (reduction_and_reconstruction.hpp:814-893) → continuous star
 * (hodge_star.hpp:396-450) → circumcentric reduction with the NBASIS×NBASIS
 * Gauss-Jordan inverse (reduction_and_reconstruction.hpp:535-647).  Used by
 * the tests to pin the closed form above on general metrics. */
static double continuous_value(const orc_geom* g, const long* idx, unsigned src, unsigned tgt)
{
    int n = g->ndim;
    unsigned all = (1u << n) - 1u;
    if ((src & tgt) || ((src | tgt) != all)) return 0.0;
    double m[ORC_MAXD * ORC_MAXD], inv[ORC_MAXD * ORC_MAXD];
    local_metric(g, idx, m);
    double det = determinant(n, m);
    if (!invert(n, m, inv)) return 0.0;
    unsigned compl_t = all & ~tgt; /* == src */
    int r[ORC_MAXD], c[ORC_MAXD], k = 0, kc = 0;
    for (int a = 0; a < n; ++a) {
        if (src & (1u << a)) r[k++] = a;
        if (compl_t & (1u << a)) c[kc++] = a;
    }
    double sub[ORC_MAXD * ORC_MAXD];
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) sub[i * k + j] = inv[r[i] * n + c[j]];
    double fact = 1.0;
    for (int i = 2; i <= k; ++i) fact *= i;
    return sqrt(fabs(det)) * concat_sign(compl_t, tgt, n) * determinant(k, sub) / fact;
}

double orc_hodge_coefficient_chain(const orc_geom* g, const long* idx, unsigned src, unsigned tgt)
{
    int n = g->ndim;
    int K = popcount_u(src);
    int Q = n - K;
    unsigned pri[16], qs[16];
    int nb = combos(n, K, pri);
    combos(n, Q, qs);
    double fact = 1.0;
    for (int i = 2; i <= K; ++i) fact *= i;
    double coeff[ORC_MAXC * ORC_MAXC] = {0}, old[ORC_MAXC * ORC_MAXC] = {0}, inv[ORC_MAXC * ORC_MAXC];
    for (int p = 0; p < nb; ++p) {
        double rj = K == 0 ? 1.0 : edge_jacobian(g, idx, pri[p], pri[p]);
        double rd = (K == 0) ? 1.0 : (fabs(rj) < 1e-14 ? 0.0 : 1.0 / rj);
        for (int s = 0; s < nb; ++s) coeff[s * nb + p] = continuous_value(g, idx, pri[p], qs[s]) * rd;
        for (int t = 0; t < nb; ++t) {
            double v = 0.0;
            unsigned all = (1u << n) - 1u;
            if (!(pri[p] & qs[t]) && ((pri[p] | qs[t]) == all)) {
                double vp = simplex_volume(g, idx, pri[p]);
                if (vp != 0.0)
                    v = concat_sign(pri[p], qs[t], n) * simplex_volume(g, idx, qs[t]) / (vp * fact);
            }
            old[t * nb + p] = v;
        }
    }
    if (!invert(nb, coeff, inv)) return 0.0;
    /* unit canonical source I: reconstructed value r_I, continuous output
     * c_s = K!·cont(I→s)·r_I, reduced_t = Σ_s (old·inv)[t][s]·c_s */
    int ii = -1, tt = -1;
    for (int p = 0; p < nb; ++p)
        if (pri[p] == src) ii = p;
    for (int t = 0; t < nb; ++t)
        if (qs[t] == tgt) tt = t;
    double rj = K == 0 ? 1.0 : edge_jacobian(g, idx, src, src);
    double rI = (K == 0) ? 1.0 : (fabs(rj) < 1e-14 ? 0.0 : 1.0 / rj);
    double out = 0.0;
    for (int s = 0; s < nb; ++s) {
        double cs = fact * continuous_value(g, idx, pri[ii], qs[s]) * rI;
        double d = 0.0;
        for (int p = 0; p < nb; ++p) d += old[tt * nb + p] * inv[p * nb + s];
        out += d * cs;
    }
    return out;
}

static orc_geom make_geom(int n, const long* ext, int pos_kind, const double* const* axis,
                          const double* nodal, int metric_kind, const double* metric)
{
    orc_geom g;
    memset(&g, 0, sizeof g);
    g.ndim = n;
    for (int a = 0; a < n; ++a) {
        g.ext[a] = ext[a];
        g.axis[a] = axis ? axis[a] : NULL;
    }
    g.pos_kind = pos_kind;
    g.nodal = nodal;
    g.metric_kind = metric_kind;
    g.metric = metric;
    return g;
}

/* fill_discrete_hodge_star (hodge_star.hpp:453-494) in the reference layout
 * [node][full source tuple, N^k row-major][antisymmetric target, C(N,N-k)]:
 * value = sign(perm)/k! · coefficient(sorted source → target), zero for
 * repeated source indices (hodge_star.hpp:275-358). */
void orc_fill_hodge_star(int n, const long* ext, int pos_kind, const double* const* axis,
                         const double* nodal, int metric_kind, const double* metric, int k,
                         int use_chain, double* out)
{
    orc_geom g = make_geom(n, ext, pos_kind, axis, nodal, metric_kind, metric);
    unsigned qs[16];
    int nq = combos(n, n - k, qs);
    long nfull = 1;
    for (int i = 0; i < k; ++i) nfull *= n;
    double fact = 1.0;
    for (int i = 2; i <= k; ++i) fact *= i;
    long nn = nodes_of(n, ext);
    for (long lin = 0; lin < nn; ++lin) {
        long idx[ORC_MAXD];
        unflatten(n, ext, lin, idx);
        for (long f = 0; f < nfull; ++f) {
            int tup[ORC_MAXD];
            long r = f;
            for (int i = k - 1; i >= 0; --i) {
                tup[i] = (int)(r % n);
                r /= n;
            }
            unsigned src = 0;
            int unique = 1, odd = 0;
            for (int i = 0; i < k; ++i) {
                if (src & (1u << tup[i])) unique = 0;
                src |= 1u << tup[i];
                for (int j = i + 1; j < k; ++j)
                    if (tup[i] > tup[j]) odd ^= 1;
            }
            for (int t = 0; t < nq; ++t) {
                double v = 0.0;
                if (unique) {
                    double c = use_chain ? orc_hodge_coefficient_chain(&g, idx, src, qs[t])
                                         : orc_hodge_coefficient(&g, idx, src, qs[t]);
                    v = (odd ? -1.0 : 1.0) / fact * c;
                }
                out[(lin * nfull + f) * nq + t] = v;
            }
        }
    }
}

/* apply ⋆ to a k-form: (⋆ω)_J(n) = Σ_I coefficient(I→J)(n)·ω_I(n)
 * (tensor_prod of the filled star with the form, codifferential.hpp:584) */
void orc_hodge_star(int n, const long* ext, int pos_kind, const double* const* axis,
                    const double* nodal, int metric_kind, const double* metric, int k,
                    const double* in, double* out)
{
    orc_geom g = make_geom(n, ext, pos_kind, axis, nodal, metric_kind, metric);
    unsigned ps[16], qs[16];
    int np = combos(n, k, ps);
    int nq = combos(n, n - k, qs);
    long nn = nodes_of(n, ext);
    for (long lin = 0; lin < nn; ++lin) {
        long idx[ORC_MAXD];
        unflatten(n, ext, lin, idx);
        for (int t = 0; t < nq; ++t) {
            double acc = 0.0;
            for (int p = 0; p < np; ++p) {
                double c = orc_hodge_coefficient(&g, idx, ps[p], qs[t]);
                if (c != 0.0) acc += c * in[lin * np + p];
            }
            out[lin * nq + t] = acc;
        }
    }
}

/* ------------------------------------------------------------------ */
/* A.4 codifferential, A.5 Laplacian                                   */
/* ------------------------------------------------------------------ */

/* δ of a k-form given on the full node domain `ext`, evaluated on out_ext
 * (codifferential.hpp:232-310, 551-632):
 *   η = ⋆ω on the full domain, G = dᵀη (zero outside), δω = s·⋆G at the
 *   output node, s = -1 iff N(k+1)+1 is odd. */
void orc_codifferential(int n, const long* ext, int pos_kind, const double* const* axis,
                        const double* nodal, int metric_kind, const double* metric, int k,
                        const double* in, const long* out_ext, double* out)
{
    orc_geom g = make_geom(n, ext, pos_kind, axis, nodal, metric_kind, metric);
    long nn = nodes_of(n, ext);
    int nd = (int)binom(n, n - k);
    int ng = (int)binom(n, n - k + 1);
    int no = (int)binom(n, k - 1);
    double* eta = (double*)malloc(sizeof(double) * nn * nd);
    orc_hodge_star(n, ext, pos_kind, axis, nodal, metric_kind, metric, k, in, eta);
    long nout = nodes_of(n, out_ext);
    double* G = (double*)malloc(sizeof(double) * nout * ng);
    orc_transposed_coboundary(n, n - k, ext, eta, out_ext, G);
    unsigned gs[16], os[16];
    combos(n, n - k + 1, gs);
    combos(n, k - 1, os);
    double s = ((n * (k + 1) + 1) % 2 == 1) ? -1.0 : 1.0;
    for (long lin = 0; lin < nout; ++lin) {
        long idx[ORC_MAXD];
        unflatten(n, out_ext, lin, idx);
        for (int t = 0; t < no; ++t) {
            double acc = 0.0;
            for (int p = 0; p < ng; ++p) {
                double c = orc_hodge_coefficient(&g, idx, gs[p], os[t]);
                if (c != 0.0) acc += c * G[lin * ng + p];
            }
            out[lin * no + t] = s * acc;
        }
    }
    free(eta);
    free(G);
}

/* Hodge Laplacian (laplacian.hpp:27-155, 599-647, 791-811, 824-951):
 *   k = 0:      Δω = δ(dω)
 *   0 < k < N:  Δω = δ(dω) + d(δω)
 *   k = N:      Δω = d(δω)
 * dω and δω live on the full node domain; the result on out_ext. */
void orc_laplacian(int n, const long* ext, int pos_kind, const double* const* axis,
                   const double* nodal, int metric_kind, const double* metric, int k,
                   const double* in, const long* out_ext, double* out)
{
    long nn = nodes_of(n, ext);
    long nout = nodes_of(n, out_ext);
    int nk = (int)binom(n, k);
    for (long i = 0; i < nout * nk; ++i) out[i] = 0.0;
    if (k < n) {
        int nk1 = (int)binom(n, k + 1);
        double* F = (double*)malloc(sizeof(double) * nn * nk1);
        orc_coboundary(n, k, ext, in, ext, F);
        double* t = (double*)malloc(sizeof(double) * nout * nk);
        orc_codifferential(n, ext, pos_kind, axis, nodal, metric_kind, metric, k + 1, F, out_ext, t);
        for (long i = 0; i < nout * nk; ++i) out[i] += t[i];
        free(t);
        free(F);
    }
    if (k > 0) {
        int nkm = (int)binom(n, k - 1);
        double* D = (double*)malloc(sizeof(double) * nn * nkm);
        orc_codifferential(n, ext, pos_kind, axis, nodal, metric_kind, metric, k, in, ext, D);
        double* t = (double*)malloc(sizeof(double) * nout * nk);
        orc_coboundary(n, k - 1, ext, D, out_ext, t);
        for (long i = 0; i < nout * nk; ++i) out[i] += t[i];
        free(t);
        free(D);
    }
}

/* ------------------------------------------------------------------ */
/* reduction / reconstruction (primal complex)                          */
/* ------------------------------------------------------------------ */

/* Reduction::run, primal complex (reduction_and_reconstruction.hpp:485-530,
 * 651-677): reduced_J = Σ_I det(J[I rows, J cols])·form_I (K! factors of
 * the full-index sum and of value() cancel). */
void orc_reduction(int n, const long* ext, int pos_kind, const double* const* axis,
                   const double* nodal, int k, const double* in, double* out)
{
    orc_geom g = make_geom(n, ext, pos_kind, axis, nodal, ORC_METRIC_IDENTITY, NULL);
    unsigned cs[16];
    int nc = combos(n, k, cs);
    long nn = nodes_of(n, ext);
    for (long lin = 0; lin < nn; ++lin) {
        long idx[ORC_MAXD];
        unflatten(n, ext, lin, idx);
        for (int t = 0; t < nc; ++t) {
            double acc = 0.0;
            for (int s = 0; s < nc; ++s) acc += edge_jacobian(&g, idx, cs[s], cs[t]) * in[lin * nc + s];
            out[lin * nc + t] = acc;
        }
    }
}

/* Reconstruction::run (reduction_and_reconstruction.hpp:814-893): the
 * diagonal inverse 1/det(J_I), zero below 1e-14. */
void orc_reconstruction(int n, const long* ext, int pos_kind, const double* const* axis,
                        const double* nodal, int k, const double* in, double* out)
{
    orc_geom g = make_geom(n, ext, pos_kind, axis, nodal, ORC_METRIC_IDENTITY, NULL);
    unsigned cs[16];
    int nc = combos(n, k, cs);
    long nn = nodes_of(n, ext);
    for (long lin = 0; lin < nn; ++lin) {
        long idx[ORC_MAXD];
        unflatten(n, ext, lin, idx);
        for (int t = 0; t < nc; ++t) {
            double d = k == 0 ? 1.0 : edge_jacobian(&g, idx, cs[t], cs[t]);
            double r = (k == 0) ? 1.0 : (fabs(d) < 1e-14 ? 0.0 : 1.0 / d);
            out[lin * nc + t] = r * in[lin * nc + t];
        }
    }
}

/* exported helpers for the tests */
long orc_binom(int n, int k) { return binom(n, k); }
int orc_combo_mask(int n, int k, int index)
{
    unsigned c[16];
    int cnt = combos(n, k, c);
    return index < cnt ? (int)c[index] : -1;
}
int orc_combo_index(int n, int k, int mask) { return combo_index(n, k, (unsigned)mask); }

/* boundary() faces of the (k+1)-simplex spanned by `vmask`, for the tests:
 * anchors (offsets from the simplex anchor), spanning masks and orientation. */
int orc_boundary(int n, int vmask, long* offsets, int* vmasks, int* negatives)
{
    orc_face faces[2 * ORC_MAXD];
    int nf = boundary_faces(n, (unsigned)vmask, faces);
    for (int f = 0; f < nf; ++f) {
        for (int a = 0; a < n; ++a) offsets[f * n + a] = faces[f].off[a];
        vmasks[f] = (int)faces[f].vmask;
        negatives[f] = faces[f].negative;
    }
    return nf;
}

/* continuous star value for a canonical (sorted) source set, including the
 * reference's 1/K! (hodge_star.hpp:396-450), for the tests */
double orc_continuous_hodge_value(int n, const long* ext, int metric_kind, const double* metric,
                                  const long* idx, int src, int tgt)
{
    const double* axis[ORC_MAXD] = {0};
    double zeros[2] = {0.0, 1.0};
    for (int a = 0; a < n; ++a) axis[a] = zeros;
    orc_geom g = make_geom(n, ext, ORC_POS_RECTILINEAR, axis, NULL, metric_kind, metric);
    return continuous_value(&g, idx, (unsigned)src, (unsigned)tgt);
}
