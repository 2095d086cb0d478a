// 2D magnetostatics on the out-of-plane potential A_z.
//
// MagnetostaticsOperator2D (onelab_interface/magnetostatics_onelab.hpp
// :1013-1522) keeps one unknown per node (row i + nx·j).  Its moments are
// the 2D induction stencils (physics/magnetostatics/
// magnetostatics_quantities.hpp:75-123):
//   B_x(s) = A_z(s+e_y) − A_z(s),   B_y(s) = A_z(s) − A_z(s+e_x),
// sampled at interior nodes s only, with no Hodge weight.  An interior row r
// gathers the material response h(s) = ∂H/∂B at the three sampled nodes whose
// stencils contain it, through the transposed stencil (the sampled rows in
// ascending order, moment x before moment y):
//   (A x)_r = h_x(r − e_y) − h_x(r) − h_y(r − e_x) + h_y(r),
// counted twice when the criterion adds the potential and the moments parts
// (each part is the same sum: the "outer" table holds the negated stencil
// and is subtracted, :1379-1432); boundary rows are the identity.  Unlike
// the 3D operator the moments read every node, boundary nodes included.
//
// Responses: h = ν·B (linear Hamiltonian), a·B (the nonlinear material at
// its own state) or a·δB + c·B0 (B0·δB) (its Jacobian), with a = f + q f′,
// c = 2(2f′ + q f″), q = |B|², f = h/b of the B-H curve on ferromagnetic
// nodes and a = ν, c = 0 elsewhere (MaterialMagnetostaticsHamiltonian,
// :957-1011).  One thread per row evaluates the three sampled nodes
// directly from the potential (a 7-node neighbourhood); the matrix-free
// application is HBM-bound at ≈ 24 B per node (x in, ν, out).
#include "apply_epilogue.hpp"
#include "bh_curve.hpp"
#include "operators.hpp"

namespace sil {

int64_t max_dot_blocks();

namespace {

constexpr int kBlock = kApplyBlock;

enum Resp2 { R2_LINEAR = 0, R2_SCALED = 1, R2_JACOBIAN = 2 };

template <typename T>
using State2 = NodeState<T>;

template <typename T>
struct Ms2Args {
    int64_t nx, ny;
    T terms;
    const T* nu;
    const State2<T>* st;
};

inline unsigned nblk(int64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

template <typename T, int RESP>
__device__ __forceinline__ void response(const Ms2Args<T>& a, int64_t s, T bx, T by, T& hx, T& hy)
{
    if (RESP == R2_LINEAR) {
        const T nu = a.nu[s];
        hx = nu * bx;
        hy = nu * by;
    } else if (RESP == R2_SCALED) {
        const T k = a.st[s].a;
        hx = k * bx;
        hy = k * by;
    } else {
        const State2<T> q = a.st[s];
        const T proj = q.c * (q.bx * bx + q.by * by);
        hx = q.a * bx + proj * q.bx;
        hy = q.a * by + proj * q.by;
    }
}

template <typename T, bool PU, int RESP>
__global__ __launch_bounds__(kBlock) void k_ms2d_apply(Ms2Args<T> a, const T* __restrict__ x, T* __restrict__ out,
                                                        double* __restrict__ partials, PUpdate<T> pu)
{
    if (PU && *pu.done) return;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t nx = a.nx, ny = a.ny;
    double dot = 0.0;
    if (t < nx * ny) {
        const PUpdate<T>* pp = PU ? &pu : nullptr;
        const int64_t i = t % nx, j = t / nx;
        const T v = p_value<T>(pp, x, t);
        T o = v;
        if (i > 0 && j > 0 && i + 1 < nx && j + 1 < ny) {
            const T ve = p_value<T>(pp, x, t + 1), vn = p_value<T>(pp, x, t + nx);
            T acc = T(0);
            if (j > 1) {  // s = r − e_y: moment x holds +A(r)
                const int64_t s = t - nx;
                const T vs = p_value<T>(pp, x, s), vse = p_value<T>(pp, x, s + 1);
                T hx, hy;
                response<T, RESP>(a, s, v - vs, vs - vse, hx, hy);
                acc += hx;
            }
            {  // s = r: moment x holds −A(r), moment y +A(r)
                T hx, hy;
                response<T, RESP>(a, t, vn - v, v - ve, hx, hy);
                acc -= hx;
                if (i > 1) {  // s = r − e_x: moment y holds −A(r)
                    const int64_t s = t - 1;
                    const T vw = p_value<T>(pp, x, s), vwn = p_value<T>(pp, x, s + nx);
                    T gx, gy;
                    response<T, RESP>(a, s, vwn - vw, vw - v, gx, gy);
                    acc -= gy;
                }
                acc += hy;
            }
            o = a.terms * acc;
        }
        out[t] = o;
        if (PU) pu.p_new[t] = v;
        dot = (double)v * (double)o;
    }
    if (partials) block_dot_store<T>(dot, partials);
}

// diagonal of the linear operator or of the Jacobian at the stored state
template <bool JAC>
__global__ void k_ms2d_diag(Ms2Args<double> a, double* __restrict__ diag)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t nx = a.nx, ny = a.ny;
    if (t >= nx * ny) return;
    const int64_t i = t % nx, j = t / nx;
    if (!(i > 0 && j > 0 && i + 1 < nx && j + 1 < ny)) {
        diag[t] = 1.0;
        return;
    }
    // Σ_s g_sᵀ J(s) g_s over the sampled nodes, g_s = ∂B(s)/∂A(r):
    // s = r − e_y: (1, 0); s = r: (−1, 1); s = r − e_x: (0, −1)
    double acc = 0.0;
    if (j > 1) {
        const int64_t s = t - nx;
        if (JAC) {
            const State2<double> q = a.st[s];
            acc += q.a + q.c * q.bx * q.bx;
        } else {
            acc += a.nu[s];
        }
    }
    if (JAC) {
        const State2<double> q = a.st[t];
        const double g = q.by - q.bx;
        acc += 2.0 * q.a + q.c * g * g;
    } else {
        acc += 2.0 * a.nu[t];
    }
    if (i > 1) {
        const int64_t s = t - 1;
        if (JAC) {
            const State2<double> q = a.st[s];
            acc += q.a + q.c * q.by * q.by;
        } else {
            acc += a.nu[s];
        }
    }
    diag[t] = a.terms * acc;
}

// the state of x at interior nodes: B, a, c (FP64 and FP32 copies)
__global__ void k_ms2d_state(int64_t nx, int64_t ny, const double* __restrict__ nu, const uint8_t* __restrict__ ferro,
                             const BhCurve* __restrict__ curve, const double* __restrict__ x,
                             State2<double>* __restrict__ st, State2<float>* __restrict__ st32)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    const int64_t i = t % nx, j = t / nx;
    double a = 0.0, c = 0.0, bx = 0.0, by = 0.0;
    if (i > 0 && j > 0 && i + 1 < nx && j + 1 < ny) {
        bx = x[t + nx] - x[t];
        by = x[t] - x[t + 1];
        if (ferro[t]) {
            const double q = bx * bx + by * by;
            double f, f1, f2;
            h_over_b_q(*curve, q, f, f1, f2);
            a = f + q * f1;
            c = 2.0 * (2.0 * f1 + q * f2);
        } else {
            a = nu[t];
        }
    }
    st[t] = State2<double>{a, c, bx, by};
    st32[t] = State2<float>{(float)a, (float)c, (float)bx, (float)by};
}

// mean of the adjacent cells' μ (magnetostatics_onelab.hpp:4467-4500)
__global__ void k_ms2d_nodal_mu(int64_t nx, int64_t ny, const double* __restrict__ cell_mu, double mu0,
                                double* __restrict__ mu)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    const int64_t i = t % nx, j = t / nx, mx = nx - 1, my = ny - 1;
    double acc = 0.0;
    int count = 0;
    for (int dj = -1; dj <= 0; ++dj)
        for (int di = -1; di <= 0; ++di) {
            const int64_t ci = i + di, cj = j + dj;
            if (ci < 0 || cj < 0 || ci >= mx || cj >= my) continue;
            acc += cell_mu[ci + mx * cj];
            ++count;
        }
    mu[t] = count == 0 ? mu0 : acc / count;
}

// ¼·area·J_z of the adjacent cells at interior nodes (:4413-4445)
__global__ void k_ms2d_rhs(int64_t nx, int64_t ny, const double* __restrict__ xc, const double* __restrict__ yc,
                           const double* __restrict__ jz, int zero, double* __restrict__ rhs)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    const int64_t i = t % nx, j = t / nx, mx = nx - 1, my = ny - 1;
    double acc = 0.0;
    if (!zero && i > 0 && j > 0 && i + 1 < nx && j + 1 < ny) {
        for (int dj = -1; dj <= 0; ++dj)
            for (int di = -1; di <= 0; ++di) {
                const int64_t ci = i + di, cj = j + dj;
                if (ci < 0 || cj < 0 || ci >= mx || cj >= my) continue;
                const double area = (xc[ci + 1] - xc[ci]) * (yc[cj + 1] - yc[cj]);
                acc += 0.25 * area * jz[ci + mx * cj];
            }
    }
    rhs[t] = acc;
}

template <typename T>
Ms2Args<T> args(const Ms2dData& d, const NlData* nl, bool fp32)
{
    Ms2Args<T> a{};
    a.nx = d.nx;
    a.ny = d.ny;
    a.terms = (T)d.terms;
    if (fp32) {
        a.nu = reinterpret_cast<const T*>(d.nu32);
        if (nl) a.st = reinterpret_cast<const State2<T>*>(nl->st_pack32);
    } else {
        a.nu = reinterpret_cast<const T*>(d.nu);
        if (nl) a.st = reinterpret_cast<const State2<T>*>(nl->st_pack);
    }
    return a;
}

template <typename T, int RESP>
int ms2_apply(const Ms2dData& d, const NlData* nl, const T* in, T* out, DotEpilogue dot, const PUpdate<T>* pu,
              int* nb_out, hipStream_t s)
{
    const unsigned nb = nblk(d.nx * d.ny);
    if (dot.partials && nb > max_dot_blocks()) {
        set_error("operator too large for the fused dot epilogue (%u blocks)", nb);
        return SIL_ERR_UNSUPPORTED;
    }
    if (nb_out) *nb_out = (int)nb;
    const Ms2Args<T> a = args<T>(d, nl, sizeof(T) == 4);
    if (pu)
        hipLaunchKernelGGL((k_ms2d_apply<T, true, RESP>), dim3(nb), dim3(kBlock), 0, s, a, in, out, dot.partials, *pu);
    else
        hipLaunchKernelGGL((k_ms2d_apply<T, false, RESP>), dim3(nb), dim3(kBlock), 0, s, a, in, out, dot.partials,
                           PUpdate<T>{});
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

}  // namespace

int ms2d_apply64(const Ms2dData& d, const double* in, double* out, DotEpilogue dot, const PUpdate<double>* pu,
                 int* nb, hipStream_t s)
{
    return ms2_apply<double, R2_LINEAR>(d, nullptr, in, out, dot, pu, nb, s);
}

int ms2d_apply32(const Ms2dData& d, const float* in, float* out, DotEpilogue dot, const PUpdate<float>* pu, int* nb,
                 hipStream_t s)
{
    return ms2_apply<float, R2_LINEAR>(d, nullptr, in, out, dot, pu, nb, s);
}

int ms2d_state_apply64(const Ms2dData& d, const NlData& nl, const double* x, double* out, hipStream_t s)
{
    return ms2_apply<double, R2_SCALED>(d, &nl, x, out, DotEpilogue{nullptr}, nullptr, nullptr, s);
}

int ms2d_jacobian_apply64(const Ms2dData& d, const NlData& nl, const double* in, double* out, DotEpilogue dot,
                          const PUpdate<double>* pu, int* nb, hipStream_t s)
{
    return ms2_apply<double, R2_JACOBIAN>(d, &nl, in, out, dot, pu, nb, s);
}

int ms2d_jacobian_apply32(const Ms2dData& d, const NlData& nl, const float* in, float* out, DotEpilogue dot,
                          const PUpdate<float>* pu, int* nb, hipStream_t s)
{
    return ms2_apply<float, R2_JACOBIAN>(d, &nl, in, out, dot, pu, nb, s);
}

int ms2d_diagonal(const Ms2dData& d, double* diag, hipStream_t s)
{
    hipLaunchKernelGGL(k_ms2d_diag<false>, dim3(nblk(d.nx * d.ny)), dim3(kBlock), 0, s,
                       args<double>(d, nullptr, false), diag);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int ms2d_jacobian_diagonal(const Ms2dData& d, const NlData& nl, double* diag, hipStream_t s)
{
    hipLaunchKernelGGL(k_ms2d_diag<true>, dim3(nblk(d.nx * d.ny)), dim3(kBlock), 0, s, args<double>(d, &nl, false),
                       diag);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int ms2d_state(const Ms2dData& d, const NlData& nl, const double* x, hipStream_t s)
{
    SIL_REQUIRE(nl.st_pack && nl.st_pack32, SIL_ERR_INVALID, "ms2d_state: no packed state arrays");
    hipLaunchKernelGGL(k_ms2d_state, dim3(nblk(d.nx * d.ny)), dim3(kBlock), 0, s, d.nx, d.ny, d.nu, nl.ferro,
                       nl.curve, x, reinterpret_cast<State2<double>*>(nl.st_pack),
                       reinterpret_cast<State2<float>*>(nl.st_pack32));
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int ms2d_nodal_mu(int64_t nx, int64_t ny, const double* cell_mu, double mu0, double* mu, hipStream_t s)
{
    hipLaunchKernelGGL(k_ms2d_nodal_mu, dim3(nblk(nx * ny)), dim3(kBlock), 0, s, nx, ny, cell_mu, mu0, mu);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int ms2d_rhs(const Ms2dData& d, const double* cell_current_z, bool zero, double* rhs, hipStream_t s)
{
    hipLaunchKernelGGL(k_ms2d_rhs, dim3(nblk(d.nx * d.ny)), dim3(kBlock), 0, s, d.nx, d.ny, d.coord[0], d.coord[1],
                       cell_current_z, zero ? 1 : 0, rhs);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

}  // namespace sil
