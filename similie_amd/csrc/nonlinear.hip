// Nonlinear magnetostatics: the material state, the Jacobian diagonal and
// the Newton solve.
//
// The reference's nonlinear branch (MaterialMagnetostaticsHamiltonian,
// onelab_interface/magnetostatics_onelab.hpp:957-1011) evaluates at every
// interior sampled node s, inside every operator and Jacobian application,
// the induction B(s) from the potential, then the generated Hamiltonian
// derivatives (∂H/∂B and ∂²H/∂B², nonlinear_magnetostatics.py:184-227 through
// generate_cpp_hamiltonian.py:306-397) of the interpolated B-H curve
// (nonlinear_magnetostatics.py:14-181).  With H = ½ q f(q), q = |B|²,
// f = h(b)/b:
//   ∂H/∂B_m        = a·B_m,                         a = f + q f′
//   ∂²H/∂B_m∂B_c   = a·δ_mc + c·B_m B_c,            c = 2 (2 f′ + q f″)
// and a = ν, c = 0 on nodes that are not ferromagnetic.  Here one pass per
// evaluation of the operator stores (a, c, B) per node — the *state* — and
// the curl-curl march applies the operator (response a·B) or its Jacobian
// at that state (response a·δB + c·B (B·δB)) without touching the curve
// again: CG's thousands of Jacobian applications read five doubles per
// node instead of re-deriving B and re-bracketing the curve each time.
//
// Newton (minimize_strong_formulation_residual.hpp:1383-1569): x₀ = 0,
// r = b − F(x); at most 30 steps while ‖r‖/‖r₀‖ > tol; J(x) δ = r by the
// graph-captured CG of krylov.hip to the same relative tolerance (identity
// or Jacobi from the Jacobian's diagonal); backtracking line search α = 1,
// halved up to 8 times until ‖r(x + αδ)‖ < ‖r(x)‖, keeping the best α tried.
#include <chrono>
#include <cmath>
#include <new>

#include "bh_curve.hpp"
#include "operators.hpp"

namespace sil {

int cg_solve(sil_operator op, const double* rhs, double* solution, const sil_solver_settings& st,
             sil_solver_diagnostics& dg, hipStream_t s);
void invalidate_jacobi(sil_operator op);
void invalidate_isai(sil_operator op);

struct NewtonBuffers {
    double *r = nullptr, *fx = nullptr, *delta = nullptr, *xc = nullptr, *rc = nullptr;
    double* partials = nullptr;
    double* host = nullptr; // pinned
};

namespace {

constexpr int kBlock = 256;
constexpr int kNormBlocks = 1024;

// ------------------------------------------------------------ B-H curve ---

// ---------------------------------------------------------------- state ---

struct StateArgs {
    int64_t nx, ny, nz;
    const double* nu;
    const uint8_t* ferro;
    const BhCurve* curve;
    double* a;
    double* c;
    double* b[3];
    float* a32;  // FP32 scaled response (a), read by the FP32 march
    float* bz32; // FP32 B_z, read by the FP32 Jacobian march (B_x, B_y are packed)
    NodeState<double>* pack;
    NodeState<float>* pack32;
};

// B at every interior node (boundary potentials dropped, the term order of
// the moment stencils), then (a, c) of its material
__global__ __launch_bounds__(kBlock) void k_nl_state(StateArgs s, const double* __restrict__ x)
{
    const int64_t nn = s.nx * s.ny * s.nz;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nn; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t % s.nx, j = (t / s.nx) % s.ny, k = t / (s.nx * s.ny);
        auto interior = [&](int64_t ii, int64_t jj, int64_t kk) {
            return ii > 0 && jj > 0 && kk > 0 && ii + 1 < s.nx && jj + 1 < s.ny && kk + 1 < s.nz;
        };
        double sa = s.nu[t], sc = 0.0, bx = 0.0, by = 0.0, bz = 0.0;
        if (interior(i, j, k)) {
            auto A = [&](int64_t ii, int64_t jj, int64_t kk, int comp) -> double {
                return interior(ii, jj, kk) ? x[3 * (ii + s.nx * (jj + s.ny * kk)) + comp] : 0.0;
            };
            const double a0x = A(i, j, k, 0), a0y = A(i, j, k, 1), a0z = A(i, j, k, 2);
            bx = A(i, j + 1, k, 2) - a0z - A(i, j, k + 1, 1) + a0y;
            by = -A(i + 1, j, k, 2) + a0z + A(i, j, k + 1, 0) - a0x;
            bz = A(i + 1, j, k, 1) - a0y - A(i, j + 1, k, 0) + a0x;
            if (s.ferro[t]) {
                const double q = bx * bx + by * by + bz * bz;
                double f, f1, f2;
                h_over_b_q(*s.curve, q, f, f1, f2);
                sa = f + q * f1;
                sc = 2.0 * (2.0 * f1 + q * f2);
            }
        }
        s.a[t] = sa;
        s.c[t] = sc;
        s.b[0][t] = bx;
        s.b[1][t] = by;
        s.b[2][t] = bz;
        s.a32[t] = (float)sa;
        s.bz32[t] = (float)bz;
        s.pack[t] = NodeState<double>{sa, sc, bx, by};
        s.pack32[t] = NodeState<float>{(float)sa, (float)sc, (float)bx, (float)by};
    }
}

// ------------------------------------------------------ Jacobian diagonal ---

struct DiagArgs {
    int64_t nx, ny, nz;
    double penalty; // 0 under the divergence gauge (its diagonal is added apart)
    double ysign;
    const double* ls[3];
    const double* inv_ls[3];
    const double* a;
    const double* c;
    const double* b[3];
};

// For row (n, C) the sampled nodes reaching it are n − e for two axes (one
// moment each) and n itself (two moments): diag = Σ_s Σ_{m,m'} c_m c_m'
// w_m(s) K_mm'(s) + P w^A_C, K = a·I + c·B Bᵀ at s.  (moment, coefficient,
// offset axis or −1 for s = n) per component, from the moment stencils.
struct Reach {
    int m, coeff, axis;
};
__constant__ Reach kReach[3][4] = {
    {{1, 1, 2}, {2, -1, 1}, {1, -1, -1}, {2, 1, -1}},
    {{0, -1, 2}, {2, 1, 0}, {0, 1, -1}, {2, -1, -1}},
    {{0, 1, 1}, {1, -1, 0}, {0, -1, -1}, {1, 1, -1}},
};

__global__ __launch_bounds__(kBlock) void k_nl_jdiag(DiagArgs d, double* __restrict__ diag)
{
    const int64_t nn = d.nx * d.ny * d.nz;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nn; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t % d.nx, j = (t / d.nx) % d.ny, k = t / (d.nx * d.ny);
        auto interior = [&](int64_t ii, int64_t jj, int64_t kk) {
            return ii > 0 && jj > 0 && kk > 0 && ii + 1 < d.nx && jj + 1 < d.ny && kk + 1 < d.nz;
        };
        if (!interior(i, j, k)) {
            diag[3 * t] = diag[3 * t + 1] = diag[3 * t + 2] = 1.0;
            continue;
        }
        // w_m(s) K_mm'(s)
        auto wk = [&](int m, int mp, int64_t ii, int64_t jj, int64_t kk) -> double {
            if (!interior(ii, jj, kk)) return 0.0;
            const int64_t s = ii + d.nx * (jj + d.ny * kk);
            const double lx = d.ls[0][ii], ly = d.ls[1][jj], lz = d.ls[2][kk];
            const double ilx = d.inv_ls[0][ii], ily = d.inv_ls[1][jj], ilz = d.inv_ls[2][kk];
            const double w = m == 0 ? lx * ily * ilz : (m == 1 ? ly * ilx * ilz : lz * ilx * ily);
            const double K = (m == mp ? d.a[s] : 0.0) + d.c[s] * d.b[m][s] * d.b[mp][s];
            return m == 1 ? w * K * d.ysign : w * K;
        };
        const double lx = d.ls[0][i], ly = d.ls[1][j], lz = d.ls[2][k];
        const double ilx = d.inv_ls[0][i], ily = d.inv_ls[1][j], ilz = d.inv_ls[2][k];
        const double wa[3] = {ly * lz * ilx, lx * lz * ily, lx * ly * ilz};
        for (int C = 0; C < 3; ++C) {
            double v = 0.0;
            for (int p = 0; p < 2; ++p) { // single-moment sampled nodes n − e_axis
                const Reach r = kReach[C][p];
                v += wk(r.m, r.m, i - (r.axis == 0), j - (r.axis == 1), k - (r.axis == 2));
            }
            for (int p = 2; p < 4; ++p) // both moments at s = n
                for (int q = 2; q < 4; ++q)
                    v += kReach[C][p].coeff * kReach[C][q].coeff * wk(kReach[C][p].m, kReach[C][q].m, i, j, k);
            diag[3 * t + C] = v + d.penalty * wa[C];
        }
    }
}

// ---------------------------------------------------------------- Newton ---

// r = b − f with per-block partial sums of r²
__global__ __launch_bounds__(kBlock) void k_residual(const double* __restrict__ b, const double* __restrict__ f,
                                                     double* __restrict__ r, int64_t n, double* __restrict__ part)
{
    double v = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double ri = b[i] - f[i];
        r[i] = ri;
        v += ri * ri;
    }
    __shared__ double sh[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) s += sh[w];
        part[blockIdx.x] = s;
    }
}

// fixed-order sum of the block partials into out[0]
__global__ __launch_bounds__(1024) void k_sum(const double* __restrict__ part, int count, double* __restrict__ out)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < count; i += blockDim.x) v += part[i];
    __shared__ double sh[16];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x / 64); ++w) s += sh[w];
        out[0] = s;
    }
}

__global__ __launch_bounds__(kBlock) void k_axpy(const double* __restrict__ x, double alpha,
                                                 const double* __restrict__ d, double* __restrict__ out, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = x[i] + alpha * d[i];
}

template <typename P>
int dalloc(P** p, int64_t n)
{
    SIL_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(p), sizeof(P) * n));
    return SIL_OK;
}

}  // namespace

namespace {
__global__ void k_ferro_flags(const double* __restrict__ frac, uint8_t* __restrict__ flags, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        flags[i] = frac[i] > 0.5 ? 1 : 0;
}
}  // namespace

// is_ferromagnetic (magnetostatics_onelab.hpp:977-983): fraction > 0.5
int ms3d_ferro_flags(const double* frac, uint8_t* flags, int64_t n, hipStream_t s)
{
    const unsigned g = static_cast<unsigned>(std::min<int64_t>((n + kBlock - 1) / kBlock, 16384));
    hipLaunchKernelGGL(k_ferro_flags, dim3(g), dim3(kBlock), 0, s, frac, flags, n);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int ms3d_state(const Ms3dData& d, const NlData& nl, const double* x, hipStream_t s)
{
    StateArgs a{d.nx,     d.ny,      d.nz,     d.nu,
                nl.ferro, nl.curve,  nl.st_a,  nl.st_c,
                {nl.st_b[0], nl.st_b[1], nl.st_b[2]},
                nl.st_a32, nl.st_b32[2],
                reinterpret_cast<NodeState<double>*>(nl.st_pack), reinterpret_cast<NodeState<float>*>(nl.st_pack32)};
    SIL_REQUIRE(nl.st_pack && nl.st_pack32 && nl.st_a && nl.st_c && nl.st_b[0] && nl.st_b[1] && nl.st_b[2] &&
                    nl.st_a32 && nl.st_b32[2],
                SIL_ERR_INVALID, "ms3d_state: state arrays missing");
    const int64_t nn = d.nx * d.ny * d.nz;
    const unsigned g = static_cast<unsigned>(std::min<int64_t>((nn + kBlock - 1) / kBlock, 16384));
    hipLaunchKernelGGL(k_nl_state, dim3(g), dim3(kBlock), 0, s, a, x);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int ms3d_jacobian_diagonal(const Ms3dData& d, const NlData& nl, double* diag, hipStream_t s)
{
    DiagArgs a{d.nx,
               d.ny,
               d.nz,
               d.divg ? 0.0 : d.penalty,
               d.ysign,
               {d.ls[0], d.ls[1], d.ls[2]},
               {d.inv_ls[0], d.inv_ls[1], d.inv_ls[2]},
               nl.st_a,
               nl.st_c,
               {nl.st_b[0], nl.st_b[1], nl.st_b[2]}};
    const int64_t nn = d.nx * d.ny * d.nz;
    const unsigned g = static_cast<unsigned>(std::min<int64_t>((nn + kBlock - 1) / kBlock, 16384));
    hipLaunchKernelGGL(k_nl_jdiag, dim3(g), dim3(kBlock), 0, s, a, diag);
    SIL_LAUNCH_CHECK();
    return ms3d_gauge_diagonal_add(d, diag, s);
}

void destroy_newton(NewtonBuffers* nb)
{
    if (!nb) return;
    for (double* p : {nb->r, nb->fx, nb->delta, nb->xc, nb->rc, nb->partials})
        if (p) (void)hipFree(p);
    if (nb->host) (void)hipHostFree(nb->host);
    delete nb;
}

namespace {

// F(x) into fx (updating the state to x), r = b − F(x) into r; returns ‖r‖
int evaluate(sil_operator op, const double* b, const double* x, double* fx, double* r, double* norm, hipStream_t s)
{
    NewtonBuffers& nb = *op->newton;
    int rc;
    if (op->kind == OP_MS2D_NONLINEAR) {
        if ((rc = ms2d_state(op->m2, op->nl, x, s)) || (rc = ms2d_state_apply64(op->m2, op->nl, x, fx, s))) return rc;
    } else if ((rc = ms3d_state(op->ms, op->nl, x, s)) || (rc = ms3d_state_apply64(op->ms, op->nl, x, fx, s))) {
        return rc;
    }
    hipLaunchKernelGGL(k_residual, dim3(kNormBlocks), dim3(kBlock), 0, s, b, fx, r, op->n, nb.partials);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, nb.partials, kNormBlocks, nb.partials + kNormBlocks);
    SIL_LAUNCH_CHECK();
    SIL_CHECK_HIP(hipMemcpyAsync(nb.host, nb.partials + kNormBlocks, sizeof(double), hipMemcpyDeviceToHost, s));
    SIL_CHECK_HIP(hipStreamSynchronize(s));
    *norm = std::sqrt(nb.host[0]);
    return SIL_OK;
}

}  // namespace

int newton_solve(sil_operator op, const double* rhs, double* x, const sil_solver_settings& st,
                 sil_solver_diagnostics& dg, hipStream_t s)
{
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t n = op->n;
    if (!op->newton) {
        op->newton = new (std::nothrow) NewtonBuffers();
        SIL_REQUIRE(op->newton != nullptr, SIL_ERR_INVALID, "newton: out of host memory");
        NewtonBuffers& nb = *op->newton;
        int rc;
        if ((rc = dalloc(&nb.r, n)) || (rc = dalloc(&nb.fx, n)) || (rc = dalloc(&nb.delta, n)) ||
            (rc = dalloc(&nb.xc, n)) || (rc = dalloc(&nb.rc, n)) || (rc = dalloc(&nb.partials, kNormBlocks + 1)))
            return rc;
        SIL_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&nb.host), sizeof(double)));
    }
    NewtonBuffers& nb = *op->newton;
    const unsigned g = static_cast<unsigned>(std::min<int64_t>((n + kBlock - 1) / kBlock, 16384));
    SIL_CHECK_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
    double r0 = 0.0;
    int rc = evaluate(op, rhs, x, nb.fx, nb.r, &r0, s);
    if (rc) return rc;
    dg.initial_residual_l2 = r0;
    dg.final_residual_l2 = r0;
    dg.final_relative_residual = r0 == 0.0 ? 0.0 : 1.0;
    dg.iterations = 0;
    dg.outer_iterations = 0;
    dg.converged = r0 == 0.0 ? 1 : 0;
    double rn = r0;
    // inner CG to the same relative tolerance; with mixed precision its
    // Jacobian applications run in FP32 and it converges on the FP64 residual
    sil_solver_settings inner = st;
    for (int it = 0; it < 30 && !dg.converged; ++it) {
        if (rn / r0 <= st.relative_tolerance) {
            dg.converged = 1;
            break;
        }
        // J(x) δ = r: the Jacobian operator reads the state of x, which is
        // the state of the last evaluation
        invalidate_jacobi(op->jac);
        invalidate_isai(op->jac);
        sil_solver_diagnostics di{};
        if ((rc = cg_solve(op->jac, nb.r, nb.delta, inner, di, s))) return rc;
        dg.iterations += di.iterations;
        double alpha = 1.0, best_alpha = 0.0, best = INFINITY, cn = 0.0;
        for (int ls = 0; ls < 8; ++ls) {
            hipLaunchKernelGGL(k_axpy, dim3(g), dim3(kBlock), 0, s, x, alpha, nb.delta, nb.xc, n);
            SIL_LAUNCH_CHECK();
            if ((rc = evaluate(op, rhs, nb.xc, nb.fx, nb.rc, &cn, s))) return rc;
            if (cn < best) {
                best_alpha = alpha;
                best = cn;
            }
            if (cn < rn) break;
            alpha *= 0.5;
        }
        if (best_alpha != alpha) {
            alpha = best_alpha;
            hipLaunchKernelGGL(k_axpy, dim3(g), dim3(kBlock), 0, s, x, alpha, nb.delta, nb.xc, n);
            SIL_LAUNCH_CHECK();
            if ((rc = evaluate(op, rhs, nb.xc, nb.fx, nb.rc, &cn, s))) return rc;
        }
        SIL_CHECK_HIP(hipMemcpyAsync(x, nb.xc, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        SIL_CHECK_HIP(hipMemcpyAsync(nb.r, nb.rc, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        rn = cn;
        dg.outer_iterations += 1;
    }
    SIL_CHECK_HIP(hipStreamSynchronize(s));
    if (!dg.converged && rn <= st.relative_tolerance * r0) dg.converged = 1; // the last step reached it
    dg.final_residual_l2 = rn;
    dg.final_relative_residual = r0 == 0.0 ? 0.0 : rn / r0;
    // x₀ = 0 gives F(x₀) = 0, so ‖r₀‖ = ‖b‖; rn is the FP64 residual of the
    // accepted iterate, evaluated on the device
    dg.true_relative_residual = dg.final_relative_residual;
    dg.verified = dg.converged;
    dg.duration = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return SIL_OK;
}

}  // namespace sil
