// Equal-size batches of 2D variable-coefficient Dirichlet Poisson problems
// (B problems of n×n nodes, one contiguous [B][n][n] array per vector) with
// three Krylov methods that share one set of kernels:
//   * Jacobi-preconditioned CG in FP64,
//   * the same CG with FP32 directions, products and operator applications
//     and FP64 iterate, residual, dots and scalars (residual replacement),
//   * flexible CG (Notay's FCG, truncation 1) whose preconditioner is a
//     two-level cycle around the learned U-Net, in either precision.
// No reference counterpart: the reference's CG (Ginkgo, src/similie/solvers/
// minimize_strong_formulation_residual.hpp:1028-1304) solves one system at a
// time.  Every solver declares convergence only on a residual recomputed in
// FP64 as b − A x, so the reported relative residual is FP64-verified.
//
// Operator (as sil_poisson2d_create with unit spacing): for an interior node
// t = i·n + j, (A u)_t = Σ over the four edges of w_e (u_t − u_nb), with
// u = 0 on the boundary ring and w the arithmetic mean of κ over the edge;
// boundary rows are zero (their unknowns stay at the zero right-hand side).
//
// Kernels run on a grid (blocks per problem, B); every reduction is a
// per-block partial sum followed by a fixed-order sum in a one-block-per-
// problem scalar kernel, so results do not depend on scheduling.  Solved
// problems skip their work; the iteration loop is captured in hipGraphs of
// R iterations (R even: the direction buffers ping-pong), the last of which
// recomputes the residual in FP64.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <new>
#include <vector>

#include "common.hpp"
#include "sil_hip.h"
#include "unet.hpp"

namespace sil {
namespace {

constexpr int NT = 256;
constexpr int kParts = 4; // partial sums per block

struct Scal {
    double rz, pq, pq_old, pr, zq, rr, bb, alpha, beta, best, rho1, cnet;
    int done, iters, stall, fallback;
};

// damping of the Jacobi smoothing around the network's correction
constexpr double kOmega = 2.0 / 3.0;

// flexible CG: a problem whose FP64 residual has not fallen by 0.8 (in norm)
// over kStallChecks consecutive replacement intervals switches from the
// network to the Jacobi preconditioner, i.e. continues as plain CG
constexpr int kStallChecks = 5;

template <class T>
__device__ __forceinline__ T apply_at(const T* __restrict__ u, const T* __restrict__ wx, const T* __restrict__ wy,
                                      size_t t, int i, int j, int n)
{
    const T uc = u[t];
    const T ue = i + 2 < n ? u[t + n] : T(0), uw = i > 1 ? u[t - n] : T(0);
    const T un = j + 2 < n ? u[t + 1] : T(0), us = j > 1 ? u[t - 1] : T(0);
    return wx[t] * (uc - ue) + wx[t - n] * (uc - uw) + wy[t] * (uc - un) + wy[t - 1] * (uc - us);
}

__device__ __forceinline__ bool interior(int i, int j, int n) { return i > 0 && j > 0 && i + 1 < n && j + 1 < n; }

// block sum of up to kParts doubles into part[(b·nbx + bx)·kParts + k]
template <int K>
__device__ __forceinline__ void block_partials(const double (&v)[K], double* __restrict__ part)
{
    __shared__ double sh[NT / 64][kParts];
    double w[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        w[k] = v[k];
        for (int off = 32; off > 0; off >>= 1) w[k] += __shfl_down(w[k], off, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) sh[wave][k] = w[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double* o = part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * kParts;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0.0;
            for (int q = 0; q < NT / 64; ++q) s += sh[q][k];
            o[k] = s;
        }
    }
}

#define NODE_LOOP                                                                                          \
    for (size_t t = blockIdx.x * (size_t)NT + threadIdx.x; t < nn; t += (size_t)gridDim.x * NT)

// ------------------------------------------------------------ set-up ---

__global__ __launch_bounds__(NT) void k_pb_weights(int n, const double* __restrict__ kappa, double* __restrict__ wx,
                                                   double* __restrict__ wy, float* __restrict__ wx32,
                                                   float* __restrict__ wy32, double* __restrict__ dinv,
                                                   float* __restrict__ dinv32, float* __restrict__ logk)
{
    const size_t nn = (size_t)n * n, base = (size_t)blockIdx.y * nn;
    NODE_LOOP
    {
        const int j = (int)(t % n), i = (int)(t / n);
        const double kc = kappa[base + t];
        const double a = i + 1 < n ? 0.5 * (kc + kappa[base + t + n]) : kc;
        const double c = j + 1 < n ? 0.5 * (kc + kappa[base + t + 1]) : kc;
        wx[base + t] = a;
        wy[base + t] = c;
        wx32[base + t] = (float)a;
        wy32[base + t] = (float)c;
        logk[base + t] = (float)log(kc);
    }
}

__global__ __launch_bounds__(NT) void k_pb_dinv(int n, const double* __restrict__ wx, const double* __restrict__ wy,
                                                int jacobi, double* __restrict__ dinv, float* __restrict__ dinv32)
{
    const size_t nn = (size_t)n * n, base = (size_t)blockIdx.y * nn;
    NODE_LOOP
    {
        const int j = (int)(t % n), i = (int)(t / n);
        double d = 0.0;
        if (interior(i, j, n)) {
            const size_t g = base + t;
            d = jacobi ? 1.0 / (wx[g] + wx[g - n] + wy[g] + wy[g - 1]) : 1.0;
        }
        dinv[base + t] = d;
        dinv32[base + t] = (float)d;
    }
}

// coefficient channel of the network input: (log κ − mean) / std per problem
__global__ __launch_bounds__(NT) void k_pb_logk_stats(int n, const float* __restrict__ logk, double* __restrict__ part)
{
    const size_t nn = (size_t)n * n, base = (size_t)blockIdx.y * nn;
    double v[2] = {0.0, 0.0};
    NODE_LOOP
    {
        const double l = logk[base + t];
        v[0] += l;
        v[1] += l * l;
    }
    block_partials<2>(v, part);
}

__global__ __launch_bounds__(NT) void k_pb_logk_norm(int n, int nbx, const double* __restrict__ part,
                                                     float* __restrict__ logk)
{
    const size_t nn = (size_t)n * n, base = (size_t)blockIdx.y * nn;
    double s = 0.0, s2 = 0.0;
    for (int q = 0; q < nbx; ++q) {
        s += part[((size_t)blockIdx.y * nbx + q) * kParts];
        s2 += part[((size_t)blockIdx.y * nbx + q) * kParts + 1];
    }
    const double mean = s / (double)nn, var = fmax(s2 / (double)nn - mean * mean, 0.0);
    const double inv = var > 0.0 ? 1.0 / sqrt(var) : 1.0;
    NODE_LOOP logk[base + t] = (float)((logk[base + t] - mean) * inv);
}

template <class T>
__global__ __launch_bounds__(NT) void k_pb_fill(size_t n, T* __restrict__ a, T v)
{
    for (size_t t = blockIdx.x * (size_t)NT + threadIdx.x; t < n; t += (size_t)gridDim.x * NT) a[t] = v;
}

__global__ __launch_bounds__(NT) void k_pb_apply64(int n, const double* __restrict__ wx,
                                                   const double* __restrict__ wy, const double* __restrict__ x,
                                                   double* __restrict__ y)
{
    const size_t nn = (size_t)n * n, base = (size_t)blockIdx.y * nn;
    NODE_LOOP
    {
        const int j = (int)(t % n), i = (int)(t / n);
        y[base + t] = interior(i, j, n) ? apply_at(x + base, wx + base, wy + base, t, i, j, n) : 0.0;
    }
}

// r = b − A x (FP64); z = D⁻¹ r (CG only); partials rz, rr, bb
template <class T>
__global__ __launch_bounds__(NT) void k_pb_residual(int n, const double* __restrict__ wx,
                                                    const double* __restrict__ wy, const double* __restrict__ b,
                                                    const double* __restrict__ x, double* __restrict__ r,
                                                    const double* __restrict__ dinv, T* __restrict__ z,
                                                    const Scal* __restrict__ sc, double* __restrict__ part)
{
    if (sc && sc[blockIdx.y].done) return;
    const size_t nn = (size_t)n * n, base = (size_t)blockIdx.y * nn;
    double v[3] = {0.0, 0.0, 0.0};
    NODE_LOOP
    {
        const int j = (int)(t % n), i = (int)(t / n);
        const size_t g = base + t;
        const double rv = interior(i, j, n) ? b[g] - apply_at(x + base, wx + base, wy + base, t, i, j, n) : 0.0;
        r[g] = rv;
        v[1] += rv * rv;
        v[2] += b[g] * b[g];
        if (z) {
            const T zv = (T)(dinv[g] * rv);
            z[g] = zv;
            v[0] += rv * (double)zv;
        }
    }
    block_partials<3>(v, part);
}

// --------------------------------------------------------------- CG ---

// p = z + β p_old (recomputed at the four neighbours), q = A p; partial pᵀq
template <class T>
__global__ __launch_bounds__(NT) void k_cg_apply(int n, const T* __restrict__ wx, const T* __restrict__ wy,
                                                 const T* __restrict__ z, const T* __restrict__ p_old,
                                                 T* __restrict__ p_new, T* __restrict__ q,
                                                 const Scal* __restrict__ sc, double* __restrict__ part)
{
    const Scal& s = sc[blockIdx.y];
    if (s.done) return;
    const T beta = (T)s.beta;
    const size_t nn = (size_t)n * n, base = (size_t)blockIdx.y * nn;
    const T* zb = z + base;
    const T* pb = p_old + base;
    auto P = [&](size_t u) { return zb[u] + beta * pb[u]; };
    double v[1] = {0.0};
    NODE_LOOP
    {
        const int j = (int)(t % n), i = (int)(t / n);
        T pv = T(0), qv = T(0);
        if (interior(i, j, n)) {
            pv = P(t);
            const T pe = i + 2 < n ? P(t + n) : T(0), pw = i > 1 ? P(t - n) : T(0);
            const T pn = j + 2 < n ? P(t + 1) : T(0), ps = j > 1 ? P(t - 1) : T(0);
            const T* ax = wx + base;
            const T* ay = wy + base;
            qv = ax[t] * (pv - pe) + ax[t - n] * (pv - pw) + ay[t] * (pv - pn) + ay[t - 1] * (pv - ps);
        }
        p_new[base + t] = pv;
        q[base + t] = qv;
        v[0] += (double)pv * (double)qv;
    }
    block_partials<1>(v, part);
}

// x += α p, r −= α q, z = D⁻¹ r; partials rᵀz, rᵀr
template <class T>
__global__ __launch_bounds__(NT) void k_cg_update(int n, const T* __restrict__ p, const T* __restrict__ q,
                                                  const double* __restrict__ dinv, double* __restrict__ x,
                                                  double* __restrict__ r, T* __restrict__ z,
                                                  const Scal* __restrict__ sc, double* __restrict__ part)
{
    const Scal& s = sc[blockIdx.y];
    if (s.done) return;
    const double alpha = s.alpha;
    const size_t nn = (size_t)n * n, base = (size_t)blockIdx.y * nn;
    double v[2] = {0.0, 0.0};
    NODE_LOOP
    {
        const size_t g = base + t;
        x[g] += alpha * (double)p[g];
        const double rv = r[g] - alpha * (double)q[g];
        r[g] = rv;
        const T zv = (T)(dinv[g] * rv);
        z[g] = zv;
        v[0] += rv * (double)zv;
        v[1] += rv * rv;
    }
    block_partials<2>(v, part);
}

// ---------------------------------------------------------------- FCG ---

// Preconditioner of flexible CG, a two-level cycle around the network:
//   s = ω D⁻¹ r                         (pre-smoothing)
//   r₁ = r − A s                        (its residual)
//   n = N(r₁ / rms(r₁), log κ)          (the network's correction)
//   c = n·r₁ / n·A n                    (its A-optimal length)
//   z = s + c n + ω D⁻¹ (r₁ − c A n)    (post-smoothing)
// A problem that has fallen back to plain CG uses z = D⁻¹ r.

// r₁ = r − A(ω D⁻¹ r) (FP32 copy); partial r₁ᵀr₁
template <class T>
__global__ __launch_bounds__(NT) void k_tl_pre(int n, const T* __restrict__ wx, const T* __restrict__ wy,
                                               const double* __restrict__ r, const double* __restrict__ dinv,
                                               float* __restrict__ r1, const Scal* __restrict__ sc,
                                               double* __restrict__ part)
{
    if (sc[blockIdx.y].done) return;
    const size_t nn = (size_t)n * n, base = (size_t)blockIdx.y * nn;
    const double* rb = r + base;
    const double* db = dinv + base;
    auto S = [&](size_t u) { return kOmega * db[u] * rb[u]; };
    double v[1] = {0.0};
    NODE_LOOP
    {
        const int j = (int)(t % n), i = (int)(t / n);
        double o = 0.0;
        if (interior(i, j, n)) {
            const double sv = S(t);
            const double se = i + 2 < n ? S(t + n) : 0.0, sw = i > 1 ? S(t - n) : 0.0;
            const double sn = j + 2 < n ? S(t + 1) : 0.0, ss = j > 1 ? S(t - 1) : 0.0;
            const T* ax = wx + base;
            const T* ay = wy + base;
            o = rb[t] - ((double)ax[t] * (sv - se) + (double)ax[t - n] * (sv - sw) + (double)ay[t] * (sv - sn) +
                         (double)ay[t - 1] * (sv - ss));
        }
        r1[base + t] = (float)o;
        v[0] += o * o;
    }
    block_partials<1>(v, part);
}

// The problems that still use the network (not converged, not fallen back
// to Jacobi) get consecutive slots of the network's batch, in problem order:
// slot[b] = rank among them or −1, *count = their number.  For the network's
// planar input each occupied slot also gets its problem (src) and the scale
// of its residual channel, 1 / rms(r₁).  One wavefront.
__global__ __launch_bounds__(64) void k_tl_compact(int B, const Scal* __restrict__ sc, int* __restrict__ slot,
                                                   int* __restrict__ count, int* __restrict__ src,
                                                   float* __restrict__ scale)
{
    const int lane = threadIdx.x;
    int next = 0;
    for (int b0 = 0; b0 < B; b0 += 64) {
        const int b = b0 + lane;
        const bool active = b < B && !sc[b].done && !sc[b].fallback;
        const unsigned long long mask = __ballot(active);
        const int rank = __popcll(mask & ((1ull << lane) - 1ull));
        if (b < B) slot[b] = active ? next + rank : -1;
        if (active && src) {
            const double rho = sc[b].rho1;
            src[next + rank] = b;
            scale[next + rank] = rho > 0.0 ? (float)(1.0 / rho) : 0.f;
        }
        next += __popcll(mask);
    }
    if (lane == 0) *count = next;
}

// network input in the problem's slot, interleaved (the planar input reads
// the same values in place): channel 0 = r₁ / rms(r₁), channel 1 =
// normalised log κ
__global__ __launch_bounds__(NT) void k_tl_input(int n, const float* __restrict__ r1, const float* __restrict__ logk,
                                                 const Scal* __restrict__ sc, const int* __restrict__ slot,
                                                 float* __restrict__ xin)
{
    const int sl = slot[blockIdx.y];
    if (sl < 0) return;
    const size_t nn = (size_t)n * n, base = (size_t)blockIdx.y * nn, sbase = (size_t)sl * nn;
    const double rho = sc[blockIdx.y].rho1;
    const float inv = rho > 0.0 ? (float)(1.0 / rho) : 0.f;
    NODE_LOOP
    {
        xin[(sbase + t) * 2] = r1[base + t] * inv;
        xin[(sbase + t) * 2 + 1] = logk[base + t];
    }
}

// A n (n, the network output in the problem's slot, zero on the boundary
// ring); partials nᵀr₁, nᵀA n
template <class T>
__global__ __launch_bounds__(NT) void k_tl_an(int n, const T* __restrict__ wx, const T* __restrict__ wy,
                                              const float* __restrict__ zn, const int* __restrict__ slot,
                                              const float* __restrict__ r1, float* __restrict__ an,
                                              const Scal* __restrict__ sc, double* __restrict__ part)
{
    const int sl = slot[blockIdx.y];
    if (sl < 0) return; // converged, or on the Jacobi fallback: no network output
    const size_t nn = (size_t)n * n, base = (size_t)blockIdx.y * nn;
    const float* nb = zn + (size_t)sl * nn;
    double v[2] = {0.0, 0.0};
    NODE_LOOP
    {
        const int j = (int)(t % n), i = (int)(t / n);
        float o = 0.f, nv = 0.f;
        if (interior(i, j, n)) {
            nv = nb[t];
            const float ne = i + 2 < n ? nb[t + n] : 0.f, nw = i > 1 ? nb[t - n] : 0.f;
            const float nnb = j + 2 < n ? nb[t + 1] : 0.f, ns = j > 1 ? nb[t - 1] : 0.f;
            const T* ax = wx + base;
            const T* ay = wy + base;
            o = (float)((double)ax[t] * (nv - ne) + (double)ax[t - n] * (nv - nw) + (double)ay[t] * (nv - nnb) +
                        (double)ay[t - 1] * (nv - ns));
        }
        an[base + t] = o;
        v[0] += (double)nv * (double)r1[base + t];
        v[1] += (double)nv * (double)o;
    }
    block_partials<2>(v, part);
}

// z = ω D⁻¹ r + c n + ω D⁻¹ (r₁ − c A n), or D⁻¹ r after the fallback;
// partial zᵀq_old
template <class T>
__global__ __launch_bounds__(NT) void k_tl_post(int n, const double* __restrict__ r, const double* __restrict__ dinv,
                                                const float* __restrict__ r1, const float* __restrict__ zn,
                                                const int* __restrict__ slot, const float* __restrict__ an,
                                                const T* __restrict__ q_old, float* __restrict__ z,
                                                const Scal* __restrict__ sc, double* __restrict__ part)
{
    const Scal& s = sc[blockIdx.y];
    if (s.done) return;
    const double c = s.cnet;
    const int sl = slot[blockIdx.y];
    const bool fb = sl < 0; // on the Jacobi fallback since the iteration began
    const size_t nn = (size_t)n * n, base = (size_t)blockIdx.y * nn;
    const float* nb = zn + (size_t)(fb ? 0 : sl) * nn;
    double v[1] = {0.0};
    NODE_LOOP
    {
        const int j = (int)(t % n), i = (int)(t / n);
        const size_t g = base + t;
        double zv = 0.0;
        if (interior(i, j, n))
            zv = fb ? dinv[g] * r[g]
                    : kOmega * dinv[g] * r[g] + c * (double)nb[t] +
                          kOmega * dinv[g] * ((double)r1[g] - c * (double)an[g]);
        z[g] = (float)zv;
        v[0] += (double)(float)zv * (double)q_old[g];
    }
    block_partials<1>(v, part);
}

// p = z − β p_old, q = A p; partials pᵀq, pᵀr
template <class T>
__global__ __launch_bounds__(NT) void k_fcg_apply(int n, const T* __restrict__ wx, const T* __restrict__ wy,
                                                  const float* __restrict__ z, const T* __restrict__ p_old,
                                                  T* __restrict__ p_new, T* __restrict__ q,
                                                  const double* __restrict__ r, const Scal* __restrict__ sc,
                                                  double* __restrict__ part)
{
    const Scal& s = sc[blockIdx.y];
    if (s.done) return;
    const T beta = (T)s.beta;
    const size_t nn = (size_t)n * n, base = (size_t)blockIdx.y * nn;
    const float* zb = z + base;
    const T* pb = p_old + base;
    auto P = [&](size_t u) { return (T)zb[u] - beta * pb[u]; };
    double v[2] = {0.0, 0.0};
    NODE_LOOP
    {
        const int j = (int)(t % n), i = (int)(t / n);
        T pv = T(0), qv = T(0);
        if (interior(i, j, n)) {
            pv = P(t);
            const T pe = i + 2 < n ? P(t + n) : T(0), pw = i > 1 ? P(t - n) : T(0);
            const T pn = j + 2 < n ? P(t + 1) : T(0), ps = j > 1 ? P(t - 1) : T(0);
            const T* ax = wx + base;
            const T* ay = wy + base;
            qv = ax[t] * (pv - pe) + ax[t - n] * (pv - pw) + ay[t] * (pv - pn) + ay[t - 1] * (pv - ps);
        }
        p_new[base + t] = pv;
        q[base + t] = qv;
        v[0] += (double)pv * (double)qv;
        v[1] += (double)pv * r[base + t];
    }
    block_partials<2>(v, part);
}

// x += α p, r −= α q; partial rᵀr (slot 1, where the residual kernel puts it)
template <class T>
__global__ __launch_bounds__(NT) void k_fcg_update(int n, const T* __restrict__ p, const T* __restrict__ q,
                                                   double* __restrict__ x, double* __restrict__ r,
                                                   const Scal* __restrict__ sc, double* __restrict__ part)
{
    const Scal& s = sc[blockIdx.y];
    if (s.done) return;
    const double alpha = s.alpha;
    const size_t nn = (size_t)n * n, base = (size_t)blockIdx.y * nn;
    double v[2] = {0.0, 0.0};
    NODE_LOOP
    {
        const size_t g = base + t;
        x[g] += alpha * (double)p[g];
        const double rv = r[g] - alpha * (double)q[g];
        r[g] = rv;
        v[1] += rv * rv;
    }
    block_partials<2>(v, part);
}

// ------------------------------------------------------------ scalars ---

enum ScalPhase {
    PH_INIT = 0,    // rz, rr, bb of the initial residual
    PH_VERIFY,      // rr, bb of a recomputed residual, for every problem
    PH_CG_ALPHA,    // pq → α = rz/pq
    PH_CG_BETA,     // rz', rr → β = rz'/rz, iteration count, convergence on a recomputed residual
    PH_TL_RHO,      // r₁ᵀr₁ → rms(r₁)
    PH_TL_C,        // nᵀr₁, nᵀA n → length of the network's correction
    PH_FCG_BETA,    // zq → β = zq / pq_old (0 on the first iteration)
    PH_FCG_ALPHA,   // pq, pr → α = pr/pq
    PH_FCG_UPDATE,  // rr → iteration count, convergence on a recomputed residual
};

__global__ __launch_bounds__(64) void k_pb_scalars(int phase, int nbx, const double* __restrict__ part, Scal* __restrict__ sc,
                             double tol2, int max_iter, int verified, int* __restrict__ done_flags,
                             double nn_total)
{
    const int b = blockIdx.x;
    Scal& s = sc[b];
    if (s.done && phase != PH_INIT && phase != PH_VERIFY) return; // uniform per workgroup
    // one wavefront sums the problem's block partials (lane-strided, then a
    // shuffle tree: a fixed order)
    double v[kParts] = {0.0, 0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < nbx; q += 64)
#pragma unroll
        for (int k = 0; k < kParts; ++k) v[k] += part[((size_t)b * nbx + q) * kParts + k];
#pragma unroll
    for (int k = 0; k < kParts; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
    if (threadIdx.x != 0) return;
    switch (phase) {
    case PH_INIT:
        s.rz = v[0];
        s.rr = v[1];
        s.bb = v[2];
        s.beta = 0.0;
        s.pq_old = 0.0;
        s.iters = 0;
        s.best = s.rr;
        s.stall = 0;
        s.fallback = 0;
        s.done = (s.rr <= tol2 * s.bb) ? 1 : 0;
        break;
    case PH_CG_ALPHA:
        s.pq = v[0];
        s.alpha = s.pq > 0.0 ? s.rz / s.pq : 0.0;
        break;
    case PH_CG_BETA:
        s.beta = s.rz != 0.0 ? v[0] / s.rz : 0.0;
        s.rz = v[0];
        s.rr = v[1];
        s.iters += 1;
        if ((verified && s.rr <= tol2 * s.bb) || s.iters >= max_iter || s.pq <= 0.0) s.done = 1;
        break;
    case PH_TL_RHO:
        s.rho1 = sqrt(v[0] / (double)nn_total);
        break;
    case PH_TL_C:
        s.cnet = v[1] > 0.0 ? v[0] / v[1] : 0.0;
        break;
    case PH_FCG_BETA:
        s.zq = v[0];
        s.beta = s.pq_old > 0.0 ? s.zq / s.pq_old : 0.0;
        break;
    case PH_FCG_ALPHA:
        s.pq = v[0];
        s.pr = v[1];
        s.alpha = s.pq > 0.0 ? s.pr / s.pq : 0.0;
        break;
    case PH_VERIFY:
        s.rr = v[1];
        s.bb = v[2];
        break;
    case PH_FCG_UPDATE:
        s.rr = v[1];
        s.pq_old = s.pq;
        s.iters += 1;
        if (verified) {
            if (s.rr < 0.64 * s.best) {
                s.best = s.rr;
                s.stall = 0;
            } else if (++s.stall >= kStallChecks) {
                s.fallback = 1;
            }
        }
        if ((verified && s.rr <= tol2 * s.bb) || s.iters >= max_iter || s.pq <= 0.0) s.done = 1;
        break;
    }
    if (done_flags) done_flags[b] = s.done;
}

// snapshots of r (FP32) and x (FP64: errors x* − x of late iterates are
// far below FP32 resolution of x) for the learned preconditioner's data
__global__ __launch_bounds__(NT) void k_pb_record(size_t total, const double* __restrict__ r,
                                                  const double* __restrict__ x, float* __restrict__ rr,
                                                  double* __restrict__ xr)
{
    for (size_t t = blockIdx.x * (size_t)NT + threadIdx.x; t < total; t += (size_t)gridDim.x * NT) {
        rr[t] = (float)r[t];
        xr[t] = x[t];
    }
}

}  // namespace
}  // namespace sil

// ------------------------------------------------------------- host ---

struct sil_poisson2d_batch_s {
    int B = 0, n = 0, nbx = 0;
    size_t nn = 0, total = 0;
    double *wx = nullptr, *wy = nullptr, *dinv = nullptr;
    float *wx32 = nullptr, *wy32 = nullptr, *dinv32 = nullptr, *logk = nullptr;
    double* r = nullptr;
    double* part = nullptr;
    sil::Scal* sc = nullptr;
    int* done = nullptr;        // device
    int* done_host = nullptr;   // pinned
    double *p64[2] = {nullptr, nullptr}, *q64[2] = {nullptr, nullptr}, *z64 = nullptr;
    float *p32[2] = {nullptr, nullptr}, *q32[2] = {nullptr, nullptr}, *z32 = nullptr;
    float* xin = nullptr;  // interleaved network input [B][n][n][2] (SIL_FCG_PLANAR=0)
    int* slot_src = nullptr;     // problem of each occupied network slot
    float* slot_scale = nullptr; // scale of its residual channel
    bool interleaved = false;    // this solve forms the interleaved input
    float* zn = nullptr;   // network output [B][n][n], problems in their slots
    int* slot = nullptr;   // network slot of each problem, −1 when it needs none
    int* nact = nullptr;   // number of occupied slots
    float* r1 = nullptr;   // residual after pre-smoothing
    float* an = nullptr;   // A · network output
    float* zf = nullptr;   // preconditioned residual of flexible CG
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    std::vector<void*> allocations;

    template <class T>
    int alloc(T*& p, size_t count)
    {
        void* q = nullptr;
        if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) {
            sil::set_error("poisson batch: out of device memory (%zu bytes)", count * sizeof(T));
            return SIL_ERR_HIP;
        }
        allocations.push_back(q);
        p = static_cast<T*>(q);
        return SIL_OK;
    }
    ~sil_poisson2d_batch_s()
    {
        for (void* q : allocations) (void)hipFree(q);
        if (done_host) (void)hipHostFree(done_host);
        if (ev) (void)hipEventDestroy(ev);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

using sil::NT;
using Batch = sil_poisson2d_batch_s;

dim3 grid_of(const Batch* h) { return dim3((unsigned)h->nbx, (unsigned)h->B); }

template <class T>
void fill(T* a, size_t n, T v, hipStream_t s)
{
    size_t g = (n + NT - 1) / NT;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(sil::k_pb_fill<T>, dim3((unsigned)g), dim3(NT), 0, s, n, a, v);
}

void scalars(const Batch* h, int phase, double tol2, int max_iter, int verified, hipStream_t s)
{
    hipLaunchKernelGGL(sil::k_pb_scalars, dim3(h->B), dim3(64), 0, s, phase, h->nbx, h->part, h->sc, tol2, max_iter,
                       verified, h->done, (double)h->nn);
}

// one CG iteration; `parity` selects the direction buffer written
template <class T>
void cg_iteration(Batch* h, T* const* p, T* const* q, T* z, const T* wx, const T* wy, const double* b, double* x,
                  int parity, bool replace, double tol2, int max_iter, hipStream_t s)
{
    hipLaunchKernelGGL(sil::k_cg_apply<T>, grid_of(h), dim3(NT), 0, s, h->n, wx, wy, z, p[parity ^ 1], p[parity],
                       q[0], h->sc, h->part);
    scalars(h, sil::PH_CG_ALPHA, tol2, max_iter, 0, s);
    hipLaunchKernelGGL(sil::k_cg_update<T>, grid_of(h), dim3(NT), 0, s, h->n, p[parity], q[0], h->dinv, x, h->r, z,
                       h->sc, h->part);
    if (replace)
        hipLaunchKernelGGL(sil::k_pb_residual<T>, grid_of(h), dim3(NT), 0, s, h->n, h->wx, h->wy, b, x, h->r,
                           h->dinv, z, h->sc, h->part);
    scalars(h, sil::PH_CG_BETA, tol2, max_iter, replace ? 1 : 0, s);
}

template <class T>
int fcg_iteration(Batch* h, sil_unet net, T* const* p, T* const* q, const T* wx, const T* wy, const double* b,
                  double* x, int parity, bool replace, double tol2, int max_iter, hipStream_t s)
{
    hipLaunchKernelGGL(sil::k_tl_pre<T>, grid_of(h), dim3(NT), 0, s, h->n, wx, wy, h->r, h->dinv, h->r1, h->sc,
                       h->part);
    scalars(h, sil::PH_TL_RHO, tol2, max_iter, 0, s);
    // the network runs only on the problems that still use it, packed into
    // the first slots of its batch
    int rc;
    if (h->interleaved) {
        hipLaunchKernelGGL(sil::k_tl_compact, dim3(1), dim3(64), 0, s, h->B, h->sc, h->slot, h->nact, (int*)nullptr,
                           (float*)nullptr);
        hipLaunchKernelGGL(sil::k_tl_input, grid_of(h), dim3(NT), 0, s, h->n, h->r1, h->logk, h->sc, h->slot,
                           h->xin);
        rc = sil::unet::forward_active(net, h->B, h->nact, h->xin, h->zn, s);
    } else {
        // the network reads r₁ and log κ in place
        hipLaunchKernelGGL(sil::k_tl_compact, dim3(1), dim3(64), 0, s, h->B, h->sc, h->slot, h->nact, h->slot_src,
                           h->slot_scale);
        sil::unet::PlanarInput in;
        in.c0 = h->r1;
        in.c1 = h->logk;
        in.src = h->slot_src;
        in.scale = h->slot_scale;
        rc = sil::unet::forward_active_planar(net, h->B, h->nact, in, h->zn, s);
    }
    if (rc != SIL_OK) return rc;
    hipLaunchKernelGGL(sil::k_tl_an<T>, grid_of(h), dim3(NT), 0, s, h->n, wx, wy, h->zn, h->slot, h->r1, h->an, h->sc,
                       h->part);
    scalars(h, sil::PH_TL_C, tol2, max_iter, 0, s);
    hipLaunchKernelGGL(sil::k_tl_post<T>, grid_of(h), dim3(NT), 0, s, h->n, h->r, h->dinv, h->r1, h->zn, h->slot,
                       h->an, q[parity ^ 1], h->zf, h->sc, h->part);
    scalars(h, sil::PH_FCG_BETA, tol2, max_iter, 0, s);
    hipLaunchKernelGGL(sil::k_fcg_apply<T>, grid_of(h), dim3(NT), 0, s, h->n, wx, wy, h->zf, p[parity ^ 1], p[parity],
                       q[parity], h->r, h->sc, h->part);
    scalars(h, sil::PH_FCG_ALPHA, tol2, max_iter, 0, s);
    hipLaunchKernelGGL(sil::k_fcg_update<T>, grid_of(h), dim3(NT), 0, s, h->n, p[parity], q[parity], x, h->r,
                       h->sc, h->part);
    if (replace)
        hipLaunchKernelGGL(sil::k_pb_residual<T>, grid_of(h), dim3(NT), 0, s, h->n, h->wx, h->wy, b, x, h->r,
                           h->dinv, (T*)nullptr, h->sc, h->part);
    scalars(h, sil::PH_FCG_UPDATE, tol2, max_iter, replace ? 1 : 0, s);
    return SIL_OK;
}

}  // namespace

extern "C" {

int sil_poisson2d_batch_create(sil_poisson2d_batch* out, int32_t batch, int32_t n, const double* kappa)
{
    SIL_REQUIRE(out != nullptr && kappa != nullptr, SIL_ERR_INVALID, "poisson batch: null argument");
    SIL_REQUIRE(batch >= 1 && n >= 3, SIL_ERR_INVALID, "poisson batch: need batch >= 1 and n >= 3");
    auto* h = new (std::nothrow) Batch();
    SIL_REQUIRE(h != nullptr, SIL_ERR_INVALID, "poisson batch: out of host memory");
    h->B = batch;
    h->n = n;
    h->nn = (size_t)n * n;
    h->total = h->nn * batch;
    int nbx = (int)((h->nn + 4 * NT - 1) / (4 * NT));
    h->nbx = nbx < 1 ? 1 : (nbx > 256 ? 256 : nbx);
    int rc = SIL_OK;
    auto A = [&](auto& p, size_t c) {
        if (rc == SIL_OK) rc = h->alloc(p, c);
    };
    A(h->wx, h->total);
    A(h->wy, h->total);
    A(h->dinv, h->total);
    A(h->wx32, h->total);
    A(h->wy32, h->total);
    A(h->dinv32, h->total);
    A(h->logk, h->total);
    A(h->r, h->total);
    A(h->part, (size_t)batch * h->nbx * sil::kParts);
    A(h->sc, (size_t)batch);
    A(h->done, (size_t)batch);
    if (rc == SIL_OK && hipHostMalloc((void**)&h->done_host, batch * sizeof(int)) != hipSuccess) {
        sil::set_error("poisson batch: pinned allocation failed");
        rc = SIL_ERR_HIP;
    }
    if (rc == SIL_OK && (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
                         hipEventCreateWithFlags(&h->ev, hipEventDisableTiming) != hipSuccess)) {
        sil::set_error("poisson batch: stream creation failed");
        rc = SIL_ERR_HIP;
    }
    if (rc != SIL_OK) {
        delete h;
        return rc;
    }
    hipStream_t s = h->stream;
    hipLaunchKernelGGL(sil::k_pb_weights, grid_of(h), dim3(NT), 0, s, n, kappa, h->wx, h->wy, h->wx32, h->wy32,
                       h->dinv, h->dinv32, h->logk);
    hipLaunchKernelGGL(sil::k_pb_logk_stats, grid_of(h), dim3(NT), 0, s, n, h->logk, h->part);
    hipLaunchKernelGGL(sil::k_pb_logk_norm, grid_of(h), dim3(NT), 0, s, n, h->nbx, h->part, h->logk);
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) {
        delete h;
        sil::set_error("poisson batch: set-up kernels failed");
        return SIL_ERR_HIP;
    }
    *out = h;
    return SIL_OK;
}

int sil_poisson2d_batch_destroy(sil_poisson2d_batch h)
{
    delete h;
    return SIL_OK;
}

int sil_poisson2d_batch_apply(sil_poisson2d_batch h, const double* x, double* y, void* stream)
{
    SIL_REQUIRE(h != nullptr && x != nullptr && y != nullptr, SIL_ERR_INVALID, "poisson batch: null argument");
    hipLaunchKernelGGL(sil::k_pb_apply64, grid_of(h), dim3(NT), 0, static_cast<hipStream_t>(stream), h->n, h->wx,
                       h->wy, x, y);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int sil_poisson2d_batch_weights(sil_poisson2d_batch h, float* wx, float* wy, float* logk, void* stream)
{
    SIL_REQUIRE(h != nullptr, SIL_ERR_INVALID, "poisson batch: null handle");
    auto s = static_cast<hipStream_t>(stream);
    const size_t bytes = h->total * sizeof(float);
    if (wx) SIL_CHECK_HIP(hipMemcpyAsync(wx, h->wx32, bytes, hipMemcpyDeviceToDevice, s));
    if (wy) SIL_CHECK_HIP(hipMemcpyAsync(wy, h->wy32, bytes, hipMemcpyDeviceToDevice, s));
    if (logk) SIL_CHECK_HIP(hipMemcpyAsync(logk, h->logk, bytes, hipMemcpyDeviceToDevice, s));
    return SIL_OK;
}

int sil_poisson2d_batch_solve(sil_poisson2d_batch h, const double* b, double* x, const sil_solver_settings* st,
                              sil_unet net, sil_batch_result* results, double* seconds,
                              const sil_solve_record* record, void* stream)
{
    SIL_REQUIRE(h != nullptr && b != nullptr && x != nullptr && st != nullptr, SIL_ERR_INVALID,
                "poisson batch: null argument");
    const bool learned = st->preconditioner == SIL_PRECOND_LEARNED;
    SIL_REQUIRE(!learned || net != nullptr, SIL_ERR_INVALID, "poisson batch: the learned preconditioner needs a network");
    if (learned) {
        int mb = 0, nh = 0, nw = 0;
        sil::unet::net_shape(net, &mb, &nh, &nw);
        SIL_REQUIRE(nh == h->n && nw == h->n && mb >= h->B, SIL_ERR_INVALID,
                    "poisson batch: network built for batch <= %d of %dx%d images, problems are %d of %dx%d", mb, nh,
                    nw, h->B, h->n, h->n);
    }
    SIL_REQUIRE(st->precision == SIL_PRECISION_FP64 || st->precision == SIL_PRECISION_MIXED, SIL_ERR_INVALID,
                "poisson batch: unknown precision %d", st->precision);
    SIL_REQUIRE(st->preconditioner == SIL_PRECOND_IDENTITY || st->preconditioner == SIL_PRECOND_JACOBI || learned,
                SIL_ERR_INVALID, "poisson batch: unknown preconditioner %d", st->preconditioner);
    SIL_REQUIRE(st->max_iterations >= 1 && st->relative_tolerance >= 0.0, SIL_ERR_INVALID,
                "poisson batch: bad iteration limit or tolerance");
    SIL_REQUIRE(record == nullptr || (record->every >= 2 && record->every % 2 == 0 && record->max_records >= 1 &&
                                      record->r != nullptr && record->x != nullptr && record->count != nullptr),
                SIL_ERR_INVALID, "poisson batch: bad record request (even interval >= 2, buffers)");
    const bool mixed = st->precision == SIL_PRECISION_MIXED;
    int R = st->replace_interval > 0 ? st->replace_interval : (learned ? 4 : 10);
    if (R % 2) ++R;
    if (record) R = record->every;
    const double tol2 = st->relative_tolerance * st->relative_tolerance;
    const int max_iter = st->max_iterations;
    hipStream_t s = h->stream;
    auto us = static_cast<hipStream_t>(stream);

    // allocate the precision's work vectors once
    int rc = SIL_OK;
    auto A = [&](auto& p, size_t c) {
        if (rc == SIL_OK && p == nullptr) rc = h->alloc(p, c);
    };
    if (mixed) {
        A(h->p32[0], h->total), A(h->p32[1], h->total), A(h->q32[0], h->total), A(h->q32[1], h->total);
        if (!learned) A(h->z32, h->total);
    } else {
        A(h->p64[0], h->total), A(h->p64[1], h->total), A(h->q64[0], h->total), A(h->q64[1], h->total);
        if (!learned) A(h->z64, h->total);
    }
    if (learned) {
        // the interleaved network input only when asked for (SIL_FCG_PLANAR=0)
        const char* e = std::getenv("SIL_FCG_PLANAR");
        h->interleaved = e != nullptr && std::atoi(e) == 0;
        if (h->interleaved) A(h->xin, h->total * 2);
        A(h->zn, h->total), A(h->r1, h->total), A(h->an, h->total), A(h->zf, h->total);
        A(h->slot, h->B), A(h->nact, 1), A(h->slot_src, h->B), A(h->slot_scale, h->B);
    }
    if (rc != SIL_OK) return rc;

    SIL_CHECK_HIP(hipEventRecord(h->ev, us));
    SIL_CHECK_HIP(hipStreamWaitEvent(s, h->ev, 0));
    SIL_CHECK_HIP(hipStreamSynchronize(s));
    const auto t0 = std::chrono::steady_clock::now();

    hipLaunchKernelGGL(sil::k_pb_dinv, grid_of(h), dim3(NT), 0, s, h->n, h->wx, h->wy,
                       st->preconditioner == SIL_PRECOND_IDENTITY ? 0 : 1, h->dinv, h->dinv32);
    if (mixed) {
        fill(h->p32[0], h->total, 0.f, s), fill(h->p32[1], h->total, 0.f, s);
        fill(h->q32[0], h->total, 0.f, s), fill(h->q32[1], h->total, 0.f, s);
    } else {
        fill(h->p64[0], h->total, 0.0, s), fill(h->p64[1], h->total, 0.0, s);
        fill(h->q64[0], h->total, 0.0, s), fill(h->q64[1], h->total, 0.0, s);
    }
    if (mixed)
        hipLaunchKernelGGL(sil::k_pb_residual<float>, grid_of(h), dim3(NT), 0, s, h->n, h->wx, h->wy, b, x, h->r,
                           h->dinv, learned ? (float*)nullptr : h->z32, (const sil::Scal*)nullptr, h->part);
    else
        hipLaunchKernelGGL(sil::k_pb_residual<double>, grid_of(h), dim3(NT), 0, s, h->n, h->wx, h->wy, b, x, h->r,
                           h->dinv, learned ? (double*)nullptr : h->z64, (const sil::Scal*)nullptr, h->part);
    scalars(h, sil::PH_INIT, tol2, max_iter, 1, s);
    SIL_LAUNCH_CHECK();

    // one graph = R iterations, the last with an FP64 residual recomputation
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    SIL_CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int crc = SIL_OK;
    for (int k = 0; k < R && crc == SIL_OK; ++k) {
        const bool rep = k == R - 1;
        const int parity = k & 1;
        if (learned) {
            if (mixed)
                crc = fcg_iteration<float>(h, net, h->p32, h->q32, h->wx32, h->wy32, b, x, parity, rep, tol2,
                                           max_iter, s);
            else
                crc = fcg_iteration<double>(h, net, h->p64, h->q64, h->wx, h->wy, b, x, parity, rep, tol2, max_iter,
                                            s);
        } else if (mixed) {
            cg_iteration<float>(h, h->p32, h->q32, h->z32, h->wx32, h->wy32, b, x, parity, rep, tol2, max_iter, s);
        } else {
            cg_iteration<double>(h, h->p64, h->q64, h->z64, h->wx, h->wy, b, x, parity, rep, tol2, max_iter, s);
        }
    }
    const hipError_t ce = hipStreamEndCapture(s, &graph);
    if (crc != SIL_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return crc;
    }
    SIL_CHECK_HIP(ce);
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    SIL_CHECK_HIP(ie);

    int launched = 0, records = 0;
    rc = SIL_OK;
    for (;;) {
        if (hipMemcpyAsync(h->done_host, h->done, h->B * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rc = SIL_ERR_HIP;
            sil::set_error("poisson batch: synchronisation failed");
            break;
        }
        int ndone = 0;
        for (int i = 0; i < h->B; ++i) ndone += h->done_host[i] ? 1 : 0;
        if (record && records < record->max_records && (launched == 0 || ndone < h->B)) {
            float* rr = record->r + (size_t)records * h->total;
            double* xr = record->x + (size_t)records * h->total;
            hipLaunchKernelGGL(sil::k_pb_record, dim3(4096), dim3(NT), 0, s, h->total, h->r, x, rr, xr);
            ++records;
        }
        if (ndone == h->B || (size_t)launched * R >= (size_t)max_iter) break;
        if (hipGraphLaunch(exec, s) != hipSuccess) {
            rc = SIL_ERR_HIP;
            sil::set_error("poisson batch: graph launch failed");
            break;
        }
        ++launched;
    }
    (void)hipGraphExecDestroy(exec);
    if (rc != SIL_OK) return rc;
    // FP64 residual b − A x of every problem, as reported
    hipLaunchKernelGGL(sil::k_pb_residual<double>, grid_of(h), dim3(NT), 0, s, h->n, h->wx, h->wy, b, x, h->r, h->dinv,
                       (double*)nullptr, (const sil::Scal*)nullptr, h->part);
    scalars(h, sil::PH_VERIFY, tol2, max_iter, 1, s);
    SIL_CHECK_HIP(hipStreamSynchronize(s));
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (record) *record->count = records;
    if (results) {
        std::vector<sil::Scal> hs(h->B);
        SIL_CHECK_HIP(hipMemcpy(hs.data(), h->sc, h->B * sizeof(sil::Scal), hipMemcpyDeviceToHost));
        for (int i = 0; i < h->B; ++i) {
            const double rel = hs[i].bb > 0.0 ? std::sqrt(hs[i].rr / hs[i].bb) : 0.0;
            results[i].iterations = hs[i].iters;
            results[i].converged = rel <= st->relative_tolerance ? 1 : 0;
            results[i].true_relative_residual = rel;
        }
    }
    SIL_CHECK_HIP(hipEventRecord(h->ev, s));
    SIL_CHECK_HIP(hipStreamWaitEvent(us, h->ev, 0));
    return SIL_OK;
}

}  // extern "C"
