// Batched persistent solver: many independent 2D Dirichlet Poisson problems
// of uneven size in one launch.
//
// No reference counterpart: SimiLie solves one operator per Ginkgo call.  A
// persistent grid of workgroups pulls problems from a device work queue
// (largest first); each workgroup runs a whole CG solve for its problem with
// workgroup-level reductions and barriers only.  The recurrence (r, p, q, and
// every operator application inside the loop) is FP32 while x, all dot
// products and the scalars are FP64; the residual is replaced by the FP64
// residual b − A x (FP64 edge weights) whenever the recurrence reports
// convergence or every `replace` iterations, and a problem is declared solved
// only on that FP64 residual.  The launch (queue reset + solve) is captured
// into a hipGraph by the host.  Workgroup size is chosen by the autotuner.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <new>
#include <numeric>
#include <vector>

#include "batched.hpp"

namespace sil {

namespace {

template <int BLOCK>
__device__ __forceinline__ double wg_sum(double v, double* sh)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads(); // previous readers of sh are done
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < BLOCK / 64; ++i) s += sh[i];
    return s; // every thread holds the sum
}

template <int BLOCK>
__device__ __forceinline__ void wg_sum2(double& a, double& b, double* sh)
{
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        sh[w] = a;
        sh[BLOCK / 64 + w] = b;
    }
    __syncthreads();
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int i = 0; i < BLOCK / 64; ++i) {
        s0 += sh[i];
        s1 += sh[BLOCK / 64 + i];
    }
    a = s0;
    b = s1;
}

// q = A p on the interior of an n×n problem (x slowest); boundary rows 0
template <typename T>
__device__ __forceinline__ T apply_row(const T* __restrict__ wx, const T* __restrict__ wy, const T* __restrict__ p,
                                       int n, int i, int j, int t)
{
    const T pv = p[t];
    const T ue = (i + 2 < n) ? p[t + n] : T(0);
    const T uw = (i > 1) ? p[t - n] : T(0);
    const T un = (j + 2 < n) ? p[t + 1] : T(0);
    const T us = (j > 1) ? p[t - 1] : T(0);
    return wx[t] * (pv - ue) + wx[t - n] * (pv - uw) + wy[t] * (pv - un) + wy[t - 1] * (pv - us);
}

struct BatchArgs {
    const BatchProblem* problems;
    const int* order;
    int nprob;
    int* queue;
    const float* wx32;
    const float* wy32;
    const double* wx64;
    const double* wy64;
    const double* b;
    double* x;
    float* r;
    float* p;
    float* q;
    const float* dinv32; // Jacobi: inverse diagonal (0 on the boundary ring)
    int jacobi;
    double tol;
    int max_it;
    int replace;
    BatchResult* results;
};

constexpr int kU = 4; // nodes per thread per pass

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_batched_cg(BatchArgs a)
{
    __shared__ double sh[2 * (BLOCK / 64)];
    __shared__ int slot;
    for (;;) {
        if (threadIdx.x == 0) slot = atomicAdd(a.queue, 1);
        __syncthreads();
        const int s = slot;
        __syncthreads();
        if (s >= a.nprob) return; // every wave reaches this once the queue is drained
        const int id = a.order[s];
        const BatchProblem P = a.problems[id];
        const int n = P.side;
        const int64_t off = P.offset;
        const int nn = n * n;
        const float* wx = a.wx32 + off;
        const float* wy = a.wy32 + off;
        const double* wx64 = a.wx64 + off;
        const double* wy64 = a.wy64 + off;
        const double* b = a.b + off;
        double* x = a.x + off;
        float* r = a.r + off;
        float* p = a.p + off;
        float* q = a.q + off;

        const float* dinv = a.dinv32 + off;
        const bool jac = a.jacobi != 0;
        // z = D⁻¹ r (Jacobi) or r
        auto zval = [&](int t, float rf) { return jac ? rf * dinv[t] : rf; };

        // x = 0, r = b (FP32), p = z; <b,b>, <r,z>
        double bb = 0.0, rz = 0.0;
        for (int t = threadIdx.x; t < nn; t += BLOCK) {
            const double bi = b[t];
            x[t] = 0.0;
            const float rf = (float)bi;
            r[t] = rf;
            const float zf = zval(t, rf);
            p[t] = zf;
            bb += bi * bi;
            rz += (double)rf * (double)zf;
        }
        wg_sum2<BLOCK>(bb, rz, sh);
        const double tol2bb = a.tol * a.tol * bb;
        double rr_ref = bb; // <r,r> at the last replacement
        int it = 0, since = 0;
        bool converged = bb == 0.0;
        double rr_true = bb;
        // The three passes of an iteration handle kU nodes per thread at a
        // time (nodes BLOCK apart, so every load stays coalesced) and issue
        // all of their loads before using any: a workgroup alone streams its
        // problem, and only loads in flight hide the memory latency.
        const bool wide = n >= 3; // node n+1 is interior: a safe address for inactive lanes
        while (!converged && it < a.max_it) {
            __syncthreads(); // p visible to the whole workgroup
            double pq = 0.0;
            for (int t0 = threadIdx.x; t0 < nn; t0 += kU * BLOCK) {
                float pc[kU], pe[kU], pw[kU], pn[kU], ps[kU], ax[kU], aw[kU], ay[kU], as[kU];
                bool in[kU], ok[kU], he[kU], hw[kU], hn[kU], hs[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const int t = t0 + u * BLOCK;
                    const int i = t / n, j = t - i * n;
                    ok[u] = t < nn;
                    in[u] = wide && ok[u] && i > 0 && j > 0 && i + 1 < n && j + 1 < n;
                    he[u] = i + 2 < n;
                    hw[u] = i > 1;
                    hn[u] = j + 2 < n;
                    hs[u] = j > 1;
                    const int tc = in[u] ? t : (wide ? n + 1 : 0);
                    pc[u] = p[ok[u] ? t : 0];
                    if (wide) {
                        pe[u] = p[tc + n];
                        pw[u] = p[tc - n];
                        pn[u] = p[tc + 1];
                        ps[u] = p[tc - 1];
                        ax[u] = wx[tc];
                        aw[u] = wx[tc - n];
                        ay[u] = wy[tc];
                        as[u] = wy[tc - 1];
                    }
                }
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    if (!ok[u]) continue;
                    float qv = 0.f;
                    if (in[u]) {
                        const float v = pc[u];
                        qv = ax[u] * (v - (he[u] ? pe[u] : 0.f)) + aw[u] * (v - (hw[u] ? pw[u] : 0.f)) +
                             ay[u] * (v - (hn[u] ? pn[u] : 0.f)) + as[u] * (v - (hs[u] ? ps[u] : 0.f));
                    }
                    q[t0 + u * BLOCK] = qv;
                    pq += (double)pc[u] * (double)qv;
                }
            }
            pq = wg_sum<BLOCK>(pq, sh);
            const double alpha = rz / pq;
            double rz_new = 0.0, rr_new = 0.0;
            for (int t0 = threadIdx.x; t0 < nn; t0 += kU * BLOCK) {
                double xv[kU];
                float pv[kU], rv[kU], qv[kU], dv[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const int t = t0 + u * BLOCK, tc = t < nn ? t : 0;
                    xv[u] = x[tc];
                    pv[u] = p[tc];
                    rv[u] = r[tc];
                    qv[u] = q[tc];
                    dv[u] = jac ? dinv[tc] : 1.f;
                }
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const int t = t0 + u * BLOCK;
                    if (t >= nn) continue;
                    x[t] = xv[u] + alpha * (double)pv[u];
                    const float rf = (float)((double)rv[u] - alpha * (double)qv[u]);
                    r[t] = rf;
                    rz_new += (double)rf * (double)(jac ? rf * dv[u] : rf);
                    rr_new += (double)rf * (double)rf;
                }
            }
            wg_sum2<BLOCK>(rz_new, rr_new, sh);
            ++it;
            ++since;
            const double rz_old = rz;
            double beta = rz_new / rz_old;
            rz = rz_new;
            // replace when the recurrence claims convergence, when its residual fell
            // 100-fold since the last replacement (FP32 drift grows with that
            // ratio), or every `replace` iterations
            if (rr_new <= tol2bb || rr_new < 1e-4 * rr_ref || since >= a.replace) {
                // FP64 residual replacement b − A x
                __syncthreads();
                double rr = 0.0, rzr = 0.0;
                for (int t = threadIdx.x; t < nn; t += BLOCK) {
                    const int i = t / n, j = t - i * n;
                    double res = 0.0;
                    if (i > 0 && j > 0 && i + 1 < n && j + 1 < n) res = b[t] - apply_row<double>(wx64, wy64, x, n, i, j, t);
                    const float rf = (float)res;
                    r[t] = rf;
                    rr += res * res;
                    rzr += (double)rf * (double)zval(t, rf);
                }
                wg_sum2<BLOCK>(rr, rzr, sh);
                rr_true = rr;
                since = 0;
                if (rr <= tol2bb) {
                    converged = true;
                    break;
                }
                beta = rzr / rz_old; // replaced <r,z> over the previous iteration's <r,z>
                rz = rzr;
                rr_ref = rr;
            }
            const float bf = (float)beta;
            for (int t0 = threadIdx.x; t0 < nn; t0 += kU * BLOCK) {
                float rv[kU], dv[kU], pv[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const int t = t0 + u * BLOCK, tc = t < nn ? t : 0;
                    rv[u] = r[tc];
                    dv[u] = jac ? dinv[tc] : 1.f;
                    pv[u] = p[tc];
                }
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const int t = t0 + u * BLOCK;
                    if (t < nn) p[t] = (jac ? rv[u] * dv[u] : rv[u]) + bf * pv[u];
                }
            }
        }
        if (!converged) {
            // report the FP64 residual of the final iterate
            __syncthreads();
            double rr = 0.0, dummy2 = 0.0;
            for (int t = threadIdx.x; t < nn; t += BLOCK) {
                const int i = t / n, j = t - i * n;
                double res = 0.0;
                if (i > 0 && j > 0 && i + 1 < n && j + 1 < n) res = b[t] - apply_row<double>(wx64, wy64, x, n, i, j, t);
                rr += res * res;
            }
            wg_sum2<BLOCK>(rr, dummy2, sh);
            rr_true = rr;
        }
        if (threadIdx.x == 0) {
            a.results[id].iterations = it;
            a.results[id].converged = converged ? 1 : 0;
            a.results[id].true_relative_residual = bb == 0.0 ? 0.0 : sqrt(rr_true / bb);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------- wide path ---

constexpr int kWideNT = 256;

enum WidePhase { W_INIT = 0, W_ALPHA, W_BETA, W_REPLACED, W_FINAL };

// block sums of two values into part[0], part[1]
__device__ __forceinline__ void wide_partials(double v0, double v1, double* __restrict__ part)
{
    __shared__ double sh[2][kWideNT / 64];
    for (int off = 32; off > 0; off >>= 1) {
        v0 += __shfl_down(v0, off, 64);
        v1 += __shfl_down(v1, off, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][w] = v0;
        sh[1][w] = v1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int i = 0; i < kWideNT / 64; ++i) {
            s0 += sh[0][i];
            s1 += sh[1][i];
        }
        part[0] = s0;
        part[1] = s1;
    }
}

// the problem of wide slot blockIdx.y and its arrays
struct WideView {
    int n, nn;
    int64_t off;
    double* part;
};

__device__ __forceinline__ WideView wide_view(const WideLaunch& L)
{
    const BatchProblem P = L.problems[L.ids[blockIdx.y]];
    WideView v;
    v.n = P.side;
    v.nn = P.side * P.side;
    v.off = P.offset;
    v.part = L.part + ((size_t)blockIdx.y * L.nbx + blockIdx.x) * 2;
    return v;
}

#define WIDE_NODES(v) for (int t = blockIdx.x * kWideNT + threadIdx.x; t < (v).nn; t += L.nbx * kWideNT)

__global__ __launch_bounds__(kWideNT) void k_wide_init(WideLaunch L)
{
    const WideView v = wide_view(L);
    const double* b = L.b + v.off;
    double* x = L.x + v.off;
    float* r = L.r + v.off;
    float* p = L.p + v.off;
    const float* dinv = L.dinv32 + v.off;
    double bb = 0.0, rz = 0.0;
    WIDE_NODES(v)
    {
        const double bi = b[t];
        x[t] = 0.0;
        const float rf = (float)bi;
        r[t] = rf;
        const float zf = L.jacobi ? rf * dinv[t] : rf;
        p[t] = zf;
        bb += bi * bi;
        rz += (double)rf * (double)zf;
    }
    wide_partials(bb, rz, v.part);
}

__global__ __launch_bounds__(kWideNT) void k_wide_apply(WideLaunch L)
{
    if (L.state[blockIdx.y].done) return; // uniform per workgroup
    const WideView v = wide_view(L);
    const int n = v.n;
    const float* wx = L.wx32 + v.off;
    const float* wy = L.wy32 + v.off;
    const float* p = L.p + v.off;
    float* q = L.q + v.off;
    double pq = 0.0;
    WIDE_NODES(v)
    {
        const int i = t / n, j = t - i * n;
        float qv = 0.f;
        if (i > 0 && j > 0 && i + 1 < n && j + 1 < n) qv = apply_row<float>(wx, wy, p, n, i, j, t);
        q[t] = qv;
        pq += (double)p[t] * (double)qv;
    }
    wide_partials(pq, 0.0, v.part);
}

__global__ __launch_bounds__(kWideNT) void k_wide_update(WideLaunch L)
{
    const WideState& S = L.state[blockIdx.y];
    if (S.done) return;
    const double alpha = S.alpha;
    const WideView v = wide_view(L);
    double* x = L.x + v.off;
    const float* p = L.p + v.off;
    float* r = L.r + v.off;
    const float* q = L.q + v.off;
    const float* dinv = L.dinv32 + v.off;
    double rz = 0.0, rr = 0.0;
    WIDE_NODES(v)
    {
        x[t] += alpha * (double)p[t];
        const float rf = (float)((double)r[t] - alpha * (double)q[t]);
        r[t] = rf;
        rz += (double)rf * (double)(L.jacobi ? rf * dinv[t] : rf);
        rr += (double)rf * (double)rf;
    }
    wide_partials(rz, rr, v.part);
}

// FP64 residual b − A x: replaces r (FP32) when the problem's scalar phase
// asked for it; with `report`, only its square norm, for problems that did
// not converge
__global__ __launch_bounds__(kWideNT) void k_wide_residual(WideLaunch L, int report)
{
    const WideState& S = L.state[blockIdx.y];
    if (report ? S.converged != 0 : (S.done || !S.replace_now)) return;
    const WideView v = wide_view(L);
    const int n = v.n;
    const double* b = L.b + v.off;
    const double* x = L.x + v.off;
    const double* wx = L.wx64 + v.off;
    const double* wy = L.wy64 + v.off;
    float* r = L.r + v.off;
    const float* dinv = L.dinv32 + v.off;
    double rr = 0.0, rz = 0.0;
    WIDE_NODES(v)
    {
        const int i = t / n, j = t - i * n;
        double res = 0.0;
        if (i > 0 && j > 0 && i + 1 < n && j + 1 < n) res = b[t] - apply_row<double>(wx, wy, x, n, i, j, t);
        rr += res * res;
        if (!report) {
            const float rf = (float)res;
            r[t] = rf;
            rz += (double)rf * (double)(L.jacobi ? rf * dinv[t] : rf);
        }
    }
    wide_partials(rr, rz, v.part);
}

__global__ __launch_bounds__(kWideNT) void k_wide_direction(WideLaunch L)
{
    const WideState& S = L.state[blockIdx.y];
    if (S.done) return;
    const float beta = (float)S.beta;
    const WideView v = wide_view(L);
    float* p = L.p + v.off;
    const float* r = L.r + v.off;
    const float* dinv = L.dinv32 + v.off;
    WIDE_NODES(v) p[t] = (L.jacobi ? r[t] * dinv[t] : r[t]) + beta * p[t];
}

// one wavefront per problem: the block partials in a fixed order, then the
// scalar recurrence of the persistent kernel
__global__ __launch_bounds__(64) void k_wide_scalars(WideLaunch L, int phase)
{
    const int slot = blockIdx.x;
    WideState& S = L.state[slot];
    if (phase == W_FINAL ? S.converged != 0 : (phase != W_INIT && S.done)) return; // uniform
    if (phase == W_REPLACED && !S.replace_now) return;
    double v0 = 0.0, v1 = 0.0;
    for (int q = threadIdx.x; q < L.nbx; q += 64) {
        v0 += L.part[((size_t)slot * L.nbx + q) * 2];
        v1 += L.part[((size_t)slot * L.nbx + q) * 2 + 1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        v0 += __shfl_down(v0, off, 64);
        v1 += __shfl_down(v1, off, 64);
    }
    if (threadIdx.x != 0) return;
    switch (phase) {
    case W_INIT:
        S.bb = v0;
        S.rz = v1;
        S.tol2bb = L.tol * L.tol * v0;
        S.rr_ref = v0;
        S.rr_true = v0;
        S.it = 0;
        S.since = 0;
        S.replace_now = 0;
        S.converged = v0 == 0.0 ? 1 : 0;
        S.done = (S.converged || L.max_it <= 0) ? 1 : 0;
        break;
    case W_ALPHA:
        S.pq = v0;
        S.alpha = S.rz / v0;
        break;
    case W_BETA:
        S.it += 1;
        S.since += 1;
        S.rz_old = S.rz;
        S.rz_new = v0;
        S.rr_new = v1;
        S.beta = v0 / S.rz_old;
        S.rz = v0;
        // replace when the recurrence claims convergence, when its residual
        // fell 100-fold since the last replacement, or every `replace`
        // iterations
        S.replace_now = (v1 <= S.tol2bb || v1 < 1e-4 * S.rr_ref || S.since >= L.replace) ? 1 : 0;
        if (!S.replace_now && S.it >= L.max_it) S.done = 1;
        break;
    case W_REPLACED:
        S.rr_true = v0;
        S.since = 0;
        S.replace_now = 0;
        if (v0 <= S.tol2bb) {
            S.converged = 1;
            S.done = 1;
        } else {
            S.beta = v1 / S.rz_old; // replaced <r,z> over the previous iteration's <r,z>
            S.rz = v1;
            S.rr_ref = v0;
            if (S.it >= L.max_it) S.done = 1;
        }
        break;
    case W_FINAL:
        S.rr_true = v0;
        break;
    }
    L.done[slot] = S.done;
}

__global__ void k_wide_results(WideLaunch L)
{
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= L.nwide) return;
    const WideState& S = L.state[slot];
    BatchResult& R = L.results[L.ids[slot]];
    R.iterations = S.it;
    R.converged = S.converged;
    R.true_relative_residual = S.bb == 0.0 ? 0.0 : sqrt(S.rr_true / S.bb);
}

__global__ void k_reset_queue(int* q) { *q = 0; }

__global__ void k_batch_weights(const BatchProblem* __restrict__ problems, int nprob, const double* __restrict__ kappa,
                                float* __restrict__ wx32, float* __restrict__ wy32, double* __restrict__ wx64,
                                double* __restrict__ wy64, float* __restrict__ dinv32)
{
    const int id = blockIdx.y;
    if (id >= nprob) return;
    const BatchProblem P = problems[id];
    const int n = P.side;
    const int64_t off = P.offset;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n * n; t += gridDim.x * blockDim.x) {
        const int i = t / n, j = t - i * n;
        const double* k = kappa + off;
        // uniform unit spacing: ⋆1 ratios are 1, edge value = mean of its two nodes
        const double ax = (i + 1 < n) ? 0.5 * (k[t] + k[t + n]) : k[t];
        const double ay = (j + 1 < n) ? 0.5 * (k[t] + k[t + 1]) : k[t];
        wx64[off + t] = ax;
        wy64[off + t] = ay;
        wx32[off + t] = (float)ax;
        wy32[off + t] = (float)ay;
        // Jacobi: the diagonal Σ of the four edge weights of an interior node
        double d = 0.0;
        if (i > 0 && j > 0 && i + 1 < n && j + 1 < n)
            d = 1.0 / (ax + 0.5 * (k[t] + k[t - n]) + ay + 0.5 * (k[t] + k[t - 1]));
        dinv32[off + t] = (float)d;
    }
}

template <int BLOCK>
int launch_block(const BatchArgs& a, int grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_batched_cg<BLOCK>, dim3(grid), dim3(BLOCK), 0, s, a);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

}  // namespace

int batched_weights(const BatchProblem* problems_dev, int nprob, int max_side, const double* kappa, float* wx32,
                    float* wy32, double* wx64, double* wy64, float* dinv32, hipStream_t s)
{
    const int64_t per = (int64_t)max_side * max_side;
    const unsigned gx = (unsigned)std::min<int64_t>((per + 255) / 256, 1024);
    hipLaunchKernelGGL(k_batch_weights, dim3(gx, nprob), dim3(256), 0, s, problems_dev, nprob, kappa, wx32, wy32,
                       wx64, wy64, dinv32);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int batched_enqueue(const BatchLaunch& L, hipStream_t s)
{
    BatchArgs a{L.problems, L.order, L.nprob, L.queue, L.wx32, L.wy32, L.wx64, L.wy64, L.b, L.x,
                L.r, L.p, L.q, L.dinv32, L.jacobi, L.tol, L.max_it, L.replace, L.results};
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(1), 0, s, L.queue);
    SIL_LAUNCH_CHECK();
    switch (L.block) {
    case 256:
        return launch_block<256>(a, L.grid, s);
    case 512:
        return launch_block<512>(a, L.grid, s);
    case 1024:
        return launch_block<1024>(a, L.grid, s);
    default:
        set_error("batched solve: unsupported workgroup size %d", L.block);
        return SIL_ERR_INVALID;
    }
}

int wide_init(const WideLaunch& L, hipStream_t s)
{
    hipLaunchKernelGGL(k_wide_init, dim3(L.nbx, L.nwide), dim3(kWideNT), 0, s, L);
    hipLaunchKernelGGL(k_wide_scalars, dim3(L.nwide), dim3(64), 0, s, L, (int)W_INIT);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int wide_iteration(const WideLaunch& L, hipStream_t s)
{
    const dim3 g(L.nbx, L.nwide);
    hipLaunchKernelGGL(k_wide_apply, g, dim3(kWideNT), 0, s, L);
    hipLaunchKernelGGL(k_wide_scalars, dim3(L.nwide), dim3(64), 0, s, L, (int)W_ALPHA);
    hipLaunchKernelGGL(k_wide_update, g, dim3(kWideNT), 0, s, L);
    hipLaunchKernelGGL(k_wide_scalars, dim3(L.nwide), dim3(64), 0, s, L, (int)W_BETA);
    hipLaunchKernelGGL(k_wide_residual, g, dim3(kWideNT), 0, s, L, 0);
    hipLaunchKernelGGL(k_wide_scalars, dim3(L.nwide), dim3(64), 0, s, L, (int)W_REPLACED);
    hipLaunchKernelGGL(k_wide_direction, g, dim3(kWideNT), 0, s, L);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int wide_finish(const WideLaunch& L, hipStream_t s)
{
    hipLaunchKernelGGL(k_wide_residual, dim3(L.nbx, L.nwide), dim3(kWideNT), 0, s, L, 1);
    hipLaunchKernelGGL(k_wide_scalars, dim3(L.nwide), dim3(64), 0, s, L, (int)W_FINAL);
    hipLaunchKernelGGL(k_wide_results, dim3((L.nwide + 255) / 256), dim3(256), 0, s, L);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

}  // namespace sil

namespace sil {

int batched_occupancy(int block, int* blocks_per_cu)
{
    hipError_t e;
    switch (block) {
    case 256:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_batched_cg<256>, 256, 0);
        break;
    case 512:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_batched_cg<512>, 512, 0);
        break;
    case 1024:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_batched_cg<1024>, 1024, 0);
        break;
    default:
        set_error("batched solve: unsupported workgroup size %d", block);
        return SIL_ERR_INVALID;
    }
    if (e != hipSuccess) {
        set_error("occupancy query failed: %s", hipGetErrorString(e));
        return SIL_ERR_HIP;
    }
    return SIL_OK;
}

}  // namespace sil
