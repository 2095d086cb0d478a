// Conjugate gradients on the GPU, with the CG scalars kept on the device.
//
// The reference hands its matrix-free operator to Ginkgo's CG
// (src/similie/solvers/minimize_strong_formulation_residual.hpp:1028-1304),
// which synchronises with the host at every iteration for its ResidualNorm
// stopping criterion.  Here one iteration is four kernels —
//   K1 p = z + β p_old fused into q = A p with per-block ⟨p, q⟩ partials,
//   K2 α = ⟨r, z⟩ / ⟨p, q⟩,
//   K3 x += α p, r −= α q with per-block ⟨r, z⟩ and ⟨r, r⟩ partials,
//   K4 β, the stopping test ‖r‖2 ≤ tol·‖b‖2 and the iteration counter —
// every scalar stays in device memory, every kernel returns at once when the
// device `done` flag is set, and a chunk of iterations is captured once into a
// hipGraph and replayed; the host reads the flag once per chunk.  Partial sums
// are reduced in a fixed order, so a solve is bit-reproducible.
//
// Mixed precision: the recurrence (r, p, q) and every operator application
// inside the loop run in FP32 while x, all dot products and the scalars are
// FP64; the residual is periodically replaced by the FP64 residual b − A x
// (van der Vorst & Ye reliable updating), and convergence is declared only
// on that FP64 residual.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <mutex>
#include <new>

#include "operators.hpp"

namespace sil {

struct CgScalars {
    double rz, rz_prev, pq, alpha, beta, rr, bb, tol2bb;
    double rr_true;
    int done, iter, max_iter, converged;
};

struct Workspace {
    int64_t n = 0;
    // FP64 recurrence
    double *x = nullptr, *b = nullptr, *r = nullptr, *q = nullptr, *p[2] = {nullptr, nullptr};
    double* inv_diag = nullptr;
    double* tmp = nullptr;
    // FP32 recurrence: r, q, p, the Jacobi inverse and the correction of x
    // accumulated since the last FP64 residual replacement
    float *r32 = nullptr, *q32 = nullptr, *p32[2] = {nullptr, nullptr};
    float *inv_diag32 = nullptr, *xc32 = nullptr;
    double* partials = nullptr; // [3][part_stride]
    int64_t part_stride = 0;    // slots per partial-sum array
    CgScalars* sc = nullptr;
    CgScalars* host_sc = nullptr; // pinned
    int apply_blocks = 0;
    bool jacobi_ready = false;
    // SpdIsai: factor G, its transpose, t = G r and z = Gᵀ t
    CsrData isai_g, isai_gt;
    double *isai_t = nullptr, *isai_z = nullptr;
    bool isai_ready = false;
    // residual and correction of the FP64 continuation of a stalled mixed solve
    double *fb_r = nullptr, *fb_d = nullptr;
    hipGraphExec_t graph[2][3] = {}; // [precision][preconditioner: identity, Jacobi, SpdIsai]
    int graph_chunk[2][3] = {};
    // mixed solves of the 3D curl-curl operator: the single-pass iteration
    bool single_pass = false;
};

namespace {

constexpr int kUpdBlock = 256;
constexpr int kUpdBlocks = 2048;
constexpr int kRedThreads = 1024;

__device__ double block_sum(double v)
{
    __shared__ double sh[kRedThreads / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x + 63) / 64; ++i) s += sh[i];
    __syncthreads();
    return s; // valid in thread 0
}

// fixed-order reduction of `count` partials by one workgroup
__device__ double reduce_partials(const double* p, int count)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < count; i += blockDim.x) v += p[i];
    return block_sum(v);
}

template <typename T>
__global__ __launch_bounds__(kUpdBlock) void k_dot2(const T* __restrict__ a, const T* __restrict__ b, int64_t n,
                                                    double* __restrict__ partials)
{
    double v = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v += (double)a[i] * (double)b[i];
    __shared__ double sh[kUpdBlock / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < kUpdBlock / 64; ++i) s += sh[i];
        partials[blockIdx.x] = s;
    }
}

// r = b, x = 0, p_old = 0; partials for <b,b>, <r,z>
// p_is_z: the first direction p = z itself (the single-pass iteration
// starts from it); otherwise p_old = 0 and the first step forms it
template <typename T>
__global__ __launch_bounds__(kUpdBlock) void k_init(const double* __restrict__ b, double* __restrict__ x,
                                                    T* __restrict__ r, T* __restrict__ pold,
                                                    const T* __restrict__ inv_diag, int64_t n,
                                                    double* __restrict__ part_bb, double* __restrict__ part_rz,
                                                    float* __restrict__ xc, bool p_is_z = false)
{
    double vb = 0.0, vz = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double bi = b[i];
        x[i] = 0.0;
        if (xc) xc[i] = 0.0f;
        r[i] = (T)bi;
        const double ri = (double)(T)bi;
        pold[i] = p_is_z ? (T)(inv_diag ? ri * (double)inv_diag[i] : ri) : T(0);
        vb += bi * bi;
        vz += ri * (inv_diag ? ri * (double)inv_diag[i] : ri);
    }
    __shared__ double sh[2][kUpdBlock / 64];
    for (int off = 32; off > 0; off >>= 1) {
        vb += __shfl_down(vb, off, 64);
        vz += __shfl_down(vz, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = vb;
        sh[1][threadIdx.x >> 6] = vz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int i = 0; i < kUpdBlock / 64; ++i) {
            s0 += sh[0][i];
            s1 += sh[1][i];
        }
        part_bb[blockIdx.x] = s0;
        part_rz[blockIdx.x] = s1;
    }
}

__global__ void k_scalars_init(CgScalars* sc, const double* part_bb, int count_bb, const double* part_rz,
                               int count_rz, double tol, int max_iter)
{
    const double bb = reduce_partials(part_bb, count_bb);
    const double rz = reduce_partials(part_rz, count_rz);
    if (threadIdx.x == 0) {
        sc->bb = bb;
        sc->rr = bb;
        sc->rz = rz;
        sc->rz_prev = rz;
        sc->beta = 0.0;
        sc->alpha = 0.0;
        sc->tol2bb = tol * tol * bb;
        sc->iter = 0;
        sc->max_iter = max_iter;
        sc->converged = 0;
        sc->rr_true = bb;
        sc->done = (bb == 0.0 || max_iter <= 0) ? 1 : 0;
    }
}

__global__ void k_alpha(CgScalars* sc, const double* part, int count)
{
    if (sc->done) return;
    const double pq = reduce_partials(part, count);
    if (threadIdx.x == 0) {
        sc->pq = pq;
        sc->alpha = sc->rz / pq;
    }
}

// XC: x accumulates in the FP32 correction xc (folded into the FP64 x at
// every residual replacement) instead of in x itself
template <typename T, bool XC = false>
__global__ __launch_bounds__(kUpdBlock) void k_update(const CgScalars* __restrict__ sc, double* __restrict__ x,
                                                      const T* __restrict__ p, T* __restrict__ r,
                                                      const T* __restrict__ q, const T* __restrict__ inv_diag,
                                                      int64_t n, double* __restrict__ part_rz,
                                                      double* __restrict__ part_rr, float* __restrict__ xc)
{
    if (sc->done) return;
    const double alpha = sc->alpha;
    const float alpha32 = (float)alpha;
    double vz = 0.0, vr = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if constexpr (XC)
            xc[i] = fmaf(alpha32, (float)p[i], xc[i]);
        else
            x[i] += alpha * (double)p[i];
        const T ri = (T)((double)r[i] - alpha * (double)q[i]);
        r[i] = ri;
        const double rd = (double)ri;
        vr += rd * rd;
        vz += rd * (inv_diag ? rd * (double)inv_diag[i] : rd);
    }
    __shared__ double sh[2][kUpdBlock / 64];
    for (int off = 32; off > 0; off >>= 1) {
        vz += __shfl_down(vz, off, 64);
        vr += __shfl_down(vr, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = vz;
        sh[1][threadIdx.x >> 6] = vr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int i = 0; i < kUpdBlock / 64; ++i) {
            s0 += sh[0][i];
            s1 += sh[1][i];
        }
        part_rz[blockIdx.x] = s0;
        part_rr[blockIdx.x] = s1;
    }
}

__global__ void k_beta(CgScalars* sc, const double* part_rz, int count_rz, const double* part_rr, int count_rr)
{
    if (sc->done) return;
    const double rz = reduce_partials(part_rz, count_rz);
    const double rr = reduce_partials(part_rr, count_rr);
    if (threadIdx.x == 0) {
        sc->beta = rz / sc->rz;
        sc->rz_prev = sc->rz;
        sc->rz = rz;
        sc->rr = rr;
        sc->iter += 1;
        if (rr <= sc->tol2bb) {
            sc->done = 1;
            sc->converged = 1;
        } else if (sc->iter >= sc->max_iter) {
            sc->done = 1;
        }
    }
}

// ---- single-pass FP32 iteration (3D curl-curl operator) ----
//
// The march that forms q = A p also sums ⟨q, z⟩ and ⟨q, D⁻¹ q⟩ (z = D⁻¹ r), so
// the next ⟨r, z⟩ is known before r is updated:
//   γ' = ⟨r − αq, D⁻¹(r − αq)⟩ = γ − 2α⟨q, z⟩ + α²⟨q, D⁻¹ q⟩,  β' = γ'/γ,
// and one streaming pass updates r, the iterate and the next direction
// p' = D⁻¹ r' + β' p together; it also sums the exact ⟨r', z'⟩ (the next α
// uses it) and ⟨r', r'⟩.  Per iteration: the march and one vector pass, where
// the standard iteration has the march, the update and the direction pass.

// α = γ / ⟨p, q⟩ and β' from the expansion; three partial arrays at stride
__global__ void k_alpha_expand(CgScalars* sc, const double* part, int64_t stride, int count)
{
    if (sc->done) return;
    const double pq = reduce_partials(part, count);
    const double qz = reduce_partials(part + stride, count);
    const double qq = reduce_partials(part + 2 * stride, count);
    if (threadIdx.x == 0) {
        const double g = sc->rz;
        const double alpha = g / pq;
        const double gn = g - 2.0 * alpha * qz + alpha * alpha * qq;
        sc->pq = pq;
        sc->alpha = alpha;
        sc->beta = gn / g;
    }
}

// r' = r − αq; xc += αp; p' = D⁻¹ r' + β' p (arithmetic of p_value); partials
// of ⟨r', z'⟩ and ⟨r', r'⟩
__global__ __launch_bounds__(kUpdBlock) void k_update_dir(const CgScalars* __restrict__ sc,
                                                          const float* __restrict__ p, float* __restrict__ pn,
                                                          float* __restrict__ r, const float* __restrict__ q,
                                                          const float* __restrict__ inv_diag, int64_t n,
                                                          double* __restrict__ part_rz, double* __restrict__ part_rr,
                                                          float* __restrict__ xc)
{
    if (sc->done) return;
    const double alpha = sc->alpha, beta = sc->beta;
    const float alpha32 = (float)alpha;
    double vz = 0.0, vr = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float pi = p[i];
        xc[i] = fmaf(alpha32, pi, xc[i]);
        const float ri = (float)((double)r[i] - alpha * (double)q[i]);
        r[i] = ri;
        const double rd = (double)ri;
        const double z = inv_diag ? rd * (double)inv_diag[i] : rd;
        pn[i] = (float)(z + beta * (double)pi);
        vr += rd * rd;
        vz += rd * z;
    }
    __shared__ double sh[2][kUpdBlock / 64];
    for (int off = 32; off > 0; off >>= 1) {
        vz += __shfl_down(vz, off, 64);
        vr += __shfl_down(vr, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = vz;
        sh[1][threadIdx.x >> 6] = vr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int i = 0; i < kUpdBlock / 64; ++i) {
            s0 += sh[0][i];
            s1 += sh[1][i];
        }
        part_rz[blockIdx.x] = s0;
        part_rr[blockIdx.x] = s1;
    }
}

// the exact ⟨r', z'⟩ and ⟨r', r'⟩: the next α, the stopping test (β' was
// taken from the expansion)
__global__ void k_gamma(CgScalars* sc, const double* part_rz, int count_rz, const double* part_rr, int count_rr)
{
    if (sc->done) return;
    const double rz = reduce_partials(part_rz, count_rz);
    const double rr = reduce_partials(part_rr, count_rr);
    if (threadIdx.x == 0) {
        sc->rz_prev = sc->rz;
        sc->rz = rz;
        sc->rr = rr;
        sc->iter += 1;
        if (rr <= sc->tol2bb) {
            sc->done = 1;
            sc->converged = 1;
        } else if (sc->iter >= sc->max_iter) {
            sc->done = 1;
        }
    }
}

// FP64 residual replacement: r64 = b − Ax (tmp = Ax) → r (T), partials of
// <r,r> (FP64, exact residual) and <r,z> of the rounded residual
template <typename T>
__global__ __launch_bounds__(kUpdBlock) void k_replace(const double* __restrict__ b, const double* __restrict__ ax,
                                                       T* __restrict__ r, const T* __restrict__ inv_diag,
                                                       int64_t n, double* __restrict__ part_rr,
                                                       double* __restrict__ part_rz)
{
    double vr = 0.0, vz = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double res = b[i] - ax[i];
        const T rt = (T)res;
        r[i] = rt;
        vr += res * res;
        const double rd = (double)rt;
        vz += rd * (inv_diag ? rd * (double)inv_diag[i] : rd);
    }
    __shared__ double sh[2][kUpdBlock / 64];
    for (int off = 32; off > 0; off >>= 1) {
        vr += __shfl_down(vr, off, 64);
        vz += __shfl_down(vz, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = vr;
        sh[1][threadIdx.x >> 6] = vz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int i = 0; i < kUpdBlock / 64; ++i) {
            s0 += sh[0][i];
            s1 += sh[1][i];
        }
        part_rr[blockIdx.x] = s0;
        part_rz[blockIdx.x] = s1;
    }
}

// After a replacement: keep the direction p and continue from the new
// residual; β of the next direction update is recomputed from the new <r,z>.
__global__ void k_after_replace(CgScalars* sc, const double* part_rr, const double* part_rz, int count)
{
    const double rr = reduce_partials(part_rr, count);
    const double rz = reduce_partials(part_rz, count);
    if (threadIdx.x == 0) {
        sc->rr_true = rr;
        sc->rr = rr;
        sc->beta = rz / sc->rz_prev;
        sc->rz = rz;
        sc->done = 0;
        sc->converged = 0;
        if (rr <= sc->tol2bb) {
            sc->done = 1;
            sc->converged = 1;
        } else if (sc->iter >= sc->max_iter) {
            sc->done = 1;
        }
    }
}

// out = b − ax
__global__ void k_sub(const double* __restrict__ b, const double* __restrict__ ax, double* __restrict__ out, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = b[i] - ax[i];
}

// x += d
__global__ void k_add(double* __restrict__ x, const double* __restrict__ d, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] += d[i];
}

// x += xc; xc = 0: the FP32 correction folded into the FP64 iterate
__global__ void k_flush(double* __restrict__ x, float* __restrict__ xc, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += (double)xc[i];
        xc[i] = 0.0f;
    }
}

__global__ void k_to_f32(const double* __restrict__ a, float* __restrict__ b, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

__global__ void k_inv(const double* __restrict__ d, double* __restrict__ inv, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        inv[i] = d[i] != 0.0 ? 1.0 / d[i] : 0.0;
}

__global__ void k_residual_f64(const double* __restrict__ b, const double* __restrict__ ax, int64_t n,
                               double* __restrict__ part)
{
    double v = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double d = b[i] - ax[i];
        v += d * d;
    }
    __shared__ double sh[kUpdBlock / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < kUpdBlock / 64; ++i) s += sh[i];
        part[blockIdx.x] = s;
    }
}

__global__ void k_store_true(CgScalars* sc, const double* part, int count)
{
    const double rr = reduce_partials(part, count);
    if (threadIdx.x == 0) sc->rr_true = rr;
}

// ------------------------------------------------------------ dispatch ---

int apply64(sil_operator op, const double* in, double* out, DotEpilogue dot, const PUpdate<double>* pu, int* nb,
            hipStream_t s)
{
    if (op->kind == OP_MAGNETOSTATICS3D) return ms3d_apply64(op->ms, in, out, dot, pu, nb, s);
    if (op->kind == OP_CSR) return csr_apply64(op->csr, in, out, dot, pu, nb, s);
    if (op->kind == OP_MS3D_JACOBIAN) return ms3d_jacobian_apply64(op->ms, op->nl, in, out, dot, pu, nb, s);
    if (op->kind == OP_MAGNETOSTATICS2D) return ms2d_apply64(op->m2, in, out, dot, pu, nb, s);
    if (op->kind == OP_MS2D_JACOBIAN) return ms2d_jacobian_apply64(op->m2, op->nl, in, out, dot, pu, nb, s);
    SIL_REQUIRE(op->kind == OP_POISSON2D, SIL_ERR_UNSUPPORTED, "operator kind %d has no linear application", op->kind);
    return p2d_apply64(op->p2, in, out, dot, pu, nb, s);
}

int apply32(sil_operator op, const float* in, float* out, DotEpilogue dot, const PUpdate<float>* pu, int* nb,
            hipStream_t s)
{
    if (op->kind == OP_MAGNETOSTATICS3D) return ms3d_apply32(op->ms, in, out, dot, pu, nb, s);
    if (op->kind == OP_MS3D_JACOBIAN) return ms3d_jacobian_apply32(op->ms, op->nl, in, out, dot, pu, nb, s);
    if (op->kind == OP_MAGNETOSTATICS2D) return ms2d_apply32(op->m2, in, out, dot, pu, nb, s);
    if (op->kind == OP_MS2D_JACOBIAN) return ms2d_jacobian_apply32(op->m2, op->nl, in, out, dot, pu, nb, s);
    SIL_REQUIRE(op->kind != OP_CSR, SIL_ERR_UNSUPPORTED, "assembled operators have FP64 values only");
    SIL_REQUIRE(op->kind == OP_POISSON2D, SIL_ERR_UNSUPPORTED, "operator kind %d has no FP32 application", op->kind);
    return p2d_apply32(op->p2, in, out, dot, pu, nb, s);
}

int diagonal(sil_operator op, double* d, hipStream_t s)
{
    if (op->kind == OP_MAGNETOSTATICS3D) return ms3d_diagonal(op->ms, d, s);
    if (op->kind == OP_CSR) return csr_diagonal(op->csr, d, s);
    if (op->kind == OP_MS3D_JACOBIAN) return ms3d_jacobian_diagonal(op->ms, op->nl, d, s);
    if (op->kind == OP_MAGNETOSTATICS2D) return ms2d_diagonal(op->m2, d, s);
    if (op->kind == OP_MS2D_JACOBIAN) return ms2d_jacobian_diagonal(op->m2, op->nl, d, s);
    SIL_REQUIRE(op->kind == OP_POISSON2D, SIL_ERR_UNSUPPORTED, "operator kind %d has no diagonal", op->kind);
    return p2d_diagonal(op->p2, d, s);
}

// z = Gᵀ (G r) with per-block ⟨r, z⟩ partials; both products return at once
// when *done is set
int isai_apply(Workspace& w, const double* r, double* part, int* nb, const int* done, hipStream_t s)
{
    int rc = csr_spmv(w.isai_g, r, w.isai_t, nullptr, nullptr, done, nullptr, s);
    if (rc) return rc;
    return csr_spmv(w.isai_gt, w.isai_t, w.isai_z, r, part, done, nb, s);
}

template <typename P>
int alloc(P** ptr, int64_t n)
{
    SIL_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(ptr), sizeof(P) * n));
    return SIL_OK;
}

int ensure_workspace(sil_operator op, bool mixed, hipStream_t s)
{
    if (!op->ws) {
        op->ws = new (std::nothrow) Workspace();
        SIL_REQUIRE(op->ws != nullptr, SIL_ERR_INVALID, "solver: out of host memory");
    }
    Workspace& w = *op->ws;
    const int64_t n = op->n;
    int rc;
    if (!w.x) {
        w.n = n;
        if ((rc = alloc(&w.x, n)) || (rc = alloc(&w.b, n)) || (rc = alloc(&w.tmp, n)) ||
            (rc = alloc(&w.partials, 3 * max_dot_blocks())) || (rc = alloc(&w.sc, 1)))
            return rc;
        w.part_stride = max_dot_blocks();
        SIL_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&w.host_sc), sizeof(CgScalars)));
        // number of blocks of one operator application
        int nb = 0;
        if (op->kind == OP_MAGNETOSTATICS3D || op->kind == OP_MS3D_JACOBIAN)
            nb = static_cast<int>((op->ms.nx * op->ms.ny * op->ms.nz + 255) / 256);
        else if (op->kind == OP_POISSON2D)
            nb = static_cast<int>((op->p2.nx * op->p2.ny + 255) / 256);
        else if (op->kind == OP_MAGNETOSTATICS2D || op->kind == OP_MS2D_JACOBIAN)
            nb = static_cast<int>((op->m2.nx * op->m2.ny + 255) / 256);
        w.apply_blocks = nb;
        (void)s;
    }
    if (!mixed && !w.r) {
        if ((rc = alloc(&w.r, n)) || (rc = alloc(&w.q, n)) || (rc = alloc(&w.p[0], n)) || (rc = alloc(&w.p[1], n)))
            return rc;
    }
    if (mixed && !w.r32) {
        if ((rc = alloc(&w.r32, n)) || (rc = alloc(&w.q32, n)) || (rc = alloc(&w.p32[0], n)) ||
            (rc = alloc(&w.p32[1], n)) || (rc = alloc(&w.xc32, n)))
            return rc;
    }
    return SIL_OK;
}

enum Precond { PC_IDENTITY = 0, PC_JACOBI = 1, PC_ISAI = 2 };

// enqueue one CG iteration; pold / pnew are the p ping-pong buffers.  With
// SpdIsai the direction is formed from z (not from r) and the update's
// ⟨r, z⟩ partials are replaced by those of the preconditioner application.
template <typename T>
int enqueue_iteration(sil_operator op, Workspace& w, T* r, T* q, T* pold, T* pnew, int pc, hipStream_t s)
{
    double* part0 = w.partials;
    double* part1 = w.partials + w.part_stride;
    double* part2 = w.partials + 2 * w.part_stride;
    const T* invd = nullptr;
    if (pc == PC_JACOBI) {
        if constexpr (sizeof(T) == 8)
            invd = w.inv_diag;
        else
            invd = w.inv_diag32;
    }
    const T* zsrc = r;
    if constexpr (sizeof(T) == 8)
        if (pc == PC_ISAI) zsrc = w.isai_z;
    if constexpr (sizeof(T) == 4) {
        if (w.single_pass) {
            // pold holds the current direction p, pnew receives the next one
            int nb = 0;
            const NlData* nl = op->kind == OP_MS3D_JACOBIAN ? &op->nl : nullptr;
            int rc = ms3d_qdots32(op->ms, nl, pold, q, r, invd, part0, w.part_stride, &w.sc->done, &nb, s);
            if (rc) return rc;
            hipLaunchKernelGGL(k_alpha_expand, dim3(1), dim3(kRedThreads), 0, s, w.sc, part0, w.part_stride, nb);
            hipLaunchKernelGGL(k_update_dir, dim3(kUpdBlocks), dim3(kUpdBlock), 0, s, w.sc, pold, pnew, r, q, invd,
                               op->n, part1, part2, w.xc32);
            hipLaunchKernelGGL(k_gamma, dim3(1), dim3(kRedThreads), 0, s, w.sc, part1, kUpdBlocks, part2, kUpdBlocks);
            SIL_LAUNCH_CHECK();
            return SIL_OK;
        }
    }
    PUpdate<T> pu{zsrc, pold, pnew, &w.sc->beta, invd, &w.sc->done};
    int nb = 0;
    int rc;
    if constexpr (sizeof(T) == 8)
        rc = apply64(op, nullptr, q, DotEpilogue{part0}, &pu, &nb, s);
    else
        rc = apply32(op, nullptr, q, DotEpilogue{part0}, &pu, &nb, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_alpha, dim3(1), dim3(kRedThreads), 0, s, w.sc, part0, nb);
    if constexpr (sizeof(T) == 4)
        hipLaunchKernelGGL((k_update<T, true>), dim3(kUpdBlocks), dim3(kUpdBlock), 0, s, w.sc, w.x, pnew, r, q, invd,
                           op->n, part1, part2, w.xc32);
    else
        hipLaunchKernelGGL((k_update<T, false>), dim3(kUpdBlocks), dim3(kUpdBlock), 0, s, w.sc, w.x, pnew, r, q, invd,
                           op->n, part1, part2, nullptr);
    int nb_rz = kUpdBlocks;
    if constexpr (sizeof(T) == 8)
        if (pc == PC_ISAI && (rc = isai_apply(w, r, part1, &nb_rz, &w.sc->done, s))) return rc;
    hipLaunchKernelGGL(k_beta, dim3(1), dim3(kRedThreads), 0, s, w.sc, part1, nb_rz, part2, kUpdBlocks);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

template <typename T>
int get_graph(sil_operator op, int prec, int pc, int chunk, T* r, T* q, T* const* p, hipStream_t s,
              hipGraphExec_t* out)
{
    Workspace& w = *op->ws;
    hipGraphExec_t& g = w.graph[prec][pc];
    if (g && w.graph_chunk[prec][pc] == chunk) {
        *out = g;
        return SIL_OK;
    }
    if (g) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    hipStream_t cs;
    SIL_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t graph;
    SIL_CHECK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    int rc = SIL_OK;
    for (int it = 0; it < chunk && rc == SIL_OK; ++it)
        rc = enqueue_iteration<T>(op, w, r, q, p[it & 1], p[(it + 1) & 1], pc, cs);
    hipError_t e = hipStreamEndCapture(cs, &graph);
    if (rc != SIL_OK) {
        (void)hipStreamDestroy(cs);
        return rc;
    }
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cs);
        set_error("graph capture failed: %s", hipGetErrorString(e));
        return SIL_ERR_HIP;
    }
    e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) {
        set_error("graph instantiation failed: %s", hipGetErrorString(e));
        return SIL_ERR_HIP;
    }
    w.graph_chunk[prec][pc] = chunk;
    (void)s;
    *out = g;
    return SIL_OK;
}

int read_scalars(Workspace& w, hipStream_t s)
{
    SIL_CHECK_HIP(hipMemcpyAsync(w.host_sc, w.sc, sizeof(CgScalars), hipMemcpyDeviceToHost, s));
    SIL_CHECK_HIP(hipStreamSynchronize(s));
    return SIL_OK;
}

int true_residual(sil_operator op, Workspace& w, hipStream_t s)
{
    int rc = apply64(op, w.x, w.tmp, DotEpilogue{nullptr}, nullptr, nullptr, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_residual_f64, dim3(kUpdBlocks), dim3(kUpdBlock), 0, s, w.b, w.tmp, op->n, w.partials);
    hipLaunchKernelGGL(k_store_true, dim3(1), dim3(kRedThreads), 0, s, w.sc, w.partials, kUpdBlocks);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

}  // namespace

// the operator's diagonal changed (a Jacobian at a new state): Jacobi is
// recomputed by the next solve
void invalidate_jacobi(sil_operator op)
{
    if (op && op->ws) op->ws->jacobi_ready = false;
}

// the operator changed (a Jacobian at a new state): SpdIsai is regenerated
// by the next solve that asks for it
void invalidate_isai(sil_operator op)
{
    if (op && op->ws) op->ws->isai_ready = false;
}

void destroy_workspace(Workspace* w)
{
    if (!w) return;
    for (double* p : {w->x, w->b, w->r, w->q, w->p[0], w->p[1], w->inv_diag, w->tmp, w->partials, w->isai_t,
                      w->isai_z, w->fb_r, w->fb_d})
        if (p) (void)hipFree(p);
    csr_free(w->isai_g);
    csr_free(w->isai_gt);
    for (float* p : {w->r32, w->q32, w->p32[0], w->p32[1], w->inv_diag32, w->xc32})
        if (p) (void)hipFree(p);
    if (w->sc) (void)hipFree(w->sc);
    if (w->host_sc) (void)hipHostFree(w->host_sc);
    for (auto& row : w->graph)
        for (auto& g : row)
            if (g) (void)hipGraphExecDestroy(g);
    delete w;
}

int cg_solve(sil_operator op, const double* rhs, double* solution, const sil_solver_settings& st,
             sil_solver_diagnostics& dg, hipStream_t s)
{
    const bool mixed = st.precision == SIL_PRECISION_MIXED;
    const bool jacobi = st.preconditioner == SIL_PRECOND_JACOBI;
    const int pc = jacobi ? PC_JACOBI : (st.preconditioner == SIL_PRECOND_SPD_ISAI ? PC_ISAI : PC_IDENTITY);
    SIL_REQUIRE(!(mixed && op->kind == OP_CSR), SIL_ERR_UNSUPPORTED,
                "mixed precision needs a matrix-free operator (assembled values are FP64)");
    SIL_REQUIRE(!(mixed && pc == PC_ISAI), SIL_ERR_UNSUPPORTED, "the SpdIsai preconditioner runs in FP64 only");
    int rc = ensure_workspace(op, mixed, s);
    if (rc) return rc;
    Workspace& w = *op->ws;
    const int64_t n = op->n;
    const auto t0 = std::chrono::steady_clock::now();

    if (pc == PC_ISAI && !w.isai_ready) {
        // the reference assembles the matrix to generate any preconditioner
        // other than the identity, also when it then applies the operator
        // matrix-free (minimize_strong_formulation_residual.hpp:1332-1353)
        CsrData assembled{};
        const CsrData* a = &op->csr;
        SIL_REQUIRE(op->kind == OP_CSR || op->kind == OP_MAGNETOSTATICS3D || op->kind == OP_MS3D_JACOBIAN ||
                        op->kind == OP_MAGNETOSTATICS2D || op->kind == OP_MS2D_JACOBIAN,
                    SIL_ERR_UNSUPPORTED,
                    "SpdIsai needs an assembled operator or a magnetostatics operator (linear or its Jacobian)");
        if (op->kind == OP_MAGNETOSTATICS3D) {
            if ((rc = csr_assemble_ms3d(op->ms, assembled, s))) return rc;
            a = &assembled;
        } else if (op->kind == OP_MS3D_JACOBIAN) {
            // the Jacobian at the state of the last evaluation, assembled at
            // every Newton step (minimize_strong_formulation_residual.hpp:1432-1433)
            if ((rc = csr_assemble_ms3d_jacobian(op->ms, op->nl, assembled, s))) return rc;
            a = &assembled;
        } else if (op->kind == OP_MAGNETOSTATICS2D) {
            if ((rc = csr_assemble_ms2d(op->m2, assembled, s))) return rc;
            a = &assembled;
        } else if (op->kind == OP_MS2D_JACOBIAN) {
            if ((rc = csr_assemble_ms2d_jacobian(op->m2, op->nl, assembled, s))) return rc;
            a = &assembled;
        }
        // graphs captured with the previous factor hold its (freed) arrays
        for (auto& row : w.graph)
            if (row[PC_ISAI]) {
                (void)hipGraphExecDestroy(row[PC_ISAI]);
                row[PC_ISAI] = nullptr;
            }
        csr_free(w.isai_g);
        csr_free(w.isai_gt);
        rc = spd_isai_generate(*a, w.isai_g, w.isai_gt, s);
        csr_free(assembled);
        if (rc) return rc;
        if (!w.isai_t && ((rc = alloc(&w.isai_t, n)) || (rc = alloc(&w.isai_z, n)))) return rc;
        w.isai_ready = true;
    }

    if (jacobi && !w.inv_diag) {
        if ((rc = alloc(&w.inv_diag, n))) return rc;
        w.jacobi_ready = false;
    }
    if (jacobi && mixed && !w.inv_diag32) {
        if ((rc = alloc(&w.inv_diag32, n))) return rc;
        w.jacobi_ready = false;
    }
    if (jacobi && !w.jacobi_ready) {
        if ((rc = diagonal(op, w.tmp, s))) return rc;
        hipLaunchKernelGGL(k_inv, dim3(kUpdBlocks), dim3(kUpdBlock), 0, s, w.tmp, w.inv_diag, n);
        if (w.inv_diag32)
            hipLaunchKernelGGL(k_to_f32, dim3(kUpdBlocks), dim3(kUpdBlock), 0, s, w.inv_diag, w.inv_diag32, n);
        SIL_LAUNCH_CHECK();
        w.jacobi_ready = true;
    }
    SIL_CHECK_HIP(hipMemcpyAsync(w.b, rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    double* part0 = w.partials;
    double* part1 = w.partials + w.part_stride;
    const double* invd = jacobi ? w.inv_diag : nullptr;
    const float* invd32 = jacobi ? w.inv_diag32 : nullptr;
    static const bool single_pass_enabled = [] {
        const char* e = std::getenv("SIL_CG_SINGLE_PASS");
        return !(e && e[0] == '0');
    }();
    w.single_pass = mixed && single_pass_enabled && pc != PC_ISAI &&
                    (op->kind == OP_MAGNETOSTATICS3D || op->kind == OP_MS3D_JACOBIAN) && ms3d_single_pass_ok(op->ms);
    if (mixed)
        hipLaunchKernelGGL((k_init<float>), dim3(kUpdBlocks), dim3(kUpdBlock), 0, s, w.b, w.x, w.r32, w.p32[0],
                           invd32, n, part0, part1, w.xc32, w.single_pass);
    else
        hipLaunchKernelGGL((k_init<double>), dim3(kUpdBlocks), dim3(kUpdBlock), 0, s, w.b, w.x, w.r, w.p[0], invd, n,
                           part0, part1, nullptr);
    int nb_rz = kUpdBlocks;
    if (pc == PC_ISAI && (rc = isai_apply(w, w.r, part1, &nb_rz, nullptr, s))) return rc;
    hipLaunchKernelGGL(k_scalars_init, dim3(1), dim3(kRedThreads), 0, s, w.sc, part0, kUpdBlocks, part1, nb_rz,
                       st.relative_tolerance, st.max_iterations);
    SIL_LAUNCH_CHECK();

    int chunk = st.check_interval > 0 ? st.check_interval : (n > (1 << 20) ? 16 : 32);
    chunk += chunk & 1; // even: the p ping-pong returns to buffer 0
    const int replace = st.replace_interval > 0 ? st.replace_interval : 200;
    hipGraphExec_t g = nullptr;
    if (mixed)
        rc = get_graph<float>(op, 1, pc, chunk, w.r32, w.q32, w.p32, s, &g);
    else
        rc = get_graph<double>(op, 0, pc, chunk, w.r, w.q, w.p, s, &g);
    if (rc) return rc;

    if ((rc = read_scalars(w, s))) return rc;
    dg.initial_residual_l2 = std::sqrt(w.host_sc->bb);
    int since_replace = 0;
    double rr_ref = w.host_sc->bb;
    // a mixed solve whose FP64-verified residual falls by less than 5 % over
    // five replacement intervals in a row has reached what the FP32 products
    // can resolve; it continues in FP64 from its iterate
    double rr_prev = w.host_sc->bb;
    int stalls = 0;
    bool stalled = false;
    while (!w.host_sc->done) {
        const int iter_before = w.host_sc->iter;
        SIL_CHECK_HIP(hipGraphLaunch(g, s));
        since_replace += chunk;
        if ((rc = read_scalars(w, s))) return rc;
        if (!mixed) continue;
        const CgScalars& h = *w.host_sc;
        // iteration `it` of a chunk writes its direction into p32[(it + 1) & 1]:
        // after an odd number of executed iterations the latest one sits in p32[1]
        const int executed = h.iter - iter_before;
        const int cur = executed & 1;
        const bool hit_max = h.iter >= h.max_iter;
        const bool want = h.converged || since_replace >= replace || h.rr < 1e-4 * rr_ref;
        if (h.done && cur && !(w.single_pass && want && !hit_max))
            SIL_CHECK_HIP(hipMemcpyAsync(w.p32[0], w.p32[1], sizeof(float) * n, hipMemcpyDeviceToDevice, s));
        if (want && !hit_max) {
            // FP64 residual replacement; keeps p (the direction buffer in use is p32[0])
            hipLaunchKernelGGL(k_flush, dim3(kUpdBlocks), dim3(kUpdBlock), 0, s, w.x, w.xc32, n);
            if ((rc = apply64(op, w.x, w.tmp, DotEpilogue{nullptr}, nullptr, nullptr, s))) return rc;
            hipLaunchKernelGGL((k_replace<float>), dim3(kUpdBlocks), dim3(kUpdBlock), 0, s, w.b, w.tmp, w.r32, invd32,
                               n, part0, part1);
            hipLaunchKernelGGL(k_after_replace, dim3(1), dim3(kRedThreads), 0, s, w.sc, part0, part1, kUpdBlocks);
            if (w.single_pass) {
                // the next direction was formed from the old residual: form it
                // again from the replaced one, z + β p with β = γ_new / γ
                PUpdate<float> pu{w.r32, w.p32[cur ^ 1], w.p32[cur], &w.sc->beta, invd32, &w.sc->done};
                if ((rc = direction_pass32(pu, n, s))) return rc;
                if (cur)
                    SIL_CHECK_HIP(
                        hipMemcpyAsync(w.p32[0], w.p32[1], sizeof(float) * n, hipMemcpyDeviceToDevice, s));
            }
            SIL_LAUNCH_CHECK();
            if ((rc = read_scalars(w, s))) return rc;
            since_replace = 0;
            rr_ref = w.host_sc->rr_true;
            if (!w.host_sc->converged) {
                stalls = rr_ref > 0.9025 * rr_prev ? stalls + 1 : 0;
                rr_prev = rr_ref;
                if (stalls >= 5) {
                    stalled = true;
                    break;
                }
            }
        } else if (h.converged && hit_max) {
            break;
        }
    }
    // the FP32 correction since the last replacement joins the iterate
    if (mixed) hipLaunchKernelGGL(k_flush, dim3(kUpdBlocks), dim3(kUpdBlock), 0, s, w.x, w.xc32, n);
    int extra_iterations = 0;
    if (stalled) {
        // FP64 CG on the correction A d = b − A x, then x + d
        const int mixed_iterations = w.host_sc->iter;
        const double bb = w.host_sc->bb, rr = w.host_sc->rr_true;
        if (!w.fb_r && ((rc = alloc(&w.fb_r, n)) || (rc = alloc(&w.fb_d, n)))) return rc;
        SIL_CHECK_HIP(hipMemcpyAsync(solution, w.x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        if ((rc = apply64(op, solution, w.tmp, DotEpilogue{nullptr}, nullptr, nullptr, s))) return rc;
        hipLaunchKernelGGL(k_sub, dim3(kUpdBlocks), dim3(kUpdBlock), 0, s, rhs, w.tmp, w.fb_r, n);
        SIL_LAUNCH_CHECK();
        sil_solver_settings fp64 = st;
        fp64.precision = SIL_PRECISION_FP64;
        fp64.max_iterations = std::max(1, st.max_iterations - mixed_iterations);
        // half the target: the FP64 recurrence residual drifts a little from the
        // true one, and the solve is judged on the true residual
        fp64.relative_tolerance = rr > 0.0 ? std::min(1.0, 0.5 * st.relative_tolerance * std::sqrt(bb / rr)) : 1.0;
        sil_solver_diagnostics dd{};
        if ((rc = cg_solve(op, w.fb_r, w.fb_d, fp64, dd, s))) return rc;
        hipLaunchKernelGGL(k_add, dim3(kUpdBlocks), dim3(kUpdBlock), 0, s, solution, w.fb_d, n);
        SIL_LAUNCH_CHECK();
        // the solve's own vectors hold the correction problem now: restore
        // b and x for the final residual
        SIL_CHECK_HIP(hipMemcpyAsync(w.b, rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        SIL_CHECK_HIP(hipMemcpyAsync(w.x, solution, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        if ((rc = read_scalars(w, s))) return rc;
        w.host_sc->bb = bb;
        w.host_sc->iter = mixed_iterations;
        SIL_CHECK_HIP(hipMemcpyAsync(w.sc, w.host_sc, sizeof(CgScalars), hipMemcpyHostToDevice, s));
        extra_iterations = dd.iterations;
    }
    if ((rc = true_residual(op, w, s))) return rc;
    if ((rc = read_scalars(w, s))) return rc;
    SIL_CHECK_HIP(hipMemcpyAsync(solution, w.x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    SIL_CHECK_HIP(hipStreamSynchronize(s));
    const CgScalars& h = *w.host_sc;
    dg.iterations = h.iter + extra_iterations;
    const double bnorm = std::sqrt(h.bb);
    dg.final_residual_l2 = std::sqrt(mixed ? h.rr_true : h.rr);
    dg.final_relative_residual = bnorm == 0.0 ? 0.0 : dg.final_residual_l2 / bnorm;
    dg.true_relative_residual = bnorm == 0.0 ? 0.0 : std::sqrt(h.rr_true) / bnorm;
    dg.converged = (bnorm == 0.0) ? 1 : (dg.true_relative_residual <= st.relative_tolerance * (1.0 + 1e-6) ? 1 : 0);
    if (!mixed) dg.converged = (bnorm == 0.0) ? 1 : h.converged;
    // the recurrence residual stops FP64 CG (as the reference's Ginkgo
    // criterion does); whether the recomputed residual meets the tolerance
    // too is reported on its own
    dg.verified = (bnorm == 0.0 || dg.true_relative_residual <= st.relative_tolerance * (1.0 + 1e-6)) ? 1 : 0;
    dg.duration = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return SIL_OK;
}

int operator_apply64(sil_operator op, const double* in, double* out, hipStream_t s)
{
    return apply64(op, in, out, DotEpilogue{nullptr}, nullptr, nullptr, s);
}

int operator_apply32(sil_operator op, const float* in, float* out, hipStream_t s)
{
    return apply32(op, in, out, DotEpilogue{nullptr}, nullptr, nullptr, s);
}

int operator_diagonal(sil_operator op, double* d, hipStream_t s) { return diagonal(op, d, s); }

// ------------------------------------------------ solve without allocation ---
//
// FP64 CG (identity or Jacobi) for callers that cannot allocate device
// memory — code under stream capture, serving loops, harnesses that forbid
// hipMalloc.  Every vector and partial sum the solve needs lives in one
// file-scope device array, `scratch`; its address on a device is looked up
// once per device (hipGetSymbolAddress, that device current) and only that
// address is handed to kernels.  The solution array is the iterate and the
// right-hand side is read in place; no graph is instantiated (iterations are
// launched in chunks on the caller's stream).  Every region a solve reads is
// written by that solve first — k_init writes r, p_old and x, the scalar
// kernels the scalars, each partial array is written before it is reduced —
// so nothing is cleared between calls.  All calls on one device share the
// array: calls that overlap on one device race.
constexpr int64_t kNoAllocMaxDofs = int64_t(1) << 19;
// partial-sum slots per array: one per workgroup of an application (a march
// workgroup covers at least one node, the other kernels 256 rows) or of the
// update kernel
constexpr int64_t kNoAllocSlots = std::max<int64_t>(kUpdBlocks, kNoAllocMaxDofs / 3 + 64);
constexpr int64_t kNoAllocVecBytes = kNoAllocMaxDofs * 8;
constexpr int kNoAllocVectors = 6; // r, q, p[0], p[1], D⁻¹, tmp
constexpr int64_t kNoAllocPartsOffset = kNoAllocVectors * kNoAllocVecBytes;
constexpr int64_t kNoAllocScalarsOffset = kNoAllocPartsOffset + 3 * kNoAllocSlots * 8;
constexpr int64_t kNoAllocBytes = kNoAllocScalarsOffset + ((int64_t(sizeof(CgScalars)) + 255) / 256) * 256;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch[kMaxDevices] = {};

// address of `scratch` on the current device (looked up on its first use)
int scratch_on_current_device(unsigned char** out)
{
    int dev = 0;
    SIL_CHECK_HIP(hipGetDevice(&dev));
    SIL_REQUIRE(dev >= 0 && dev < kMaxDevices, SIL_ERR_UNSUPPORTED, "device %d outside 0..%d", dev, kMaxDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            set_error("scratch array address on device %d: %s", dev, hipGetErrorString(e));
            return SIL_ERR_HIP;
        }
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[dev];
    return SIL_OK;
}
}  // namespace

int64_t noalloc_max_dofs() { return kNoAllocMaxDofs; }

int cg_solve_noalloc(sil_operator op, const double* rhs, double* solution, const sil_solver_settings& st,
                     sil_solver_diagnostics& dg, hipStream_t s)
{
    SIL_REQUIRE(st.precision == SIL_PRECISION_FP64, SIL_ERR_UNSUPPORTED,
                "solve without allocation: FP64 only (mixed precision needs FP32 vectors)");
    SIL_REQUIRE(st.preconditioner == SIL_PRECOND_IDENTITY || st.preconditioner == SIL_PRECOND_JACOBI,
                SIL_ERR_UNSUPPORTED, "solve without allocation: identity or Jacobi preconditioning");
    SIL_REQUIRE(op->kind == OP_POISSON2D || op->kind == OP_MAGNETOSTATICS3D || op->kind == OP_MAGNETOSTATICS2D,
                SIL_ERR_UNSUPPORTED, "solve without allocation: linear matrix-free operators (Poisson 2D, "
                                     "magnetostatics 2D / 3D) only");
    const int64_t n = op->n;
    SIL_REQUIRE(n <= kNoAllocMaxDofs, SIL_ERR_UNSUPPORTED,
                "solve without allocation: %lld unknowns exceed the %lld the scratch array holds", (long long)n,
                (long long)kNoAllocMaxDofs);
    unsigned char* base = nullptr;
    int rc = scratch_on_current_device(&base);
    if (rc) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    const int pc = st.preconditioner == SIL_PRECOND_JACOBI ? PC_JACOBI : PC_IDENTITY;
    Workspace w{};
    w.n = n;
    w.x = solution;
    w.b = const_cast<double*>(rhs); // read only
    auto vec = [&](int i) { return reinterpret_cast<double*>(base + i * kNoAllocVecBytes); };
    w.r = vec(0);
    w.q = vec(1);
    w.p[0] = vec(2);
    w.p[1] = vec(3);
    w.inv_diag = vec(4);
    w.tmp = vec(5);
    w.partials = reinterpret_cast<double*>(base + kNoAllocPartsOffset);
    w.part_stride = kNoAllocSlots;
    w.sc = reinterpret_cast<CgScalars*>(base + kNoAllocScalarsOffset);
    CgScalars host{}; // pageable copy of the device scalars
    w.host_sc = &host;

    const double* invd = nullptr;
    if (pc == PC_JACOBI) {
        if ((rc = diagonal(op, w.tmp, s))) return rc;
        hipLaunchKernelGGL(k_inv, dim3(kUpdBlocks), dim3(kUpdBlock), 0, s, w.tmp, w.inv_diag, n);
        SIL_LAUNCH_CHECK();
        invd = w.inv_diag;
    }
    double* part0 = w.partials;
    double* part1 = w.partials + w.part_stride;
    hipLaunchKernelGGL((k_init<double>), dim3(kUpdBlocks), dim3(kUpdBlock), 0, s, w.b, w.x, w.r, w.p[0], invd, n,
                       part0, part1, nullptr);
    hipLaunchKernelGGL(k_scalars_init, dim3(1), dim3(kRedThreads), 0, s, w.sc, part0, kUpdBlocks, part1, kUpdBlocks,
                       st.relative_tolerance, st.max_iterations);
    SIL_LAUNCH_CHECK();
    int chunk = st.check_interval > 0 ? st.check_interval : (n > (1 << 20) ? 16 : 32);
    chunk += chunk & 1; // even: the p ping-pong returns to buffer 0
    if ((rc = read_scalars(w, s))) return rc;
    dg.initial_residual_l2 = std::sqrt(host.bb);
    while (!host.done) {
        for (int it = 0; it < chunk; ++it)
            if ((rc = enqueue_iteration<double>(op, w, w.r, w.q, w.p[it & 1], w.p[(it + 1) & 1], pc, s))) return rc;
        if ((rc = read_scalars(w, s))) return rc;
    }
    if ((rc = true_residual(op, w, s))) return rc;
    if ((rc = read_scalars(w, s))) return rc;
    dg.iterations = host.iter;
    const double bnorm = std::sqrt(host.bb);
    dg.final_residual_l2 = std::sqrt(host.rr);
    dg.final_relative_residual = bnorm == 0.0 ? 0.0 : dg.final_residual_l2 / bnorm;
    dg.true_relative_residual = bnorm == 0.0 ? 0.0 : std::sqrt(host.rr_true) / bnorm;
    dg.converged = bnorm == 0.0 ? 1 : host.converged;
    dg.verified = (bnorm == 0.0 || dg.true_relative_residual <= st.relative_tolerance * (1.0 + 1e-6)) ? 1 : 0;
    dg.outer_iterations = 0;
    dg.duration = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return SIL_OK;
}

}  // namespace sil
