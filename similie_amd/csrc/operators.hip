// Matrix-free operators for the CG solve.
//
// Magnetostatics 3D (onelab_interface/magnetostatics_onelab.hpp:2241-2396):
// for an interior DOF r = (n, c)
//   (A x)_r = Σ_s Σ_m C_m(s; r) · w^B_m(s) · ν(s) · B_m(s) + P · w^A_c(n) · x_r
// with B = (dA)_yz, −(dA)_xz, (dA)_xy at interior nodes s (boundary DOFs
// dropped), w^B_x = ℓx/(ℓy ℓz), w^A_x = ℓy ℓz/ℓx (cyclic), ℓ = local_spacing
// (central, one-sided at the ends, :856-933); boundary rows are the identity.
// The reference gathers through ≤24 precomputed transposed-moment slots per
// row; here each thread evaluates the curl of its node's neighbourhood and the
// transposed curl directly: out = curlᵀ(W ν curl x).
//
// Poisson 2D: d0ᵀ ⋆1^κ d0 on interior nodes with zero Dirichlet values.
//
// Both come in FP64 and FP32 and can (a) emit per-block partial sums of
// ⟨in, out⟩ and (b) form the CG direction p = z + β p_old on the fly.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdlib>

#include "apply_epilogue.hpp"
#include "operators.hpp"

namespace sil {

namespace {

constexpr int kBlock = kApplyBlock;
constexpr int64_t kMaxDotBlocks = 1 << 18;

// --------------------------------------------------------- magnetostatics ---

template <typename T>
struct MsArgs {
    int64_t nx, ny, nz;
    T penalty;
    const T* ls[3];
    const T* inv_ls[3];
    const T* nu;
    // nonlinear material state: a, c and B0 per node
    const T* st_a;
    const T* st_c;
    const T* st_b[3];
    const NodeState<T>* st4; // (a, c, B0_x, B0_y) per node
};

// material response of the curl-curl march: ν·B (linear), a·B (the
// nonlinear operator at the state's own x) or a·δB + c·B0·(B0·δB) (its
// Jacobian at the state)
enum MsResponse { RESP_LINEAR = 0, RESP_SCALED = 1, RESP_JACOBIAN = 2 };

// p_new = z + β p_old with z = D⁻¹ r (Jacobi) or r: the CG direction, in
// the arithmetic of p_value
template <typename T>
__global__ __launch_bounds__(kBlock) void k_direction(PUpdate<T> pu, int64_t n)
{
    if (*pu.done) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        pu.p_new[i] = p_value<T>(&pu, nullptr, i);
}

template <typename T, bool PU>
__global__ __launch_bounds__(kBlock) void k_ms3d_apply(MsArgs<T> a, const T* __restrict__ x, T* __restrict__ out,
                                                       double* __restrict__ partials, PUpdate<T> pu)
{
    if (pu.done && *pu.done) return;
    const int64_t nn = a.nx * a.ny * a.nz;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double dot = 0.0;
    if (t < nn) {
        const int64_t i = t % a.nx, j = (t / a.nx) % a.ny, k = t / (a.nx * a.ny);
        const int64_t sx = 1, sy = a.nx, sz = a.nx * a.ny;
        auto interior = [&](int64_t ii, int64_t jj, int64_t kk) {
            return ii > 0 && jj > 0 && kk > 0 && ii + 1 < a.nx && jj + 1 < a.ny && kk + 1 < a.nz;
        };
        const PUpdate<T>* pp = PU ? &pu : nullptr;
        // potential component c at node (ii,jj,kk); boundary DOFs are dropped
        auto A = [&](int64_t ii, int64_t jj, int64_t kk, int c) -> T {
            if (!interior(ii, jj, kk)) return T(0);
            return p_value<T>(pp, x, 3 * (ii + a.nx * (jj + a.ny * kk)) + c);
        };
        // H_m(s) = w^B_m(s) ν(s) B_m(s) at interior s, else 0
        auto Hx = [&](int64_t ii, int64_t jj, int64_t kk) -> T {
            if (!interior(ii, jj, kk)) return T(0);
            const T B = A(ii, jj + 1, kk, 2) - A(ii, jj, kk, 2) - A(ii, jj, kk + 1, 1) + A(ii, jj, kk, 1);
            const T w = a.ls[0][ii] * a.inv_ls[1][jj] * a.inv_ls[2][kk];
            return w * a.nu[ii + a.nx * (jj + a.ny * kk)] * B;
        };
        auto Hy = [&](int64_t ii, int64_t jj, int64_t kk) -> T {
            if (!interior(ii, jj, kk)) return T(0);
            const T B = -A(ii + 1, jj, kk, 2) + A(ii, jj, kk, 2) + A(ii, jj, kk + 1, 0) - A(ii, jj, kk, 0);
            const T w = a.ls[1][jj] * a.inv_ls[0][ii] * a.inv_ls[2][kk];
            return w * a.nu[ii + a.nx * (jj + a.ny * kk)] * B;
        };
        auto Hz = [&](int64_t ii, int64_t jj, int64_t kk) -> T {
            if (!interior(ii, jj, kk)) return T(0);
            const T B = A(ii + 1, jj, kk, 1) - A(ii, jj, kk, 1) - A(ii, jj + 1, kk, 0) + A(ii, jj, kk, 0);
            const T w = a.ls[2][kk] * a.inv_ls[0][ii] * a.inv_ls[1][jj];
            return w * a.nu[ii + a.nx * (jj + a.ny * kk)] * B;
        };
        T o[3];
        T pv[3];
        for (int c = 0; c < 3; ++c) pv[c] = p_value<T>(pp, x, 3 * t + c);
        if (!interior(i, j, k)) {
            o[0] = pv[0];
            o[1] = pv[1];
            o[2] = pv[2];
        } else {
            const T hx0 = Hx(i, j, k), hy0 = Hy(i, j, k), hz0 = Hz(i, j, k);
            const T lx = a.ls[0][i], ly = a.ls[1][j], lz = a.ls[2][k];
            const T ilx = a.inv_ls[0][i], ily = a.inv_ls[1][j], ilz = a.inv_ls[2][k];
            o[0] = Hy(i, j, k - 1) - hy0 - Hz(i, j - 1, k) + hz0 + a.penalty * (ly * lz * ilx) * pv[0];
            o[1] = -Hx(i, j, k - 1) + hx0 + Hz(i - 1, j, k) - hz0 + a.penalty * (lx * lz * ily) * pv[1];
            o[2] = Hx(i, j - 1, k) - hx0 - Hy(i - 1, j, k) + hy0 + a.penalty * (lx * ly * ilz) * pv[2];
        }
        (void)sx;
        (void)sy;
        (void)sz;
        for (int c = 0; c < 3; ++c) {
            out[3 * t + c] = o[c];
            if (PU) pu.p_new[3 * t + c] = pv[c];
            dot += (double)pv[c] * (double)o[c];
        }
    }
    if (partials) block_dot_store<T>(dot, partials);
}

// The same operator by a march along z: a workgroup owns 64 (x, one
// wavefront) × kMsTY (y) nodes including a one-node halo and walks a chunk of
// kMsChunk z-planes.  Each node's input value is read once per tile (PU: the
// direction p = z + β p_old is formed instead and written to p_new by the
// node's owner); the x/y neighbours of A and of H = w^B ν B go through LDS,
// the z neighbours stay in registers, so the ~40 neighbour reads of the
// direct kernel become one staged read per node.  Boundary rows are the
// identity; boundary DOFs are dropped from every H.
constexpr int kMsTY = 8;
constexpr int kMsChunk = 32;

// NT: the product leaves through non-temporal stores
template <typename T, bool PU, int RESP = RESP_LINEAR, bool NT = false>
__global__ __launch_bounds__(64 * kMsTY) void k_ms3d_march(MsArgs<T> a, const T* __restrict__ x, T* __restrict__ out,
                                                           double* __restrict__ partials, PUpdate<T> pu, int ntx,
                                                           int nty)
{
    __shared__ T sA[3][kMsTY][64];
    __shared__ T sH[3][kMsTY][64];
    __shared__ double sdot[kMsTY];
    if (pu.done && *pu.done) return; // uniform
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    const int tx = blockIdx.x % ntx, ty = (blockIdx.x / ntx) % nty, tk = blockIdx.x / (ntx * nty);
    const int64_t nx = a.nx, ny = a.ny, nz = a.nz, plane = nx * ny;
    const int64_t i = (int64_t)tx * 62 - 1 + lane, j = (int64_t)ty * (kMsTY - 2) - 1 + row;
    const int64_t ka = (int64_t)tk * kMsChunk, kb = min(ka + (int64_t)kMsChunk, nz);
    const bool inx = i >= 0 && i < nx, iny = j >= 0 && j < ny;
    const bool writer = lane >= 1 && lane <= 62 && row >= 1 && row <= kMsTY - 2 && inx && iny;
    const bool in_ij = i > 0 && j > 0 && i + 1 < nx && j + 1 < ny;
    const int64_t ic = min(max(i, (int64_t)0), nx - 1), jc = min(max(j, (int64_t)0), ny - 1);
    const int64_t lin = ic + nx * jc;
    const T lx = a.ls[0][ic], ly = a.ls[1][jc], ilx = a.inv_ls[0][ic], ily = a.inv_ls[1][jc];
    const PUpdate<T>* pp = PU ? &pu : nullptr;

    // raw direction value (3 components) of this lane's node in plane k
    auto load = [&](int64_t k, T v[3]) {
        const int64_t idx = 3 * (lin + plane * min(max(k, (int64_t)0), nz - 1));
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = p_value<T>(pp, x, idx + c);
    };
    auto interior = [&](int64_t k) { return in_ij && k > 0 && k + 1 < nz; };

    T P0[3], P1[3], P2[3], Hp[3] = {T(0), T(0), T(0)};
    double dot = 0.0;
    int64_t k = ka > 0 ? ka - 1 : ka;
    load(k, P0);
    load(k + 1, P1);

    // H at plane k from A(k) (LDS for the +x / +y neighbours) and A(k+1);
    // leaves A(k) in sA and H(k) in sH
    auto h_plane = [&](T H[3]) {
        const bool ik = interior(k), ik1 = interior(k + 1);
        T A0[3], A1[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            A0[c] = ik ? P0[c] : T(0);
            A1[c] = ik1 ? P1[c] : T(0);
            sA[c][row][lane] = A0[c];
        }
        __syncthreads();
        const int lp = lane + 1 < 64 ? lane + 1 : lane, rp = row + 1 < kMsTY ? row + 1 : row;
        const T ax1 = sA[1][row][lp], ax2 = sA[2][row][lp]; // A(i+1, j, k)
        const T ay0 = sA[0][rp][lane], ay2 = sA[2][rp][lane]; // A(i, j+1, k)
        H[0] = H[1] = H[2] = T(0);
        if (ik) {
            const int64_t kk = k;
            const T lz = a.ls[2][kk], ilz = a.inv_ls[2][kk];
            const T bx = ay2 - A0[2] - A1[1] + A0[1];
            const T by = -ax2 + A0[2] + A1[0] - A0[0];
            const T bz = ax1 - A0[1] - ay0 + A0[0];
            const int64_t node = lin + plane * kk;
            if constexpr (RESP == RESP_LINEAR) {
                const T nu = a.nu[node];
                H[0] = (lx * ily * ilz) * nu * bx;
                H[1] = (ly * ilx * ilz) * nu * by;
                H[2] = (lz * ilx * ily) * nu * bz;
            } else if constexpr (RESP == RESP_SCALED) {
                const T sa = a.st_a[node];
                H[0] = (lx * ily * ilz) * (bx * sa);
                H[1] = (ly * ilx * ilz) * (by * sa);
                H[2] = (lz * ilx * ily) * (bz * sa);
            } else {
                const NodeState<T> q = a.st4[node];
                const T sa = q.a, sc = q.c, b0x = q.bx, b0y = q.by, b0z = a.st_b[2][node];
                const T cd = sc * (b0x * bx + b0y * by + b0z * bz);
                H[0] = (lx * ily * ilz) * (sa * bx + cd * b0x);
                H[1] = (ly * ilx * ilz) * (sa * by + cd * b0y);
                H[2] = (lz * ilx * ily) * (sa * bz + cd * b0z);
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) sH[c][row][lane] = H[c];
        __syncthreads();
    };

    if (ka > 0) {
        // plane ka−1: its H only
        h_plane(Hp);
        ++k;
#pragma unroll
        for (int c = 0; c < 3; ++c) P0[c] = P1[c];
        load(k + 1, P1);
    }
    for (; k < kb; ++k) {
        load(k + 2, P2); // next step's A(k+1)
        T H[3];
        h_plane(H);
        if (writer) {
            const int lm = lane - 1, rm = row - 1; // in the tile for writers
            T o[3];
            if (interior(k)) {
                const T lz = a.ls[2][k], ilz = a.inv_ls[2][k];
                o[0] = Hp[1] - H[1] - sH[2][rm][lane] + H[2] + a.penalty * (ly * lz * ilx) * P0[0];
                o[1] = -Hp[0] + H[0] + sH[2][row][lm] - H[2] + a.penalty * (lx * lz * ily) * P0[1];
                o[2] = sH[0][rm][lane] - H[0] - sH[1][row][lm] + H[1] + a.penalty * (lx * ly * ilz) * P0[2];
            } else {
                o[0] = P0[0];
                o[1] = P0[1];
                o[2] = P0[2];
            }
            const int64_t idx = 3 * (lin + plane * k);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if constexpr (NT)
                    __builtin_nontemporal_store(o[c], out + idx + c);
                else
                    out[idx + c] = o[c];
                if (PU) pu.p_new[idx + c] = P0[c];
                dot += (double)P0[c] * (double)o[c];
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Hp[c] = H[c];
            P0[c] = P1[c];
            P1[c] = P2[c];
        }
    }
    if (partials) {
        for (int off = 32; off > 0; off >>= 1) dot += __shfl_down(dot, off, 64);
        if (lane == 0) sdot[row] = dot;
        __syncthreads();
        if (threadIdx.x == 0) {
            double sum = 0.0;
            for (int r = 0; r < kMsTY; ++r) sum += sdot[r];
            partials[blockIdx.x] = sum;
        }
    }
}

__global__ void k_ms3d_diag(MsArgs<double> a, double* __restrict__ diag)
{
    const int64_t nn = a.nx * a.ny * a.nz;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nn) return;
    const int64_t i = t % a.nx, j = (t / a.nx) % a.ny, k = t / (a.nx * a.ny);
    auto interior = [&](int64_t ii, int64_t jj, int64_t kk) {
        return ii > 0 && jj > 0 && kk > 0 && ii + 1 < a.nx && jj + 1 < a.ny && kk + 1 < a.nz;
    };
    if (!interior(i, j, k)) {
        diag[3 * t] = diag[3 * t + 1] = diag[3 * t + 2] = 1.0;
        return;
    }
    auto W = [&](int m, int64_t ii, int64_t jj, int64_t kk) -> double {
        if (!interior(ii, jj, kk)) return 0.0;
        const double nu = a.nu[ii + a.nx * (jj + a.ny * kk)];
        if (m == 0) return a.ls[0][ii] * a.inv_ls[1][jj] * a.inv_ls[2][kk] * nu;
        if (m == 1) return a.ls[1][jj] * a.inv_ls[0][ii] * a.inv_ls[2][kk] * nu;
        return a.ls[2][kk] * a.inv_ls[0][ii] * a.inv_ls[1][jj] * nu;
    };
    const double lx = a.ls[0][i], ly = a.ls[1][j], lz = a.ls[2][k];
    const double ilx = a.inv_ls[0][i], ily = a.inv_ls[1][j], ilz = a.inv_ls[2][k];
    diag[3 * t + 0] = W(1, i, j, k - 1) + W(1, i, j, k) + W(2, i, j - 1, k) + W(2, i, j, k) + a.penalty * ly * lz * ilx;
    diag[3 * t + 1] = W(0, i, j, k - 1) + W(0, i, j, k) + W(2, i - 1, j, k) + W(2, i, j, k) + a.penalty * lx * lz * ily;
    diag[3 * t + 2] = W(0, i, j - 1, k) + W(0, i, j, k) + W(1, i - 1, j, k) + W(1, i, j, k) + a.penalty * lx * ly * ilz;
}

__global__ void k_ms3d_nodal_mu(int64_t nx, int64_t ny, int64_t nz, const double* __restrict__ cell_mu, double mu0,
                                double* __restrict__ mu)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nx * ny * nz) return;
    const int64_t i = t % nx, j = (t / nx) % ny, k = t / (nx * ny);
    const int64_t cx = nx - 1, cy = ny - 1, cz = nz - 1;
    double acc = 0.0;
    int cnt = 0;
    for (int dk = -1; dk <= 0; ++dk)
        for (int dj = -1; dj <= 0; ++dj)
            for (int di = -1; di <= 0; ++di) {
                const int64_t ci = i + di, cj = j + dj, ck = k + dk;
                if (ci < 0 || cj < 0 || ck < 0 || ci >= cx || cj >= cy || ck >= cz) continue;
                acc += cell_mu[ci + cx * (cj + cy * ck)];
                ++cnt;
            }
    mu[t] = cnt == 0 ? mu0 : acc / (double)cnt;
}

__global__ void k_ms3d_rhs(int64_t nx, int64_t ny, int64_t nz, const double* __restrict__ xc,
                           const double* __restrict__ yc, const double* __restrict__ zc,
                           const double* __restrict__ cell_j, double* __restrict__ rhs)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nx * ny * nz) return;
    const int64_t i = t % nx, j = (t / nx) % ny, k = t / (nx * ny);
    const int64_t cx = nx - 1, cy = ny - 1, cz = nz - 1;
    const bool boundary = i == 0 || j == 0 || k == 0 || i + 1 == nx || j + 1 == ny || k + 1 == nz;
    for (int c = 0; c < 3; ++c) {
        double acc = 0.0;
        if (!boundary) {
            for (int dk = -1; dk <= 0; ++dk)
                for (int dj = -1; dj <= 0; ++dj)
                    for (int di = -1; di <= 0; ++di) {
                        const int64_t ci = i + di, cj = j + dj, ck = k + dk;
                        if (ci < 0 || cj < 0 || ck < 0 || ci >= cx || cj >= cy || ck >= cz) continue;
                        const double wx = xc[ci + 1] - xc[ci], wy = yc[cj + 1] - yc[cj], wz = zc[ck + 1] - zc[ck];
                        const double face = c == 0 ? wy * wz : (c == 1 ? wx * wz : wx * wy);
                        acc += 0.125 * face * cell_j[3 * (ci + cx * (cj + cy * ck)) + c];
                    }
        }
        rhs[3 * t + c] = acc;
    }
}

// ------------------------------------------------------------- Poisson 2D ---

template <typename T, bool PU>
__global__ __launch_bounds__(kBlock) void k_p2d_apply(int64_t nx, int64_t ny, const T* __restrict__ wx,
                                                      const T* __restrict__ wy, const T* __restrict__ x,
                                                      T* __restrict__ out, double* __restrict__ partials, PUpdate<T> pu)
{
    if (PU && *pu.done) return;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double dot = 0.0;
    if (t < nx * ny) {
        const int64_t j = t % ny, i = t / ny;
        const PUpdate<T>* pp = PU ? &pu : nullptr;
        const bool inner = i > 0 && j > 0 && i + 1 < nx && j + 1 < ny;
        T pv = inner ? p_value<T>(pp, x, t) : T(0);
        T o = T(0);
        if (inner) {
            const T ue = (i + 2 < nx) ? p_value<T>(pp, x, t + ny) : T(0);
            const T uw = (i > 1) ? p_value<T>(pp, x, t - ny) : T(0);
            const T un = (j + 2 < ny) ? p_value<T>(pp, x, t + 1) : T(0);
            const T us = (j > 1) ? p_value<T>(pp, x, t - 1) : T(0);
            o = wx[t] * (pv - ue) + wx[t - ny] * (pv - uw) + wy[t] * (pv - un) + wy[t - 1] * (pv - us);
        }
        out[t] = o;
        if (PU) pu.p_new[t] = pv;
        dot = (double)pv * (double)o;
    }
    if (partials) block_dot_store<T>(dot, partials);
}

__global__ void k_p2d_weights(int64_t nx, int64_t ny, const double* __restrict__ hx, const double* __restrict__ hy,
                              const double* __restrict__ kappa, double* __restrict__ wx, double* __restrict__ wy,
                              float* __restrict__ wx32, float* __restrict__ wy32)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    const int64_t j = t % ny, i = t / ny;
    // ⋆1 of the edge at node (i,j): h_y/h_x along x, h_x/h_y along y (per-node rule)
    double kx = 1.0, ky = 1.0;
    if (kappa) {
        kx = (i + 1 < nx) ? 0.5 * (kappa[t] + kappa[t + ny]) : kappa[t];
        ky = (j + 1 < ny) ? 0.5 * (kappa[t] + kappa[t + 1]) : kappa[t];
    }
    const double a = kx * hy[j] / hx[i];
    const double b = ky * hx[i] / hy[j];
    wx[t] = a;
    wy[t] = b;
    wx32[t] = (float)a;
    wy32[t] = (float)b;
}

__global__ void k_p2d_diag(int64_t nx, int64_t ny, const double* __restrict__ wx, const double* __restrict__ wy,
                           double* __restrict__ diag)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    const int64_t j = t % ny, i = t / ny;
    const bool inner = i > 0 && j > 0 && i + 1 < nx && j + 1 < ny;
    diag[t] = inner ? wx[t] + wx[t - ny] + wy[t] + wy[t - 1] : 1.0;
}

__global__ void k_nu_from_mu(const double* __restrict__ mu, double* __restrict__ nu, float* __restrict__ nu32,
                             int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = 1.0 / mu[i];
        nu[i] = v;
        nu32[i] = (float)v;
    }
}

inline unsigned nblk(int64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

template <typename T>
MsArgs<T> ms_args(const Ms3dData& d);

template <>
MsArgs<double> ms_args<double>(const Ms3dData& d)
{
    MsArgs<double> a{};
    a.nx = d.nx;
    a.ny = d.ny;
    a.nz = d.nz;
    a.penalty = d.penalty;
    for (int q = 0; q < 3; ++q) {
        a.ls[q] = d.ls[q];
        a.inv_ls[q] = d.inv_ls[q];
    }
    a.nu = d.nu;
    return a;
}

template <>
MsArgs<float> ms_args<float>(const Ms3dData& d)
{
    MsArgs<float> a{};
    a.nx = d.nx;
    a.ny = d.ny;
    a.nz = d.nz;
    a.penalty = (float)d.penalty;
    for (int q = 0; q < 3; ++q) {
        a.ls[q] = d.ls32[q];
        a.inv_ls[q] = d.inv_ls32[q];
    }
    a.nu = d.nu32;
    return a;
}

template <typename T, int RESP = RESP_LINEAR>
int ms_apply(const Ms3dData& d, const T* in, T* out, DotEpilogue dot, const PUpdate<T>* pu, int* nb_out,
             hipStream_t s, const NlData* nl = nullptr)
{
    const int64_t nn = d.nx * d.ny * d.nz;
    const unsigned nb = nblk(nn);
    if (dot.partials && nb > kMaxDotBlocks) {
        // the dot epilogue needs one slot per block
        set_error("operator too large for the fused dot epilogue (%u blocks)", nb);
        return SIL_ERR_UNSUPPORTED;
    }
    if (nb_out) *nb_out = (int)nb;
    MsArgs<T> a = ms_args<T>(d);
    if (nl) {
        if constexpr (sizeof(T) == 8) {
            a.st_a = nl->st_a;
            a.st_c = nl->st_c;
            for (int q = 0; q < 3; ++q) a.st_b[q] = nl->st_b[q];
            a.st4 = reinterpret_cast<const NodeState<T>*>(nl->st_pack);
        } else {
            a.st_a = nl->st_a32;
            a.st_c = nl->st_c32;
            for (int q = 0; q < 3; ++q) a.st_b[q] = nl->st_b32[q];
            a.st4 = reinterpret_cast<const NodeState<T>*>(nl->st_pack32);
        }
    }
    SIL_REQUIRE(RESP == RESP_LINEAR || (d.nx >= 3 && d.ny >= 3 && d.nz >= 3), SIL_ERR_UNSUPPORTED,
                "nonlinear magnetostatics needs at least 3 nodes per axis");
    if (d.nx >= 3 && d.ny >= 3 && d.nz >= 3) {
        // march: tiles of 62 × (kMsTY − 2) output nodes, chunks of kMsChunk planes
        const int ntx = (int)((d.nx + 61) / 62), nty = (int)((d.ny + kMsTY - 3) / (kMsTY - 2));
        const int nkc = (int)((d.nz + kMsChunk - 1) / kMsChunk);
        const int64_t nbm = (int64_t)ntx * nty * nkc;
        SIL_REQUIRE(!dot.partials || nbm <= kMaxDotBlocks, SIL_ERR_UNSUPPORTED,
                    "operator too large for the fused dot epilogue (%lld workgroups)", (long long)nbm);
        if (nb_out) *nb_out = (int)nbm;
        const T* src = in;
        PUpdate<T> skip{};
        if (pu) {
            // the CG direction in its own streaming pass (forming it while the
            // march stages its tile reads three strided arrays per node and
            // is slower than the pass)
            const int64_t ndof = 3 * nn;
            const unsigned g = (unsigned)std::min<int64_t>((ndof + kBlock - 1) / kBlock, 8192);
            hipLaunchKernelGGL((k_direction<T>), dim3(g), dim3(kBlock), 0, s, *pu, ndof);
            src = pu->p_new;
            skip.done = pu->done;
        }
        static const bool nt = [] {
            const char* e = std::getenv("SIL_MS3D_NT");
            return !(e && e[0] == '0');
        }();
        if (nt)
            hipLaunchKernelGGL((k_ms3d_march<T, false, RESP, true>), dim3((unsigned)nbm), dim3(64 * kMsTY), 0, s, a,
                               src, out, dot.partials, skip, ntx, nty);
        else
            hipLaunchKernelGGL((k_ms3d_march<T, false, RESP>), dim3((unsigned)nbm), dim3(64 * kMsTY), 0, s, a, src,
                               out, dot.partials, skip, ntx, nty);
    } else if (pu) {
        // the direction is formed in its own pass: formed inside the
        // direct kernel, every one of the ~40 neighbour reads of a node would
        // read r, D⁻¹ and p_old instead of p
        const int64_t ndof = 3 * nn;
        const unsigned g = (unsigned)std::min<int64_t>((ndof + kBlock - 1) / kBlock, 8192);
        hipLaunchKernelGGL((k_direction<T>), dim3(g), dim3(kBlock), 0, s, *pu, ndof);
        PUpdate<T> skip{};
        skip.done = pu->done;
        hipLaunchKernelGGL((k_ms3d_apply<T, false>), dim3(nb), dim3(kBlock), 0, s, a, pu->p_new, out, dot.partials,
                           skip);
    } else {
        hipLaunchKernelGGL((k_ms3d_apply<T, false>), dim3(nb), dim3(kBlock), 0, s, a, in, out, dot.partials,
                           PUpdate<T>{});
    }
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

template <typename T>
int p2_apply(const Poisson2dData& d, const T* wx, const T* wy, const T* in, T* out, DotEpilogue dot,
             const PUpdate<T>* pu, int* nb_out, hipStream_t s)
{
    const int64_t nn = d.nx * d.ny;
    const unsigned nb = nblk(nn);
    if (dot.partials && nb > kMaxDotBlocks) {
        set_error("operator too large for the fused dot epilogue (%u blocks)", nb);
        return SIL_ERR_UNSUPPORTED;
    }
    if (nb_out) *nb_out = (int)nb;
    if (pu)
        hipLaunchKernelGGL((k_p2d_apply<T, true>), dim3(nb), dim3(kBlock), 0, s, d.nx, d.ny, wx, wy, in, out,
                           dot.partials, *pu);
    else
        hipLaunchKernelGGL((k_p2d_apply<T, false>), dim3(nb), dim3(kBlock), 0, s, d.nx, d.ny, wx, wy, in, out,
                           dot.partials, PUpdate<T>{});
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

}  // namespace

int64_t max_dot_blocks() { return kMaxDotBlocks; }

int ms3d_state_apply64(const Ms3dData& d, const NlData& nl, const double* x, double* out, hipStream_t s)
{
    return ms_apply<double, RESP_SCALED>(d, x, out, DotEpilogue{nullptr}, nullptr, nullptr, s, &nl);
}

int ms3d_jacobian_apply64(const Ms3dData& d, const NlData& nl, const double* in, double* out, DotEpilogue dot,
                          const PUpdate<double>* pu, int* nb, hipStream_t s)
{
    return ms_apply<double, RESP_JACOBIAN>(d, in, out, dot, pu, nb, s, &nl);
}

int ms3d_jacobian_apply32(const Ms3dData& d, const NlData& nl, const float* in, float* out, DotEpilogue dot,
                          const PUpdate<float>* pu, int* nb, hipStream_t s)
{
    return ms_apply<float, RESP_JACOBIAN>(d, in, out, dot, pu, nb, s, &nl);
}

int direction_pass64(const PUpdate<double>& pu, int64_t n, hipStream_t s)
{
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 8192));
    hipLaunchKernelGGL((k_direction<double>), dim3(g), dim3(kBlock), 0, s, pu, n);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int ms3d_apply64(const Ms3dData& d, const double* in, double* out, DotEpilogue dot, const PUpdate<double>* pu,
                 int* nb, hipStream_t s)
{
    return ms_apply<double>(d, in, out, dot, pu, nb, s);
}

int ms3d_apply32(const Ms3dData& d, const float* in, float* out, DotEpilogue dot, const PUpdate<float>* pu, int* nb,
                 hipStream_t s)
{
    return ms_apply<float>(d, in, out, dot, pu, nb, s);
}

int ms3d_diagonal(const Ms3dData& d, double* diag, hipStream_t s)
{
    const int64_t nn = d.nx * d.ny * d.nz;
    hipLaunchKernelGGL(k_ms3d_diag, dim3(nblk(nn)), dim3(kBlock), 0, s, ms_args<double>(d), diag);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int ms3d_nodal_mu(int64_t nx, int64_t ny, int64_t nz, const double* cell_mu, double mu0, double* mu, hipStream_t s)
{
    hipLaunchKernelGGL(k_ms3d_nodal_mu, dim3(nblk(nx * ny * nz)), dim3(kBlock), 0, s, nx, ny, nz, cell_mu, mu0, mu);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int ms3d_rhs(const Ms3dData& d, const double* cell_current, double* rhs, hipStream_t s)
{
    hipLaunchKernelGGL(k_ms3d_rhs, dim3(nblk(d.nx * d.ny * d.nz)), dim3(kBlock), 0, s, d.nx, d.ny, d.nz, d.coord[0],
                       d.coord[1], d.coord[2], cell_current, rhs);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int ms3d_nu_from_mu(const double* mu, double* nu, float* nu32, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_nu_from_mu, dim3(1024), dim3(kBlock), 0, s, mu, nu, nu32, n);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int p2d_apply64(const Poisson2dData& d, const double* in, double* out, DotEpilogue dot, const PUpdate<double>* pu,
                int* nb, hipStream_t s)
{
    return p2_apply<double>(d, d.wx, d.wy, in, out, dot, pu, nb, s);
}

int p2d_apply32(const Poisson2dData& d, const float* in, float* out, DotEpilogue dot, const PUpdate<float>* pu,
                int* nb, hipStream_t s)
{
    return p2_apply<float>(d, d.wx32, d.wy32, in, out, dot, pu, nb, s);
}

int p2d_diagonal(const Poisson2dData& d, double* diag, hipStream_t s)
{
    hipLaunchKernelGGL(k_p2d_diag, dim3(nblk(d.nx * d.ny)), dim3(kBlock), 0, s, d.nx, d.ny, d.wx, d.wy, diag);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int p2d_weights(Poisson2dData& d, const double* hx_dev, const double* hy_dev, const double* kappa, hipStream_t s)
{
    hipLaunchKernelGGL(k_p2d_weights, dim3(nblk(d.nx * d.ny)), dim3(kBlock), 0, s, d.nx, d.ny, hx_dev, hy_dev, kappa,
                       d.wx, d.wy, d.wx32, d.wy32);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

}  // namespace sil
