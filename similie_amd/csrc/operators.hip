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
#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "apply_epilogue.hpp"
#include "buffer_io.hpp"
#include "operators.hpp"

namespace sil {

namespace {

constexpr int kBlock = kApplyBlock;
constexpr int64_t kMaxDotBlocks = 1 << 18;

// --------------------------------------------------------- magnetostatics ---

template <typename T>
struct MsArgs {
    int64_t nx, ny, nz;
    T penalty;  // 0 under the divergence gauge (added by its own pass)
    T ysign;    // sign of the y moment's response
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
            return w * a.nu[ii + a.nx * (jj + a.ny * kk)] * B * a.ysign;
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
        H[1] *= a.ysign;
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

// Pipelined march (the default on grids with ≥ 3 nodes per axis): the same
// tiles and arithmetic as k_ms3d_march with every global access through a
// raw buffer resource (buffer_io.hpp), so no branch stands between a load and
// its consumer.  Plane k+2 of A and the material state of plane k+1 are
// requested at the top of step k and consumed one step later, and the LDS
// planes are double-buffered by plane parity, so a step has one barrier:
//   step k: H(k) from A(k) (LDS: x+1 / y+1 neighbours) and A(k+1);
//           write H(k) and A(k+1) to LDS; barrier;
//           out(k) from H(k−1), H(k) and H(k) of the x−1 / y−1 neighbours.
constexpr int kMsPipeChunkMax = 64;

template <typename T, int TY>
struct MsPipeLds {
    T a[2][3][TY][64];
    T h[2][3][TY][64];
    T lz[kMsPipeChunkMax + 2];
    T ilz[kMsPipeChunkMax + 2];
    double dot[TY];
};

// material record per node: ν (linear), a (the operator at its state) or
// (a, c, B0x, B0y, B0z) (its Jacobian)
template <int RESP>
constexpr int ms_state_width()
{
    return RESP == RESP_JACOBIAN ? 5 : 1;
}

// Input modes: PIPE_X — x; PIPE_PU — the CG direction p = z + β p_old
// (z = D⁻¹ r or r), formed from r, D⁻¹ and p_old as each plane arrives (in
// the arithmetic of p_value) and written to p_new by the node's owner;
// PIPE_QDOTS — x is the direction p, and with ⟨p, q⟩ the block also sums
// ⟨q, z⟩ and ⟨q, D⁻¹ q⟩ (z = D⁻¹ r of the owned nodes) into partials + stride
// and partials + 2·stride: the dots that give the next ⟨r, z⟩ before r is
// updated (the single-pass CG update).
enum PipeMode { PIPE_X = 0, PIPE_PU = 1, PIPE_QDOTS = 2 };

template <typename T, int RESP, int TY, int MODE>
__global__ __launch_bounds__(64 * TY) void k_ms3d_pipe(MsArgs<T> a, const T* __restrict__ x, T* __restrict__ out,
                                                       double* __restrict__ partials, PUpdate<T> pu, int ntx,
                                                       int nty, int chunk, int64_t pstride)
{
    constexpr bool PU = MODE == PIPE_PU, QD = MODE == PIPE_QDOTS;
    __shared__ MsPipeLds<T, TY> s;
    if (pu.done && *pu.done) return; // uniform
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;

    // XCD-contiguous tile order: hardware places workgroup b on XCD b % 8
    int logical;
    {
        const int nb = gridDim.x, b = blockIdx.x;
        const int per = (nb + 7) / 8, rem = nb % 8;
        const int xcd = b % 8, slot = b / 8;
        const int full = rem == 0 ? 8 : rem;
        logical = (xcd < full) ? xcd * per + slot : full * per + (xcd - full) * (per - 1) + slot;
    }
    const int tx = logical % ntx, ty = (logical / ntx) % nty, tk = logical / (ntx * nty);
    const int64_t nx = a.nx, ny = a.ny, nz = a.nz, plane = nx * ny;
    const int64_t i = (int64_t)tx * 62 - 1 + lane, j = (int64_t)ty * (TY - 2) - 1 + row;
    const int64_t ka = (int64_t)tk * chunk, kb = min(ka + (int64_t)chunk, nz);
    const int64_t k0 = ka > 0 ? ka - 1 : ka;   // first plane of the walk
    const int64_t qe = min(kb + 1, nz);        // planes read: k0 … qe−1
    const bool inij = i >= 0 && i < nx && j >= 0 && j < ny;
    const bool writer = lane >= 1 && lane <= 62 && row >= 1 && row <= TY - 2 && inij;
    const bool in_ij = i > 0 && j > 0 && i + 1 < nx && j + 1 < ny;
    const int64_t ic = min(max(i, (int64_t)0), nx - 1), jc = min(max(j, (int64_t)0), ny - 1);
    const T lx = a.ls[0][ic], ly = a.ls[1][jc], ilx = a.inv_ls[0][ic], ily = a.inv_ls[1][jc];

    for (int t = threadIdx.x; t < (int)(qe - k0); t += 64 * TY) {
        s.lz[t] = a.ls[2][k0 + t];
        s.ilz[t] = a.inv_ls[2][k0 + t];
    }

    constexpr int SW = ms_state_width<RESP>();
    const uint32_t node = inij ? static_cast<uint32_t>(i + nx * j) : 0u;
    const int64_t in_bytes = (qe - k0) * plane * 3 * (int64_t)sizeof(T);
    const int64_t out_bytes = (kb - ka) * plane * 3 * (int64_t)sizeof(T);
    const __amdgpu_buffer_rsrc_t rx = buffer_of((PU ? pu.r : x) + 3 * plane * k0, in_bytes);
    const __amdgpu_buffer_rsrc_t ro = buffer_of(out + 3 * plane * ka, out_bytes);
    // PU: D⁻¹ (absent: identity), p_old, and p_new of the owned planes;
    // QDOTS: r and D⁻¹ of the owned nodes
    const bool jac = (PU || QD) && pu.inv_diag != nullptr;
    const __amdgpu_buffer_rsrc_t rdi = buffer_of(jac ? pu.inv_diag + 3 * plane * k0 : x, jac ? in_bytes : 0);
    const __amdgpu_buffer_rsrc_t rpo = buffer_of(PU ? pu.p_old + 3 * plane * k0 : x, PU ? in_bytes : 0);
    const __amdgpu_buffer_rsrc_t rrr = buffer_of(QD ? pu.r + 3 * plane * k0 : x, QD ? in_bytes : 0);
    const __amdgpu_buffer_rsrc_t rpn = buffer_of(PU ? pu.p_new + 3 * plane * ka : out, PU ? out_bytes : 0);
    const double beta = PU ? *pu.beta : 0.0;
    const T* st0 = RESP == RESP_LINEAR ? a.nu : (RESP == RESP_SCALED ? a.st_a : reinterpret_cast<const T*>(a.st4));
    const int64_t st0_width = RESP == RESP_JACOBIAN ? 4 : 1;
    const __amdgpu_buffer_rsrc_t rs0 =
        buffer_of(st0 + st0_width * plane * k0, (kb - k0) * plane * st0_width * (int64_t)sizeof(T));
    const __amdgpu_buffer_rsrc_t rs1 = buffer_of(RESP == RESP_JACOBIAN ? a.st_b[2] + plane * k0 : st0,
                                                 RESP == RESP_JACOBIAN ? (kb - k0) * plane * (int64_t)sizeof(T) : 0);

    // raw input of this lane's node in plane q (zero outside the planes
    // read): x, or (r, D⁻¹, p_old) under PU
    constexpr int RW = (PU || QD) ? 9 : 3;
    auto load = [&](int64_t q, T v[RW]) {
        const uint32_t off = (inij && q < qe)
                                 ? static_cast<uint32_t>(((q - k0) * plane + node) * 3 * (int64_t)sizeof(T))
                                 : kOutside;
        buffer_load3<T>(rx, off, v);
        if constexpr (PU) {
            if (jac)
                buffer_load3<T>(rdi, off, v + 3);
            buffer_load3<T>(rpo, off, v + 6);
        }
        if constexpr (QD) {
            // owned nodes of the owned planes only: no traffic for halos
            const uint32_t own = (writer && q >= ka && q < kb) ? off : kOutside;
            buffer_load3<T>(rrr, own, v + 3);
            if (jac) buffer_load3<T>(rdi, own, v + 6);
        }
    };
    // the operator's input from a raw record
    auto input = [&](const T v[RW], T A[3]) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if constexpr (PU) {
                const double z = jac ? (double)v[c] * (double)v[3 + c] : (double)v[c];
                A[c] = (T)(z + beta * (double)v[6 + c]);
            } else {
                A[c] = v[c];
            }
        }
    };
    // material record of plane q (only planes k0 … kb−1 have an H)
    auto load_state = [&](int64_t q, T m[SW]) {
        const bool ok = inij && q < kb;
        const int64_t n = (q - k0) * plane + node;
        if constexpr (RESP == RESP_JACOBIAN) {
            buffer_load4<T>(rs0, ok ? static_cast<uint32_t>(n * 4 * (int64_t)sizeof(T)) : kOutside, m);
            m[4] = buffer_load1<T>(rs1, ok ? static_cast<uint32_t>(n * (int64_t)sizeof(T)) : kOutside);
        } else {
            m[0] = buffer_load1<T>(rs0, ok ? static_cast<uint32_t>(n * (int64_t)sizeof(T)) : kOutside);
        }
    };
    auto interior = [&](int64_t k) { return in_ij && k > 0 && k + 1 < nz; };

    // Register sets rotate with period 3 (the loop is unrolled by 3): at
    // phase u, PB[u] holds A(k), PB[u+1] A(k+1), and A(k+2) is requested into
    // PB[u+2]; MB[u] holds the material record of plane k and plane k+1's is
    // requested into MB[u+1].  No register holding a load in flight is
    // copied, so the wavefront never waits for a load before its consumer.
    // RB: raw records in flight; PC: inputs formed from them (PC[u] = A(k))
    T RB[3][RW], PC[3][3], MB[3][SW];
    load(k0, RB[0]);
    load(k0 + 1, RB[1]);
    load_state(k0, MB[0]);
    input(RB[0], PC[0]);
    {
        const bool ik = interior(k0);
#pragma unroll
        for (int c = 0; c < 3; ++c) s.a[0][c][row][lane] = ik ? PC[0][c] : T(0);
    }
    __syncthreads();

    const int lp = lane + 1 < 64 ? lane + 1 : lane, rp = row + 1 < TY ? row + 1 : row;
    const int lm = lane > 0 ? lane - 1 : lane, rm = row > 0 ? row - 1 : row;
    T Hp[3] = {T(0), T(0), T(0)};
    double dot = 0.0, dqz = 0.0, dqq = 0.0;
    auto step = [&](auto U, int64_t k) {
        constexpr int u = decltype(U)::value, u1 = (u + 1) % 3, u2 = (u + 2) % 3;
        const T* M0 = MB[u];
        const int buf = static_cast<int>(k - k0) & 1;
        load(k + 2, RB[u2]);
        load_state(k + 1, MB[u1]);
        input(RB[u1], PC[u1]);
        const T* P0 = PC[u];
        const T* P1 = PC[u1];

        const bool ik = interior(k), ik1 = interior(k + 1);
        T A0[3], A1[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            A0[c] = ik ? P0[c] : T(0);
            A1[c] = ik1 ? P1[c] : T(0);
        }
        const T ax1 = s.a[buf][1][row][lp], ax2 = s.a[buf][2][row][lp]; // A(i+1, j, k)
        const T ay0 = s.a[buf][0][rp][lane], ay2 = s.a[buf][2][rp][lane]; // A(i, j+1, k)
        const T lz = s.lz[k - k0], ilz = s.ilz[k - k0];
        const T bx = ay2 - A0[2] - A1[1] + A0[1];
        const T by = -ax2 + A0[2] + A1[0] - A0[0];
        const T bz = ax1 - A0[1] - ay0 + A0[0];
        const T cx = lx * ily * ilz, cy = ly * ilx * ilz, cz = lz * ilx * ily;
        T H[3];
        if constexpr (RESP == RESP_LINEAR) {
            const T nu = M0[0];
            H[0] = cx * nu * bx;
            H[1] = cy * nu * by;
            H[2] = cz * nu * bz;
        } else if constexpr (RESP == RESP_SCALED) {
            const T sa = M0[0];
            H[0] = cx * (bx * sa);
            H[1] = cy * (by * sa);
            H[2] = cz * (bz * sa);
        } else {
            const T sa = M0[0], sc = M0[1], b0x = M0[2], b0y = M0[3], b0z = M0[4];
            const T cd = sc * (b0x * bx + b0y * by + b0z * bz);
            H[0] = cx * (sa * bx + cd * b0x);
            H[1] = cy * (sa * by + cd * b0y);
            H[2] = cz * (sa * bz + cd * b0z);
        }
        H[1] *= a.ysign;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            H[c] = ik ? H[c] : T(0);
            s.h[buf][c][row][lane] = H[c];
            s.a[buf ^ 1][c][row][lane] = A1[c];
        }
        __syncthreads();

        if (k >= ka) { // uniform
            const T pen = a.penalty;
            T o[3];
            const T hym2 = s.h[buf][2][rm][lane], hym0 = s.h[buf][0][rm][lane]; // H(i, j−1, k)
            const T hxm2 = s.h[buf][2][row][lm], hxm1 = s.h[buf][1][row][lm]; // H(i−1, j, k)
            o[0] = Hp[1] - H[1] - hym2 + H[2] + pen * (ly * lz * ilx) * P0[0];
            o[1] = -Hp[0] + H[0] + hxm2 - H[2] + pen * (lx * lz * ily) * P0[1];
            o[2] = hym0 - H[0] - hxm1 + H[1] + pen * (lx * ly * ilz) * P0[2];
#pragma unroll
            for (int c = 0; c < 3; ++c) o[c] = ik ? o[c] : P0[c];
            const uint32_t off =
                writer ? static_cast<uint32_t>(((k - ka) * plane + node) * 3 * (int64_t)sizeof(T)) : kOutside;
            buffer_store3<T, 2>(ro, off, o);
            if constexpr (PU) buffer_store3<T>(rpn, off, P0);
            const double dd = (double)P0[0] * (double)o[0] + (double)P0[1] * (double)o[1] +
                              (double)P0[2] * (double)o[2];
            dot += writer ? dd : 0.0;
            if constexpr (QD) {
                const T* R = RB[u];
                double qz = 0.0, qq = 0.0;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double z = jac ? (double)R[3 + c] * (double)R[6 + c] : (double)R[3 + c];
                    qz += (double)o[c] * z;
                    qq += jac ? (double)o[c] * (double)R[6 + c] * (double)o[c] : (double)o[c] * (double)o[c];
                }
                dqz += writer ? qz : 0.0;
                dqq += writer ? qq : 0.0;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) Hp[c] = H[c];
    };
    for (int64_t k = k0; k < kb; k += 3) {
        step(std::integral_constant<int, 0>{}, k);
        if (k + 1 < kb) step(std::integral_constant<int, 1>{}, k + 1);
        if (k + 2 < kb) step(std::integral_constant<int, 2>{}, k + 2);
    }
    if (partials) {
        auto block_sum_to = [&](double v, double* dst) {
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            __syncthreads(); // s.dot free
            if (lane == 0) s.dot[row] = v;
            __syncthreads();
            if (threadIdx.x == 0) {
                double sum = 0.0;
                for (int r = 0; r < TY; ++r) sum += s.dot[r];
                dst[blockIdx.x] = sum;
            }
        };
        block_sum_to(dot, partials);
        if constexpr (QD) {
            block_sum_to(dqz, partials + pstride);
            block_sum_to(dqq, partials + 2 * pstride);
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
        if (m == 1) return a.ls[1][jj] * a.inv_ls[0][ii] * a.inv_ls[2][kk] * nu * a.ysign;
        return a.ls[2][kk] * a.inv_ls[0][ii] * a.inv_ls[1][jj] * nu;
    };
    const double lx = a.ls[0][i], ly = a.ls[1][j], lz = a.ls[2][k];
    const double ilx = a.inv_ls[0][i], ily = a.inv_ls[1][j], ilz = a.inv_ls[2][k];
    diag[3 * t + 0] = W(1, i, j, k - 1) + W(1, i, j, k) + W(2, i, j - 1, k) + W(2, i, j, k) + a.penalty * ly * lz * ilx;
    diag[3 * t + 1] = W(0, i, j, k - 1) + W(0, i, j, k) + W(2, i - 1, j, k) + W(2, i, j, k) + a.penalty * lx * lz * ily;
    diag[3 * t + 2] = W(0, i, j - 1, k) + W(0, i, j, k) + W(1, i - 1, j, k) + W(1, i, j, k) + a.penalty * lx * ly * ilz;
}

// ------------------------------------------------------ divergence gauge ---
//
// magnetostatics_onelab.hpp:1893-1941, 2370-2392: with the divergence gauge an
// interior row (m, c) adds factor · Σ_{n ∈ {m, m+e_c}} w(n) ρ_c(n) (δx)(n) in
// place of the penalty term, where
//   (δx)(n) = Σ_d [x_d(n−e_d) / (e_d(n) e_d(n−e_d)) − x_d(n) / e_d(n)²]
// over interior DOFs is the codifferential of codifferential.hpp:191-310 at an
// identity metric on a rectilinear grid (e: edge of the Hodge star's cell,
// forward, backward at the last node), w = ℓx ℓy ℓz (local spacings) and
// ρ_c(m) = −1/e_c(m)², ρ_c(m+e_c) = 1/(e_c(m+e_c) e_c(m)).  A pass of its own
// after the curl-curl march (which then carries no penalty term); it also
// forms the CG dot ⟨x, Ax⟩ over every DOF.
struct GaugeArgs {
    int64_t nx, ny, nz;
    double factor;
    const double* coord[3];
    const double* ls[3];
};

__device__ __forceinline__ double cell_edge(const double* c, int64_t n, int64_t i)
{
    return i + 1 < n ? c[i + 1] - c[i] : c[i] - c[i - 1];
}

__device__ __forceinline__ bool gauge_interior(const GaugeArgs& g, int64_t i, int64_t j, int64_t k)
{
    return i > 0 && j > 0 && k > 0 && i + 1 < g.nx && j + 1 < g.ny && k + 1 < g.nz;
}

// (δx)(n): entries (n−e_x, x), (n−e_y, y), (n−e_z, z), (n, x), (n, y), (n, z)
template <typename T>
__device__ double codiff_at(const GaugeArgs& g, const T* __restrict__ x, int64_t i, int64_t j, int64_t k)
{
    const int64_t idx[3] = {i, j, k}, ext[3] = {g.nx, g.ny, g.nz};
    double e[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) e[d] = cell_edge(g.coord[d], ext[d], idx[d]);
    double v = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int64_t q[3] = {i, j, k};
        q[d] -= 1;
        if (q[d] < 0 || !gauge_interior(g, q[0], q[1], q[2])) continue;
        const double cf = 1.0 / (e[d] * cell_edge(g.coord[d], ext[d], q[d]));
        v += cf * (double)x[3 * (q[0] + g.nx * (q[1] + g.ny * q[2])) + d];
    }
    if (gauge_interior(g, i, j, k)) {
        const int64_t n = i + g.nx * (j + g.ny * k);
#pragma unroll
        for (int d = 0; d < 3; ++d) v += (-1.0 / (e[d] * e[d])) * (double)x[3 * n + d];
    }
    return v;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_ms3d_gauge(GaugeArgs g, const T* __restrict__ x, T* __restrict__ out,
                                                       double* __restrict__ partials, const int* __restrict__ done)
{
    if (done && *done) return;
    const int64_t nn = g.nx * g.ny * g.nz;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double dot = 0.0;
    if (t < nn) {
        const int64_t i = t % g.nx, j = (t / g.nx) % g.ny, k = t / (g.nx * g.ny);
        if (gauge_interior(g, i, j, k)) {
            const int64_t idx[3] = {i, j, k}, ext[3] = {g.nx, g.ny, g.nz};
            const double d0 = codiff_at(g, x, i, j, k);
            const double w0 = g.ls[0][i] * g.ls[1][j] * g.ls[2][k];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                int64_t q[3] = {i, j, k};
                q[c] += 1;
                const double ec = cell_edge(g.coord[c], ext[c], idx[c]);
                const double r0 = -1.0 / (ec * ec), r1 = 1.0 / (cell_edge(g.coord[c], ext[c], q[c]) * ec);
                const double d1 = codiff_at(g, x, q[0], q[1], q[2]);
                const double w1 = g.ls[0][q[0]] * g.ls[1][q[1]] * g.ls[2][q[2]];
                double gr = 0.0;
                gr += g.factor * w0 * r0 * d0;
                gr += g.factor * w1 * r1 * d1;
                out[3 * t + c] = (T)((double)out[3 * t + c] + gr);
            }
        }
        if (partials)
#pragma unroll
            for (int c = 0; c < 3; ++c) dot += (double)x[3 * t + c] * (double)out[3 * t + c];
    }
    if (partials) block_dot_store<T>(dot, partials);
}

// the gauge's diagonal: factor · Σ_n w(n) ρ_c(n)²
__global__ __launch_bounds__(kBlock) void k_ms3d_gauge_diag(GaugeArgs g, double* __restrict__ diag)
{
    const int64_t nn = g.nx * g.ny * g.nz;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nn) return;
    const int64_t i = t % g.nx, j = (t / g.nx) % g.ny, k = t / (g.nx * g.ny);
    if (!gauge_interior(g, i, j, k)) return;
    const int64_t idx[3] = {i, j, k}, ext[3] = {g.nx, g.ny, g.nz};
    const double w0 = g.ls[0][i] * g.ls[1][j] * g.ls[2][k];
    for (int c = 0; c < 3; ++c) {
        int64_t q[3] = {i, j, k};
        q[c] += 1;
        const double ec = cell_edge(g.coord[c], ext[c], idx[c]);
        const double r0 = -1.0 / (ec * ec), r1 = 1.0 / (cell_edge(g.coord[c], ext[c], q[c]) * ec);
        const double w1 = g.ls[0][q[0]] * g.ls[1][q[1]] * g.ls[2][q[2]];
        double gd = 0.0;
        gd += g.factor * w0 * r0 * r0;
        gd += g.factor * w1 * r1 * r1;
        diag[3 * t + c] += gd;
    }
}

GaugeArgs gauge_args(const Ms3dData& d)
{
    GaugeArgs g{};
    g.nx = d.nx;
    g.ny = d.ny;
    g.nz = d.nz;
    g.factor = d.gauge_factor;
    for (int q = 0; q < 3; ++q) {
        g.coord[q] = d.coord[q];
        g.ls[q] = d.ls[q];
    }
    return g;
}

__global__ void k_zero(double* __restrict__ v, int64_t n)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < n) v[t] = 0.0;
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
    a.penalty = d.divg ? 0.0 : d.penalty;
    a.ysign = d.ysign;
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
    a.penalty = d.divg ? 0.0f : (float)d.penalty;
    a.ysign = (float)d.ysign;
    for (int q = 0; q < 3; ++q) {
        a.ls[q] = d.ls32[q];
        a.inv_ls[q] = d.inv_ls32[q];
    }
    a.nu = d.nu32;
    return a;
}

// pipelined-march shape: rows per tile (16 or 8; 0 selects the earlier
// march) and planes per chunk, from SIL_MS3D_PIPE="ty,chunk"
struct MsPipeShape {
    int ty, chunk;
};

MsPipeShape ms_pipe_shape()
{
    static const MsPipeShape shape = [] {
        MsPipeShape p{16, 24};
        if (const char* e = std::getenv("SIL_MS3D_PIPE")) {
            int ty = 16, ch = 24;
            if (std::sscanf(e, "%d,%d", &ty, &ch) >= 1) {
                p.ty = (ty == 0 || ty == 8) ? ty : 16;
                p.chunk = std::min(std::max(ch, 1), kMsPipeChunkMax);
            }
        }
        return p;
    }();
    return shape;
}

template <typename T, int RESP = RESP_LINEAR>
int ms_apply_curl(const Ms3dData& d, const T* in, T* out, DotEpilogue dot, const PUpdate<T>* pu, int* nb_out,
                  hipStream_t s, const NlData* nl);

// the operator: the curl-curl march (with the penalty term), or under the
// divergence gauge the march without it followed by the gauge pass, which
// then forms the dot partials
template <typename T, int RESP = RESP_LINEAR>
int ms_apply(const Ms3dData& d, const T* in, T* out, DotEpilogue dot, const PUpdate<T>* pu, int* nb_out,
             hipStream_t s, const NlData* nl = nullptr)
{
    if (!d.divg) return ms_apply_curl<T, RESP>(d, in, out, dot, pu, nb_out, s, nl);
    const int64_t nn = d.nx * d.ny * d.nz;
    const unsigned nb = nblk(nn);
    SIL_REQUIRE(!dot.partials || nb <= kMaxDotBlocks, SIL_ERR_UNSUPPORTED,
                "operator too large for the fused dot epilogue (%u blocks)", nb);
    int rc = ms_apply_curl<T, RESP>(d, in, out, DotEpilogue{nullptr}, pu, nullptr, s, nl);
    if (rc) return rc;
    const T* x = pu ? pu->p_new : in;
    hipLaunchKernelGGL((k_ms3d_gauge<T>), dim3(nb), dim3(kBlock), 0, s, gauge_args(d), x, out, dot.partials,
                       pu ? pu->done : nullptr);
    if (nb_out) *nb_out = (int)nb;
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

template <typename T, int RESP>
int ms_apply_curl(const Ms3dData& d, const T* in, T* out, DotEpilogue dot, const PUpdate<T>* pu, int* nb_out,
                  hipStream_t s, const NlData* nl)
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
        static const bool nt = [] {
            const char* e = std::getenv("SIL_MS3D_NT");
            return !(e && e[0] == '0');
        }();
        const MsPipeShape ps = ms_pipe_shape();
        const int64_t chunk_bytes = (int64_t)(ps.chunk + 2) * d.nx * d.ny * 4 * (int64_t)sizeof(T);
        if (ps.ty > 0 && chunk_bytes <= kMaxBufferBytes) {
            const int pty = ps.ty - 2;
            const int ptx = (int)((d.nx + 61) / 62), pny = (int)((d.ny + pty - 1) / pty);
            const int pkc = (int)((d.nz + ps.chunk - 1) / ps.chunk);
            const int64_t nbp = (int64_t)ptx * pny * pkc;
            SIL_REQUIRE(!dot.partials || nbp <= kMaxDotBlocks, SIL_ERR_UNSUPPORTED,
                        "operator too large for the fused dot epilogue (%lld workgroups)", (long long)nbp);
            if (nb_out) *nb_out = (int)nbp;
            // FP32: the direction is formed inside the march; FP64: in its own
            // pass (the nine-value records of the fused form spill the
            // march's FP64 registers)
            const bool fused = pu && sizeof(T) == 4;
            if (pu && !fused) {
                const int64_t ndof = 3 * nn;
                const unsigned gd = (unsigned)std::min<int64_t>((ndof + kBlock - 1) / kBlock, 8192);
                hipLaunchKernelGGL((k_direction<T>), dim3(gd), dim3(kBlock), 0, s, *pu, ndof);
            }
            PUpdate<T> pp{};
            if (pu) pp = *pu;
            const T* src = (pu && !fused) ? pu->p_new : in;
            const dim3 g((unsigned)nbp);
            if (ps.ty == 16) {
                if (fused)
                    hipLaunchKernelGGL((k_ms3d_pipe<T, RESP, 16, PIPE_PU>), g, dim3(64 * 16), 0, s, a, src, out,
                                       dot.partials, pp, ptx, pny, ps.chunk, (int64_t)0);
                else
                    hipLaunchKernelGGL((k_ms3d_pipe<T, RESP, 16, PIPE_X>), g, dim3(64 * 16), 0, s, a, src, out,
                                       dot.partials, pp, ptx, pny, ps.chunk, (int64_t)0);
            } else {
                if (fused)
                    hipLaunchKernelGGL((k_ms3d_pipe<T, RESP, 8, PIPE_PU>), g, dim3(64 * 8), 0, s, a, src, out,
                                       dot.partials, pp, ptx, pny, ps.chunk, (int64_t)0);
                else
                    hipLaunchKernelGGL((k_ms3d_pipe<T, RESP, 8, PIPE_X>), g, dim3(64 * 8), 0, s, a, src, out,
                                       dot.partials, pp, ptx, pny, ps.chunk, (int64_t)0);
            }
            SIL_LAUNCH_CHECK();
            return SIL_OK;
        }
        if (pu) {
            // the CG direction in its own streaming pass (forming it while the
            // earlier march stages its tile reads three strided arrays per
            // node and is slower than the pass)
            const int64_t ndof = 3 * nn;
            const unsigned g = (unsigned)std::min<int64_t>((ndof + kBlock - 1) / kBlock, 8192);
            hipLaunchKernelGGL((k_direction<T>), dim3(g), dim3(kBlock), 0, s, *pu, ndof);
            src = pu->p_new;
            skip.done = pu->done;
        }
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

int direction_pass32(const PUpdate<float>& pu, int64_t n, hipStream_t s)
{
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 8192));
    hipLaunchKernelGGL((k_direction<float>), dim3(g), dim3(kBlock), 0, s, pu, n);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
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

namespace {
template <int RESP>
int ms_qdots32(const Ms3dData& d, const NlData* nl, const float* p, float* q, const float* r, const float* inv_diag,
               double* partials, int64_t pstride, const int* done, int* nb_out, hipStream_t s)
{
    const MsPipeShape ps = ms_pipe_shape();
    const int pty = ps.ty - 2;
    const int ptx = (int)((d.nx + 61) / 62), pny = (int)((d.ny + pty - 1) / pty);
    const int pkc = (int)((d.nz + ps.chunk - 1) / ps.chunk);
    const int64_t nbp = (int64_t)ptx * pny * pkc;
    SIL_REQUIRE(nbp <= kMaxDotBlocks && nbp <= pstride, SIL_ERR_UNSUPPORTED,
                "operator too large for the fused dot epilogue (%lld workgroups)", (long long)nbp);
    MsArgs<float> a = ms_args<float>(d);
    if (nl) {
        a.st_a = nl->st_a32;
        a.st_c = nl->st_c32;
        for (int c = 0; c < 3; ++c) a.st_b[c] = nl->st_b32[c];
        a.st4 = reinterpret_cast<const NodeState<float>*>(nl->st_pack32);
    }
    PUpdate<float> pu{};
    pu.r = r;
    pu.inv_diag = inv_diag;
    pu.done = done;
    const dim3 g((unsigned)nbp);
    if (ps.ty == 16)
        hipLaunchKernelGGL((k_ms3d_pipe<float, RESP, 16, PIPE_QDOTS>), g, dim3(64 * 16), 0, s, a, p, q, partials, pu,
                           ptx, pny, ps.chunk, pstride);
    else
        hipLaunchKernelGGL((k_ms3d_pipe<float, RESP, 8, PIPE_QDOTS>), g, dim3(64 * 8), 0, s, a, p, q, partials, pu,
                           ptx, pny, ps.chunk, pstride);
    if (nb_out) *nb_out = (int)nbp;
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}
}  // namespace

bool ms3d_single_pass_ok(const Ms3dData& d)
{
    const MsPipeShape ps = ms_pipe_shape();
    const int64_t chunk_bytes = (int64_t)(ps.chunk + 2) * d.nx * d.ny * 4 * (int64_t)sizeof(float);
    return !d.divg && ps.ty > 0 && d.nx >= 3 && d.ny >= 3 && d.nz >= 3 && chunk_bytes <= kMaxBufferBytes;
}

int ms3d_qdots32(const Ms3dData& d, const NlData* nl, const float* p, float* q, const float* r,
                 const float* inv_diag, double* partials, int64_t pstride, const int* done, int* nb_out,
                 hipStream_t s)
{
    SIL_REQUIRE(ms3d_single_pass_ok(d), SIL_ERR_UNSUPPORTED, "single-pass CG step: operator not supported");
    if (nl)
        return ms_qdots32<RESP_JACOBIAN>(d, nl, p, q, r, inv_diag, partials, pstride, done, nb_out, s);
    return ms_qdots32<RESP_LINEAR>(d, nullptr, p, q, r, inv_diag, partials, pstride, done, nb_out, s);
}

int ms3d_diagonal(const Ms3dData& d, double* diag, hipStream_t s)
{
    const int64_t nn = d.nx * d.ny * d.nz;
    hipLaunchKernelGGL(k_ms3d_diag, dim3(nblk(nn)), dim3(kBlock), 0, s, ms_args<double>(d), diag);
    if (d.divg) hipLaunchKernelGGL(k_ms3d_gauge_diag, dim3(nblk(nn)), dim3(kBlock), 0, s, gauge_args(d), diag);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int ms3d_gauge_diagonal_add(const Ms3dData& d, double* diag, hipStream_t s)
{
    if (!d.divg) return SIL_OK;
    hipLaunchKernelGGL(k_ms3d_gauge_diag, dim3(nblk(d.nx * d.ny * d.nz)), dim3(kBlock), 0, s, gauge_args(d), diag);
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
    // the potential-temporal-derivative criterion has a zero right-hand side
    // (magnetostatics_onelab.hpp:4930-4935)
    if (d.criterion == 0)
        hipLaunchKernelGGL(k_zero, dim3(nblk(3 * d.nx * d.ny * d.nz)), dim3(kBlock), 0, s, rhs, 3 * d.nx * d.ny * d.nz);
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
