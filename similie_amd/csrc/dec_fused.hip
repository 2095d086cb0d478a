// Fused single-pass kernels for the hot shapes of the DEC operators on
// rectilinear grids with identity or diagonal metric (index-last arrays):
//
//  * 3D Hodge Laplacian of a 1-form, Δω = δdω + dδω
//    (src/similie/exterior/laplacian.hpp:599-647 runs it as six kernels on two
//    execution-space partitions with five HBM buffers and 90 precomputed ⋆
//    doubles per node).  Here one pass reads ω and the diagonal metric once
//    and writes Δω once: 72 algorithmic bytes per node.
//  * 2D Laplacian of a 0-form (laplacian.hpp:27-155).
//  * 3D coboundary of a 0-form (coboundary.hpp:614-654).
//
// Per node m, with s_a = sqrt(g_aa)·h_a (h_a the forward edge, backward at the
// last node: volume.hpp:55-82) the discrete stars are volume ratios
// (reduction_and_reconstruction.hpp:403-432):
//   ⋆(YZ→X) = sx/(sy sz), ⋆(XZ→Y) = −sy/(sx sz), ⋆(XY→Z) = sz/(sx sy)
//   ⋆(X→YZ) = sy sz/sx,   ⋆(Y→XZ) = −sx sz/sy,  ⋆(Z→XY) = sx sy/sz
//   ⋆(XYZ→∅) = 1/(sx sy sz)
// and all seven come from one division 1/(sx sy sz).  Nodes the reference
// treats as degenerate (a reconstruction Jacobian below 1e-14, a zero metric
// entry) get zero coefficients, as in the reference.
//
// The 3D kernels march along x (the slowest axis): a workgroup owns a tile of
// rows (y) × lanes (z) including a one-node halo, keeps the planes x−1, x,
// x+1 in registers, has the next planes' loads in flight while it computes,
// and exchanges the y/z neighbours of ω, ⋆ω, ⋆dω and δω through LDS.  Each
// node's coefficients are computed once per tile.  Workgroups are mapped to
// tiles XCD-contiguously so the halos shared by neighbouring tiles are served
// from the same L2.  The pipelined variant addresses memory through buffer
// resources so that no branch separates a load from its consumer: the
// compiler's wait counts then let a step's loads stay in flight across the
// whole step instead of being drained at its top.
#include "buffer_io.hpp"
#include "dec_kernels.hpp"

#include <utility>

namespace sil {

namespace {

struct Coefs {
    double a1, a2, a3; // ⋆ 2-set → 1-set magnitudes (YZ→X, XZ→Y, XY→Z)
    double b1, b2, b3; // ⋆ 1-set → 2-set magnitudes (X→YZ, Y→XZ, Z→XY)
    double c;          // ⋆ XYZ → ∅
};

// 1/sqrt(v) for v > 0: the hardware estimate refined by two Newton–Raphson
// steps y ← y + y (1/2 − (v y / 2) y), accurate to a few ulp.  Outside
// [1e-290, 1e290] the estimate's range is not relied on and the quotient
// 1/(sqrt(X) sqrt(Y) sqrt(Z)) of the caller's factors is used instead.
__device__ __forceinline__ double inv_volume(double X, double Y, double Z)
{
    const double v = X * Y * Z;
    if (v == 0.0) return 0.0;
    if (v >= 1e-290 && v <= 1e290) {
        double y = __builtin_amdgcn_rsq(v);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const double h = 0.5 * v * y;
            y = fma(y, fma(-h, y, 0.5), y);
        }
        return y;
    }
    return 1.0 / (sqrt(X) * sqrt(Y) * sqrt(Z));
}

__device__ __forceinline__ Coefs node_coefs(double gx, double gy, double gz, double hx, double hy, double hz)
{
    Coefs k;
    const bool metric_ok = (gx != 0.0) && (gy != 0.0) && (gz != 0.0);
    const double ax = fabs(hx), ay = fabs(hy), az = fabs(hz);
    const bool ok1 = metric_ok && ax >= 1e-14 && ay >= 1e-14 && az >= 1e-14;
    const bool ok2 = metric_ok && ax * ay >= 1e-14 && ax * az >= 1e-14 && ay * az >= 1e-14;
    const bool ok3 = metric_ok && ax * ay * az >= 1e-14;
    // squared edge measures s_a² = |g_aa| h_a²; every coefficient is a
    // product of them with c = 1/(sx sy sz)
    const double X = fabs(gx) * ax * ax, Y = fabs(gy) * ay * ay, Z = fabs(gz) * az * az;
    const double c = inv_volume(X, Y, Z);
    const double c1 = ok1 ? c : 0.0, c2 = ok2 ? c : 0.0;
    k.a1 = X * c2;
    k.a2 = Y * c2;
    k.a3 = Z * c2;
    k.b1 = (Y * Z) * c1;
    k.b2 = (X * Z) * c1;
    k.b3 = (X * Y) * c1;
    k.c = ok3 ? c : 0.0;
    return k;
}

// ------------------------------------------------------ 3D 1-form, direct ---

// One thread per output node, every neighbour read through the caches.
template <bool DIAG>
__global__ __launch_bounds__(256) void k_lap3d_direct(const double* __restrict__ w, const double* __restrict__ g,
                                                      const double* __restrict__ hx_, const double* __restrict__ hy_,
                                                      const double* __restrict__ hz_, double* __restrict__ out,
                                                      int64_t Nx, int64_t Ny, int64_t Nz, int64_t Ox, int64_t Oy,
                                                      int64_t Oz)
{
    const int64_t total = Ox * Oy * Oz;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t z = t % Oz, y = (t / Oz) % Oy, x = t / (Oz * Oy);
        auto inside = [&](int64_t i, int64_t j, int64_t k) {
            return i >= 0 && j >= 0 && k >= 0 && i < Nx && j < Ny && k < Nz;
        };
        auto W = [&](int64_t i, int64_t j, int64_t k, int c) {
            i = min(i, Nx - 1);
            j = min(j, Ny - 1);
            k = min(k, Nz - 1);
            return w[((i * Ny + j) * Nz + k) * 3 + c];
        };
        auto K = [&](int64_t i, int64_t j, int64_t k) {
            const int64_t lin = (i * Ny + j) * Nz + k;
            double gx = 1.0, gy = 1.0, gz = 1.0;
            if (DIAG) {
                gx = g[lin * 3 + 0];
                gy = g[lin * 3 + 1];
                gz = g[lin * 3 + 2];
            }
            return node_coefs(gx, gy, gz, hx_[i], hy_[j], hz_[k]);
        };
        // η = ⋆dω at m (zero outside)
        auto eta = [&](int64_t i, int64_t j, int64_t k, double e[3]) {
            if (!inside(i, j, k)) {
                e[0] = e[1] = e[2] = 0.0;
                return;
            }
            const Coefs c = K(i, j, k);
            const double w0 = W(i, j, k, 0), w1 = W(i, j, k, 1), w2 = W(i, j, k, 2);
            const double Fxy = (W(i + 1, j, k, 1) - w1) - (W(i, j + 1, k, 0) - w0);
            const double Fxz = (W(i + 1, j, k, 2) - w2) - (W(i, j, k + 1, 0) - w0);
            const double Fyz = (W(i, j + 1, k, 2) - w2) - (W(i, j, k + 1, 1) - w1);
            e[0] = c.a1 * Fyz;
            e[1] = -c.a2 * Fxz;
            e[2] = c.a3 * Fxy;
        };
        // ⋆ω components (YZ, XZ, XY) at m (zero outside)
        auto etap = [&](int64_t i, int64_t j, int64_t k, double e[3]) {
            if (!inside(i, j, k)) {
                e[0] = e[1] = e[2] = 0.0;
                return;
            }
            const Coefs c = K(i, j, k);
            e[0] = c.b1 * W(i, j, k, 0);
            e[1] = -c.b2 * W(i, j, k, 1);
            e[2] = c.b3 * W(i, j, k, 2);
        };
        // δω at m, clamped into the domain
        auto dw = [&](int64_t i, int64_t j, int64_t k) {
            i = min(i, Nx - 1);
            j = min(j, Ny - 1);
            k = min(k, Nz - 1);
            double h0[3], hx[3], hy[3], hz[3];
            etap(i, j, k, h0);
            etap(i - 1, j, k, hx);
            etap(i, j - 1, k, hy);
            etap(i, j, k - 1, hz);
            const double H = (h0[0] - hx[0]) - (h0[1] - hy[1]) + (h0[2] - hz[2]);
            return -K(i, j, k).c * H;
        };
        double e0[3], ex[3], ey[3], ez[3];
        eta(x, y, z, e0);
        eta(x - 1, y, z, ex);
        eta(x, y - 1, z, ey);
        eta(x, y, z - 1, ez);
        const double Gxy = (e0[1] - ex[1]) - (e0[0] - ey[0]);
        const double Gxz = (e0[2] - ex[2]) - (e0[0] - ez[0]);
        const double Gyz = (e0[2] - ey[2]) - (e0[1] - ez[1]);
        const Coefs cn = K(min(x, Nx - 1), min(y, Ny - 1), min(z, Nz - 1));
        const double d0 = dw(x, y, z);
        double* o = out + ((x * Oy + y) * Oz + z) * 3;
        o[0] = cn.a1 * Gyz + (dw(x + 1, y, z) - d0);
        o[1] = -cn.a2 * Gxz + (dw(x, y + 1, z) - d0);
        o[2] = cn.a3 * Gxy + (dw(x, y, z + 1) - d0);
    }
}

// ------------------------------------------------------ 3D 1-form, march ---

// Registers hold planes p, p+1 and the prefetched p+2; loads are guarded by
// the grid bounds.

template <int TY>
struct MarchLds {
    double w[3][TY][64];   // ω(p) for the y+1 / z+1 neighbours of dω
    double ep[2][TY][64];  // ⋆ω(p+1): XZ and XY components for δω's y−1 / z−1 terms
    double eta[3][TY][64]; // ⋆dω(p) for dᵀ's y−1 / z−1 terms
    double dw[2][TY][64];  // δω ping-pong: planes p and p+1
};

template <int TY, bool DIAG>
__global__ __launch_bounds__(64 * TY) void k_lap3d_march(const double* __restrict__ w, const double* __restrict__ g,
                                                         const double* __restrict__ hxs, const double* __restrict__ hys,
                                                         const double* __restrict__ hzs, double* __restrict__ out,
                                                         int64_t Nx, int64_t Ny, int64_t Nz, int64_t Ox, int64_t Oy,
                                                         int64_t Oz, int ntz, int nty, int nxc, int xchunk)
{
    __shared__ MarchLds<TY> s;
    const int lane = threadIdx.x & 63;
    const int row = threadIdx.x >> 6;

    // XCD-contiguous tile order: hardware places workgroup b on XCD b % 8.
    const int nb = gridDim.x;
    const int b = blockIdx.x;
    const int per = (nb + 7) / 8;
    const int rem = nb % 8;
    int logical;
    {
        const int xcd = b % 8, slot = b / 8;
        // XCDs [0, rem) own `per` logical blocks, the rest `per - 1` (rem == 0: all `per`)
        const int full = rem == 0 ? 8 : rem;
        logical = (xcd < full) ? xcd * per + slot : full * per + (xcd - full) * (per - 1) + slot;
    }
    const int tz = logical % ntz;
    const int ty = (logical / ntz) % nty;
    const int tx = logical / (ntz * nty);
    if (tx >= nxc) return; // uniform per workgroup

    const int64_t z = (int64_t)tz * 62 - 1 + lane;
    const int64_t y = (int64_t)ty * (TY - 2) - 1 + row;
    const int64_t xa = (int64_t)tx * xchunk;
    const int64_t xb = min(xa + xchunk, Ox);
    const bool pos_in = (z >= 0 && z < Nz && y >= 0 && y < Ny);
    const bool writer = lane >= 1 && lane <= 62 && row >= 1 && row <= TY - 2 && z < Oz && y < Oy;
    const bool has_y1 = (y + 1 < Ny), has_z1 = (z + 1 < Nz);
    const double hy = pos_in ? hys[y] : 0.0;
    const double hz = pos_in ? hzs[z] : 0.0;
    const int64_t plane = Ny * Nz;
    const int64_t lin_yz = y * Nz + z;

    auto load = [&](int64_t q, double W[3], double G[3], double& hxq) {
        if (pos_in && q >= 0 && q < Nx) {
            const int64_t lin = q * plane + lin_yz;
            W[0] = w[lin * 3 + 0];
            W[1] = w[lin * 3 + 1];
            W[2] = w[lin * 3 + 2];
            if (DIAG) {
                G[0] = g[lin * 3 + 0];
                G[1] = g[lin * 3 + 1];
                G[2] = g[lin * 3 + 2];
            } else {
                G[0] = G[1] = G[2] = 1.0;
            }
            hxq = hxs[q];
        } else {
            W[0] = W[1] = W[2] = 0.0;
            G[0] = G[1] = G[2] = 0.0; // outside: zero coefficients (zero-extension)
            hxq = 0.0;
        }
    };
    auto coefs_of = [&](const double G[3], double hxq, bool valid) {
        if (!valid) return Coefs{0, 0, 0, 0, 0, 0, 0};
        return node_coefs(G[0], G[1], G[2], hxq, hy, hz);
    };

    // state: c = plane p, n = plane p+1, f = prefetched plane p+2
    double Wc[3], Wn[3], Wf[3], Gc[3], Gn[3], Gf[3], hc, hn, hf;
    int64_t p = xa - 1;
    load(p, Wc, Gc, hc);
    load(p + 1, Wn, Gn, hn);
    Coefs Kc = coefs_of(Gc, hc, pos_in && p >= 0 && p < Nx);
    Coefs Kn = coefs_of(Gn, hn, pos_in && p + 1 < Nx);
    double etaPrev[3] = {0.0, 0.0, 0.0};
    // ⋆ω(p) own: YZ, XZ, XY with signs
    double epc0 = Kc.b1 * Wc[0], epc1 = -Kc.b2 * Wc[1], epc2 = Kc.b3 * Wc[2];
    (void)epc1;
    (void)epc2;
    double dwc = 0.0;
    int buf = 0;

    for (; p < xb; ++p) {
        load(p + 2, Wf, Gf, hf); // prefetch, consumed at the end of the step

        // ⋆ω(p+1) own
        const double epn0 = Kn.b1 * Wn[0], epn1 = -Kn.b2 * Wn[1], epn2 = Kn.b3 * Wn[2];
        s.w[0][row][lane] = Wc[0];
        s.w[1][row][lane] = Wc[1];
        s.w[2][row][lane] = Wc[2];
        s.ep[0][row][lane] = epn1;
        s.ep[1][row][lane] = epn2;
        __syncthreads();

        // dω(p) own, clamped forward differences
        const double* Wx = (p + 1 < Nx) ? Wn : Wc;
        const double wy0 = has_y1 && row + 1 < TY ? s.w[0][row + 1][lane] : Wc[0];
        const double wy2 = has_y1 && row + 1 < TY ? s.w[2][row + 1][lane] : Wc[2];
        const double wz0 = has_z1 && lane + 1 < 64 ? s.w[0][row][lane + 1] : Wc[0];
        const double wz1 = has_z1 && lane + 1 < 64 ? s.w[1][row][lane + 1] : Wc[1];
        const double Fxy = (Wx[1] - Wc[1]) - (wy0 - Wc[0]);
        const double Fxz = (Wx[2] - Wc[2]) - (wz0 - Wc[0]);
        const double Fyz = (wy2 - Wc[2]) - (wz1 - Wc[1]);
        const double eta0 = Kc.a1 * Fyz, eta1 = -Kc.a2 * Fxz, eta2 = Kc.a3 * Fxy;

        // δω(p+1) own: −⋆(dᵀ ⋆ω) with zero outside
        const double ep_y1 = row > 0 ? s.ep[0][row - 1][lane] : 0.0;
        const double ep_z2 = lane > 0 ? s.ep[1][row][lane - 1] : 0.0;
        const double H = (epn0 - epc0) - (epn1 - ep_y1) + (epn2 - ep_z2);
        const double dwn = -Kn.c * H;

        s.eta[0][row][lane] = eta0;
        s.eta[1][row][lane] = eta1;
        s.eta[2][row][lane] = eta2;
        s.dw[buf ^ 1][row][lane] = dwn;
        __syncthreads();

        // plane p+2 (requested at the top of this step) is consumed before
        // the stores, so waiting for it never waits for them
        const Coefs Knext = coefs_of(Gf, hf, pos_in && p + 2 >= 0 && p + 2 < Nx);
        asm volatile("" ::"v"(Wf[0]), "v"(Wf[1]), "v"(Wf[2]));

        if (p >= xa && writer) {
            const double ey0 = s.eta[0][row - 1][lane], ey2 = s.eta[2][row - 1][lane];
            const double ez0 = s.eta[0][row][lane - 1], ez1 = s.eta[1][row][lane - 1];
            const double Gxy = (eta1 - etaPrev[1]) - (eta0 - ey0);
            const double Gxz = (eta2 - etaPrev[2]) - (eta0 - ez0);
            const double Gyz = (eta2 - ey2) - (eta1 - ez1);
            const double ddx = (p + 1 < Nx) ? dwn - dwc : 0.0;
            const double ddy = has_y1 ? s.dw[buf][row + 1][lane] - dwc : 0.0;
            const double ddz = has_z1 ? s.dw[buf][row][lane + 1] - dwc : 0.0;
            double* o = out + ((p * Oy + y) * Oz + z) * 3;
            o[0] = Kc.a1 * Gyz + ddx;
            o[1] = -Kc.a2 * Gxz + ddy;
            o[2] = Kc.a3 * Gxy + ddz;
        }

        // rotate
        etaPrev[0] = eta0;
        etaPrev[1] = eta1;
        etaPrev[2] = eta2;
        epc0 = epn0;
        dwc = dwn;
        buf ^= 1;
        Wc[0] = Wn[0];
        Wc[1] = Wn[1];
        Wc[2] = Wn[2];
        Kc = Kn;
        Wn[0] = Wf[0];
        Wn[1] = Wf[1];
        Wn[2] = Wf[2];
        Kn = Knext;
    }
}


// ------------------------------------------ 3D 1-form, pipelined march ---

constexpr int kMaxXChunk = 256;

// three consecutive doubles at byte offset off (16 + 8 byte requests)
template <int AUX = 0>
__device__ __forceinline__ void load3(__amdgpu_buffer_rsrc_t r, uint32_t off, double v[3])
{
    const U32x4 a = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, AUX);
    const U32x2 b = __builtin_amdgcn_raw_buffer_load_b64(r, off + 16, 0, AUX);
    v[0] = as_double(a[0], a[1]);
    v[1] = as_double(a[2], a[3]);
    v[2] = as_double(b[0], b[1]);
}

// AUX: the cache policy of a buffer access (0 default, 2 non-temporal)
template <int AUX = 0>
__device__ __forceinline__ void store3(__amdgpu_buffer_rsrc_t r, uint32_t off, double a, double b, double c)
{
    const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b),
                   uc = __builtin_bit_cast(uint64_t, c);
    const U32x4 lo = {static_cast<unsigned>(ua), static_cast<unsigned>(ua >> 32), static_cast<unsigned>(ub),
                      static_cast<unsigned>(ub >> 32)};
    const U32x2 hi = {static_cast<unsigned>(uc), static_cast<unsigned>(uc >> 32)};
    __builtin_amdgcn_raw_buffer_store_b128(lo, r, off, 0, AUX);
    __builtin_amdgcn_raw_buffer_store_b64(hi, r, off + 16, 0, AUX);
}

// A copy into a register distinct from the source.  A loaded ω value lives
// three steps (in flight, plane p+1, plane p); moving it out of its load
// registers when it is consumed lets the next load reuse them, so the loop
// unrolled by two needs no register copies of data still in flight (a copy
// would make the wavefront wait for it).
__device__ __forceinline__ double moved(double x)
{
    double r;
    asm volatile("v_mov_b64 %0, %1" : "=&v"(r) : "v"(x));
    return r;
}

// f(integral_constant<int, 0>) … f(integral_constant<int, N−1>), unrolled
template <int N, class F>
__device__ __forceinline__ void unrolled(F&& f)
{
    [&]<int... K>(std::integer_sequence<int, K...>) {
        (f(std::integral_constant<int, K>{}), ...);
    }(std::make_integer_sequence<int, N>{});
}

template <int TY, int ZW>
struct PipeLds {
    double w[3][TY][ZW];   // ω(p) for the y+1 / z+1 neighbours of dω
    double ep[2][TY][ZW];  // ⋆ω(p+1): XZ and XY components for δω's y−1 / z−1 terms
    double eta[3][TY][ZW]; // ⋆dω(p) for dᵀ's y−1 / z−1 terms
    double dw[2][TY][ZW];  // δω ping-pong: planes p and p+1
    double hx[kMaxXChunk + 4]; // x edge lengths of planes xa−1 … xb+2 (0 outside)
};

// Number of tiles of `width` nodes along an axis with n nodes and o outputs.
// A tile's first and last nodes are halos (their outputs belong to the
// neighbouring tiles) except at the grid's ends, where no neighbour exists:
// tile t starts at node t·(width−2).
__host__ __device__ inline int64_t halo_tiles(int64_t n, int64_t o, int width)
{
    const int64_t covered = o - 1 - (o == n ? 1 : 0);
    return std::max<int64_t>(1, (covered + width - 3) / (width - 2));
}

// The march of k_lap3d_march with every memory access made branch-free
// through buffer resources, so the compiler's wait counts follow the data:
// plane p+1+D is requested at the top of step p and consumed D−1 steps
// later, D planes stay in flight per wavefront, and a step never waits for
// its predecessors' stores.  A workgroup's resources cover only the planes it
// reads (xa−1 … xb) and writes (xa … xb−1), which the host keeps below 2 GiB.
//
// A workgroup is TY rows (y) of ZW lanes (z, ZW/64 wavefronts).  Lines of
// ZW ≥ Nz nodes read whole contiguous z-lines, and consecutive rows are
// adjacent in memory, so each plane of a tile is one contiguous block.
template <int TY, int ZW, bool DIAG, int D, int SPOL = 0>
__global__ __launch_bounds__(ZW * TY) void k_lap3d_pipe(const double* __restrict__ w, const double* __restrict__ g,
                                                        const double* __restrict__ hxs, const double* __restrict__ hys,
                                                        const double* __restrict__ hzs, double* __restrict__ out,
                                                        int64_t Nx, int64_t Ny, int64_t Nz, int64_t Ox, int64_t Oy,
                                                        int64_t Oz, int ntz, int nty, int nxc, int xchunk)
{
    __shared__ PipeLds<TY, ZW> s;
    const int lane = threadIdx.x % ZW;
    const int row = threadIdx.x / ZW;

    // XCD-contiguous tile order: hardware places workgroup b on XCD b % 8.
    const int nb = gridDim.x;
    const int b = blockIdx.x;
    const int per = (nb + 7) / 8;
    const int rem = nb % 8;
    int logical;
    {
        const int xcd = b % 8, slot = b / 8;
        const int full = rem == 0 ? 8 : rem;
        logical = (xcd < full) ? xcd * per + slot : full * per + (xcd - full) * (per - 1) + slot;
    }
    const int tz = logical % ntz;
    const int ty = (logical / ntz) % nty;
    const int tx = logical / (ntz * nty);
    if (tx >= nxc) return; // uniform per workgroup

    const int64_t z = (int64_t)tz * (ZW - 2) + lane;
    const int64_t y = (int64_t)ty * (TY - 2) + row;
    const int64_t xa = (int64_t)tx * xchunk;
    const int64_t xb = min(xa + xchunk, Ox);
    const bool pos_in = (z < Nz && y < Ny);
    // a lane writes unless it is a halo: the first / last of its row or
    // column, except at the grid's first node and last node
    const bool wz = (lane >= 1 || z == 0) && (lane <= ZW - 2 || z == Nz - 1) && z < Oz;
    const bool wy = (row >= 1 || y == 0) && (row <= TY - 2 || y == Ny - 1) && y < Oy;
    const bool writer = wz && wy;
    const bool has_y1 = (y + 1 < Ny), has_z1 = (z + 1 < Nz);
    const double hy = pos_in ? hys[y] : 0.0;
    const double hz = pos_in ? hzs[z] : 0.0;
    const int64_t x0 = xa - 1; // plane of s.hx[0]

    // planes read: qa … qe−1; planes written: xa … xb−1
    const int64_t qa = max(x0, (int64_t)0), qe = min(xb + 1, Nx);
    const int64_t plane_bytes = Ny * Nz * 24;
    const __amdgpu_buffer_rsrc_t rw = buffer_of(w + qa * Ny * Nz * 3, (qe - qa) * plane_bytes);
    const __amdgpu_buffer_rsrc_t rg = buffer_of(DIAG ? g + qa * Ny * Nz * 3 : w, DIAG ? (qe - qa) * plane_bytes : 0);
    const __amdgpu_buffer_rsrc_t ro = buffer_of(out + xa * Oy * Oz * 3, (xb - xa) * Oy * Oz * 24);
    const uint32_t yz_bytes = pos_in ? static_cast<uint32_t>((y * Nz + z) * 24) : 0u;
    const uint32_t oyz_bytes = writer ? static_cast<uint32_t>((y * Oz + z) * 24) : 0u;
    const uint32_t oplane_bytes = static_cast<uint32_t>(Oy * Oz * 24);

    for (int i = threadIdx.x; i < (int)(xb - xa) + 4; i += ZW * TY) {
        const int64_t q = x0 + i;
        s.hx[i] = (q >= 0 && q < qe) ? hxs[q] : 0.0;
    }

    // byte offset of this lane's node in plane q (outside: beyond the range)
    auto offset = [&](int64_t q) {
        const bool ok = pos_in && q >= qa && q < qe;
        return ok ? static_cast<uint32_t>((q - qa) * plane_bytes) + yz_bytes : kOutside;
    };
    auto load = [&](int64_t q, double W[3], double G[3]) {
        const uint32_t off = offset(q);
        load3(rw, off, W);
        if (DIAG) {
            load3(rg, off, G);
        } else {
            G[0] = G[1] = G[2] = 1.0;
        }
    };
    // nodes outside the grid or the workgroup's planes have zero edge
    // lengths in the tables, hence zero coefficients; their ω reads as zero
    auto coefs_of = [&](const double G[3], int64_t q) { return node_coefs(G[0], G[1], G[2], s.hx[q - x0], hy, hz); };

    // planes: c = p, n = p+1 (ready); in-flight register set j holds plane
    // p+1+j for j = 1 … D−1 at the top of a step, set 0 is requested in it
    double Wc[3], Wn[3], Gc[3], Gn[3], WF[D][3], GF[D][3];
    int64_t p = x0;
    load(p, Wc, Gc);
    load(p + 1, Wn, Gn);
    unrolled<D - 1>([&](auto J) { load(p + 2 + J(), WF[J() + 1], GF[J() + 1]); });
    __syncthreads(); // s.hx
    Coefs Kc = coefs_of(Gc, p);
    Coefs Kn = coefs_of(Gn, p + 1);
    double etaPrev[3] = {0.0, 0.0, 0.0};
    double epc0 = Kc.b1 * Wc[0]; // ⋆ω(p) YZ component
    double dwc = 0.0;

    // neighbour indices inside the tile; reads at a clamped index are
    // replaced by the boundary value afterwards
    const int ry = row + 1 < TY ? row + 1 : row, lz = lane + 1 < ZW ? lane + 1 : lane;
    const int rm = row > 0 ? row - 1 : row, lm = lane > 0 ? lane - 1 : lane;
    const bool ny = has_y1 && row + 1 < TY, nz = has_z1 && lane + 1 < ZW;

    // step p in the k-th position of an unrolled group of D: requests plane
    // p+1+D into set k and consumes plane p+2 from set (k+1) mod D
    auto step = [&](auto K) {
        constexpr int k = K(), u = (K() + 1) % D;
        const int buf = static_cast<int>(p - x0) & 1;
        load(p + 1 + D, WF[k], GF[k]);

        // ⋆ω(p+1) own
        const double epn0 = Kn.b1 * Wn[0], epn1 = -Kn.b2 * Wn[1], epn2 = Kn.b3 * Wn[2];
        s.w[0][row][lane] = Wc[0];
        s.w[1][row][lane] = Wc[1];
        s.w[2][row][lane] = Wc[2];
        s.ep[0][row][lane] = epn1;
        s.ep[1][row][lane] = epn2;
        __syncthreads();

        // dω(p) own, clamped forward differences
        const bool has_x1 = p + 1 < Nx;
        const double wx1 = has_x1 ? Wn[1] : Wc[1], wx2 = has_x1 ? Wn[2] : Wc[2];
        const double sy0 = s.w[0][ry][lane], sy2 = s.w[2][ry][lane];
        const double sz0 = s.w[0][row][lz], sz1 = s.w[1][row][lz];
        const double wy0 = ny ? sy0 : Wc[0];
        const double wy2 = ny ? sy2 : Wc[2];
        const double wz0 = nz ? sz0 : Wc[0];
        const double wz1 = nz ? sz1 : Wc[1];
        const double Fxy = (wx1 - Wc[1]) - (wy0 - Wc[0]);
        const double Fxz = (wx2 - Wc[2]) - (wz0 - Wc[0]);
        const double Fyz = (wy2 - Wc[2]) - (wz1 - Wc[1]);
        const double eta0 = Kc.a1 * Fyz, eta1 = -Kc.a2 * Fxz, eta2 = Kc.a3 * Fxy;

        // δω(p+1) own: −⋆(dᵀ ⋆ω); neighbours below the tile read as zero
        // (outside the grid, or a halo whose output is not stored)
        const double sep0 = s.ep[0][rm][lane], sep1 = s.ep[1][row][lm];
        const double ep_y1 = row > 0 ? sep0 : 0.0;
        const double ep_z2 = lane > 0 ? sep1 : 0.0;
        const double H = (epn0 - epc0) - (epn1 - ep_y1) + (epn2 - ep_z2);
        const double dwn = -Kn.c * H;

        s.eta[0][row][lane] = eta0;
        s.eta[1][row][lane] = eta1;
        s.eta[2][row][lane] = eta2;
        s.dw[buf ^ 1][row][lane] = dwn;
        __syncthreads();

        // plane p+2, requested D−1 steps ago
        const Coefs Knn = coefs_of(GF[u], p + 2);

        // Δω(p); every lane computes it, writers store it
        const double sey0 = s.eta[0][rm][lane], sey2 = s.eta[2][rm][lane];
        const double sez0 = s.eta[0][row][lm], sez1 = s.eta[1][row][lm];
        const double ey0 = row > 0 ? sey0 : 0.0, ey2 = row > 0 ? sey2 : 0.0;
        const double ez0 = lane > 0 ? sez0 : 0.0, ez1 = lane > 0 ? sez1 : 0.0;
        const double Gxy = (eta1 - etaPrev[1]) - (eta0 - ey0);
        const double Gxz = (eta2 - etaPrev[2]) - (eta0 - ez0);
        const double Gyz = (eta2 - ey2) - (eta1 - ez1);
        const double sdy = s.dw[buf][ry][lane], sdz = s.dw[buf][row][lz];
        const double ddx = has_x1 ? dwn - dwc : 0.0;
        const double ddy = has_y1 ? sdy - dwc : 0.0;
        const double ddz = has_z1 ? sdz - dwc : 0.0;
        const bool store = writer && p >= xa;
        const uint32_t ooff = store ? static_cast<uint32_t>(p - xa) * oplane_bytes + oyz_bytes : kOutside;
        store3<SPOL ? 2 : 0>(ro, ooff, Kc.a1 * Gyz + ddx, -Kc.a2 * Gxz + ddy, Kc.a3 * Gxy + ddz);

        // rotate
        etaPrev[0] = eta0;
        etaPrev[1] = eta1;
        etaPrev[2] = eta2;
        epc0 = epn0;
        dwc = dwn;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Wc[c] = Wn[c];
            Wn[c] = moved(WF[u][c]);
        }
        Kc = Kn;
        Kn = Knn;
        ++p;
    };

    // drain the prologue's loads once (vmcnt(0)) so the loop's wait counts
    // depend only on the loop's own memory operations
    __builtin_amdgcn_s_waitcnt(0x0F70);
    // steps p = xa−1 … xb−1 in groups of D, then the remainder
    const int steps = static_cast<int>(xb - x0);
    int done = 0;
    for (; done + D <= steps; done += D) unrolled<D>(step);
    const int rest = steps - done;
    unrolled<D - 1>([&](auto K) {
        if (K() < rest) step(K);
    });
}

// ----------------------------------- 3D 1-form, staged stores and DPP z ---

// values of the neighbouring lanes of a 64-lane wavefront through DPP wave
// shifts: from_up(v) is lane l+1's v (lane 63 keeps its own), from_down(v)
// lane l−1's v (lane 0 reads zero)
__device__ __forceinline__ double from_up(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const int lo = static_cast<int>(static_cast<uint32_t>(u)), hi = static_cast<int>(static_cast<uint32_t>(u >> 32));
    const int l2 = __builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xF, 0xF, false); // wave_shl:1
    const int h2 = __builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xF, 0xF, false);
    return __builtin_bit_cast(double, static_cast<uint64_t>(static_cast<uint32_t>(l2)) |
                                          (static_cast<uint64_t>(static_cast<uint32_t>(h2)) << 32));
}

__device__ __forceinline__ double from_down(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const int lo = static_cast<int>(static_cast<uint32_t>(u)), hi = static_cast<int>(static_cast<uint32_t>(u >> 32));
    const int l2 = __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xF, 0xF, false); // wave_shr:1
    const int h2 = __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xF, 0xF, false);
    return __builtin_bit_cast(double, static_cast<uint64_t>(static_cast<uint32_t>(l2)) |
                                          (static_cast<uint64_t>(static_cast<uint32_t>(h2)) << 32));
}

// LDS ordering of a wavefront's own accesses around a cross-lane exchange
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int TY>
struct StageLds {
    // per row: ω_x(p), ω_z(p) (y+1 neighbours of dω) and ⋆ω_XZ(p+1) (y−1
    // neighbour of δω); after the step's second barrier the row's 192
    // doubles are free and stage the row's output for contiguous stores
    double wep[TY][3][64];
    double eta[2][TY][64]; // ⋆dω(p) components 0 and 2: y−1 neighbours
    double dw[2][TY][64];  // δω ping-pong: planes p and p+1 (y+1 neighbour)
    double hx[kMaxXChunk + 4];
};

// The march of k_lap3d_pipe (64-lane rows, one wavefront each) with the z
// neighbours exchanged inside the wavefront by DPP wave shifts — LDS only
// carries the y neighbours — and every row's output written as contiguous
// 8-byte words: the row's 62 output nodes (186 doubles) are staged in LDS in
// node order and stored by three instructions that each cover 512
// consecutive bytes, instead of 24-byte records at a 24-byte lane stride.
template <int TY, bool DIAG, int D, int SPOL = 0>
__global__ __launch_bounds__(64 * TY) void k_lap3d_stage(const double* __restrict__ w, const double* __restrict__ g,
                                                         const double* __restrict__ hxs, const double* __restrict__ hys,
                                                         const double* __restrict__ hzs, double* __restrict__ out,
                                                         int64_t Nx, int64_t Ny, int64_t Nz, int64_t Ox, int64_t Oy,
                                                         int64_t Oz, int ntz, int nty, int nxc, int xchunk)
{
    constexpr int ZW = 64;
    __shared__ StageLds<TY> s;
    const int lane = threadIdx.x % ZW;
    const int row = threadIdx.x / ZW;

    const int nb = gridDim.x;
    const int b = blockIdx.x;
    const int per = (nb + 7) / 8;
    const int rem = nb % 8;
    int logical;
    {
        const int xcd = b % 8, slot = b / 8;
        const int full = rem == 0 ? 8 : rem;
        logical = (xcd < full) ? xcd * per + slot : full * per + (xcd - full) * (per - 1) + slot;
    }
    const int tz = logical % ntz;
    const int ty = (logical / ntz) % nty;
    const int tx = logical / (ntz * nty);
    if (tx >= nxc) return; // uniform per workgroup

    const int64_t z0 = (int64_t)tz * (ZW - 2);
    const int64_t z = z0 + lane;
    const int64_t y = (int64_t)ty * (TY - 2) + row;
    const int64_t xa = (int64_t)tx * xchunk;
    const int64_t xb = min(xa + xchunk, Ox);
    const bool pos_in = (z < Nz && y < Ny);
    const bool wy = (row >= 1 || y == 0) && (row <= TY - 2 || y == Ny - 1) && y < Oy;
    const bool has_y1 = (y + 1 < Ny), has_z1 = (z + 1 < Nz);
    const double hy = pos_in ? hys[y] : 0.0;
    const double hz = pos_in ? hzs[z] : 0.0;
    const int64_t x0 = xa - 1;

    const int64_t qa = max(x0, (int64_t)0), qe = min(xb + 1, Nx);
    const int64_t plane_bytes = Ny * Nz * 24;
    const __amdgpu_buffer_rsrc_t rw = buffer_of(w + qa * Ny * Nz * 3, (qe - qa) * plane_bytes);
    const __amdgpu_buffer_rsrc_t rg = buffer_of(DIAG ? g + qa * Ny * Nz * 3 : w, DIAG ? (qe - qa) * plane_bytes : 0);
    const __amdgpu_buffer_rsrc_t ro = buffer_of(out + xa * Oy * Oz * 3, (xb - xa) * Oy * Oz * 24);
    const uint32_t yz_bytes = pos_in ? static_cast<uint32_t>((y * Nz + z) * 24) : 0u;
    const uint32_t oplane_bytes = static_cast<uint32_t>(Oy * Oz * 24);
    // the staged output row: doubles j = 64 k + lane of the row's segment
    // (node z0 + j/3), stored when that node is an output of this tile
    uint32_t seg_off[3];
    bool seg_st[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int j = 64 * k + lane;
        const int nl = j / 3;
        const int64_t zn = z0 + nl;
        seg_st[k] = wy && (nl >= 1 || zn == 0) && (nl <= ZW - 2 || zn == Nz - 1) && zn < Oz;
        seg_off[k] = seg_st[k] ? static_cast<uint32_t>((y * Oz + z0) * 24 + 8 * j) : 0u;
    }

    for (int i = threadIdx.x; i < (int)(xb - xa) + 4; i += ZW * TY) {
        const int64_t q = x0 + i;
        s.hx[i] = (q >= 0 && q < qe) ? hxs[q] : 0.0;
    }

    auto offset = [&](int64_t q) {
        const bool ok = pos_in && q >= qa && q < qe;
        return ok ? static_cast<uint32_t>((q - qa) * plane_bytes) + yz_bytes : kOutside;
    };
    auto load = [&](int64_t q, double W[3], double G[3]) {
        const uint32_t off = offset(q);
        load3(rw, off, W);
        if (DIAG) {
            load3(rg, off, G);
        } else {
            G[0] = G[1] = G[2] = 1.0;
        }
    };
    auto coefs_of = [&](const double G[3], int64_t q) { return node_coefs(G[0], G[1], G[2], s.hx[q - x0], hy, hz); };

    double Wc[3], Wn[3], Gc[3], Gn[3], WF[D][3], GF[D][3];
    int64_t p = x0;
    load(p, Wc, Gc);
    load(p + 1, Wn, Gn);
    unrolled<D - 1>([&](auto J) { load(p + 2 + J(), WF[J() + 1], GF[J() + 1]); });
    __syncthreads(); // s.hx
    Coefs Kc = coefs_of(Gc, p);
    Coefs Kn = coefs_of(Gn, p + 1);
    double etaPrev[3] = {0.0, 0.0, 0.0};
    double epc0 = Kc.b1 * Wc[0];
    double dwc = 0.0;

    const int ry = row + 1 < TY ? row + 1 : row;
    const int rm = row > 0 ? row - 1 : row;
    const bool ny = has_y1 && row + 1 < TY, nz = has_z1 && lane + 1 < ZW;
    double* stage = &s.wep[row][0][0];

    auto step = [&](auto K) {
        constexpr int k = K(), u = (K() + 1) % D;
        const int buf = static_cast<int>(p - x0) & 1;
        load(p + 1 + D, WF[k], GF[k]);

        const double epn0 = Kn.b1 * Wn[0], epn1 = -Kn.b2 * Wn[1], epn2 = Kn.b3 * Wn[2];
        s.wep[row][0][lane] = Wc[0];
        s.wep[row][1][lane] = Wc[2];
        s.wep[row][2][lane] = epn1;
        // z neighbours inside the wavefront
        const double sz0 = from_up(Wc[0]), sz1 = from_up(Wc[1]);
        const double ep_z2 = from_down(epn2);
        __syncthreads();

        const bool has_x1 = p + 1 < Nx;
        const double wx1 = has_x1 ? Wn[1] : Wc[1], wx2 = has_x1 ? Wn[2] : Wc[2];
        const double sy0 = s.wep[ry][0][lane], sy2 = s.wep[ry][1][lane];
        const double wy0 = ny ? sy0 : Wc[0];
        const double wy2 = ny ? sy2 : Wc[2];
        const double wz0 = nz ? sz0 : Wc[0];
        const double wz1 = nz ? sz1 : Wc[1];
        const double Fxy = (wx1 - Wc[1]) - (wy0 - Wc[0]);
        const double Fxz = (wx2 - Wc[2]) - (wz0 - Wc[0]);
        const double Fyz = (wy2 - Wc[2]) - (wz1 - Wc[1]);
        const double eta0 = Kc.a1 * Fyz, eta1 = -Kc.a2 * Fxz, eta2 = Kc.a3 * Fxy;

        const double sep0 = s.wep[rm][2][lane];
        const double ep_y1 = row > 0 ? sep0 : 0.0;
        const double H = (epn0 - epc0) - (epn1 - ep_y1) + (epn2 - ep_z2);
        const double dwn = -Kn.c * H;

        s.eta[0][row][lane] = eta0;
        s.eta[1][row][lane] = eta2;
        s.dw[buf ^ 1][row][lane] = dwn;
        const double ez0 = from_down(eta0), ez1 = from_down(eta1);
        const double sdz = from_up(dwc);
        __syncthreads();

        const Coefs Knn = coefs_of(GF[u], p + 2);

        const double sey0 = s.eta[0][rm][lane], sey2 = s.eta[1][rm][lane];
        const double ey0 = row > 0 ? sey0 : 0.0, ey2 = row > 0 ? sey2 : 0.0;
        const double Gxy = (eta1 - etaPrev[1]) - (eta0 - ey0);
        const double Gxz = (eta2 - etaPrev[2]) - (eta0 - ez0);
        const double Gyz = (eta2 - ey2) - (eta1 - ez1);
        const double sdy = s.dw[buf][ry][lane];
        const double ddx = has_x1 ? dwn - dwc : 0.0;
        const double ddy = has_y1 ? sdy - dwc : 0.0;
        const double ddz = has_z1 ? (lane + 1 < ZW ? sdz : dwc) - dwc : 0.0;
        // this row's ω / ⋆ω staging words were last read before the barrier
        // above: the row's output goes through them
        stage[3 * lane + 0] = Kc.a1 * Gyz + ddx;
        stage[3 * lane + 1] = -Kc.a2 * Gxz + ddy;
        stage[3 * lane + 2] = Kc.a3 * Gxy + ddz;
        wave_lds_sync();
        const bool store = p >= xa;
        const uint32_t pbase = store ? static_cast<uint32_t>(p - xa) * oplane_bytes : 0u;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const uint64_t v = __builtin_bit_cast(uint64_t, stage[64 * q + lane]);
            const U32x2 vv = {static_cast<unsigned>(v), static_cast<unsigned>(v >> 32)};
            __builtin_amdgcn_raw_buffer_store_b64(vv, ro, (store && seg_st[q]) ? pbase + seg_off[q] : kOutside, 0,
                                                  SPOL ? 2 : 0);
        }
        wave_lds_sync();

        etaPrev[0] = eta0;
        etaPrev[1] = eta1;
        etaPrev[2] = eta2;
        epc0 = epn0;
        dwc = dwn;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Wc[c] = Wn[c];
            Wn[c] = moved(WF[u][c]);
        }
        Kc = Kn;
        Kn = Knn;
        ++p;
    };

    __builtin_amdgcn_s_waitcnt(0x0F70);
    const int steps = static_cast<int>(xb - x0);
    int done = 0;
    for (; done + D <= steps; done += D) unrolled<D>(step);
    const int rest = steps - done;
    unrolled<D - 1>([&](auto K) {
        if (K() < rest) step(K);
    });
}

// ------------------------------------------------------ 2D 0-form, direct ---

// Δf(n) = −(1/(sx sy)) [ (sy/sx)(n)(f(n+x)−f(n)) − (sy/sx)(n−x)(f(n)−f(n−x))
//                       + (sx/sy)(n)(f(n+y)−f(n)) − (sx/sy)(n−y)(f(n)−f(n−y)) ]
template <bool DIAG>
__global__ __launch_bounds__(256) void k_lap2d_0form(const double* __restrict__ f, const double* __restrict__ g,
                                                     const double* __restrict__ hxs, const double* __restrict__ hys,
                                                     double* __restrict__ out, int64_t Nx, int64_t Ny, int64_t Ox,
                                                     int64_t Oy)
{
    const int64_t total = Ox * Oy;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t % Oy, i = t / Oy;
        // per-node coefficients: r = sy/sx (⋆ X→Y), q = sx/sy (⋆ Y→X magnitude), c = 1/(sx sy)
        auto coef = [&](int64_t a, int64_t b, double& r, double& q, double& c) {
            const int64_t lin = a * Ny + b;
            const double gx = DIAG ? g[lin * 2 + 0] : 1.0, gy = DIAG ? g[lin * 2 + 1] : 1.0;
            const double hx = hxs[a], hy = hys[b];
            const bool mok = gx != 0.0 && gy != 0.0;
            const bool ok1 = mok && fabs(hx) >= 1e-14 && fabs(hy) >= 1e-14;
            const bool ok2 = mok && fabs(hx * hy) >= 1e-14;
            const double sx = sqrt(fabs(gx)) * fabs(hx), sy = sqrt(fabs(gy)) * fabs(hy);
            const double v = sx * sy;
            const double inv = v != 0.0 ? 1.0 / v : 0.0;
            r = ok1 ? sy * sy * inv : 0.0;
            q = ok1 ? sx * sx * inv : 0.0;
            c = ok2 ? inv : 0.0;
        };
        const int64_t ic = min(i, Nx - 1), jc = min(j, Ny - 1);
        const double fc = f[ic * Ny + jc];
        const double fe = f[min(ic + 1, Nx - 1) * Ny + jc];
        const double fn = f[ic * Ny + min(jc + 1, Ny - 1)];
        double r0, q0, c0;
        coef(ic, jc, r0, q0, c0);
        double G = r0 * (fe - fc) + q0 * (fn - fc);
        if (ic > 0) {
            double r1, q1, c1;
            coef(ic - 1, jc, r1, q1, c1);
            G -= r1 * (fc - f[(ic - 1) * Ny + jc]);
        }
        if (jc > 0) {
            double r1, q1, c1;
            coef(ic, jc - 1, r1, q1, c1);
            G -= q1 * (fc - f[ic * Ny + jc - 1]);
        }
        out[i * Oy + j] = -c0 * G;
    }
}

// ------------------------------------------------ 3D coboundary of 0-form ---

__global__ __launch_bounds__(256) void k_cob3d_0form(const double* __restrict__ f, double* __restrict__ out,
                                                     int64_t Nx, int64_t Ny, int64_t Nz)
{
    const int64_t total = Nx * Ny * Nz;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t z = t % Nz, y = (t / Nz) % Ny, x = t / (Nz * Ny);
        const double c = f[t];
        const double fx = x + 1 < Nx ? f[t + Ny * Nz] : c;
        const double fy = y + 1 < Ny ? f[t + Nz] : c;
        const double fz = z + 1 < Nz ? f[t + 1] : c;
        out[t * 3 + 0] = fx - c;
        out[t * 3 + 1] = fy - c;
        out[t * 3 + 2] = fz - c;
    }
}

__global__ __launch_bounds__(256) void k_stream_copy(const double2* __restrict__ in, double2* __restrict__ out,
                                                     int64_t n2)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n2; t += (int64_t)gridDim.x * blockDim.x)
        out[t] = in[t];
}

inline unsigned blocks_for(int64_t total, int block, int64_t cap = 1 << 20)
{
    int64_t b = (total + block - 1) / block;
    if (b > cap) b = cap;
    if (b < 1) b = 1;
    return static_cast<unsigned>(b);
}

template <int TY>
int launch_march(const FusedGeom& g, const double* in, double* out, const int64_t* o, int xchunk,
                  hipStream_t stream)
{
    const int ntz = static_cast<int>((o[2] + 61) / 62);
    const int nty = static_cast<int>((o[1] + TY - 3) / (TY - 2));
    const int nxc = static_cast<int>((o[0] + xchunk - 1) / xchunk);
    const int64_t nb = (int64_t)ntz * nty * nxc;
    SIL_REQUIRE(nb < (1LL << 31), SIL_ERR_UNSUPPORTED, "laplacian: grid too large for one launch");
    if (g.metric)
        hipLaunchKernelGGL((k_lap3d_march<TY, true>), dim3((unsigned)nb), dim3(64 * TY), 0, stream, in, g.metric,
                           g.spacing[0], g.spacing[1], g.spacing[2], out, g.ext[0], g.ext[1], g.ext[2], o[0], o[1],
                           o[2], ntz, nty, nxc, xchunk);
    else
        hipLaunchKernelGGL((k_lap3d_march<TY, false>), dim3((unsigned)nb), dim3(64 * TY), 0, stream, in, g.metric,
                           g.spacing[0], g.spacing[1], g.spacing[2], out, g.ext[0], g.ext[1], g.ext[2], o[0], o[1],
                           o[2], ntz, nty, nxc, xchunk);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

template <int TY, int ZW, int D, int SPOL = 0>
int launch_pipe(const FusedGeom& g, const double* in, double* out, const int64_t* o, int xchunk, hipStream_t stream)
{
    const int64_t ntz = halo_tiles(g.ext[2], o[2], ZW), nty = halo_tiles(g.ext[1], o[1], TY);
    // a workgroup's buffer ranges (xchunk + 1 input planes, xchunk output
    // planes) stay below 2 GiB: shorter chunks for large cross-sections
    const int64_t in_plane = g.ext[1] * g.ext[2] * 24, out_plane = o[1] * o[2] * 24;
    const int64_t fit = std::min(kMaxBufferBytes / in_plane - 1, kMaxBufferBytes / std::max<int64_t>(out_plane, 1));
    SIL_REQUIRE(fit >= 1, SIL_ERR_UNSUPPORTED, "laplacian: a %lld x %lld cross-section exceeds the pipelined kernel",
                (long long)g.ext[1], (long long)g.ext[2]);
    xchunk = static_cast<int>(std::min<int64_t>({(int64_t)xchunk, fit, (int64_t)kMaxXChunk}));
    SIL_REQUIRE(xchunk >= 1, SIL_ERR_INVALID, "laplacian: x chunk %d outside [1, %d]", xchunk, kMaxXChunk);
    const int64_t nxc = (o[0] + xchunk - 1) / xchunk;
    const int64_t nb = ntz * nty * nxc;
    SIL_REQUIRE(nb < (1LL << 31), SIL_ERR_UNSUPPORTED, "laplacian: grid too large for one launch");
    if (g.metric)
        hipLaunchKernelGGL((k_lap3d_pipe<TY, ZW, true, D, SPOL>), dim3((unsigned)nb), dim3(ZW * TY), 0, stream, in, g.metric,
                           g.spacing[0], g.spacing[1], g.spacing[2], out, g.ext[0], g.ext[1], g.ext[2], o[0], o[1],
                           o[2], (int)ntz, (int)nty, (int)nxc, xchunk);
    else
        hipLaunchKernelGGL((k_lap3d_pipe<TY, ZW, false, D, SPOL>), dim3((unsigned)nb), dim3(ZW * TY), 0, stream, in, g.metric,
                           g.spacing[0], g.spacing[1], g.spacing[2], out, g.ext[0], g.ext[1], g.ext[2], o[0], o[1],
                           o[2], (int)ntz, (int)nty, (int)nxc, xchunk);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

template <int TY, int D, int SPOL = 0>
int launch_stage(const FusedGeom& g, const double* in, double* out, const int64_t* o, int xchunk, hipStream_t stream)
{
    const int64_t ntz = halo_tiles(g.ext[2], o[2], 64), nty = halo_tiles(g.ext[1], o[1], TY);
    const int64_t in_plane = g.ext[1] * g.ext[2] * 24, out_plane = o[1] * o[2] * 24;
    const int64_t fit = std::min(kMaxBufferBytes / in_plane - 1, kMaxBufferBytes / std::max<int64_t>(out_plane, 1));
    SIL_REQUIRE(fit >= 1, SIL_ERR_UNSUPPORTED, "laplacian: a %lld x %lld cross-section exceeds the staged kernel",
                (long long)g.ext[1], (long long)g.ext[2]);
    xchunk = static_cast<int>(std::min<int64_t>({(int64_t)xchunk, fit, (int64_t)kMaxXChunk}));
    SIL_REQUIRE(xchunk >= 1, SIL_ERR_INVALID, "laplacian: x chunk %d outside [1, %d]", xchunk, kMaxXChunk);
    const int64_t nxc = (o[0] + xchunk - 1) / xchunk;
    const int64_t nb = ntz * nty * nxc;
    SIL_REQUIRE(nb < (1LL << 31), SIL_ERR_UNSUPPORTED, "laplacian: grid too large for one launch");
    if (g.metric)
        hipLaunchKernelGGL((k_lap3d_stage<TY, true, D, SPOL>), dim3((unsigned)nb), dim3(64 * TY), 0, stream, in,
                           g.metric, g.spacing[0], g.spacing[1], g.spacing[2], out, g.ext[0], g.ext[1], g.ext[2], o[0],
                           o[1], o[2], (int)ntz, (int)nty, (int)nxc, xchunk);
    else
        hipLaunchKernelGGL((k_lap3d_stage<TY, false, D, SPOL>), dim3((unsigned)nb), dim3(64 * TY), 0, stream, in,
                           g.metric, g.spacing[0], g.spacing[1], g.spacing[2], out, g.ext[0], g.ext[1], g.ext[2], o[0],
                           o[1], o[2], (int)ntz, (int)nty, (int)nxc, xchunk);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

}  // namespace

int launch_laplacian_3d_1form(const FusedGeom& g, const double* in, double* out, const int64_t* o,
                              const FusedTile& tile, hipStream_t stream)
{
    if (o[0] * o[1] * o[2] == 0) return SIL_OK;
    if (tile.variant == FUSED_DIRECT) {
        const int block = 256;
        const unsigned nb = blocks_for(o[0] * o[1] * o[2], block);
        if (g.metric)
            hipLaunchKernelGGL((k_lap3d_direct<true>), dim3(nb), dim3(block), 0, stream, in, g.metric, g.spacing[0],
                               g.spacing[1], g.spacing[2], out, g.ext[0], g.ext[1], g.ext[2], o[0], o[1], o[2]);
        else
            hipLaunchKernelGGL((k_lap3d_direct<false>), dim3(nb), dim3(block), 0, stream, in, g.metric, g.spacing[0],
                               g.spacing[1], g.spacing[2], out, g.ext[0], g.ext[1], g.ext[2], o[0], o[1], o[2]);
        SIL_LAUNCH_CHECK();
        return SIL_OK;
    }
    const int xchunk = tile.xchunk > 0 ? tile.xchunk : 32;
    if (tile.variant == FUSED_PIPE) {
        // tile.tz: lanes per row (64, 128 or 256); tile.ty: rows
        const int zw = tile.tz;
        // the 64 × 12 tile streams its output past L2 (non-temporal stores:
        // +1 % on 256³; non-temporal loads lose the halo rows' L2 hits, −18 %)
        if (tile.nt && zw == 64 && tile.ty == 12) return launch_pipe<12, 64, 2, 1>(g, in, out, o, xchunk, stream);
#define SIL_PIPE(ZW, TY) if (zw == ZW && tile.ty == TY) return launch_pipe<TY, ZW, 2>(g, in, out, o, xchunk, stream)
        SIL_PIPE(64, 8);
        SIL_PIPE(64, 12);
        SIL_PIPE(64, 16);
        SIL_PIPE(128, 4);
        SIL_PIPE(128, 8);
        SIL_PIPE(256, 4);
#undef SIL_PIPE
        set_error("laplacian: unsupported pipelined tile %d x %d", zw, tile.ty);
        return SIL_ERR_INVALID;
    }
    if (tile.variant == FUSED_STAGE) {
        // 64-lane rows; tile.ty rows; tile.tz: planes of loads in flight (1–3;
        // other values: 2); tile.nt: non-temporal output stores
        const int depth = (tile.tz >= 1 && tile.tz <= 3) ? tile.tz : 2;
#define SIL_STAGE(TY, D)                                                                     \
    if (tile.ty == TY && depth == D)                                                         \
        return tile.nt ? launch_stage<TY, D, 1>(g, in, out, o, xchunk, stream)               \
                       : launch_stage<TY, D, 0>(g, in, out, o, xchunk, stream)
        SIL_STAGE(8, 2);
        SIL_STAGE(12, 2);
        SIL_STAGE(16, 1);
        SIL_STAGE(16, 2);
        SIL_STAGE(16, 3);
#undef SIL_STAGE
        set_error("laplacian: unsupported staged tile (%d rows, depth %d)", tile.ty, depth);
        return SIL_ERR_INVALID;
    }
    SIL_REQUIRE(tile.variant == FUSED_MARCH, SIL_ERR_INVALID, "laplacian: unknown kernel variant %d", tile.variant);
    switch (tile.ty) {
    case 4: return launch_march<4>(g, in, out, o, xchunk, stream);
    case 8: return launch_march<8>(g, in, out, o, xchunk, stream);
    case 12: return launch_march<12>(g, in, out, o, xchunk, stream);
    case 16: return launch_march<16>(g, in, out, o, xchunk, stream);
    default:
        set_error("laplacian: unsupported tile height %d (4, 8, 12 or 16)", tile.ty);
        return SIL_ERR_INVALID;
    }
}

int launch_laplacian_2d_0form(const FusedGeom& g, const double* in, double* out, int64_t ox, int64_t oy,
                              hipStream_t stream)
{
    if (ox * oy == 0) return SIL_OK;
    const int block = 256;
    const unsigned nb = blocks_for(ox * oy, block);
    if (g.metric)
        hipLaunchKernelGGL((k_lap2d_0form<true>), dim3(nb), dim3(block), 0, stream, in, g.metric, g.spacing[0],
                           g.spacing[1], out, g.ext[0], g.ext[1], ox, oy);
    else
        hipLaunchKernelGGL((k_lap2d_0form<false>), dim3(nb), dim3(block), 0, stream, in, g.metric, g.spacing[0],
                           g.spacing[1], out, g.ext[0], g.ext[1], ox, oy);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int launch_coboundary_3d_0form(const double* in, double* out, const int64_t* ext, hipStream_t stream)
{
    const int64_t total = ext[0] * ext[1] * ext[2];
    if (total == 0) return SIL_OK;
    const int block = 256;
    hipLaunchKernelGGL(k_cob3d_0form, dim3(blocks_for(total, block)), dim3(block), 0, stream, in, out, ext[0], ext[1],
                       ext[2]);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int launch_stream_copy(const double* in, double* out, int64_t n, hipStream_t stream)
{
    SIL_REQUIRE(n % 2 == 0, SIL_ERR_INVALID, "stream copy: even element count required");
    const int block = 256;
    hipLaunchKernelGGL(k_stream_copy, dim3(blocks_for(n / 2, block, 4096 * 8)), dim3(block), 0, stream,
                       reinterpret_cast<const double2*>(in), reinterpret_cast<double2*>(out), n / 2);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

}  // namespace sil
