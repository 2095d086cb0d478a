// Learned preconditioner: a 4-level convolutional U-Net (32 channels, 3×3
// kernels, ReLU) with BF16 weights and activations and FP32 accumulation on
// the gfx950 matrix cores, forward and backward, trained on the GPU.
//
// No reference counterpart.  Layers (NHWC, C = 32, H = W = 256 at level 0):
//   a0 = relu(conv_in(x))          x = [r/rms(r), (log κ − mean)/std]  (2 ch, FP32)
//   e0 = relu(conv(a0))                                    skip 0   (256²)
//   a1, e1 = relu(conv(pool(e0))), relu(conv(a1))          skip 1   (128²)
//   a2, e2 = relu(conv(pool(e1))), relu(conv(a2))          skip 2   (64²)
//   a3, b3 = relu(conv(pool(e2))), relu(conv(a3))                   (32²)
//   c2, f2 = relu(conv(up(b3) + e2)), relu(conv(c2))                (64²)
//   c1, f1 = relu(conv(up(f2) + e1)), relu(conv(c1))                (128²)
//   c0     = relu(conv(up(f1) + e0))                                (256²)
//   out    = conv_out(c0)           (1 ch, FP32, linear)
// The 32→32 convolutions are implicit GEMMs on v_mfma_f32_32x32x16_bf16 in
// persistent workgroups that walk 8×32-pixel output tiles: the (8+2)×(32+2)
// input tile is staged in LDS (pixel stride padded to 40 elements so
// 16-byte fragment reads are bank-conflict free; double-buffered, the next
// tile's loads in flight), the 9·32·32 weights stay in registers as A
// fragments (weights as A, pixels as B, so a lane's results are 16
// channels of one pixel), and each wavefront accumulates two 32-pixel
// rows × 32 output channels over K = 9 taps × 32 input channels.  The 2→32
// input layer runs on two-term BF16 splits of its FP32 operands (fused
// with layer 1 for inference, k_conv_in33); the 32→1 output head reduces
// channels first (k_head).  The input gradient is the same kernel on
// the flipped, channel-transposed weights with the ReLU mask applied while
// the tile is staged; the weight gradient is a second MFMA GEMM over pixels
// with per-workgroup partial sums reduced in a fixed order.
#include <cmath>
#include <cstdint>
#include <new>
#include <type_traits>
#include <vector>

#include "buffer_io.hpp"
#include "mfma_layout.hpp"
#include "unet.hpp"

namespace sil {

namespace {

using unet::PlanarInput;

constexpr int C = 32;      // channels
constexpr int CP = 40;     // padded channel stride in LDS
constexpr int TH = 8;      // output tile rows
constexpr int TW = 32;     // output tile columns
constexpr int NT = 256;    // threads per workgroup

using bf16 = __bf16;

enum StageMode { ST_PLAIN = 0, ST_UPADD = 1, ST_POOL = 2, ST_MASK = 3 };

// persistent convolution workgroups per compute unit: two, one for the
// pooling mode (its prefetch of four raw chunks per staged chunk needs more
// than half of a SIMD's registers)
constexpr int conv_wgs_per_cu(int mode) { return mode == ST_POOL ? 1 : 2; }
constexpr long kComputeUnits = 256;

struct alignas(16) Chunk8 {
    bf16 v[8];
};

__device__ __forceinline__ Chunk8 zero_chunk()
{
    Chunk8 c;
#pragma unroll
    for (int i = 0; i < 8; ++i) c.v[i] = to_bf16(0.f);
    return c;
}

// a D fragment of the transposed products (lane: one pixel, the channels
// d_row(l, r)) after bias, as BF16 with optional ReLU (on the rounded
// pairs), zeros when !keep, into the pixel's LDS row px_row (32 channels)
__device__ __forceinline__ void put_channels(bf16* __restrict__ px_row, const f32x16& acc, bool relu, bool keep)
{
    const int h = (threadIdx.x & 63) >> 5;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        unsigned w0 = pack_bf16x2(acc[4 * g], acc[4 * g + 1]), w1 = pack_bf16x2(acc[4 * g + 2], acc[4 * g + 3]);
        if (relu) {
            w0 = relu_bf16x2(w0);
            w1 = relu_bf16x2(w1);
        }
        *reinterpret_cast<U32x2*>(px_row + 8 * g + 4 * h) = keep ? U32x2{w0, w1} : U32x2{0u, 0u};
    }
}

// Persistent workgroups take tiles by XCD.  The hardware dispatches
// workgroup b to XCD b mod 8, so the workgroups of XCD x share one
// contiguous range of tiles (the tiles split in proportion to each XCD's
// workgroup count) and walk it together, the j-th of them taking
// start + j + k·n (n the XCD's workgroup count).  Neighbouring tiles, which
// read each other's halo rows, then run on one XCD and share its L2.
struct TileWalk {
    int first, stride, end;
};

__device__ __forceinline__ TileWalk xcd_walk(int ntiles)
{
    const int G = gridDim.x, b = blockIdx.x, x = b % 8, j = b / 8;
    auto before = [&](int xx) { return (G / 8) * xx + min(G % 8, xx); }; // workgroups on XCDs < xx
    const int n = G / 8 + (x < G % 8 ? 1 : 0);
    const int start = (int)((int64_t)ntiles * before(x) / G);
    const int end = (int)((int64_t)ntiles * before(x + 1) / G);
    return {start + j, n, end};
}

// ------------------------------------------------------ 3×3 32→32 conv ---

// The input tile of a convolution is staged from one of four sources, so the
// parameter-free layers of the network are never materialised:
//   ST_PLAIN  X                              [B][H][W][32]
//   ST_UPADD  X + up(X2)                     X the skip [B][H][W], X2 [B][H/2][W/2]
//   ST_POOL   2x2 average of X               X [B][2H][2W]
//   ST_MASK   X ⊙ [X2 > 0]                   input gradient: X = dY, X2 = forward output
// Values are rounded to BF16 exactly as a stored tensor would be.
// The raw chunks one staged chunk is made of (the loads of a prefetch), and
// their combination (done when the chunk is written to LDS, so that the
// loads of the next tile stay in flight while the current one is computed).
template <int MODE>
struct RawChunk {
    static constexpr int n = MODE == ST_POOL ? 4 : (MODE == ST_PLAIN ? 1 : 2);
    Chunk8 c[n];
};

template <int MODE>
__device__ __forceinline__ void load_raw(RawChunk<MODE>& r, const bf16* __restrict__ X, const bf16* __restrict__ X2,
                                         int b, int gy, int gx, int part, int H, int W)
{
    if (MODE == ST_POOL) {
        const int W2 = 2 * W;
        const size_t base = (((size_t)b * 2 * H + 2 * gy) * W2 + 2 * gx) * C + part * 8;
        r.c[0] = *reinterpret_cast<const Chunk8*>(&X[base]);
        r.c[1] = *reinterpret_cast<const Chunk8*>(&X[base + C]);
        r.c[2] = *reinterpret_cast<const Chunk8*>(&X[base + (size_t)W2 * C]);
        r.c[3] = *reinterpret_cast<const Chunk8*>(&X[base + (size_t)W2 * C + C]);
        return;
    }
    const size_t off = (((size_t)b * H + gy) * W + gx) * C + part * 8;
    r.c[0] = *reinterpret_cast<const Chunk8*>(&X[off]);
    if (MODE == ST_UPADD) {
        const size_t q = (((size_t)b * (H / 2) + gy / 2) * (W / 2) + gx / 2) * C + part * 8;
        r.c[1] = *reinterpret_cast<const Chunk8*>(&X2[q]);
    } else if (MODE == ST_MASK) {
        r.c[1] = *reinterpret_cast<const Chunk8*>(&X2[off]);
    }
}

template <int MODE>
__device__ __forceinline__ void zero_raw(RawChunk<MODE>& r)
{
#pragma unroll
    for (int i = 0; i < RawChunk<MODE>::n; ++i) r.c[i] = zero_chunk();
}

// the combination on packed words (two BF16 channels per word, the lower
// channel in the low half), each sum formed in FP32 and rounded once
template <int MODE>
__device__ __forceinline__ U32x4 combine_words(const U32x4 (&w)[RawChunk<MODE>::n])
{
    if constexpr (MODE == ST_PLAIN) {
        return w[0];
    } else {
        unsigned o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const unsigned a = w[0][k], b = w[1][k];
            if constexpr (MODE == ST_POOL) {
                const unsigned c = w[2][k], d = w[3][k];
                o[k] = pack_bf16x2(0.25f * (bf16_lo(a) + bf16_lo(b) + bf16_lo(c) + bf16_lo(d)),
                                   0.25f * (bf16_hi(a) + bf16_hi(b) + bf16_hi(c) + bf16_hi(d)));
            } else if constexpr (MODE == ST_UPADD) {
                o[k] = pack_bf16x2(bf16_lo(b) + bf16_lo(a), bf16_hi(b) + bf16_hi(a));
            } else { // ST_MASK: a channel whose mask value is not > 0 becomes +0
                o[k] = a & ((bf16_lo(b) > 0.f ? 0x0000FFFFu : 0u) | (bf16_hi(b) > 0.f ? 0xFFFF0000u : 0u));
            }
        }
        return U32x4{o[0], o[1], o[2], o[3]};
    }
}

__device__ __forceinline__ U32x4 words_of(const Chunk8& c)
{
    unsigned w[4];
    __builtin_memcpy(w, &c, sizeof(w));
    return U32x4{w[0], w[1], w[2], w[3]};
}

__device__ __forceinline__ Chunk8 chunk_of(U32x4 v)
{
    const unsigned w[4] = {v.x, v.y, v.z, v.w};
    Chunk8 c;
    __builtin_memcpy(&c, w, sizeof(c));
    return c;
}

template <int MODE>
__device__ __forceinline__ Chunk8 combine(const RawChunk<MODE>& r)
{
    U32x4 w[RawChunk<MODE>::n];
#pragma unroll
    for (int i = 0; i < RawChunk<MODE>::n; ++i) w[i] = words_of(r.c[i]);
    return chunk_of(combine_words<MODE>(w));
}

template <int MODE>
__device__ __forceinline__ Chunk8 stage_chunk(const bf16* __restrict__ X, const bf16* __restrict__ X2, int b, int gy,
                                              int gx, int part, int H, int W)
{
    RawChunk<MODE> r;
    load_raw<MODE>(r, X, X2, b, gy, gx, part, H, W);
    return combine<MODE>(r);
}

// the (TH+2)×(TW+2) input tile with origin (oy0−1, ox0−1), zero outside
template <int MODE>
__device__ __forceinline__ void stage_tile(bf16* __restrict__ sX, const bf16* __restrict__ X,
                                           const bf16* __restrict__ X2, int b, int oy0, int ox0, int H, int W)
{
    for (int c = threadIdx.x; c < (TH + 2) * (TW + 2) * 4; c += NT) {
        const int pix = c >> 2, part = c & 3;
        const int ty = pix / (TW + 2), tx = pix % (TW + 2);
        const int gy = oy0 - 1 + ty, gx = ox0 - 1 + tx;
        Chunk8 v = zero_chunk();
        if (gy >= 0 && gy < H && gx >= 0 && gx < W) v = stage_chunk<MODE>(X, X2, b, gy, gx, part, H, W);
        *reinterpret_cast<Chunk8*>(&sX[pix * CP + part * 8]) = v;
    }
}

// The raw chunks of one staged chunk through buffer resources over one image
// of X and X2: a chunk outside the image (in == false) reads zeros and needs
// no branch, so the compiler's wait before the chunks are combined counts
// only the loads issued before them.
template <int MODE>
struct RawBuf {
    U32x4 c[RawChunk<MODE>::n];
};

template <int MODE>
__device__ __forceinline__ void load_raw_buf(RawBuf<MODE>& r, __amdgpu_buffer_rsrc_t rx, __amdgpu_buffer_rsrc_t rx2,
                                             int gy, int gx, int part, int W, bool in)
{
    if constexpr (MODE == ST_POOL) {
        const int W2 = 2 * W;
        const uint32_t base = (uint32_t)(((2 * gy) * W2 + 2 * gx) * C + part * 8) * 2;
        const uint32_t row = (uint32_t)W2 * C * 2, px = C * 2;
        r.c[0] = __builtin_amdgcn_raw_buffer_load_b128(rx, in ? base : kOutside, 0, 0);
        r.c[1] = __builtin_amdgcn_raw_buffer_load_b128(rx, in ? base + px : kOutside, 0, 0);
        r.c[2] = __builtin_amdgcn_raw_buffer_load_b128(rx, in ? base + row : kOutside, 0, 0);
        r.c[3] = __builtin_amdgcn_raw_buffer_load_b128(rx, in ? base + row + px : kOutside, 0, 0);
    } else {
        const uint32_t off = (uint32_t)((gy * W + gx) * C + part * 8) * 2;
        r.c[0] = __builtin_amdgcn_raw_buffer_load_b128(rx, in ? off : kOutside, 0, 0);
        if constexpr (MODE == ST_UPADD) {
            const uint32_t q = (uint32_t)(((gy / 2) * (W / 2) + gx / 2) * C + part * 8) * 2;
            r.c[1] = __builtin_amdgcn_raw_buffer_load_b128(rx2, in ? q : kOutside, 0, 0);
        } else if constexpr (MODE == ST_MASK) {
            r.c[1] = __builtin_amdgcn_raw_buffer_load_b128(rx2, in ? off : kOutside, 0, 0);
        }
    }
}

template <int MODE>
__device__ __forceinline__ Chunk8 combine_buf(const RawBuf<MODE>& r)
{
    return chunk_of(combine_words<MODE>(r.c));
}

// bytes of one image of X and of X2 for a staging mode on H×W outputs
__host__ __device__ constexpr int64_t x_image_bytes(int mode, int H, int W)
{
    return (mode == ST_POOL ? 4 : 1) * (int64_t)H * W * C * 2;
}
__host__ __device__ constexpr int64_t x2_image_bytes(int mode, int H, int W)
{
    return mode == ST_UPADD ? (int64_t)(H / 2) * (W / 2) * C * 2 : (mode == ST_MASK ? (int64_t)H * W * C * 2 : 0);
}

// Y = act(conv(staged input) + bias) on H×W images: the forward pass, or
// (ST_MASK, flipped weights, no bias, no ReLU) the input gradient.
//
// Persistent workgroups walk the 8×32-pixel output tiles (xcd_walk: the
// workgroups of one XCD share a contiguous range, each taking every G-th
// tile of it, G their count).  Each lane keeps its B fragments (the weights of its output
// channel, 9 taps × 32 input channels) in registers for the whole walk; the
// input tile is double-buffered in LDS.  One step for tile t: the 36 MFMAs
// per wavefront on the staged input of t; its results into the consumed
// buffer; the raw chunks of tile t + G (loaded during the step before)
// combined into the other buffer; the loads of tile t + 2G issued; then the
// stores of t, every one a 16-byte vector store of 8 channels.  The loads are
// issued before the step's stores and waited for one step later, so neither
// the loads nor the stores of a step stall the next, and all global
// accesses go through buffer resources (no branch between a load and its
// use).
constexpr int kStageChunks = (TH + 2) * (TW + 2) * 4; // 16-byte chunks of one input tile
// threads per workgroup: the pooling mode stages four raw chunks per chunk
// and runs one workgroup per compute unit, so it uses eight wavefronts (one
// output row each) to keep as many in flight as two workgroups of four
constexpr int conv_threads(int mode) { return mode == ST_POOL ? 2 * NT : NT; }
static_assert(TH * TW * 4 % (2 * NT) == 0 && TH * TW % NT == 0, "the output stores of a tile are straight-line");

template <int MODE>
__global__ __launch_bounds__(conv_threads(MODE), conv_wgs_per_cu(MODE)) void k_conv33(const bf16* __restrict__ X, const bf16* __restrict__ X2,
                                               const bf16* __restrict__ Wt, const float* __restrict__ bias,
                                               bf16* __restrict__ Y, int Bmax, const int* __restrict__ dB, int H,
                                               int W, int relu)
{
    constexpr int NTH = conv_threads(MODE), MR = TH / (NTH / 64); // threads; output rows per wavefront
    constexpr int kPerThread = (kStageChunks + NTH - 1) / NTH;
    // images: the first *dB of the batch when a device-side count is given
    const int B = dB ? min(*dB, Bmax) : Bmax;
    __shared__ __attribute__((aligned(16))) bf16 sX[2][(TH + 2) * (TW + 2) * CP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_x = (W + TW - 1) / TW, tiles_y = (H + TH - 1) / TH;
    const int ntiles = tiles_x * tiles_y * B;
    const TileWalk tw = xcd_walk(ntiles);
    const int G = tw.stride, tend = tw.end;
    int tile = tw.first;
    if (tile >= tend) return; // uniform per workgroup

    const int col = Mfma32::a_row(lane); // output column (pixel) of the A fragment
    const int k0 = Mfma32::a_k0(lane);
    const int co_b = Mfma32::b_col(lane);
    bf16x8 bw[9][2];
#pragma unroll
    for (int tap = 0; tap < 9; ++tap)
#pragma unroll
        for (int kh = 0; kh < 2; ++kh)
            bw[tap][kh] = *reinterpret_cast<const bf16x8*>(&Wt[(tap * C + co_b) * C + kh * 16 + k0]);
    // the products are transposed (weights as A, pixels as B): the D
    // fragment of a lane is pixel `col` of its row with the 16 output
    // channels d_row(lane, r), so its results leave as 8-byte runs of four
    // channels
    float bv[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) bv[r] = bias ? bias[Mfma32::d_row(lane, r)] : 0.f;
    const int64_t xb = x_image_bytes(MODE, H, W), x2b = x2_image_bytes(MODE, H, W);

    auto coords = [&](int t, int& b, int& oy0, int& ox0) {
        b = t / (tiles_x * tiles_y);
        oy0 = ((t / tiles_x) % tiles_y) * TH;
        ox0 = (t % tiles_x) * TW;
    };
    // fetch issues the loads of tile t (zeros for t beyond the last tile);
    // put combines them into LDS
    auto fetch = [&](int t, RawBuf<MODE> (&pre)[kPerThread]) {
        int b, oy0, ox0;
        coords(min(t, tend - 1), b, oy0, ox0);
        const __amdgpu_buffer_rsrc_t rx = buffer_of(X + (size_t)b * (xb / 2), xb);
        const __amdgpu_buffer_rsrc_t rx2 = buffer_of(x2b ? X2 + (size_t)b * (x2b / 2) : X, x2b);
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) {
            const int c = tid + k * NTH;
            const int pix = c >> 2, part = c & 3;
            const int gy = oy0 - 1 + pix / (TW + 2), gx = ox0 - 1 + pix % (TW + 2);
            const bool in = t < tend && c < kStageChunks && gy >= 0 && gy < H && gx >= 0 && gx < W;
            load_raw_buf<MODE>(pre[k], rx, rx2, gy, gx, part, W, in);
        }
    };
    auto put = [&](bf16* dst, const RawBuf<MODE> (&pre)[kPerThread]) {
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) {
            const int c = tid + k * NTH;
            if (c < kStageChunks)
                *reinterpret_cast<Chunk8*>(&dst[(c >> 2) * CP + (c & 3) * 8]) = combine_buf<MODE>(pre[k]);
        }
    };

    RawBuf<MODE> pre[kPerThread];
    fetch(tile, pre);
    put(sX[0], pre);
    fetch(tile + G, pre);
    __syncthreads();
    // the weights and bias are in registers before the walk: no wait inside
    // it is for them
    __builtin_amdgcn_s_waitcnt(0x0F70);
    int buf = 0;
    for (; tile < tend; tile += G) {
        f32x16 acc[MR];
#pragma unroll
        for (int m = 0; m < MR; ++m)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[m][r] = 0.f;
        const bf16* cur = sX[buf];
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int ky = tap / 3, kx = tap % 3;
#pragma unroll
            for (int kh = 0; kh < 2; ++kh) {
#pragma unroll
                for (int m = 0; m < MR; ++m) {
                    const int tr = MR * wave + m + ky, tc = col + kx;
                    const bf16x8 ax = *reinterpret_cast<const bf16x8*>(&cur[(tr * (TW + 2) + tc) * CP + kh * 16 + k0]);
                    acc[m] = Mfma32::mma(bw[tap][kh], ax, acc[m]);
                }
            }
        }
        __syncthreads(); // every wavefront is done with the input tile in sX[buf]

        int b, oy0, ox0;
        coords(tile, b, oy0, ox0);
        // results → sX[buf] as [8 rows][32 pixels][32 channels] (stride CP)
        bf16* out_s = sX[buf];
#pragma unroll
        for (int m = 0; m < MR; ++m) {
            f32x16 v;
#pragma unroll
            for (int r = 0; r < 16; ++r) v[r] = acc[m][r] + bv[r];
            put_channels(&out_s[((MR * wave + m) * TW + col) * CP], v, relu != 0, true);
        }
        put(sX[buf ^ 1], pre);
        fetch(tile + 2 * G, pre);
        __syncthreads();

        const __amdgpu_buffer_rsrc_t ry = buffer_of(Y + (size_t)b * H * W * C, (int64_t)H * W * C * 2);
#pragma unroll
        for (int k = 0; k < TH * TW * 4 / NTH; ++k) {
            const int c = tid + k * NTH;
            const int pix = c >> 2, part = c & 3;
            const int gy = oy0 + pix / TW, gx = ox0 + pix % TW;
            const U32x4 v = *reinterpret_cast<const U32x4*>(&out_s[pix * CP + part * 8]);
            __builtin_amdgcn_raw_buffer_store_b128(
                v, ry, (gy < H && gx < W) ? (uint32_t)((gy * W + gx) * C + part * 8) * 2 : kOutside, 0, 0);
        }
        buf ^= 1;
    }
}

// weight gradient: dW[tap][co][ci] += Σ_q dZ[q][co] · X[q + off(tap)][ci],
// dZ = dY ⊙ [Yfwd > 0]; db[co] += Σ_q dZ[q][co].  One partial per workgroup
// (grid-stride over tiles), reduced later in a fixed order.
template <int MODE>
__global__ __launch_bounds__(NT) void k_conv33_wgrad(const bf16* __restrict__ X, const bf16* __restrict__ X2,
                                                     const bf16* __restrict__ dY, const bf16* __restrict__ Yf,
                                                     float* __restrict__ part_w, float* __restrict__ part_b, int B,
                                                     int H, int W)
{
    __shared__ __attribute__((aligned(16))) bf16 sX[(TH + 2) * (TW + 2) * CP];
    __shared__ __attribute__((aligned(16))) bf16 sZ[C * (TH * TW + 8)]; // transposed dZ: [co][pixel]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_x = (W + TW - 1) / TW, tiles_y = (H + TH - 1) / TH;
    const int ntiles = tiles_x * tiles_y * B;
    // taps per wave: 0:{0,4,8} 1:{1,5} 2:{2,6} 3:{3,7}
    f32x16 acc[3];
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][r] = 0.f;
    float bsum = 0.f; // channel tid & 31, pixels strided by 8 rows of threads
    const int zs = TH * TW + 8;

    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int bx = tile % tiles_x, by = (tile / tiles_x) % tiles_y, b = tile / (tiles_x * tiles_y);
        const int oy0 = by * TH, ox0 = bx * TW;
        const size_t img = (size_t)b * H * W * C;
        __syncthreads();
        stage_tile<MODE>(sX, X, X2, b, oy0, ox0, H, W);
        for (int c = tid; c < TH * TW * 4; c += NT) {
            const int pix = c >> 2, part = c & 3;
            const int gy = oy0 + pix / TW, gx = ox0 + pix % TW;
            Chunk8 d = zero_chunk(), y = zero_chunk();
            if (gy < H && gx < W) {
                const size_t off = img + ((size_t)gy * W + gx) * C + part * 8;
                d = *reinterpret_cast<const Chunk8*>(&dY[off]);
                y = *reinterpret_cast<const Chunk8*>(&Yf[off]);
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const float z = from_bf16(y.v[i]) > 0.f ? from_bf16(d.v[i]) : 0.f;
                sZ[(part * 8 + i) * zs + pix] = to_bf16(z);
            }
        }
        __syncthreads();
        {
            const int ch = tid & 31;
            for (int pix = tid >> 5; pix < TH * TW; pix += NT / 32) bsum += from_bf16(sZ[ch * zs + pix]);
        }
        const int co_a = Mfma32::a_row(lane), kk = Mfma32::a_k0(lane);
        const int ci_b = Mfma32::b_col(lane);
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const int tap = wave + 4 * t;
            if (tap > 8) continue;
            const int ky = tap / 3, kx = tap % 3;
            for (int s = 0; s < TH * TW / 16; ++s) {
                const int p0 = 16 * s + kk; // 8 consecutive pixels of one tile row
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sZ[co_a * zs + p0]);
                bf16x8 bb;
                const int row = p0 / TW, c0 = p0 % TW;
#pragma unroll
                for (int i = 0; i < 8; ++i) bb[i] = sX[((row + ky) * (TW + 2) + c0 + i + kx) * CP + ci_b];
                acc[t] = Mfma32::mma(a, bb, acc[t]);
            }
        }
    }
    float* pw = part_w + (size_t)blockIdx.x * 9 * C * C;
    const int ci = Mfma32::d_col(lane);
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const int tap = wave + 4 * t;
        if (tap > 8) continue;
#pragma unroll
        for (int r = 0; r < 16; ++r) pw[(tap * C + Mfma32::d_row(lane, r)) * C + ci] = acc[t][r];
    }
    __shared__ float sb[NT];
    sb[tid] = bsum;
    __syncthreads();
    if (tid < C) {
        float s = 0.f;
        for (int j = tid; j < NT; j += C) s += sb[j];
        part_b[blockIdx.x * C + tid] = s;
    }
}

// sum of `parts` partial vectors of length n: out[i] = Σ_p part[p*n + i]
__global__ void k_reduce_parts(const float* __restrict__ part, int parts, int n, float* __restrict__ out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int p = 0; p < parts; ++p) s += part[(size_t)p * n + i];
        out[i] = s;
    }
}

// -------------------------------------------------- input / output convs ---

// a0 = relu(conv_in(x)): x FP32 [B][H][W][2], weights [tap][co][ci=2].
//
// The input layer on the BF16 matrix cores with FP32 operands carried as
// two-term BF16 sums: v ≈ v_hi + v_lo, v_hi = bf16(v), v_lo = bf16(v − v_hi)
// (v − v_hi is exact in FP32, so the pair holds v to 16 significant bits,
// |v − v_hi − v_lo| ≤ 2⁻¹⁶|v|), and x·w as the four BF16 products
// x_hi·w_hi + x_hi·w_lo + x_lo·w_hi + x_lo·w_lo, each exact in FP32, summed
// by v_mfma_f32_32x32x16_bf16 in FP32.  One tap's two input channels fill
// eight consecutive K rows:
//   B (x, per pixel)   [x0h, x0h, x0l, x0l, x1h, x1h, x1l, x1l]
//   A (w, per channel) [w0h, w0l, w0h, w0l, w1h, w1l, w1h, w1l]
// so an instruction covers two taps (lanes 0–31 tap 2q, lanes 32–63 tap
// 2q + 1, K = 16) and five cover the nine; the upper half of the fifth has
// zero weights (its B is tap 8's again, finite).  x is staged split: one
// 16-byte B row per pixel (split_pair), the tile's row stride XW pixels.
// The weights are A (32 output channels) and the pixels B, so the result
// holds, per lane, pixel l % 32 and the 16 output channels Mfma32::d_row(l,
// r) — four runs of four consecutive channels, each one 8-byte write.
// Accumulation starts at the bias and adds taps (0,1), (2,3), (4,5), (6,7),
// (8) in that order.  Every kernel that forms a0 uses this sequence, so a0
// is the same wherever it is computed.  Nine FP32 instructions of K = 2
// (64 cycles each) did the same work in 576 matrix-core cycles; these five
// take 160.
__device__ __forceinline__ unsigned split_bf16(float v) // hi in the low half, lo in the high half
{
    // v − hi stays a subtraction of the rounded v: fused with a multiply
    // that forms v (the planar input's scale), it would see the unrounded
    // product and give another lo than a stored copy of v does
#pragma clang fp contract(off)
    const bf16 hi = to_bf16(v);
    const bf16 lo = to_bf16(v - from_bf16(hi));
    return (unsigned)__builtin_bit_cast(unsigned short, hi) | ((unsigned)__builtin_bit_cast(unsigned short, lo) << 16);
}

// the staged B row of one x pixel (x0, x1)
__device__ __forceinline__ U32x4 split_pair(float x0, float x1)
{
#pragma clang fp contract(off) // as in split_bf16
    const unsigned h = pack_bf16x2(x0, x1);
    const unsigned l = pack_bf16x2(x0 - bf16_lo(h), x1 - bf16_hi(h));
    const unsigned h0 = h & 0xFFFFu, h1 = h >> 16, l0 = l & 0xFFFFu, l1 = l >> 16;
    return U32x4{h0 | (h0 << 16), l0 | (l0 << 16), h1 | (h1 << 16), l1 | (l1 << 16)};
}

constexpr int kInMfma = 5; // instructions of the input layer per 32-pixel block

// the lane's A operands of the input layer: output channel l % 32, taps 2q + l / 32
__device__ __forceinline__ void conv_in_split_weights(const float* __restrict__ w, bf16x8 (&wi)[kInMfma])
{
    const int lane = threadIdx.x & 63;
    const int co = Mfma32::a_row(lane), h = Mfma32::a_k0(lane) / 8;
#pragma unroll
    for (int q = 0; q < kInMfma; ++q) {
        const int t = 2 * q + h;
        U32x4 v{0u, 0u, 0u, 0u};
        if (t < 9) {
            const unsigned s0 = split_bf16(w[(t * C + co) * 2]), s1 = split_bf16(w[(t * C + co) * 2 + 1]);
            v = U32x4{s0, s0, s1, s1};
        }
        wi[q] = __builtin_bit_cast(bf16x8, v);
    }
}

template <int XW>
__device__ __forceinline__ f32x16 conv_in_mma(const U32x4* __restrict__ sx, int py, int px,
                                              const bf16x8 (&wi)[kInMfma], const f32x16& init)
{
    const int h = Mfma32::b_k0(threadIdx.x & 63) / 8;
    f32x16 acc = init;
#pragma unroll
    for (int q = 0; q < kInMfma; ++q) {
        const int t = min(2 * q + h, 8);
        acc = Mfma32::mma(wi[q], __builtin_bit_cast(bf16x8, sx[(py + t / 3) * XW + px + t % 3]), acc);
    }
    return acc;
}

// the biases of the lane's 16 D rows (four runs of four channels) from a
// 32-channel array in LDS
__device__ __forceinline__ f32x16 row_biases(const float* __restrict__ sb)
{
    const int h = (threadIdx.x & 63) >> 5;
    f32x16 v;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const float4 q = *reinterpret_cast<const float4*>(&sb[8 * g + 4 * h]);
        v[4 * g] = q.x;
        v[4 * g + 1] = q.y;
        v[4 * g + 2] = q.z;
        v[4 * g + 3] = q.w;
    }
    return v;
}

// A workgroup stages the (TH+2)×(TW+2) tile of x around an 8×32-pixel output
// tile; each wavefront computes two 32-pixel rows, and the results go through
// LDS so that every global store is 8 channels (16 bytes).
__global__ __launch_bounds__(NT) void k_conv_in(const float* __restrict__ x, const float* __restrict__ w,
                                                const float* __restrict__ bias, bf16* __restrict__ y, int H, int W,
                                                const int* __restrict__ dB)
{
    __shared__ U32x4 sx[(TH + 2) * (TW + 2)];
    __shared__ __attribute__((aligned(16))) bf16 sY[TH * TW * CP];
    const int b = blockIdx.z, oy0 = blockIdx.y * TH, ox0 = blockIdx.x * TW;
    if (dB && b >= *dB) return; // uniform per workgroup
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < (TH + 2) * (TW + 2); i += NT) {
        const int gy = oy0 - 1 + i / (TW + 2), gx = ox0 - 1 + i % (TW + 2);
        const bool ok = gy >= 0 && gy < H && gx >= 0 && gx < W;
        const size_t q = ((size_t)b * H + gy) * W + gx;
        sx[i] = split_pair(ok ? x[q * 2] : 0.f, ok ? x[q * 2 + 1] : 0.f);
    }
    bf16x8 wi[kInMfma];
    conv_in_split_weights(w, wi);
    float b16[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) b16[r] = bias[Mfma32::d_row(lane, r)];
    __syncthreads();
    const int px = Mfma32::b_col(lane);
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const int row = 2 * wave + m;
        f32x16 init;
#pragma unroll
        for (int r = 0; r < 16; ++r) init[r] = b16[r];
        const f32x16 acc = conv_in_mma<TW + 2>(sx, row, px, wi, init);
        put_channels(&sY[(row * TW + px) * CP], acc, true, true);
    }
    __syncthreads();
    for (int c = tid; c < TH * TW * 4; c += NT) {
        const int pix = c >> 2, part = c & 3;
        const int gy = oy0 + pix / TW, gx = ox0 + pix % TW;
        if (gy < H && gx < W)
            *reinterpret_cast<Chunk8*>(&y[(((size_t)b * H + gy) * W + gx) * C + part * 8]) =
                *reinterpret_cast<const Chunk8*>(&sY[pix * CP + part * 8]);
    }
}

// e0 = relu(conv(relu(conv_in(x)))): the input layer and the first 32→32
// layer in one persistent kernel, for passes that keep no activations (a0 is
// never written).  Per output tile a workgroup stages the (TH+4)×(TW+4) tile
// of x (each pixel's FP32 pair as its split B row), forms a0 on the
// (TH+2)×(TW+2) halo in LDS with conv_in_mma — eleven 32-pixel blocks,
// zero outside the image as the
// layer's zero padding requires — and runs the 32→32 layer on it exactly as
// k_conv33 does (same instructions in the same order), so the output equals
// conv_in followed by k_conv33 bit for bit.
//
// One step for tile t: layer 1 of t (from a0 of t in one LDS buffer) with
// a0 of tile t + G (the workgroup's next tile, xcd_walk) formed into the other buffer in the same
// phase, interleaved in segments so that both matrix chains and the VALU
// epilogues overlap; a barrier; the x tile of t + 2G written to LDS, the
// loads of t + 3G issued, layer 1's outputs of t stored straight from the
// registers; a barrier.  The x tiles are read through a buffer resource
// (pixels outside the image read zero, no branch between a load and its
// use) a whole step before they are written to LDS.
constexpr int XTH = TH + 4, XTW = TW + 4;           // x tile of the fused layers
constexpr int kXPairs = XTH * XTW;                  // 432 x pixels
constexpr int kXPerThread = (kXPairs + NT - 1) / NT;
constexpr int kHalo = (TH + 2) * (TW + 2);          // 340 a0 pixels
constexpr int kHaloBlocks = (kHalo + 31) / 32;      // 11
static_assert(TH * TW * 4 % NT == 0, "the output stores of a tile are straight-line");
static_assert(kHaloBlocks <= 3 * (NT / 64) && (3 * (NT / 64) - 1) * 32 >= kHalo,
              "three blocks per wavefront cover the halo; the twelfth lies wholly past it");

// With PLANAR the input is two FP32 planes read through a slot map (image
// b of the batch is source image src[b]): channel 0 = c0[src[b]] ·
// scale[b], channel 1 = c1[src[b]] — the same values an interleaved input
// formed by one multiply per pixel would hold.  The (src, scale) pair of a
// tile is loaded one step before the tile's pixels, so the address of those
// loads never waits on a memory round trip.
template <bool PLANAR>
__global__ __launch_bounds__(NT, 2) void k_conv_in33(const float* __restrict__ x, PlanarInput pin,
                                                     const float* __restrict__ w_in, const float* __restrict__ b_in,
                                                     const bf16* __restrict__ Wt, const float* __restrict__ bias,
                                                     bf16* __restrict__ Y, int Bmax, const int* __restrict__ dB, int H,
                                                     int W)
{
    const int B = dB ? min(*dB, Bmax) : Bmax;
    __shared__ __attribute__((aligned(16))) bf16 sX[2][kHalo * CP];
    __shared__ U32x4 sx[kXPairs]; // x pixels split for the BF16 products
    __shared__ __attribute__((aligned(16))) float sbias[2][C]; // input layer, layer 1
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_x = (W + TW - 1) / TW, tiles_y = (H + TH - 1) / TH;
    const int ntiles = tiles_x * tiles_y * B;
    const TileWalk tw = xcd_walk(ntiles);
    const int G = tw.stride, tend = tw.end;
    int tile = tw.first;
    if (tile >= tend) return; // uniform per workgroup

    // the biases live in LDS (read per block and per epilogue) to leave the
    // registers to the weights and accumulators
    bf16x8 wi[kInMfma];
    conv_in_split_weights(w_in, wi);
    if (tid < 2 * C) sbias[tid / C][tid % C] = tid < C ? b_in[tid] : bias[tid - C];
    const int col = Mfma32::a_row(lane);
    const int k0 = Mfma32::a_k0(lane);
    const int co_b = Mfma32::b_col(lane);
    bf16x8 bw[9][2];
#pragma unroll
    for (int tap = 0; tap < 9; ++tap)
#pragma unroll
        for (int kh = 0; kh < 2; ++kh)
            bw[tap][kh] = *reinterpret_cast<const bf16x8*>(&Wt[(tap * C + co_b) * C + kh * 16 + k0]);

    auto coords = [&](int t, int& b, int& oy0, int& ox0) {
        b = t / (tiles_x * tiles_y);
        oy0 = ((t / tiles_x) % tiles_y) * TH;
        ox0 = (t % tiles_x) * TW;
    };
    // the x pairs of tile t this thread stages (zero outside the image; a
    // tile beyond the last reads zeros and is never used); xs = the scale of
    // channel 0 (planar input)
    // the (src, scale) pair of tile t as two uniform vector loads
    const __amdgpu_buffer_rsrc_t rsrc = buffer_of(pin.src, PLANAR ? (int64_t)Bmax * 4 : 0);
    const __amdgpu_buffer_rsrc_t rscl = buffer_of(pin.scale, PLANAR ? (int64_t)Bmax * 4 : 0);
    auto meta = [&](int t, U32x2& m) {
        const int b = min(t, tend - 1) / (tiles_x * tiles_y);
        m = U32x2{__builtin_amdgcn_raw_buffer_load_b32(rsrc, (uint32_t)b * 4, 0, 0),
                  __builtin_amdgcn_raw_buffer_load_b32(rscl, (uint32_t)b * 4, 0, 0)};
    };
    auto fetch = [&](int t, U32x2 (&xr)[kXPerThread], float& xs, const U32x2& m) {
        int b, oy0, ox0;
        coords(min(t, tend - 1), b, oy0, ox0); // t may lie beyond the last tile
        if constexpr (PLANAR) {
            const unsigned m0 = m.x, m1 = m.y;
            const int src = __builtin_amdgcn_readfirstlane((int)m0);
            xs = f32_of(m1);
            const __amdgpu_buffer_rsrc_t r0 = buffer_of(pin.c0 + (size_t)src * H * W, (int64_t)H * W * 4);
            const __amdgpu_buffer_rsrc_t r1 = buffer_of(pin.c1 + (size_t)src * H * W, (int64_t)H * W * 4);
#pragma unroll
            for (int k = 0; k < kXPerThread; ++k) {
                const int c = tid + k * NT;
                const int gy = oy0 - 2 + c / XTW, gx = ox0 - 2 + c % XTW;
                const bool in = c < kXPairs && t < tend && gy >= 0 && gy < H && gx >= 0 && gx < W;
                const uint32_t off = in ? (uint32_t)((gy * W + gx) * 4) : kOutside;
                xr[k] = U32x2{__builtin_amdgcn_raw_buffer_load_b32(r0, off, 0, 0),
                              __builtin_amdgcn_raw_buffer_load_b32(r1, off, 0, 0)};
            }
        } else {
            xs = 1.f;
            const __amdgpu_buffer_rsrc_t rs = buffer_of(x + (size_t)b * H * W * 2, (int64_t)H * W * 8);
#pragma unroll
            for (int k = 0; k < kXPerThread; ++k) {
                const int c = tid + k * NT;
                const int gy = oy0 - 2 + c / XTW, gx = ox0 - 2 + c % XTW;
                const bool in = c < kXPairs && t < tend && gy >= 0 && gy < H && gx >= 0 && gx < W;
                xr[k] = __builtin_amdgcn_raw_buffer_load_b64(rs, in ? (uint32_t)((gy * W + gx) * 8) : kOutside, 0, 0);
            }
        }
    };
    auto put_x = [&](const U32x2 (&xr)[kXPerThread], float xs) {
#pragma unroll
        for (int k = 0; k < kXPerThread; ++k) {
            const int c = tid + k * NT;
            if (c < kXPairs) {
                if constexpr (PLANAR)
                    sx[c] = split_pair(f32_of(xr[k].x) * xs, f32_of(xr[k].y));
                else
                    sx[c] = split_pair(f32_of(xr[k].x), f32_of(xr[k].y));
            }
        }
    };
    // a0 of tile t on its halo, from the x tile in sx, into dst
    // block j (32 halo pixels) of a0 of the tile at (oy0, ox0), from the x
    // tile in sx, into dst; the last block runs past the halo: its extra
    // lanes read the last pixel and write nothing
    auto a0_block = [&](int j, int oy0, int ox0, bf16* __restrict__ dst) {
        const int hp = j * 32 + Mfma32::b_col(lane);
        const int ha = min(hp, kHalo - 1);
        const int hy = ha / (TW + 2), hx = ha % (TW + 2);
        const f32x16 acc = conv_in_mma<XTW>(sx, hy, hx, wi, row_biases(sbias[0]));
        const int gy = oy0 - 1 + hy, gx = ox0 - 1 + hx;
        const bool inside = gy >= 0 && gy < H && gx >= 0 && gx < W;
        if (hp < kHalo) put_channels(&dst[hp * CP], acc, true, inside);
    };

    // prologue: a0 of the first tile in sX[0], x of the second tile in sx,
    // x of the third in flight
    U32x2 xr[kXPerThread], mt{0u, 0u};
    float xs;
    if constexpr (PLANAR) meta(tile, mt);
    fetch(tile, xr, xs, mt);
    put_x(xr, xs);
    if constexpr (PLANAR) meta(tile + G, mt);
    fetch(tile + G, xr, xs, mt);
    if constexpr (PLANAR) meta(tile + 2 * G, mt);
    __syncthreads();
    {
        int b, oy0, ox0;
        coords(tile, b, oy0, ox0);
        for (int j = wave; j < kHaloBlocks; j += NT / 64) a0_block(j, oy0, ox0, sX[0]);
    }
    __syncthreads();
    put_x(xr, xs);
    fetch(tile + 2 * G, xr, xs, mt);
    if constexpr (PLANAR) meta(tile + 3 * G, mt);
    __syncthreads();
    // the weights and biases are in registers before the walk: no wait
    // inside it is for them
    __builtin_amdgcn_s_waitcnt(0x0F70);
    int buf = 0;
    for (; tile < tend; tile += G) {
        const int next = tile + G;
        int nb, noy0, nox0;
        coords(min(next, tend - 1), nb, noy0, nox0);

        // one phase of matrix work: layer 1 of this tile from sX[buf], in
        // three segments of three taps, each followed by one a0 block of
        // the next tile (from sx into sX[buf ^ 1]) — independent chains, so
        // the VALU epilogue of a block runs while layer 1's products do.
        // The blocks are unconditional, so the compiler can interleave
        // their products with layer 1's: past the last tile a block forms
        // a0 of the last tile again into the buffer no later phase reads,
        // and block 11 (the third of wavefront 3) has no pixel in the halo
        // and writes nothing.
        f32x16 acc[2];
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[m][r] = 0.f;
        const bf16* cur = sX[buf];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
#pragma unroll
            for (int tap = 3 * q; tap < 3 * q + 3; ++tap) {
                const int ky = tap / 3, kx = tap % 3;
#pragma unroll
                for (int kh = 0; kh < 2; ++kh) {
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        const int tr = 2 * wave + m + ky, tc = col + kx;
                        const bf16x8 ax =
                            *reinterpret_cast<const bf16x8*>(&cur[(tr * (TW + 2) + tc) * CP + kh * 16 + k0]);
                        acc[m] = Mfma32::mma(bw[tap][kh], ax, acc[m]);
                    }
                }
            }
            const int j = wave + (NT / 64) * q;
            a0_block(j, noy0, nox0, sX[buf ^ 1]);
        }
        __syncthreads(); // sX[buf] consumed, a0 of the next tile complete, sx consumed

        put_x(xr, xs);   // x of tile next + G (zeros beyond the last tile)
        fetch(next + 2 * G, xr, xs, mt);
        if constexpr (PLANAR) meta(next + 3 * G, mt);

        // layer 1's outputs straight from the registers: the lane's pixel,
        // four runs of four channels as 8-byte stores (dropped outside the
        // image)
        int b, oy0, ox0;
        coords(tile, b, oy0, ox0);
        {
            const __amdgpu_buffer_rsrc_t ry = buffer_of(Y + (size_t)b * H * W * C, (int64_t)H * W * C * 2);
            const int h = lane >> 5;
            const f32x16 bv = row_biases(sbias[1]);
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const int gy = oy0 + 2 * wave + m, gx = ox0 + col;
                const bool in = gy < H && gx < W;
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const unsigned w0 = relu_bf16x2(
                        pack_bf16x2(acc[m][4 * g] + bv[4 * g], acc[m][4 * g + 1] + bv[4 * g + 1]));
                    const unsigned w1 = relu_bf16x2(
                        pack_bf16x2(acc[m][4 * g + 2] + bv[4 * g + 2], acc[m][4 * g + 3] + bv[4 * g + 3]));
                    __builtin_amdgcn_raw_buffer_store_b64(
                        U32x2{w0, w1}, ry, in ? (uint32_t)(((gy * W + gx) * C + 8 * g + 4 * h) * 2) : kOutside, 0, 0);
                }
            }
        }
        __syncthreads(); // sx holds the x tile of next + G for the next phase
        buf ^= 1;
    }
}

// weight gradient of conv_in: partials per block of dW [9][32][2] and db [32]
__global__ void k_conv_in_wgrad(const float* __restrict__ x, const bf16* __restrict__ dY, const bf16* __restrict__ Yf,
                                float* __restrict__ part, int B, int H, int W)
{
    // thread (co, tap-chunk); loop over a strided set of pixels
    const int co = threadIdx.x & 31;
    const int lanegrp = threadIdx.x >> 5; // 8 groups of pixels
    float acc[9 * 2 + 1];
    for (int i = 0; i < 19; ++i) acc[i] = 0.f;
    const size_t n = (size_t)B * H * W;
    for (size_t p = blockIdx.x * 8 + lanegrp; p < n; p += (size_t)gridDim.x * 8) {
        const float y = from_bf16(Yf[p * C + co]);
        const float dz = y > 0.f ? from_bf16(dY[p * C + co]) : 0.f;
        if (dz == 0.f) continue;
        const int xx = p % W, yy = (p / W) % H;
        const size_t b = p / ((size_t)W * H);
        for (int t = 0; t < 9; ++t) {
            const int gy = yy + t / 3 - 1, gx = xx + t % 3 - 1;
            if (gy < 0 || gy >= H || gx < 0 || gx >= W) continue;
            const size_t q = (b * H + gy) * W + gx;
            acc[t * 2] += dz * x[q * 2];
            acc[t * 2 + 1] += dz * x[q * 2 + 1];
        }
        acc[18] += dz;
    }
    __shared__ float s[256][19];
    for (int i = 0; i < 19; ++i) s[threadIdx.x][i] = acc[i];
    __syncthreads();
    if (threadIdx.x < 32) {
        float* out = part + (size_t)blockIdx.x * (9 * C * 2 + C);
        for (int i = 0; i < 19; ++i) {
            float v = 0.f;
            for (int g = 0; g < 8; ++g) v += s[g * 32 + threadIdx.x][i];
            if (i < 18) out[((i / 2) * C + threadIdx.x) * 2 + (i & 1)] = v;
            else out[9 * C * 2 + threadIdx.x] = v;
        }
    }
}

// head weights w[tap][ci] (FP32) → B matrix [tap][co][ci] (BF16) with
// co = 0: hi = bf16(w), co = 1: lo = bf16(w − hi), other columns zero
// ------------------------------------------------------------ output head ---
// out = conv_out(x): 32 → 1 channel, 3×3, BF16 x, FP32 weights, bias and
// output.
//
// Channels first, taps second.  Per pixel p and tap t the channel sum
// g_t(p) = Σ_c w[t][c]·x[p][c] is one row of a matrix product with the
// weights as A — rows 2t and 2t+1 hold the BF16 hi and lo parts of tap t's
// weights (w ≈ hi + lo to 16 bits), rows 18–31 are zero, K = the 32
// channels — and 32 pixels as B: two v_mfma_f32_32x32x16_bf16 per 32
// pixels.  Then out(q) = b + Σ_t g_t(q + off_t), taps in order, from the g
// of the tile's halo in LDS.  x is read once per halo pixel, straight into
// the B registers through a buffer resource (pixels outside the image read
// zero, the layer's zero padding); no input tile is staged, and the matrix
// work is 1/9 of a 32→32 layer's.
constexpr int HTH = 16, HTW = 32;                       // output tile
constexpr int kHHalo = (HTH + 2) * (HTW + 2);           // 612 halo pixels
constexpr int kHBlocks = (kHHalo + 31) / 32;            // 20 blocks of 32 pixels
constexpr int kHPerWave = kHBlocks / (NT / 64);         // 5
constexpr int kHTaps = 9;                               // g values per halo pixel in LDS
static_assert(kHBlocks % (NT / 64) == 0 && HTH * HTW % NT == 0, "the head's loops are straight-line");

__global__ __launch_bounds__(NT) void k_head(const bf16* __restrict__ x, const bf16* __restrict__ Wa,
                                             const float* __restrict__ bias, float* __restrict__ y, int H, int W,
                                             const int* __restrict__ dB)
{
    __shared__ float sg[kHHalo * kHTaps];
    const int b = blockIdx.z, oy0 = blockIdx.y * HTH, ox0 = blockIdx.x * HTW;
    if (dB && b >= *dB) return; // uniform per workgroup
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = Mfma32::b_col(lane), h = lane >> 5;
    const __amdgpu_buffer_rsrc_t rx = buffer_of(x + (size_t)b * H * W * C, (int64_t)H * W * C * 2);
    // every load of the wavefront's blocks first: lane l supplies channels
    // 16·kh + 8·(l / 32) + 0..7 of pixel l % 32
    U32x4 xr[kHPerWave][2];
#pragma unroll
    for (int i = 0; i < kHPerWave; ++i) {
        const int hp = (wave + (NT / 64) * i) * 32 + n;
        const int gy = oy0 - 1 + hp / (HTW + 2), gx = ox0 - 1 + hp % (HTW + 2);
        const bool in = hp < kHHalo && gy >= 0 && gy < H && gx >= 0 && gx < W;
        const uint32_t off = (uint32_t)((gy * W + gx) * C + Mfma32::b_k0(lane)) * 2;
#pragma unroll
        for (int kh = 0; kh < 2; ++kh)
            xr[i][kh] = __builtin_amdgcn_raw_buffer_load_b128(rx, in ? off + kh * 32 : kOutside, 0, 0);
    }
    bf16x8 wa[2];
#pragma unroll
    for (int kh = 0; kh < 2; ++kh)
        wa[kh] = *reinterpret_cast<const bf16x8*>(&Wa[Mfma32::a_row(lane) * C + kh * 16 + Mfma32::a_k0(lane)]);
    const float b0 = bias[0];
#pragma unroll
    for (int i = 0; i < kHPerWave; ++i) {
        const int hp = (wave + (NT / 64) * i) * 32 + n;
        f32x16 acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = 0.f;
        acc = Mfma32::mma(wa[0], __builtin_bit_cast(bf16x8, xr[i][0]), acc);
        acc = Mfma32::mma(wa[1], __builtin_bit_cast(bf16x8, xr[i][1]), acc);
        // D row 8g + 4h + j is part j & 1 of tap 4g + 2h + (j >> 1)
        if (hp < kHHalo) {
            float* g = &sg[hp * kHTaps];
            g[2 * h] = acc[0] + acc[1];
            g[2 * h + 1] = acc[2] + acc[3];
            g[4 + 2 * h] = acc[4] + acc[5];
            g[5 + 2 * h] = acc[6] + acc[7];
            if (h == 0) g[8] = acc[8] + acc[9];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < HTH * HTW / NT; ++k) {
        const int q = tid + k * NT;
        const int py = q / HTW, px = q % HTW;
        float v = b0;
#pragma unroll
        for (int t = 0; t < 9; ++t) v += sg[((py + t / 3) * (HTW + 2) + px + t % 3) * kHTaps + t];
        const int gy = oy0 + py, gx = ox0 + px;
        if (gy < H && gx < W) y[((size_t)b * H + gy) * W + gx] = v;
    }
}

// the head's A matrix: row 2t + s (t < 9) = part s (0 hi, 1 lo) of tap t's
// 32 weights, rows 18–31 zero
__global__ void k_pack_head(const float* __restrict__ w, bf16* __restrict__ Wa)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C * C) return;
    const int m = i / C, c = i % C, t = m / 2;
    const unsigned split = t < 9 ? split_bf16(w[t * C + c]) : 0u;
    Wa[i] = __builtin_bit_cast(bf16, (unsigned short)(m % 2 == 0 ? split & 0xFFFFu : split >> 16));
}

// backward of conv_out: dX[q][ci] = Σ_t dout[q − off(t)] · w[t][ci] (BF16),
// and per-block partials of dW [9][32] and db
__global__ void k_conv_out_bwd(const bf16* __restrict__ x, const float* __restrict__ dout, const float* __restrict__ w,
                               bf16* __restrict__ dx, float* __restrict__ part, int B, int H, int W)
{
    __shared__ float sw[9 * C];
    for (int i = threadIdx.x; i < 9 * C; i += blockDim.x) sw[i] = w[i];
    __syncthreads();
    const int ci = threadIdx.x & 31, grp = threadIdx.x >> 5;
    float gw[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    float gb = 0.f;
    const size_t n = (size_t)B * H * W;
    for (size_t p = blockIdx.x * 8 + grp; p < n; p += (size_t)gridDim.x * 8) {
        const int xx = p % W, yy = (p / W) % H;
        const size_t b = p / ((size_t)W * H);
        float s = 0.f;
        const float d0 = dout[p];
        if (ci == 0) gb += d0;
        for (int t = 0; t < 9; ++t) {
            const int dy = t / 3 - 1, dx_ = t % 3 - 1;
            // input gradient: q = p, output pixel p − off
            const int oy = yy - dy, ox = xx - dx_;
            if (oy >= 0 && oy < H && ox >= 0 && ox < W) s += dout[(b * H + oy) * W + ox] * sw[t * C + ci];
            // weight gradient: output pixel p, input p + off
            const int iy = yy + dy, ix = xx + dx_;
            if (iy >= 0 && iy < H && ix >= 0 && ix < W) gw[t] += d0 * from_bf16(x[((b * H + iy) * W + ix) * C + ci]);
        }
        dx[p * C + ci] = to_bf16(s);
    }
    __shared__ float sp[256][10];
    for (int t = 0; t < 9; ++t) sp[threadIdx.x][t] = gw[t];
    sp[threadIdx.x][9] = gb;
    __syncthreads();
    if (threadIdx.x < 32) {
        float* out = part + (size_t)blockIdx.x * (9 * C + 1);
        for (int t = 0; t < 9; ++t) {
            float v = 0.f;
            for (int g = 0; g < 8; ++g) v += sp[g * 32 + threadIdx.x][t];
            out[t * C + threadIdx.x] = v;
        }
        if (threadIdx.x == 0) {
            float v = 0.f;
            for (int g = 0; g < 8; ++g) v += sp[g * 32][9];
            out[9 * C] = v;
        }
    }
}

// ----------------------------------------------------- pooling, up-add ---

// dX (fine) += dY (coarse)/4 on each of its 4 children; 8 channels per thread
__global__ void k_pool_bwd(const bf16* __restrict__ dy, bf16* __restrict__ dx, int B, int H, int W)
{
    const size_t n = (size_t)B * H * W * 4;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int part = (int)(i & 3);
        const size_t p = i >> 2;
        const int xx = (int)(p % W), yy = (int)((p / W) % H);
        const size_t b = p / ((size_t)W * H);
        const size_t q = ((b * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + part * 8;
        Chunk8 d = *reinterpret_cast<Chunk8*>(&dx[p * C + part * 8]);
        const Chunk8 c = *reinterpret_cast<const Chunk8*>(&dy[q]);
#pragma unroll
        for (int k = 0; k < 8; ++k) d.v[k] = to_bf16(from_bf16(d.v[k]) + 0.25f * from_bf16(c.v[k]));
        *reinterpret_cast<Chunk8*>(&dx[p * C + part * 8]) = d;
    }
}

// d xc (coarse) = Σ of the 4 children of dY (fine); 8 channels per thread
__global__ void k_up_bwd(const bf16* __restrict__ dy, bf16* __restrict__ dxc, int B, int h, int w)
{
    const int H = 2 * h, W = 2 * w;
    const size_t n = (size_t)B * h * w * 4;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int part = (int)(i & 3);
        const size_t p = i >> 2;
        const int xx = (int)(p % w), yy = (int)((p / w) % h);
        const size_t b = p / ((size_t)w * h);
        const size_t base = ((b * H + 2 * yy) * W + 2 * xx) * C + part * 8;
        const Chunk8 a0 = *reinterpret_cast<const Chunk8*>(&dy[base]);
        const Chunk8 a1 = *reinterpret_cast<const Chunk8*>(&dy[base + C]);
        const Chunk8 a2 = *reinterpret_cast<const Chunk8*>(&dy[base + (size_t)W * C]);
        const Chunk8 a3 = *reinterpret_cast<const Chunk8*>(&dy[base + (size_t)W * C + C]);
        Chunk8 o;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            o.v[k] = to_bf16(from_bf16(a0.v[k]) + from_bf16(a1.v[k]) + from_bf16(a2.v[k]) + from_bf16(a3.v[k]));
        *reinterpret_cast<Chunk8*>(&dxc[p * C + part * 8]) = o;
    }
}

// ------------------------------------------------------------- loss ---

// Direction loss of the preconditioner output N against the error e of the
// residual r = A e (A the Dirichlet-zero 5-point operator with edge weights
// wx, wy): L = 1 − cos²_A(e, N) = 1 − (rᵀN)² / ((eᵀr)(NᵀAN)), the relative
// A-norm error left after the best step along N.  Stage 1 stores AN and the
// per-block partial sums of rᵀN, NᵀAN and eᵀr.
__device__ __forceinline__ float apply5(const float* __restrict__ u, const float* __restrict__ wx,
                                        const float* __restrict__ wy, size_t t, int i, int j, int n)
{
    if (!(i > 0 && j > 0 && i + 1 < n && j + 1 < n)) return 0.f;
    const float uc = u[t];
    const float ue = i + 2 < n ? u[t + n] : 0.f, uw = i > 1 ? u[t - n] : 0.f;
    const float un = j + 2 < n ? u[t + 1] : 0.f, us = j > 1 ? u[t - 1] : 0.f;
    return wx[t] * (uc - ue) + wx[t - n] * (uc - uw) + wy[t] * (uc - un) + wy[t - 1] * (uc - us);
}

template <int NT_>
__device__ __forceinline__ void block_sum3(float v0, float v1, float v2, float* out)
{
    __shared__ float sh[3][NT_];
    sh[0][threadIdx.x] = v0;
    sh[1][threadIdx.x] = v1;
    sh[2][threadIdx.x] = v2;
    __syncthreads();
    for (int st = NT_ / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + st];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + st];
            sh[2][threadIdx.x] += sh[2][threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sh[0][0];
        out[1] = sh[1][0];
        out[2] = sh[2][0];
    }
}

// grid (nbx, B); r is read with stride rs (the network input holds it
// interleaved with the coefficient channel)
__global__ __launch_bounds__(256) void k_cos_loss_stats(const float* __restrict__ N, const float* __restrict__ r,
                                                        int rs, const float* __restrict__ e,
                                                        const float* __restrict__ wx, const float* __restrict__ wy,
                                                        float* __restrict__ AN, float* __restrict__ part, int n)
{
    const int b = blockIdx.y;
    const size_t nn = (size_t)n * n, base = (size_t)b * nn;
    float sa = 0.f, sb = 0.f, se = 0.f;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nn; t += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(t % n), i = (int)(t / n);
        const float an = apply5(N + base, wx + base, wy + base, t, i, j, n);
        AN[base + t] = an;
        const float rv = r[(base + t) * rs];
        sa += rv * N[base + t];
        sb += N[base + t] * an;
        se += e[base + t] * rv;
    }
    block_sum3<256>(sa, sb, se, part + ((size_t)b * gridDim.x + blockIdx.x) * 3);
}

// dN = (2a / (E b B)) (−r + (a/b) AN); loss[b] = 1 − a² / (E b)
__global__ __launch_bounds__(256) void k_cos_loss_grad(const float* __restrict__ r, int rs,
                                                       const float* __restrict__ AN, const float* __restrict__ part,
                                                       int parts, float* __restrict__ dN, float* __restrict__ loss,
                                                       int n, int B)
{
    const int b = blockIdx.y;
    const size_t nn = (size_t)n * n, base = (size_t)b * nn;
    double a = 0.0, bb = 0.0, E = 0.0;
    for (int q = 0; q < parts; ++q) {
        const float* pq = part + ((size_t)b * parts + q) * 3;
        a += pq[0];
        bb += pq[1];
        E += pq[2];
    }
    const bool ok = bb > 0.0 && E > 0.0;
    const float f = ok ? (float)(2.0 * a / (E * bb * B)) : 0.f;
    const float g = ok ? (float)(a / bb) : 0.f;
    if (blockIdx.x == 0 && threadIdx.x == 0) loss[b] = ok ? (float)(1.0 - a * a / (E * bb)) : 1.f;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nn; t += (size_t)gridDim.x * blockDim.x)
        dN[base + t] = f * (-r[(base + t) * rs] + g * AN[base + t]);
}

// ------------------------------------------------------------ Adam ---

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// BF16 copies of one 32→32 layer's weights: forward [tap][co][ci] and the
// input-gradient form W'[t][a][b] = W[8 − t][b][a]
__global__ void k_pack33(const float* __restrict__ w, bf16* __restrict__ wf, bf16* __restrict__ wb)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 9 * C * C; i += gridDim.x * blockDim.x) {
        const int t = i / (C * C), a = (i / C) % C, b = i % C;
        wf[i] = to_bf16(w[i]);
        wb[i] = to_bf16(w[((8 - t) * C + b) * C + a]);
    }
}

__global__ void k_bf16_to_f32(const bf16* __restrict__ a, float* __restrict__ b, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

inline unsigned grid1(size_t n, int block = 256, size_t cap = 8192)
{
    size_t g = (n + block - 1) / block;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace

// -------------------------------------------------------------- host ---

namespace unet {

namespace {
template <class F>
int by_mode(int mode, F&& f)
{
    switch (mode) {
    case ST_PLAIN: return f(std::integral_constant<int, ST_PLAIN>());
    case ST_UPADD: return f(std::integral_constant<int, ST_UPADD>());
    case ST_POOL: return f(std::integral_constant<int, ST_POOL>());
    case ST_MASK: return f(std::integral_constant<int, ST_MASK>());
    default: set_error("unet: unknown staging mode %d", mode); return SIL_ERR_INVALID;
    }
}
}  // namespace

int conv33(int mode, const bf16* X, const bf16* X2, const bf16* Wt, const float* bias, bf16* Y, int B, int H, int W,
           bool relu, hipStream_t s, const int* dB)
{
    // one buffer resource spans one image of each operand
    SIL_REQUIRE(x_image_bytes(mode, H, W) <= kMaxBufferBytes, SIL_ERR_UNSUPPORTED,
                "unet: %dx%d images exceed one buffer resource per image", H, W);
    // persistent grid (the LDS of a workgroup is 54 KB), capped by the
    // number of tiles
    const long ntiles = (long)((W + TW - 1) / TW) * ((H + TH - 1) / TH) * B;
    return by_mode(mode, [&](auto m) -> int {
        constexpr int M = decltype(m)::value;
        const long cap = kComputeUnits * conv_wgs_per_cu(M);
        const unsigned grid = (unsigned)(ntiles < cap ? ntiles : cap);
        hipLaunchKernelGGL((k_conv33<M>), dim3(grid), dim3(conv_threads(M)), 0, s, X, X2, Wt, bias, Y, B, dB, H,
                           W, relu ? 1 : 0);
        SIL_LAUNCH_CHECK();
        return (int)SIL_OK;
    });
}

int conv33_wgrad(int mode, const bf16* X, const bf16* X2, const bf16* dY, const bf16* Yf, float* part_w,
                 float* part_b, int parts, float* dW, float* db, int B, int H, int W, hipStream_t s)
{
    const int ntiles = ((W + TW - 1) / TW) * ((H + TH - 1) / TH) * B;
    const int grid = ntiles < parts ? ntiles : parts;
    const int rc = by_mode(mode, [&](auto m) -> int {
        hipLaunchKernelGGL(k_conv33_wgrad<decltype(m)::value>, dim3(grid), dim3(NT), 0, s, X, X2, dY, Yf, part_w,
                           part_b, B, H, W);
        SIL_LAUNCH_CHECK();
        return (int)SIL_OK;
    });
    if (rc != SIL_OK) return rc;
    hipLaunchKernelGGL(k_reduce_parts, dim3(36), dim3(256), 0, s, part_w, grid, 9 * C * C, dW);
    hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(32), 0, s, part_b, grid, C, db);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int conv_in(const float* x, const float* w, const float* bias, bf16* y, int B, int H, int W, hipStream_t s,
            const int* dB)
{
    hipLaunchKernelGGL(k_conv_in, dim3((W + TW - 1) / TW, (H + TH - 1) / TH, B), dim3(NT), 0, s, x, w, bias, y, H, W,
                       dB);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int conv_in33(const float* x, const PlanarInput* planar, const float* w_in, const float* b_in, const bf16* Wt,
              const float* bias, bf16* Y, int B, int H, int W, hipStream_t s, const int* dB)
{
    SIL_REQUIRE((int64_t)H * W * C * 2 <= kMaxBufferBytes, SIL_ERR_UNSUPPORTED,
                "unet: %dx%d images exceed one buffer resource per image", H, W);
    const long ntiles = (long)((W + TW - 1) / TW) * ((H + TH - 1) / TH) * B;
    const long cap = kComputeUnits * 2;
    const unsigned grid = (unsigned)(ntiles < cap ? ntiles : cap);
    if (planar) {
        SIL_REQUIRE(planar->c0 && planar->c1 && planar->src && planar->scale, SIL_ERR_INVALID,
                    "unet: incomplete planar input");
        hipLaunchKernelGGL(k_conv_in33<true>, dim3(grid), dim3(NT), 0, s, nullptr, *planar, w_in, b_in, Wt, bias, Y, B,
                           dB, H, W);
    } else {
        hipLaunchKernelGGL(k_conv_in33<false>, dim3(grid), dim3(NT), 0, s, x, PlanarInput{}, w_in, b_in, Wt, bias, Y,
                           B, dB, H, W);
    }
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int conv_in_wgrad(const float* x, const bf16* dY, const bf16* Yf, float* part, int parts, float* dWb, int B, int H,
                  int W, hipStream_t s)
{
    hipLaunchKernelGGL(k_conv_in_wgrad, dim3(parts), dim3(256), 0, s, x, dY, Yf, part, B, H, W);
    hipLaunchKernelGGL(k_reduce_parts, dim3(4), dim3(256), 0, s, part, parts, 9 * C * 2 + C, dWb);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

// 3×3 conv 32→1 (the output head, no activation) with the A matrix of
// k_pack_head: the FP32 weights split into w = hi + lo (BF16 rows 2t and
// 2t+1), so the two rows sum to the product with the weights to about 16
// bits, and the activations (BF16 already) are multiplied exactly with FP32
// accumulation.
int conv_out(const bf16* x, const bf16* Wt, const float* bias, float* y, int B, int H, int W, hipStream_t s,
             const int* dB)
{
    SIL_REQUIRE(x_image_bytes(ST_PLAIN, H, W) <= kMaxBufferBytes, SIL_ERR_UNSUPPORTED,
                "unet: %dx%d images exceed one buffer resource per image", H, W);
    SIL_REQUIRE(B <= 65535, SIL_ERR_UNSUPPORTED, "unet: at most 65535 images per batch");
    hipLaunchKernelGGL(k_head, dim3((W + HTW - 1) / HTW, (H + HTH - 1) / HTH, B), dim3(NT), 0, s, x, Wt, bias, y,
                       H, W, dB);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int conv_out_bwd(const bf16* x, const float* dout, const float* w, bf16* dx, float* part, int parts, float* dWb, int B,
                 int H, int W, hipStream_t s)
{
    hipLaunchKernelGGL(k_conv_out_bwd, dim3(parts), dim3(256), 0, s, x, dout, w, dx, part, B, H, W);
    hipLaunchKernelGGL(k_reduce_parts, dim3(2), dim3(256), 0, s, part, parts, 9 * C + 1, dWb);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int pool_bwd_accumulate(const bf16* dy, bf16* dx, int B, int H, int W, hipStream_t s)
{
    hipLaunchKernelGGL(k_pool_bwd, dim3(grid1((size_t)B * H * W * 4)), dim3(256), 0, s, dy, dx, B, H, W);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int up_bwd(const bf16* dy, bf16* dxc, int B, int h, int w, hipStream_t s)
{
    hipLaunchKernelGGL(k_up_bwd, dim3(grid1((size_t)B * h * w * 4)), dim3(256), 0, s, dy, dxc, B, h, w);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int cos_loss(const float* N, const float* r, int rs, const float* e, const float* wx, const float* wy, float* AN,
             float* part, int parts, float* dN, float* loss, int B, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_cos_loss_stats, dim3(parts, B), dim3(256), 0, s, N, r, rs, e, wx, wy, AN, part, n);
    hipLaunchKernelGGL(k_cos_loss_grad, dim3(parts, B), dim3(256), 0, s, r, rs, AN, part, parts, dN, loss, n, B);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int adam(float* w, const float* g, float* m, float* v, int n, float lr, int step, hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(grid1(n)), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2, 1e-8f, c1, c2);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int pack33(const float* w, bf16* wf, bf16* wb, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack33, dim3(36), dim3(256), 0, s, w, wf, wb);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int to_f32(const bf16* a, float* b, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_bf16_to_f32, dim3(grid1((size_t)n)), dim3(256), 0, s, a, b, n);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int pack_head(const float* w, bf16* Wt, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack_head, dim3(C * C / 256), dim3(256), 0, s, w, Wt);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

}  // namespace unet

}  // namespace sil
