// Kernels of the learned U-Net preconditioner (BF16 activations and
// weights, FP32 accumulation).  Tensors are NHWC with 32 channels; every
// launcher returns a sil_status and records the message on failure.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "common.hpp"

struct sil_unet_s;

namespace sil::unet {

using bf16 = __bf16;

constexpr int kChannels = 32;
constexpr int kTileH = 8;
constexpr int kTileW = 32;

// Input staging of a 3x3 convolution: X itself, X + nearest-upsampled X2
// (skip + coarse), the 2x2 average of X (X twice the size), or X ⊙ [X2 > 0]
enum StageModeHost { STAGE_PLAIN = 0, STAGE_UPADD = 1, STAGE_POOL = 2, STAGE_MASK = 3 };

// Y = act(conv3x3(staged input) + bias), 32 → 32 channels on H×W images, zero padding
// With dB, only the first *dB (device memory, read by the kernels) of the B
// images are computed.
int conv33(int mode, const bf16* X, const bf16* X2, const bf16* Wt, const float* bias, bf16* Y, int B, int H, int W,
           bool relu, hipStream_t s, const int* dB = nullptr);
// dW[tap][co][ci], db[co] of conv33 from its staged input, output gradient dY
// and forward output Yf (ReLU mask); part_w holds parts·9·32·32 floats, part_b parts·32
int conv33_wgrad(int mode, const bf16* X, const bf16* X2, const bf16* dY, const bf16* Yf, float* part_w,
                 float* part_b, int parts, float* dW, float* db, int B, int H, int W, hipStream_t s);
// 2 → 32 input convolution (FP32 input) and its weight gradient (dWb = [dW(9·32·2) | db(32)])
int conv_in(const float* x, const float* w, const float* bias, bf16* y, int B, int H, int W, hipStream_t s,
            const int* dB = nullptr);
// The network input as two FP32 planes [images][H][W] read through a slot
// map: image b of the batch is source image src[b], with channel 0 =
// c0 · scale[b] and channel 1 = c1 (src and scale: device memory, one entry
// per image of the batch).
struct PlanarInput {
    const float* c0 = nullptr;
    const float* c1 = nullptr;
    const int* src = nullptr;
    const float* scale = nullptr;
};
// relu(conv(relu(conv_in(x)))) in one kernel (a0 is formed in LDS and not
// written): equal, bit for bit, to conv_in followed by conv33 in STAGE_PLAIN;
// the input is x ([B][H][W][2]) or, when planar is given, its planes
int conv_in33(const float* x, const PlanarInput* planar, const float* w_in, const float* b_in, const bf16* Wt,
              const float* bias, bf16* Y, int B, int H, int W, hipStream_t s, const int* dB = nullptr);
int conv_in_wgrad(const float* x, const bf16* dY, const bf16* Yf, float* part, int parts, float* dWb, int B, int H,
                  int W, hipStream_t s);
// 32 → 1 output convolution (FP32 output) and its backward (dWb = [dW(9·32) | db])
int conv_out(const bf16* x, const bf16* Wt, const float* bias, float* y, int B, int H, int W, hipStream_t s,
             const int* dB = nullptr);
int conv_out_bwd(const bf16* x, const float* dout, const float* w, bf16* dx, float* part, int parts, float* dWb, int B,
                 int H, int W, hipStream_t s);
// backward of the 2x2 average pooling of an H×W tensor (accumulates into dx)
int pool_bwd_accumulate(const bf16* dy, bf16* dx, int B, int H, int W, hipStream_t s);
// backward of nearest upsampling to the coarse h×w tensor
int up_bwd(const bf16* dy, bf16* dxc, int B, int h, int w, hipStream_t s);
// direction loss 1 − cos²_A(e, N) per image and its gradient dN (batch mean);
// r is read with element stride rs; part holds B·parts·3 floats
int cos_loss(const float* N, const float* r, int rs, const float* e, const float* wx, const float* wy, float* AN,
             float* part, int parts, float* dN, float* loss, int B, int n, hipStream_t s);
int adam(float* w, const float* g, float* m, float* v, int n, float lr, int step, hipStream_t s);
int to_f32(const bf16* a, float* b, int64_t n, hipStream_t s);
// BF16 forward weights and the flipped, channel-transposed input-gradient weights of one 32→32 layer
int pack33(const float* w, bf16* wf, bf16* wb, hipStream_t s);
// output-head weights [tap][ci] → the head's BF16 A matrix (32 × 32: rows 2t,
// 2t+1 the hi and lo parts of tap t, rows 18–31 zero)
int pack_head(const float* w, bf16* Wt, hipStream_t s);

// maximum batch and image size of a network handle (defined with the handle)
// Forward pass over the first *d_count (device memory, at most `batch`) of
// the images in x; the rest of `out` is left as it is.  Used inside captured
// solver iterations, where the number of images still needing the network is
// only known on the device.
// The pass keeps no activations for a backward pass: the input layer and the
// first 32→32 layer run fused.
int forward_active(sil_unet_s* net, int batch, const int* d_count, const float* x, float* out, hipStream_t s);
// the same pass on a planar input (no interleaved copy of it is formed)
int forward_active_planar(sil_unet_s* net, int batch, const int* d_count, const PlanarInput& in, float* out,
                          hipStream_t s);
void net_shape(const struct ::sil_unet_s* net, int* max_batch, int* height, int* width);

}  // namespace sil::unet
