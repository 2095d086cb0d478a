// The learned preconditioner's network object: FP32 master parameters with
// Adam state, BF16 packed weights, and the activation and gradient buffers of
// a 4-level U-Net for a maximum batch.  No reference counterpart.
//
// Parameter vector (FP32, in this order):
//   conv_in   W[9][32][2], b[32]
//   conv 1…12 W[9][32][32] ([tap][out][in], tap = 3·dy + dx), b[32]
//   conv_out  W[9][32], b[1]
// Layer order: 1 a0→e0, 2 p1→a1, 3 a1→e1, 4 p2→a2, 5 a2→e2, 6 p3→a3,
// 7 a3→b3, 8 u2→c2, 9 c2→f2, 10 u1→c1, 11 c1→f1, 12 u0→c0.
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "sil_hip.h"
#include "unet.hpp"

namespace {

using sil::unet::bf16;
constexpr int C = sil::unet::kChannels;
constexpr int64_t kIn = 9 * C * 2 + C;      // 608
constexpr int64_t kLayer = 9 * C * C + C;   // 9248
constexpr int64_t kOut = 9 * C + 1;         // 289
constexpr int kLayers = 12;
constexpr int64_t kParams = kIn + kLayers * kLayer + kOut;
constexpr int kParts = 512; // weight-gradient partial sums per layer

inline int64_t layer_off(int l) { return kIn + (int64_t)(l - 1) * kLayer; } // l = 1…12
constexpr int64_t kOutOff = kIn + kLayers * kLayer;

}  // namespace

struct sil_unet_s {
    int max_batch = 0, H = 0, W = 0;
    int Hl[4] = {0, 0, 0, 0}, Wl[4] = {0, 0, 0, 0};
    float* params = nullptr;
    float* grad = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    bf16* wf = nullptr; // kLayers × 9·32·32 forward weights
    bf16* wb = nullptr; // kLayers × 9·32·32 input-gradient weights
    bf16* wh = nullptr; // 32·32 output-head A matrix (hi / lo rows per tap)
    // activations
    // (pooled and upsampled-plus-skip inputs are formed while staging)
    bf16 *a0 = nullptr, *e0 = nullptr, *c0 = nullptr;
    bf16 *a1 = nullptr, *e1 = nullptr, *c1 = nullptr, *f1 = nullptr;
    bf16 *a2 = nullptr, *e2 = nullptr, *c2 = nullptr, *f2 = nullptr;
    bf16 *a3 = nullptr, *b3 = nullptr;
    // gradient buffers
    bf16 *g0a = nullptr, *g0b = nullptr, *g1a = nullptr, *g1b = nullptr, *g2a = nullptr, *g2b = nullptr,
         *g3a = nullptr, *g3b = nullptr;
    float* out = nullptr;   // network output [B][H][W]
    float* dout = nullptr;  // its gradient
    float* an = nullptr;    // A·out of the loss
    float* part_w = nullptr;
    float* part_b = nullptr;
    float* part_s = nullptr; // small partials (conv_in / conv_out / loss)
    int adam_step = 0;
    std::vector<void*> allocations;

    template <class T>
    int alloc(T*& p, size_t count)
    {
        void* q = nullptr;
        if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) {
            sil::set_error("unet: out of device memory (%zu bytes)", count * sizeof(T));
            return SIL_ERR_HIP;
        }
        allocations.push_back(q);
        p = static_cast<T*>(q);
        return SIL_OK;
    }
    ~sil_unet_s()
    {
        for (void* q : allocations) (void)hipFree(q);
    }
    size_t act(int level) const { return (size_t)max_batch * Hl[level] * Wl[level] * C; }
};

namespace {

#define UNET_TRY(expr)                   \
    do {                                 \
        const int rc_ = (expr);          \
        if (rc_ != SIL_OK) return rc_;   \
    } while (0)

int repack(sil_unet_s* u, hipStream_t s)
{
    for (int l = 1; l <= kLayers; ++l)
        UNET_TRY(sil::unet::pack33(u->params + layer_off(l), u->wf + (size_t)(l - 1) * 9 * C * C,
                                   u->wb + (size_t)(l - 1) * 9 * C * C, s));
    UNET_TRY(sil::unet::pack_head(u->params + kOutOff, u->wh, s));
    return SIL_OK;
}

const bf16* WF(const sil_unet_s* u, int l) { return u->wf + (size_t)(l - 1) * 9 * C * C; }
const bf16* WB(const sil_unet_s* u, int l) { return u->wb + (size_t)(l - 1) * 9 * C * C; }
const float* BIAS(const sil_unet_s* u, int l) { return u->params + layer_off(l) + 9 * C * C; }

// keep = false: inference only — a0 is not stored (the input layer and
// layer 1 run as one kernel), the other activations are
int forward(sil_unet_s* u, int B, const float* x, float* out, hipStream_t s, const int* dB = nullptr, bool keep = true,
            const sil::unet::PlanarInput* planar = nullptr)
{
    namespace un = sil::unet;
    constexpr int P = un::STAGE_PLAIN, UP = un::STAGE_UPADD, PO = un::STAGE_POOL;
    const int *H = u->Hl, *W = u->Wl;
    if (keep) {
        UNET_TRY(un::conv_in(x, u->params, u->params + 9 * C * 2, u->a0, B, H[0], W[0], s, dB));
        UNET_TRY(un::conv33(P, u->a0, nullptr, WF(u, 1), BIAS(u, 1), u->e0, B, H[0], W[0], true, s, dB));
    } else {
        UNET_TRY(un::conv_in33(x, planar, u->params, u->params + 9 * C * 2, WF(u, 1), BIAS(u, 1), u->e0, B, H[0],
                               W[0], s, dB));
    }
    UNET_TRY(un::conv33(PO, u->e0, nullptr, WF(u, 2), BIAS(u, 2), u->a1, B, H[1], W[1], true, s, dB));
    UNET_TRY(un::conv33(P, u->a1, nullptr, WF(u, 3), BIAS(u, 3), u->e1, B, H[1], W[1], true, s, dB));
    UNET_TRY(un::conv33(PO, u->e1, nullptr, WF(u, 4), BIAS(u, 4), u->a2, B, H[2], W[2], true, s, dB));
    UNET_TRY(un::conv33(P, u->a2, nullptr, WF(u, 5), BIAS(u, 5), u->e2, B, H[2], W[2], true, s, dB));
    UNET_TRY(un::conv33(PO, u->e2, nullptr, WF(u, 6), BIAS(u, 6), u->a3, B, H[3], W[3], true, s, dB));
    UNET_TRY(un::conv33(P, u->a3, nullptr, WF(u, 7), BIAS(u, 7), u->b3, B, H[3], W[3], true, s, dB));
    UNET_TRY(un::conv33(UP, u->e2, u->b3, WF(u, 8), BIAS(u, 8), u->c2, B, H[2], W[2], true, s, dB));
    UNET_TRY(un::conv33(P, u->c2, nullptr, WF(u, 9), BIAS(u, 9), u->f2, B, H[2], W[2], true, s, dB));
    UNET_TRY(un::conv33(UP, u->e1, u->f2, WF(u, 10), BIAS(u, 10), u->c1, B, H[1], W[1], true, s, dB));
    UNET_TRY(un::conv33(P, u->c1, nullptr, WF(u, 11), BIAS(u, 11), u->f1, B, H[1], W[1], true, s, dB));
    UNET_TRY(un::conv33(UP, u->e0, u->f1, WF(u, 12), BIAS(u, 12), u->c0, B, H[0], W[0], true, s, dB));
    UNET_TRY(un::conv_out(u->c0, u->wh, u->params + kOutOff + 9 * C, out, B, H[0], W[0], s, dB));
    return SIL_OK;
}

// weight gradient + input gradient of layer l: (mode, X, X2) its staged
// input, Y its forward output, dY the gradient wrt Y (pre-mask); dX receives
// the input gradient (may be null)
int layer_bwd(sil_unet_s* u, int l, int mode, const bf16* X, const bf16* X2, const bf16* Y, const bf16* dY, bf16* dX,
              int B, int level, float* grad, hipStream_t s)
{
    namespace un = sil::unet;
    const int H = u->Hl[level], W = u->Wl[level];
    UNET_TRY(un::conv33_wgrad(mode, X, X2, dY, Y, u->part_w, u->part_b, kParts, grad + layer_off(l),
                              grad + layer_off(l) + 9 * C * C, B, H, W, s));
    if (dX) UNET_TRY(un::conv33(un::STAGE_MASK, dY, Y, WB(u, l), nullptr, dX, B, H, W, false, s));
    return SIL_OK;
}

int backward(sil_unet_s* u, int B, const float* x, const float* dout, float* grad, hipStream_t s)
{
    namespace un = sil::unet;
    constexpr int P = un::STAGE_PLAIN, UP = un::STAGE_UPADD, PO = un::STAGE_POOL;
    const int *H = u->Hl, *W = u->Wl;
    UNET_TRY(un::conv_out_bwd(u->c0, dout, u->params + kOutOff, u->g0a, u->part_s, kParts, grad + kOutOff, B, H[0],
                              W[0], s));
    UNET_TRY(layer_bwd(u, 12, UP, u->e0, u->f1, u->c0, u->g0a, u->g0b, B, 0, grad, s)); // g0b = du0 = de0 (skip)
    UNET_TRY(un::up_bwd(u->g0b, u->g1a, B, H[1], W[1], s));                              // g1a = df1
    UNET_TRY(layer_bwd(u, 11, P, u->c1, nullptr, u->f1, u->g1a, u->g1b, B, 1, grad, s)); // g1b = dc1
    UNET_TRY(layer_bwd(u, 10, UP, u->e1, u->f2, u->c1, u->g1b, u->g1a, B, 1, grad, s));  // g1a = du1 = de1
    UNET_TRY(un::up_bwd(u->g1a, u->g2a, B, H[2], W[2], s));                              // g2a = df2
    UNET_TRY(layer_bwd(u, 9, P, u->c2, nullptr, u->f2, u->g2a, u->g2b, B, 2, grad, s));  // g2b = dc2
    UNET_TRY(layer_bwd(u, 8, UP, u->e2, u->b3, u->c2, u->g2b, u->g2a, B, 2, grad, s));   // g2a = du2 = de2
    UNET_TRY(un::up_bwd(u->g2a, u->g3a, B, H[3], W[3], s));                              // g3a = db3
    UNET_TRY(layer_bwd(u, 7, P, u->a3, nullptr, u->b3, u->g3a, u->g3b, B, 3, grad, s));  // g3b = da3
    UNET_TRY(layer_bwd(u, 6, PO, u->e2, nullptr, u->a3, u->g3b, u->g3a, B, 3, grad, s)); // g3a = dp3
    UNET_TRY(un::pool_bwd_accumulate(u->g3a, u->g2a, B, H[2], W[2], s));                 // de2 += pool'(dp3)
    UNET_TRY(layer_bwd(u, 5, P, u->a2, nullptr, u->e2, u->g2a, u->g2b, B, 2, grad, s));  // g2b = da2
    UNET_TRY(layer_bwd(u, 4, PO, u->e1, nullptr, u->a2, u->g2b, u->g2a, B, 2, grad, s)); // g2a = dp2
    UNET_TRY(un::pool_bwd_accumulate(u->g2a, u->g1a, B, H[1], W[1], s));                 // de1 += pool'(dp2)
    UNET_TRY(layer_bwd(u, 3, P, u->a1, nullptr, u->e1, u->g1a, u->g1b, B, 1, grad, s));  // g1b = da1
    UNET_TRY(layer_bwd(u, 2, PO, u->e0, nullptr, u->a1, u->g1b, u->g1a, B, 1, grad, s)); // g1a = dp1
    UNET_TRY(un::pool_bwd_accumulate(u->g1a, u->g0b, B, H[0], W[0], s));                 // de0 += pool'(dp1)
    UNET_TRY(layer_bwd(u, 1, P, u->a0, nullptr, u->e0, u->g0b, u->g0a, B, 0, grad, s));  // g0a = da0
    UNET_TRY(un::conv_in_wgrad(x, u->g0a, u->a0, u->part_s, kParts, grad, B, H[0], W[0], s));
    return SIL_OK;
}

int check_batch(const sil_unet_s* u, int B)
{
    if (B < 1 || B > u->max_batch) {
        sil::set_error("unet: batch %d outside [1, %d]", B, u->max_batch);
        return SIL_ERR_INVALID;
    }
    return SIL_OK;
}

}  // namespace

int sil::unet::forward_active(sil_unet_s* net, int batch, const int* d_count, const float* x, float* out, hipStream_t s)
{
    SIL_REQUIRE(net != nullptr && d_count != nullptr && x != nullptr && out != nullptr, SIL_ERR_INVALID,
                "unet: null argument");
    UNET_TRY(check_batch(net, batch));
    return forward(net, batch, x, out, s, d_count, false);
}

int sil::unet::forward_active_planar(sil_unet_s* net, int batch, const int* d_count, const PlanarInput& in,
                                     float* out, hipStream_t s)
{
    SIL_REQUIRE(net != nullptr && d_count != nullptr && out != nullptr, SIL_ERR_INVALID, "unet: null argument");
    UNET_TRY(check_batch(net, batch));
    return forward(net, batch, nullptr, out, s, d_count, false, &in);
}

void sil::unet::net_shape(const sil_unet_s* net, int* max_batch, int* height, int* width)
{
    *max_batch = net->max_batch;
    *height = net->H;
    *width = net->W;
}

extern "C" {

int64_t sil_unet_param_count(void) { return kParams; }

int sil_unet_create(sil_unet* out, int32_t max_batch, int32_t height, int32_t width, uint64_t seed)
{
    SIL_REQUIRE(out != nullptr, SIL_ERR_INVALID, "unet: null handle");
    SIL_REQUIRE(max_batch >= 1 && height >= 8 && width >= 8, SIL_ERR_INVALID, "unet: bad shape");
    SIL_REQUIRE(height % 8 == 0 && width % 8 == 0, SIL_ERR_UNSUPPORTED,
                "unet: height and width must be multiples of 8 (three 2x2 poolings), got %dx%d", height, width);
    auto* u = new (std::nothrow) sil_unet_s();
    SIL_REQUIRE(u != nullptr, SIL_ERR_INVALID, "unet: out of host memory");
    u->max_batch = max_batch;
    u->H = height;
    u->W = width;
    for (int l = 0; l < 4; ++l) {
        u->Hl[l] = height >> l;
        u->Wl[l] = width >> l;
    }
    const size_t s0 = u->act(0), s1 = u->act(1), s2 = u->act(2), s3 = u->act(3);
    const size_t npx = (size_t)max_batch * height * width;
    int rc = SIL_OK;
    auto A = [&](auto& p, size_t n) {
        if (rc == SIL_OK) rc = u->alloc(p, n);
    };
    A(u->params, kParams);
    A(u->grad, kParams);
    A(u->adam_m, kParams);
    A(u->adam_v, kParams);
    A(u->wf, (size_t)kLayers * 9 * C * C);
    A(u->wb, (size_t)kLayers * 9 * C * C);
    A(u->wh, (size_t)C * C);
    for (bf16** p : {&u->a0, &u->e0, &u->c0, &u->g0a, &u->g0b}) A(*p, s0);
    for (bf16** p : {&u->a1, &u->e1, &u->c1, &u->f1, &u->g1a, &u->g1b}) A(*p, s1);
    for (bf16** p : {&u->a2, &u->e2, &u->c2, &u->f2, &u->g2a, &u->g2b}) A(*p, s2);
    for (bf16** p : {&u->a3, &u->b3, &u->g3a, &u->g3b}) A(*p, s3);
    A(u->out, npx);
    A(u->dout, npx);
    A(u->an, npx);
    A(u->part_w, (size_t)kParts * 9 * C * C);
    A(u->part_b, (size_t)kParts * C);
    A(u->part_s, (size_t)kParts * kIn);
    if (rc != SIL_OK) {
        delete u;
        return rc;
    }
    // He-normal initialisation (fan-in of each layer), zero biases
    std::vector<float> h(kParams, 0.f);
    std::mt19937_64 gen(seed);
    auto fill = [&](int64_t off, int64_t count, double fan_in, double gain) {
        std::normal_distribution<double> nd(0.0, std::sqrt(gain / fan_in));
        for (int64_t i = 0; i < count; ++i) h[off + i] = (float)nd(gen);
    };
    fill(0, 9 * C * 2, 18, 2.0);
    for (int l = 1; l <= kLayers; ++l) fill(layer_off(l), 9 * C * C, 9 * C, 2.0);
    fill(kOutOff, 9 * C, 9 * C, 1.0);
    if (hipMemcpy(u->params, h.data(), kParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(u->adam_m, 0, kParams * sizeof(float)) != hipSuccess ||
        hipMemset(u->adam_v, 0, kParams * sizeof(float)) != hipSuccess) {
        delete u;
        sil::set_error("unet: initialisation copy failed");
        return SIL_ERR_HIP;
    }
    rc = repack(u, nullptr);
    if (rc == SIL_OK && hipDeviceSynchronize() != hipSuccess) rc = SIL_ERR_HIP;
    if (rc != SIL_OK) {
        delete u;
        return rc;
    }
    *out = u;
    return SIL_OK;
}

int sil_unet_destroy(sil_unet u)
{
    delete u;
    return SIL_OK;
}

int sil_unet_get_params(sil_unet u, float* dst, void* stream)
{
    SIL_REQUIRE(u != nullptr && dst != nullptr, SIL_ERR_INVALID, "unet: null argument");
    SIL_CHECK_HIP(hipMemcpyAsync(dst, u->params, kParams * sizeof(float), hipMemcpyDeviceToDevice,
                                 static_cast<hipStream_t>(stream)));
    return SIL_OK;
}

int sil_unet_set_params(sil_unet u, const float* src, void* stream)
{
    SIL_REQUIRE(u != nullptr && src != nullptr, SIL_ERR_INVALID, "unet: null argument");
    auto s = static_cast<hipStream_t>(stream);
    SIL_CHECK_HIP(hipMemcpyAsync(u->params, src, kParams * sizeof(float), hipMemcpyDeviceToDevice, s));
    return repack(u, s);
}

int sil_unet_forward(sil_unet u, int32_t batch, const float* x, float* out, void* stream)
{
    SIL_REQUIRE(u != nullptr && x != nullptr && out != nullptr, SIL_ERR_INVALID, "unet: null argument");
    UNET_TRY(check_batch(u, batch));
    return forward(u, batch, x, out, static_cast<hipStream_t>(stream));
}

int sil_unet_infer(sil_unet u, int32_t batch, const float* x, float* out, void* stream)
{
    SIL_REQUIRE(u != nullptr && x != nullptr && out != nullptr, SIL_ERR_INVALID, "unet: null argument");
    UNET_TRY(check_batch(u, batch));
    return forward(u, batch, x, out, static_cast<hipStream_t>(stream), nullptr, false);
}

int sil_unet_activation(sil_unet u, int32_t which, int32_t batch, float* out, void* stream)
{
    SIL_REQUIRE(u != nullptr && out != nullptr, SIL_ERR_INVALID, "unet: null argument");
    UNET_TRY(check_batch(u, batch));
    // graph order: a0 e0 a1 e1 a2 e2 a3 b3 c2 f2 c1 f1 c0, levels 0 0 1 1 2 2 3 3 2 2 1 1 0
    const bf16* acts[13] = {u->a0, u->e0, u->a1, u->e1, u->a2, u->e2, u->a3, u->b3, u->c2, u->f2, u->c1, u->f1, u->c0};
    static constexpr int level[13] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0};
    SIL_REQUIRE(which >= 0 && which < 13, SIL_ERR_INVALID, "unet: activation %d outside 0..12", which);
    const int64_t n = (int64_t)batch * u->Hl[level[which]] * u->Wl[level[which]] * C;
    return sil::unet::to_f32(acts[which], out, n, sil::as_stream(stream));
}

int sil_unet_backward(sil_unet u, int32_t batch, const float* x, const float* dout, float* grad, void* stream)
{
    SIL_REQUIRE(u != nullptr && x != nullptr && dout != nullptr && grad != nullptr, SIL_ERR_INVALID,
                "unet: null argument");
    UNET_TRY(check_batch(u, batch));
    return backward(u, batch, x, dout, grad, static_cast<hipStream_t>(stream));
}

int sil_unet_train_step(sil_unet u, int32_t batch, const float* x, const float* err, const float* wx,
                        const float* wy, float lr, float* loss, void* stream)
{
    SIL_REQUIRE(u != nullptr && x != nullptr && err != nullptr && wx != nullptr && wy != nullptr && loss != nullptr,
                SIL_ERR_INVALID, "unet: null argument");
    SIL_REQUIRE(u->H == u->W, SIL_ERR_UNSUPPORTED, "unet: the training loss needs square images");
    UNET_TRY(check_batch(u, batch));
    auto s = static_cast<hipStream_t>(stream);
    UNET_TRY(forward(u, batch, x, u->out, s));
    const int parts = 64;
    UNET_TRY(sil::unet::cos_loss(u->out, x, 2, err, wx, wy, u->an, u->part_s, parts, u->dout, loss, batch, u->H, s));
    UNET_TRY(backward(u, batch, x, u->dout, u->grad, s));
    u->adam_step += 1;
    UNET_TRY(sil::unet::adam(u->params, u->grad, u->adam_m, u->adam_v, (int)kParams, lr, u->adam_step, s));
    return repack(u, s);
}

int sil_unet_loss(sil_unet u, int32_t batch, const float* out, const float* x, const float* err, const float* wx,
                  const float* wy, float* dout, float* loss, void* stream)
{
    SIL_REQUIRE(u != nullptr && out != nullptr && x != nullptr && err != nullptr && wx != nullptr && wy != nullptr &&
                    dout != nullptr && loss != nullptr,
                SIL_ERR_INVALID, "unet: null argument");
    SIL_REQUIRE(u->H == u->W, SIL_ERR_UNSUPPORTED, "unet: the training loss needs square images");
    UNET_TRY(check_batch(u, batch));
    return sil::unet::cos_loss(out, x, 2, err, wx, wy, u->an, u->part_s, 64, dout, loss, batch, u->H,
                               static_cast<hipStream_t>(stream));
}

}  // extern "C"
