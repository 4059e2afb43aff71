// C ABI of the learned disparity proposal: the network object (parameters, Adam
// state, activations), its kernel-level convolution entry points, and the
// learned-proposal census matcher that runs it in front of the work-list matcher.
#include <cmath>
#include <vector>

#include "capi_internal.hpp"

using namespace capi;

namespace {

struct LayerDef {
    int cin, cout, dil;      // padded channel counts used by the kernels
    int cin_real, cout_real; // channels of the public parameter layout
};

// (2,16,1) (16,16,2) (16,16,4) (16,16,2) (16,2,1); the 2-channel ends are padded
// to 16 channels with zero weights, so every convolution runs on whole MFMA blocks.
constexpr LayerDef kLayers[5] = {{16, 16, 1, 2, 16}, {16, 16, 2, 16, 16}, {16, 16, 4, 16, 16}, {16, 16, 2, 16, 16},
                                 {16, 16, 1, 16, 2}};
constexpr int kNumLayers = 5;

size_t layer_w(int l) { return (size_t)kLayers[l].cout * 9 * kLayers[l].cin; }
size_t layer_n(int l) { return layer_w(l) + kLayers[l].cout; }
size_t layer_off(int l)
{
    size_t o = 0;
    for (int j = 0; j < l; ++j) o += layer_n(j);
    return o;
}
size_t padded_total() { return layer_off(kNumLayers); }

constexpr float kLambda = 0.5f;  // weight of the confidence loss
constexpr float kTau = 2.0f;     // centre within 2 quarter pixels (8 pixels) counts as confident
constexpr float kBeta1 = 0.9f, kBeta2 = 0.999f, kEps = 1e-8f;

}  // namespace

struct stereo_cnn {
    stereo_context* ctx = nullptr;
    int W4 = 0, H4 = 0, maxn = 0;
    float* param = nullptr;  // padded FP32 master parameters
    float* grad = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    uint16_t* wf = nullptr;  // bf16 [cout][9][cin] per layer (at the parameter offsets)
    uint16_t* wb = nullptr;  // bf16 [cin][9][cout] per layer
    uint16_t* act[kNumLayers] = {};  // act[0] = input (16 ch), act[l] = output of layer l-1 (32 ch)
    float* out = nullptr;            // [P][16] FP32
    uint16_t* g[2] = {};             // gradient ping-pong [P][32] bf16
    float* dfill = nullptr;
    float* labels = nullptr;
    float* clabels = nullptr;  // confidence targets (0, 1 or NaN)
    float* dwpart = nullptr;
    int dwparts = 0;
    float* lossp = nullptr;
    int step = 0;

    size_t P() const { return (size_t)maxn * W4 * H4; }
};

namespace {

void cnn_release(stereo_cnn* c)
{
    for (void* p : {(void*)c->param, (void*)c->grad, (void*)c->m, (void*)c->v, (void*)c->wf, (void*)c->wb, (void*)c->out,
                    (void*)c->dfill, (void*)c->labels, (void*)c->clabels, (void*)c->dwpart, (void*)c->lossp, (void*)c->g[0], (void*)c->g[1]})
        if (p) (void)hipFree(p);
    for (uint16_t* a : c->act)
        if (a) (void)hipFree(a);
}

stereo_status refresh_weights(stereo_cnn* c, float lr, hipStream_t s)
{
    float c1 = 1.f, c2 = 1.f;
    if (lr != 0.f) {
        ++c->step;
        c1 = 1.f / (1.f - std::pow(kBeta1, (float)c->step));
        c2 = 1.f / (1.f - std::pow(kBeta2, (float)c->step));
    }
    for (int l = 0; l < kNumLayers; ++l) {
        const size_t o = layer_off(l);
        HIP_TRY(c->ctx,
                stereo::launch_cnn_adam(c->param + o, c->grad + o, c->m + o, c->v + o, (int)layer_n(l), kLayers[l].cin,
                                        kLayers[l].cout, lr, kBeta1, kBeta2, kEps, c1, c2, c->wf + o, c->wb + o, s),
                "adam launch");
    }
    return STEREO_OK;
}

// Forward over slots [0, n).  Training keeps every layer's activation for the backward
// pass and runs the layers one launch each; inference (fused) runs the whole network in
// one launch whose output is the same, bit for bit.
stereo_status cnn_forward(stereo_cnn* c, int n, hipStream_t s, bool fused = false)
{
    if (fused) {
        stereo::CnnFusedArgs a;
        a.x = c->act[0];
        for (int l = 0; l < kNumLayers; ++l) {
            a.w[l] = c->wf + layer_off(l);
            a.bias[l] = c->param + layer_off(l) + layer_w(l);
        }
        a.out = c->out;
        a.N = n;
        a.H = c->H4;
        a.W = c->W4;
        HIP_TRY(c->ctx, stereo::launch_cnn_fused_forward(a, s), "fused network launch");
        return STEREO_OK;
    }
    for (int l = 0; l < kNumLayers; ++l) {
        const size_t o = layer_off(l);
        stereo::CnnConvArgs a;
        a.x = c->act[l];
        a.w = c->wf + o;
        a.bias = c->param + o + layer_w(l);
        a.mask = nullptr;
        a.y = l + 1 < kNumLayers ? (void*)c->act[l + 1] : (void*)c->out;
        a.N = n;
        a.H = c->H4;
        a.W = c->W4;
        a.dil = kLayers[l].dil;
        a.sign = 1;
        HIP_TRY(c->ctx,
                stereo::launch_cnn_conv(kLayers[l].cin, kLayers[l].cout,
                                        l + 1 < kNumLayers ? stereo::CNN_EPI_RELU_BF16 : stereo::CNN_EPI_F32, a, s),
                "conv launch");
    }
    return STEREO_OK;
}

// Host <-> padded parameter layouts.
void to_padded(const float* host, std::vector<float>& pad)
{
    pad.assign(padded_total(), 0.f);
    size_t k = 0;
    for (int l = 0; l < kNumLayers; ++l) {
        const LayerDef& L = kLayers[l];
        const size_t o = layer_off(l);
        for (int co = 0; co < L.cout_real; ++co)
            for (int t = 0; t < 9; ++t)
                for (int ci = 0; ci < L.cin_real; ++ci) pad[o + ((size_t)co * 9 + t) * L.cin + ci] = host[k++];
        for (int co = 0; co < L.cout_real; ++co) pad[o + layer_w(l) + co] = host[k++];
    }
}

void from_padded(const std::vector<float>& pad, float* host)
{
    size_t k = 0;
    for (int l = 0; l < kNumLayers; ++l) {
        const LayerDef& L = kLayers[l];
        const size_t o = layer_off(l);
        for (int co = 0; co < L.cout_real; ++co)
            for (int t = 0; t < 9; ++t)
                for (int ci = 0; ci < L.cin_real; ++ci) host[k++] = pad[o + ((size_t)co * 9 + t) * L.cin + ci];
        for (int co = 0; co < L.cout_real; ++co) host[k++] = pad[o + layer_w(l) + co];
    }
}

stereo_status download_params(stereo_cnn* c, const float* dev, float* host, size_t n)
{
    if (n != STEREO_CNN_PARAMS) return fail(c->ctx, STEREO_ERROR_INVALID_ARGUMENT, "expected %d parameters, got %zu",
                                            STEREO_CNN_PARAMS, n);
    std::vector<float> pad(padded_total());
    HIP_TRY(c->ctx, hipMemcpy(pad.data(), dev, pad.size() * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    from_padded(pad, host);
    return STEREO_OK;
}

}  // namespace

extern "C" {

stereo_status stereo_cnn_create(stereo_context* ctx, stereo_cnn** out, int32_t W4, int32_t H4, int32_t max_frames)
{
    if (!ctx || !out) return STEREO_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (W4 < 1 || H4 < 1 || W4 > 4096 || max_frames < 1 || (size_t)W4 * H4 * max_frames > ((size_t)1 << 28))
        return fail(ctx, STEREO_ERROR_INVALID_ARGUMENT, "stereo_cnn_create: bad size %dx%d x %d", W4, H4, max_frames);
    stereo_status st = set_device(ctx);
    if (st) return st;
    GUARD(ctx, "stereo_cnn_create");
    stereo_cnn* c = new stereo_cnn();
    c->ctx = ctx;
    c->W4 = W4;
    c->H4 = H4;
    c->maxn = max_frames;
    const size_t P = c->P();
    const size_t np = padded_total();
    auto bail = [&](hipError_t e) {
        cnn_release(c);
        delete c;
        return hip_fail(ctx, e, "hipMalloc");
    };
    hipError_t e;
    if ((e = hipMalloc((void**)&c->param, np * 4)) || (e = hipMalloc((void**)&c->grad, np * 4)) ||
        (e = hipMalloc((void**)&c->m, np * 4)) || (e = hipMalloc((void**)&c->v, np * 4)) ||
        (e = hipMalloc((void**)&c->wf, np * 2)) || (e = hipMalloc((void**)&c->wb, np * 2)) ||
        (e = hipMalloc((void**)&c->out, P * 16 * 4)) || (e = hipMalloc((void**)&c->dfill, P * 4)) ||
        (e = hipMalloc((void**)&c->labels, P * 4)) || (e = hipMalloc((void**)&c->clabels, P * 4)) ||
        (e = hipMalloc((void**)&c->g[0], P * 32 * 2)) ||
        (e = hipMalloc((void**)&c->g[1], P * 32 * 2)) || (e = hipMalloc((void**)&c->lossp, ((P + 255) / 256) * 16)))
        return bail(e);
    if ((e = hipMalloc((void**)&c->act[0], P * 16 * 2))) return bail(e);
    for (int l = 1; l < kNumLayers; ++l)
        if ((e = hipMalloc((void**)&c->act[l], P * 32 * 2))) return bail(e);
    stereo::CnnDwArgs da{};
    da.N = max_frames;
    da.H = H4;
    da.W = W4;
    c->dwparts = stereo::cnn_dw_parts(da);
    if ((e = hipMalloc((void**)&c->dwpart, (size_t)c->dwparts * (32 * 9 * 32 + 32) * 4))) return bail(e);
    if ((e = hipMemset(c->param, 0, np * 4)) || (e = hipMemset(c->grad, 0, np * 4)) || (e = hipMemset(c->m, 0, np * 4)) ||
        (e = hipMemset(c->v, 0, np * 4)) || (e = hipMemset(c->labels, 0xFF, P * 4)) ||
        (e = hipMemset(c->clabels, 0xFF, P * 4)) || (e = hipMemset(c->dfill, 0, P * 4)) ||
        (e = hipMemset(c->act[0], 0, P * 16 * 2)))
        return bail(e);
    st = refresh_weights(c, 0.f, nullptr);
    if (st) {
        cnn_release(c);
        delete c;
        return st;
    }
    *out = c;
    return STEREO_OK;
}

void stereo_cnn_destroy(stereo_cnn* c)
{
    if (!c) return;
    cnn_release(c);
    delete c;
}

stereo_status stereo_cnn_set_params(stereo_cnn* c, const float* host, size_t n)
{
    if (!c || !host) return STEREO_ERROR_INVALID_ARGUMENT;
    if (n != STEREO_CNN_PARAMS)
        return fail(c->ctx, STEREO_ERROR_INVALID_ARGUMENT, "expected %d parameters, got %zu", STEREO_CNN_PARAMS, n);
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(host[i])) return fail(c->ctx, STEREO_ERROR_INVALID_ARGUMENT, "parameter %zu is not finite", i);
    stereo_status st = set_device(c->ctx);
    if (st) return st;
    std::vector<float> pad;
    to_padded(host, pad);
    const size_t np = padded_total();
    HIP_TRY(c->ctx, hipMemcpy(c->param, pad.data(), np * 4, hipMemcpyHostToDevice), "hipMemcpy");
    HIP_TRY(c->ctx, hipMemset(c->m, 0, np * 4), "hipMemset");
    HIP_TRY(c->ctx, hipMemset(c->v, 0, np * 4), "hipMemset");
    HIP_TRY(c->ctx, hipMemset(c->grad, 0, np * 4), "hipMemset");
    c->step = 0;
    if ((st = refresh_weights(c, 0.f, nullptr))) return st;
    HIP_TRY(c->ctx, hipDeviceSynchronize(), "hipDeviceSynchronize");
    return STEREO_OK;
}

stereo_status stereo_cnn_get_params(stereo_cnn* c, float* host, size_t n)
{
    if (!c || !host) return STEREO_ERROR_INVALID_ARGUMENT;
    HIP_TRY(c->ctx, hipDeviceSynchronize(), "hipDeviceSynchronize");
    return download_params(c, c->param, host, n);
}

stereo_status stereo_cnn_get_grads(stereo_cnn* c, float* host, size_t n)
{
    if (!c || !host) return STEREO_ERROR_INVALID_ARGUMENT;
    HIP_TRY(c->ctx, hipDeviceSynchronize(), "hipDeviceSynchronize");
    return download_params(c, c->grad, host, n);
}

size_t stereo_cnn_state_size(void) { return 16 + 3 * (size_t)STEREO_CNN_PARAMS * sizeof(float); }

stereo_status stereo_cnn_save_state(stereo_cnn* c, void* host, size_t n)
{
    if (!c || !host) return STEREO_ERROR_INVALID_ARGUMENT;
    if (n < stereo_cnn_state_size())
        return fail(c->ctx, STEREO_ERROR_INVALID_ARGUMENT, "state buffer of %zu bytes, need %zu", n, stereo_cnn_state_size());
    HIP_TRY(c->ctx, hipDeviceSynchronize(), "hipDeviceSynchronize");
    uint8_t* b = static_cast<uint8_t*>(host);
    std::memcpy(b, "CLSTCNN1", 8);
    const uint32_t version = 1, step = (uint32_t)c->step;
    std::memcpy(b + 8, &version, 4);
    std::memcpy(b + 12, &step, 4);
    float* f = reinterpret_cast<float*>(b + 16);  // the caller's buffer; 16-byte header keeps 4-byte alignment
    stereo_status st;
    if ((st = download_params(c, c->param, f, STEREO_CNN_PARAMS))) return st;
    if ((st = download_params(c, c->m, f + STEREO_CNN_PARAMS, STEREO_CNN_PARAMS))) return st;
    return download_params(c, c->v, f + 2 * (size_t)STEREO_CNN_PARAMS, STEREO_CNN_PARAMS);
}

stereo_status stereo_cnn_load_state(stereo_cnn* c, const void* host, size_t n)
{
    if (!c || !host) return STEREO_ERROR_INVALID_ARGUMENT;
    const uint8_t* b = static_cast<const uint8_t*>(host);
    uint32_t version = 0, step = 0;
    if (n < stereo_cnn_state_size() || std::memcmp(b, "CLSTCNN1", 8) != 0)
        return fail(c->ctx, STEREO_ERROR_INVALID_ARGUMENT, "not a network state blob");
    std::memcpy(&version, b + 8, 4);
    std::memcpy(&step, b + 12, 4);
    if (version != 1) return fail(c->ctx, STEREO_ERROR_INVALID_ARGUMENT, "network state version %u not supported", version);
    std::vector<float> f(3 * (size_t)STEREO_CNN_PARAMS);
    std::memcpy(f.data(), b + 16, f.size() * sizeof(float));
    stereo_status st = stereo_cnn_set_params(c, f.data(), STEREO_CNN_PARAMS);  // also resets the moments
    if (st) return st;
    std::vector<float> pad;
    to_padded(f.data() + STEREO_CNN_PARAMS, pad);
    HIP_TRY(c->ctx, hipMemcpy(c->m, pad.data(), pad.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
    to_padded(f.data() + 2 * (size_t)STEREO_CNN_PARAMS, pad);
    HIP_TRY(c->ctx, hipMemcpy(c->v, pad.data(), pad.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
    c->step = (int)step;
    return STEREO_OK;
}

stereo_status stereo_cnn_set_input(stereo_cnn* c, int32_t index, const stereo_image* L4, const stereo_image* R4,
                                   const stereo_image* coarse, stereo_stream stream)
{
    if (!c || !L4 || !R4 || !coarse) return STEREO_ERROR_INVALID_ARGUMENT;
    if (index < 0 || index >= c->maxn)
        return fail(c->ctx, STEREO_ERROR_INVALID_ARGUMENT, "input slot %d outside [0, %d)", index, c->maxn);
    stereo_status st;
    if ((st = check_image(c->ctx, L4, c->W4, c->H4, STEREO_PIXEL_LUM_U8, "L4"))) return st;
    if ((st = check_image(c->ctx, R4, c->W4, c->H4, STEREO_PIXEL_LUM_U8, "R4"))) return st;
    if ((st = check_image(c->ctx, coarse, c->W4, c->H4, STEREO_PIXEL_LUM_I16, "coarse disparity"))) return st;
    HIP_TRY(c->ctx,
            stereo::launch_cnn_input(view_of(L4), view_of(R4), view_of(coarse), index, c->act[0], c->dfill,
                                     (hipStream_t)stream),
            "cnn input launch");
    return STEREO_OK;
}

stereo_status stereo_cnn_set_labels(stereo_cnn* c, int32_t index, const stereo_image* disp, stereo_stream stream)
{
    if (!c || !disp) return STEREO_ERROR_INVALID_ARGUMENT;
    if (index < 0 || index >= c->maxn)
        return fail(c->ctx, STEREO_ERROR_INVALID_ARGUMENT, "label slot %d outside [0, %d)", index, c->maxn);
    stereo_status st;
    if ((st = check_image(c->ctx, disp, 4 * c->W4, 4 * c->H4, STEREO_PIXEL_LUM_I16, "label map"))) return st;
    HIP_TRY(c->ctx, stereo::launch_cnn_labels(view_of(disp), index, c->W4, c->H4, c->labels, (hipStream_t)stream),
            "labels launch");
    return STEREO_OK;
}

stereo_status stereo_cnn_set_conf_labels(stereo_cnn* c, int32_t index, const stereo_image* conf, stereo_stream stream)
{
    if (!c || !conf || index < 0 || index >= c->maxn) return STEREO_ERROR_INVALID_ARGUMENT;
    stereo_status st;
    if ((st = check_image(c->ctx, conf, c->W4, c->H4, STEREO_PIXEL_LUM_U8, "confidence labels"))) return st;
    if ((st = set_device(c->ctx))) return st;
    HIP_TRY(c->ctx, stereo::launch_cnn_conf_labels(view_of(conf), index, c->clabels, (hipStream_t)stream),
            "confidence labels launch");
    return STEREO_OK;
}

stereo_status stereo_cnn_forward(stereo_cnn* c, int32_t n, float* out, stereo_stream stream)
{
    if (!c) return STEREO_ERROR_INVALID_ARGUMENT;
    if (n < 1 || n > c->maxn) return fail(c->ctx, STEREO_ERROR_INVALID_ARGUMENT, "frame count %d outside [1, %d]", n, c->maxn);
    const hipStream_t s = (hipStream_t)stream;
    stereo_status st = cnn_forward(c, n, s, true);
    if (st) return st;
    if (out) {
        const size_t P = (size_t)n * c->W4 * c->H4;
        HIP_TRY(c->ctx, hipMemcpy2DAsync(out, 2 * sizeof(float), c->out, 16 * sizeof(float), 2 * sizeof(float), P,
                                         hipMemcpyDeviceToDevice, s),
                "hipMemcpy2DAsync");
    }
    return STEREO_OK;
}

stereo_status stereo_cnn_train_step(stereo_cnn* c, int32_t n, float lr, float* loss, stereo_stream stream)
{
    if (!c) return STEREO_ERROR_INVALID_ARGUMENT;
    if (n < 1 || n > c->maxn) return fail(c->ctx, STEREO_ERROR_INVALID_ARGUMENT, "frame count %d outside [1, %d]", n, c->maxn);
    if (!(lr >= 0.f) || !std::isfinite(lr)) return fail(c->ctx, STEREO_ERROR_INVALID_ARGUMENT, "bad learning rate");
    const hipStream_t s = (hipStream_t)stream;
    stereo_status st = cnn_forward(c, n, s);
    if (st) return st;
    const int P = n * c->W4 * c->H4;
    const int nblk = (P + 255) / 256;
    HIP_TRY(c->ctx,
            stereo::launch_cnn_loss(c->out, c->labels, c->clabels, c->dfill, c->g[0], c->lossp, P, 1.0f / (float)P, kLambda, kTau, s),
            "loss launch");
    // backward, last layer first; g[cur] holds the gradient of the current layer's output
    int cur = 0;
    for (int l = kNumLayers - 1; l >= 0; --l) {
        const LayerDef& L = kLayers[l];
        const size_t o = layer_off(l);
        stereo::CnnDwArgs da;
        da.x = c->act[l];
        da.dy = c->g[cur];
        da.part = c->dwpart;
        da.N = n;
        da.H = c->H4;
        da.W = c->W4;
        da.dil = L.dil;
        HIP_TRY(c->ctx, stereo::launch_cnn_dw(L.cin, L.cout, da, c->grad + o, s), "weight gradient launch");
        if (l > 0) {
            stereo::CnnConvArgs a;
            a.x = c->g[cur];
            a.w = c->wb + o;
            a.bias = nullptr;
            a.mask = c->act[l];
            a.y = c->g[cur ^ 1];
            a.N = n;
            a.H = c->H4;
            a.W = c->W4;
            a.dil = L.dil;
            a.sign = -1;
            HIP_TRY(c->ctx, stereo::launch_cnn_conv(L.cout, L.cin, stereo::CNN_EPI_MASK_BF16, a, s), "data gradient launch");
            cur ^= 1;
        }
    }
    if ((st = refresh_weights(c, lr, s))) return st;
    if (loss) {
        std::vector<float> part((size_t)nblk * 4);
        HIP_TRY(c->ctx, hipMemcpyAsync(part.data(), c->lossp, part.size() * 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        HIP_TRY(c->ctx, hipStreamSynchronize(s), "hipStreamSynchronize");
        double acc[4] = {0, 0, 0, 0};
        for (int b = 0; b < nblk; ++b)
            for (int k = 0; k < 4; ++k) acc[k] += part[(size_t)b * 4 + k];
        const double cnt = acc[2] > 0 ? acc[2] : 1.0;
        loss[0] = (float)(acc[0] / cnt);
        loss[1] = (float)(acc[1] / cnt);
        loss[2] = (float)acc[2];
        loss[3] = (float)(acc[3] / cnt);
    }
    return STEREO_OK;
}

stereo_status stereo_conv3x3_bf16(stereo_context* ctx, const uint16_t* x, int32_t N, int32_t H, int32_t W, int32_t cin,
                                  const uint16_t* w, const float* bias, int32_t cout, int32_t dil, int32_t relu, void* y,
                                  stereo_stream stream)
{
    if (!ctx || !x || !w || !y || N < 1 || H < 1 || W < 1 || dil < 1) return STEREO_ERROR_INVALID_ARGUMENT;
    if ((cin != 16 && cin != 32) || (cout != 16 && cout != 32))
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "channels %d -> %d not supported (16 or 32)", cin, cout);
    if (dil > stereo::CNN_MAX_DIL)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "dilation %d above %d", dil, stereo::CNN_MAX_DIL);
    if (((uintptr_t)x | (uintptr_t)w | (uintptr_t)y) & 15)
        return fail(ctx, STEREO_ERROR_INVALID_ARGUMENT, "x, w and y must be 16-byte aligned device pointers");
    stereo::CnnConvArgs a;
    a.x = x;
    a.w = w;
    a.bias = bias;
    a.mask = nullptr;
    a.y = y;
    a.N = N;
    a.H = H;
    a.W = W;
    a.dil = dil;
    a.sign = 1;
    const hipError_t e = stereo::launch_cnn_conv(cin, cout, relu ? stereo::CNN_EPI_RELU_BF16 : stereo::CNN_EPI_F32, a,
                                                 (hipStream_t)stream);
    if (e == hipErrorInvalidValue)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "conv %d -> %d relu=%d has no kernel", cin, cout, relu);
    HIP_TRY(ctx, e, "conv launch");
    return STEREO_OK;
}

stereo_status stereo_conv3x3_bwd_bf16(stereo_context* ctx, const uint16_t* x, int32_t N, int32_t H, int32_t W,
                                      int32_t cin, const uint16_t* wt, const uint16_t* dy, int32_t cout, int32_t dil,
                                      float* dx, float* dwdb, stereo_stream stream)
{
    if (!ctx || !x || !dy || !dwdb || N < 1 || H < 1 || W < 1 || dil < 1 || (dx && !wt))
        return STEREO_ERROR_INVALID_ARGUMENT;
    if ((cin != 16 && cin != 32) || (cout != 16 && cout != 32))
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "channels %d -> %d not supported", cin, cout);
    if (dil > stereo::CNN_MAX_DIL)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "dilation %d above %d", dil, stereo::CNN_MAX_DIL);
    if (((uintptr_t)x | (uintptr_t)wt | (uintptr_t)dy | (uintptr_t)dx | (uintptr_t)dwdb) & 15)
        return fail(ctx, STEREO_ERROR_INVALID_ARGUMENT, "device pointers must be 16-byte aligned");
    const hipStream_t s = (hipStream_t)stream;
    if (dx) {
        stereo::CnnConvArgs a;
        a.x = dy;
        a.w = wt;
        a.bias = nullptr;
        a.mask = nullptr;
        a.y = dx;
        a.N = N;
        a.H = H;
        a.W = W;
        a.dil = dil;
        a.sign = -1;
        const hipError_t e = stereo::launch_cnn_conv(cout, cin, stereo::CNN_EPI_F32, a, s);
        if (e == hipErrorInvalidValue)
            return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "data gradient %d -> %d has no kernel", cout, cin);
        HIP_TRY(ctx, e, "conv launch");
    }
    stereo::CnnDwArgs da;
    da.x = x;
    da.dy = dy;
    da.N = N;
    da.H = H;
    da.W = W;
    da.dil = dil;
    const int parts = stereo::cnn_dw_parts(da);
    float* part = nullptr;
    HIP_TRY(ctx, hipMalloc((void**)&part, (size_t)parts * (cout * 9 * cin + cout) * 4), "hipMalloc");
    da.part = part;
    hipError_t e = stereo::launch_cnn_dw(cin, cout, da, dwdb, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(part);
    HIP_TRY(ctx, e, "weight gradient");
    return STEREO_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------
// learned-proposal census matcher
// ------------------------------------------------------------------------------------
namespace {

struct LearnedExt : DriverExt {
    stereo_learned_params lp{};
    stereo_context* ctx = nullptr;
    stereo_driver* coarse = nullptr;
    stereo_cnn* cnn = nullptr;
    stereo_image L4{}, R4{}, dc{};
    uint8_t* padL = nullptr;
    uint8_t* padR = nullptr;
    size_t pad_pitch = 0;
    stereo::BmFrame* frame_dev = nullptr;
    stereo::BmFrame frame{};  // last frame table written to the device
    stereo::BmTile* tiles = nullptr;
    stereo::BmItem* items = nullptr;
    int* nitems = nullptr;
    int* tile_evals = nullptr;
    uint32_t* tile_mask = nullptr;
    int* counter = nullptr;  // item counter of the list matcher (reset by its finish)
    // the band search's key planes (left, right; W x H), zero at every matcher launch:
    // cleared at configure and by every merge; planes_dirty when a matcher ran without it
    uint32_t* lplane = nullptr;
    uint32_t* rplane = nullptr;
    bool planes_dirty = false;
    int ntx = 0, nty = 0, max_items = 0;
    hipEvent_t ev[6] = {};
    hipEvent_t ev_done = nullptr;  // recorded after every process call's last launch
    bool done_recorded = false;
    bool timed = false;
    // the launch sequence captured for a given pair of input images
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* graph_key[2] = {nullptr, nullptr};

    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_key[0] = graph_key[1] = nullptr;
    }

    ~LearnedExt() override
    {
        drop_graph();
        if (cap) (void)hipStreamDestroy(cap);
        if (coarse) stereo_driver_destroy(coarse);
        if (cnn) stereo_cnn_destroy(cnn);
        for (stereo_image* im : {&L4, &R4, &dc})
            if (im->data) (void)hipFree(im->data);
        for (void* p : {(void*)padL, (void*)padR, (void*)frame_dev, (void*)tiles, (void*)items, (void*)nitems,
                        (void*)tile_evals, (void*)tile_mask, (void*)counter, (void*)lplane, (void*)rplane})
            if (p) (void)hipFree(p);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (ev_done) (void)hipEventDestroy(ev_done);
    }
};

}  // namespace

extern "C" {

stereo_status stereo_learned_configure(stereo_context* ctx, stereo_driver** out, const stereo_image_format* fmt,
                                       const stereo_learned_params* p, const float* cnn_params, size_t n)
{
    if (!ctx || !out || !fmt || !p || !cnn_params) return STEREO_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!format_valid(fmt)) return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "Invalid image format");
    if (fmt->pixel_type != STEREO_PIXEL_LUM_U8)
        return fail(ctx, STEREO_ERROR_UNSUPPORTED_PIXEL_TYPE, "learned matcher: LUM_U8 images only");
    if (fmt->width > stereo::BM_LIST_MAX_WIDTH)
        return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "learned matcher: width above %d", stereo::BM_LIST_MAX_WIDTH);
    if (fmt->width % 4 || fmt->height % 4 || fmt->width < 128 || fmt->height < 96)
        return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "learned matcher: size %dx%d must be multiples of 4, at least 128x96",
                    fmt->width, fmt->height);
    if (p->matcher.algorithm != STEREO_ALGO_CENSUS)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "learned matcher: the full-resolution matcher is the census matcher");
    char msg[256];
    stereo_status st = stereo_matcher_validate(fmt, &p->matcher, msg, sizeof(msg));
    if (st) return fail(ctx, st, "%s", msg);
    const int D = p->matcher.num_disparities;
    int Dc = p->coarse_disparities ? p->coarse_disparities : ((D / 4 + 7) & ~7);
    if (Dc < 8) Dc = 8;
    if (Dc % 8 || Dc > 256 || p->band < 8 || p->band % 8 || p->chunk < 8 || p->chunk % 8)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "learned matcher: coarse D %d, band %d, chunk %d must be multiples of 8",
                    Dc, p->band, p->chunk);
    if ((p->centre_residual != 0 && p->centre_residual != 1) || (p->graph_replay != 0 && p->graph_replay != 1) ||
        p->reserved[0] || p->reserved[1])
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "learned matcher: centre_residual and graph_replay are 0 or 1, "
                    "reserved fields 0");
    if ((st = set_device(ctx))) return st;

    stereo_driver* d = new stereo_driver(stereo_driver::LEARNED);
    d->ctx = ctx;
    d->fmt = *fmt;
    d->mp = p->matcher;
    d->name = "Learned Proposal Census Matcher";
    LearnedExt* x = new LearnedExt();
    d->ext = x;
    x->lp = *p;
    x->lp.coarse_disparities = Dc;
    x->ctx = ctx;
    auto cleanup = [&](stereo_status s2) {
        driver_release(d);
        delete d;
        return s2;
    };
    if ((st = driver_events(d))) return cleanup(st);
    for (hipEvent_t& e : x->ev) HIP_TRY(ctx, hipEventCreate(&e), "hipEventCreate");

    const int W = fmt->width, H = fmt->height, W4 = W / 4, H4 = H / 4;
    stereo_image_format f4{W4, H4, STEREO_PIXEL_LUM_U8};
    stereo_matcher_params cp = p->matcher;
    cp.num_disparities = Dc;
    cp.use_disp_input = 0;
    if ((st = stereo_matcher_configure(ctx, &x->coarse, &f4, &cp))) return cleanup(st);
    if ((st = stereo_cnn_create(ctx, &x->cnn, W4, H4, 1))) return cleanup(st);
    if ((st = stereo_cnn_set_params(x->cnn, cnn_params, n))) return cleanup(st);
    if ((st = stereo_image_alloc(ctx, &x->L4, W4, H4, STEREO_PIXEL_LUM_U8))) return cleanup(st);
    if ((st = stereo_image_alloc(ctx, &x->R4, W4, H4, STEREO_PIXEL_LUM_U8))) return cleanup(st);
    if ((st = stereo_image_alloc(ctx, &x->dc, W4, H4, STEREO_PIXEL_LUM_I16))) return cleanup(st);
    if (W4 > stereo::CNN_MAX_W4)
        return cleanup(fail(ctx, STEREO_ERROR_INVALID_CONFIG, "learned matcher: quarter width %d above %d", W4,
                            stereo::CNN_MAX_W4));
    // the coarse matcher's merge writes the network input (the work of stereo_cnn_set_input)
    // row by row, right after each row of its checked map
    x->coarse->finish_cnn = stereo::CnnInputTarget{view_of(&x->L4), view_of(&x->R4), x->cnn->act[0], x->cnn->dfill, 0};

    d->bx = p->matcher.window_w / 2 + p->matcher.census_w / 2;
    d->by = p->matcher.window_h / 2 + p->matcher.census_h / 2;
    x->ntx = (W + stereo::BM_TILE_W - 1) / stereo::BM_TILE_W;
    x->nty = (H + stereo::BM_TILE_H - 1) / stereo::BM_TILE_H;
    // at most 16 runs of bins per tile, each cut into pieces of at most `chunk`
    x->max_items = x->ntx * x->nty * (16 + (D + p->chunk - 1) / p->chunk);
    const size_t cols = (size_t)stereo::BM_PAD_LEFT + W + stereo::BM_PAD_RIGHT;
    x->pad_pitch = (cols * 8 + 255) & ~(size_t)255;
    hipError_t e;
    const int nt = x->ntx * x->nty;
    if ((e = hipMalloc((void**)&x->padL, x->pad_pitch * H)) || (e = hipMalloc((void**)&x->padR, x->pad_pitch * H)) ||
        (e = hipMalloc((void**)&x->frame_dev, sizeof(stereo::BmFrame))) ||
        (e = hipMalloc((void**)&x->tiles, nt * sizeof(stereo::BmTile))) ||
        (e = hipMalloc((void**)&x->items, x->max_items * sizeof(stereo::BmItem))) ||
        (e = hipMalloc((void**)&x->nitems, sizeof(int))) || (e = hipMalloc((void**)&x->tile_evals, nt * sizeof(int))) ||
        (e = hipMalloc((void**)&x->tile_mask, nt * sizeof(uint32_t))) ||
        (e = hipMalloc((void**)&x->counter, sizeof(int))) ||
        (e = hipStreamCreateWithFlags(&x->cap, hipStreamNonBlocking)) || (e = hipEventCreate(&x->ev_done)) ||
        (e = hipMalloc((void**)&x->lplane, (size_t)W * H * 4)) || (e = hipMalloc((void**)&x->rplane, (size_t)W * H * 4)))
        return cleanup(hip_fail(ctx, e, "hipMalloc"));
    if ((e = hipMemset(x->padL, 0, x->pad_pitch * H)) || (e = hipMemset(x->padR, 0, x->pad_pitch * H)) ||
        (e = hipMemset(x->lplane, 0, (size_t)W * H * 4)) || (e = hipMemset(x->rplane, 0, (size_t)W * H * 4)) ||
        (e = hipMemset(x->counter, 0, sizeof(int))) || (e = hipDeviceSynchronize()))
        return cleanup(hip_fail(ctx, e, "hipMemset"));
    stereo::BmFrame& f = x->frame;
    std::memset(&f, 0, sizeof(f));
    f.L.ptr = x->padL + (size_t)stereo::BM_PAD_LEFT * 8;
    f.R.ptr = x->padR + (size_t)stereo::BM_PAD_LEFT * 8;
    f.L.pitch = f.R.pitch = x->pad_pitch;
    f.L.w = f.R.w = W;
    f.L.h = f.R.h = H;
    f.W = W;
    f.H = H;
    f.D = D;
    f.bx = d->bx;
    f.by = d->by;
    f.tol = p->matcher.lr_tolerance;
    f.ntx = x->ntx;
    f.nty = x->nty;
    if ((e = hipMemcpy(x->frame_dev, &f, sizeof(f), hipMemcpyHostToDevice))) return cleanup(hip_fail(ctx, e, "hipMemcpy"));
    *out = d;
    return STEREO_OK;
}

namespace {

// The whole learned pipeline on stream s; with `timed`, events between the stages.
stereo_status learned_enqueue(stereo_driver* d, LearnedExt* x, const stereo_image* left, const stereo_image* right,
                              hipStream_t s, bool timed)
{
    stereo_context* ctx = d->ctx;
    const int W = d->fmt.width, H = d->fmt.height;
    const stereo::BmFrame& f = x->frame;
    stereo_status st;
    if (timed) HIP_TRY(ctx, hipEventRecord(x->ev[0], s), "hipEventRecord");
    const View vl = view_of(left), vr = view_of(right);
    const View q0 = view_of(&x->L4), q1 = view_of(&x->R4);
    HIP_TRY(ctx, stereo::launch_census_pair(vl, f.L, &vr, &f.R, d->mp.census_w, d->mp.census_h, s, &q0, &q1),
            "census launch");
    if (timed) HIP_TRY(ctx, hipEventRecord(x->ev[1], s), "hipEventRecord");
    if ((st = stereo_matcher_process(x->coarse, &x->dc, nullptr, nullptr, &x->L4, &x->R4, s))) return st;
    if (timed) HIP_TRY(ctx, hipEventRecord(x->ev[2], s), "hipEventRecord");
    if ((st = stereo_cnn_forward(x->cnn, 1, nullptr, s))) return st;
    if (timed) HIP_TRY(ctx, hipEventRecord(x->ev[3], s), "hipEventRecord");
    stereo::CnnBandArgs ba;
    ba.out = x->cnn->out;
    ba.dfill = x->cnn->dfill;
    ba.tiles = x->tiles;
    ba.items = x->items;
    ba.nitems = x->nitems;
    ba.list_counter = x->counter;
    ba.tile_evals = x->tile_evals;
    ba.tile_mask = x->tile_mask;
    ba.centre_residual = x->lp.centre_residual;
    ba.W4 = W / 4;
    ba.H4 = H / 4;
    ba.ntx = x->ntx;
    ba.nty = x->nty;
    ba.D = d->mp.num_disparities;
    ba.band = x->lp.band;
    ba.chunk = x->lp.chunk;
    ba.logit_thresh = x->lp.logit_thresh;
    HIP_TRY(ctx, stereo::launch_band_items(ba, s), "band launch");
    stereo::BmListParams lp;
    lp.frames = x->frame_dev;
    lp.tiles = x->tiles;
    lp.items = x->items;
    lp.nitems = x->nitems;
    lp.sched = nullptr;
    lp.counter = x->counter;
    lp.lplane = x->lplane;
    lp.rplane = x->rplane;
    lp.nframes = 1;
    lp.total_rows = H;
    const int grid = x->max_items < 768 ? x->max_items : 768;
    x->planes_dirty = true;  // until the merge is enqueued behind the matcher
    HIP_TRY(ctx, stereo::launch_bm_list(stereo::BM_COST_CENSUS, d->mp.window_w, lp, grid, s), "list matcher launch");
    if (timed) HIP_TRY(ctx, hipEventRecord(x->ev[4], s), "hipEventRecord");
    HIP_TRY(ctx, stereo::launch_bm_list_finish(lp, W, s), "list finish launch");
    x->planes_dirty = false;
    if (timed) HIP_TRY(ctx, hipEventRecord(x->ev[5], s), "hipEventRecord");
    return STEREO_OK;
}

}  // namespace

stereo_status stereo_learned_process(stereo_driver* d, stereo_image* disp, stereo_image* disp_lr, stereo_image* disp_rl,
                                     const stereo_image* left, const stereo_image* right, stereo_stream stream)
{
    if (!d || d->kind != stereo_driver::LEARNED || !left || !right) return STEREO_ERROR_INVALID_ARGUMENT;
    LearnedExt* x = static_cast<LearnedExt*>(d->ext);
    stereo_context* ctx = d->ctx;
    const int W = d->fmt.width, H = d->fmt.height;
    stereo_status st;
    if ((st = check_image(ctx, left, W, H, STEREO_PIXEL_LUM_U8, "left image"))) return st;
    if ((st = check_image(ctx, right, W, H, STEREO_PIXEL_LUM_U8, "right image"))) return st;
    if (disp && (st = check_image(ctx, disp, W, H, STEREO_PIXEL_LUM_I16, "disparity output"))) return st;
    if (disp_lr && (st = check_image(ctx, disp_lr, W, H, STEREO_PIXEL_LUM_I16, "left-to-right output"))) return st;
    if (disp_rl && (st = check_image(ctx, disp_rl, W, H, STEREO_PIXEL_LUM_I16, "right-to-left output"))) return st;
    if ((st = set_device(ctx))) return st;
    GUARD(ctx, "stereo_learned_process");
    const hipStream_t s = (hipStream_t)stream;

    // outputs live in the frame table; rewrite it (synchronously) when they change
    stereo::BmFrame f = x->frame;
    f.disp = disp ? view_of(disp) : null_view();
    f.disp_l = disp_lr ? view_of(disp_lr) : null_view();
    f.disp_r = disp_rl ? view_of(disp_rl) : null_view();
    f.dispid = null_view();
    if (std::memcmp(&f, &x->frame, sizeof(f)) != 0) {
        // the previous call's launches read the table and may still run on another stream
        if (x->done_recorded) HIP_TRY(ctx, hipEventSynchronize(x->ev_done), "hipEventSynchronize");
        HIP_TRY(ctx, hipStreamSynchronize(s), "hipStreamSynchronize");
        HIP_TRY(ctx, hipMemcpy(x->frame_dev, &f, sizeof(f), hipMemcpyHostToDevice), "hipMemcpy");
        x->frame = f;
    }

    if (x->planes_dirty) {
        // an earlier frame's band search ran without its merge: clear its keys
        const size_t n = (size_t)W * H;
        HIP_TRY(ctx, stereo::launch_bm_clear(x->lplane, n, s), "clear launch");
        HIP_TRY(ctx, stereo::launch_bm_clear(x->rplane, n, s), "clear launch");
        x->planes_dirty = false;
    }

    // the stages run directly (timed calls with events between them); with graph_replay
    // the sequence is captured once per pair of input images and replayed as a graph
    x->timed = ctx->timing != 0;
    if (x->timed || !x->lp.graph_replay) {
        if ((st = learned_enqueue(d, x, left, right, s, x->timed))) return st;
        HIP_TRY(ctx, hipEventRecord(x->ev_done, s), "hipEventRecord");
        x->done_recorded = true;
        return STEREO_OK;
    }
    if (!x->exec || x->graph_key[0] != left->data || x->graph_key[1] != right->data) {
        x->drop_graph();
        HIP_TRY(ctx, hipStreamBeginCapture(x->cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        const stereo_status cst = learned_enqueue(d, x, left, right, x->cap, false);
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(x->cap, &g);
        if (cst) {
            if (g) (void)hipGraphDestroy(g);
            return cst;
        }
        HIP_TRY(ctx, ee, "hipStreamEndCapture");
        x->graph = g;
        HIP_TRY(ctx, hipGraphInstantiate(&x->exec, x->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        x->graph_key[0] = left->data;
        x->graph_key[1] = right->data;
    }
    HIP_TRY(ctx, hipGraphLaunch(x->exec, s), "hipGraphLaunch");
    HIP_TRY(ctx, hipEventRecord(x->ev_done, s), "hipEventRecord");
    x->done_recorded = true;
    return STEREO_OK;
}

stereo_status stereo_learned_stats(stereo_driver* d, int64_t* evals, int32_t* tiles, int32_t* fallback_tiles,
                                   double* stage_ms, int32_t n_stages)
{
    if (!d || d->kind != stereo_driver::LEARNED) return STEREO_ERROR_INVALID_ARGUMENT;
    LearnedExt* x = static_cast<LearnedExt*>(d->ext);
    stereo_context* ctx = d->ctx;
    HIP_TRY(ctx, hipDeviceSynchronize(), "hipDeviceSynchronize");
    const int nt = x->ntx * x->nty;
    std::vector<int> te(nt);
    HIP_TRY(ctx, hipMemcpy(te.data(), x->tile_evals, nt * sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
    const int W = d->fmt.width, H = d->fmt.height;
    int64_t ev = 0;
    int fb = 0;
    for (int t = 0; t < nt; ++t) {
        const int tx = t % x->ntx, ty = t / x->ntx;
        const int w = std::min(stereo::BM_TILE_W, W - tx * stereo::BM_TILE_W);
        const int h = std::min(stereo::BM_TILE_H, H - ty * stereo::BM_TILE_H);
        ev += (int64_t)(te[t] & 0xFFFF) * w * h;
        fb += (te[t] >> 16) & 1;
    }
    if (evals) *evals = ev;
    if (tiles) *tiles = nt;
    if (fallback_tiles) *fallback_tiles = fb;
    if (stage_ms && n_stages > 0) {
        if (!x->timed) return fail(ctx, STEREO_ERROR_NOT_READY, "%s: no timed process call", d->name);
        for (int i = 0; i < n_stages && i < 5; ++i) {
            float t = 0.f;
            HIP_TRY(ctx, hipEventElapsedTime(&t, x->ev[i], x->ev[i + 1]), "hipEventElapsedTime");
            stage_ms[i] = t;
        }
    }
    return STEREO_OK;
}

stereo_status stereo_learned_tile_masks(stereo_driver* d, uint32_t* masks, int32_t n)
{
    if (!d || d->kind != stereo_driver::LEARNED || !masks) return STEREO_ERROR_INVALID_ARGUMENT;
    LearnedExt* x = static_cast<LearnedExt*>(d->ext);
    const int nt = x->ntx * x->nty;
    if (n < nt) return fail(d->ctx, STEREO_ERROR_INVALID_ARGUMENT, "need %d entries", nt);
    HIP_TRY(d->ctx, hipDeviceSynchronize(), "hipDeviceSynchronize");
    HIP_TRY(d->ctx, hipMemcpy(masks, x->tile_mask, nt * sizeof(uint32_t), hipMemcpyDeviceToHost), "hipMemcpy");
    return STEREO_OK;
}

stereo_status stereo_learned_tile_ranges(stereo_driver* d, int32_t* lohi, int32_t n)
{
    if (!d || d->kind != stereo_driver::LEARNED || !lohi) return STEREO_ERROR_INVALID_ARGUMENT;
    LearnedExt* x = static_cast<LearnedExt*>(d->ext);
    const int nt = x->ntx * x->nty;
    if (n < 2 * nt) return fail(d->ctx, STEREO_ERROR_INVALID_ARGUMENT, "need %d entries", 2 * nt);
    HIP_TRY(d->ctx, hipDeviceSynchronize(), "hipDeviceSynchronize");
    std::vector<stereo::BmTile> t(nt);
    HIP_TRY(d->ctx, hipMemcpy(t.data(), x->tiles, nt * sizeof(stereo::BmTile), hipMemcpyDeviceToHost), "hipMemcpy");
    for (int i = 0; i < nt; ++i) {
        lohi[2 * i] = t[i].lo;
        lohi[2 * i + 1] = t[i].hi;
    }
    return STEREO_OK;
}

}  // extern "C"
