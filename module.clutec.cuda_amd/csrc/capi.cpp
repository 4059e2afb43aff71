// C ABI of the gfx950 stereo engine: contexts, device images and operator drivers.
// The declarations (and the reference interface each entry point replaces) are in
// include/clutec_stereo.h.
#include "capi_internal.hpp"
#include "synth.h"

using namespace capi;

extern "C" {

int stereo_abi_version(void) { return STEREO_ABI_VERSION; }

const char* stereo_status_string(stereo_status s)
{
    switch (s) {
    case STEREO_OK: return "ok";
    case STEREO_ERROR_INVALID_ARGUMENT: return "invalid argument";
    case STEREO_ERROR_INVALID_CONFIG: return "invalid algorithm configuration";
    case STEREO_ERROR_UNSUPPORTED_PIXEL_TYPE: return "pixel types of given images not supported";
    case STEREO_ERROR_FORMAT_MISMATCH: return "image formats do not match";
    case STEREO_ERROR_INVALID_DEVICE: return "invalid device";
    case STEREO_ERROR_INVALID_FORMAT: return "invalid image format";
    case STEREO_ERROR_HIP: return "HIP runtime error";
    case STEREO_ERROR_OUT_OF_MEMORY: return "out of device memory";
    case STEREO_ERROR_NOT_READY: return "not ready";
    }
    return "unknown status";
}

stereo_status stereo_create(stereo_context** out, int device)
{
    if (!out) return fail(nullptr, STEREO_ERROR_INVALID_ARGUMENT, "stereo_create: null output");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return fail(nullptr, STEREO_ERROR_INVALID_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= n)
        return fail(nullptr, STEREO_ERROR_INVALID_DEVICE, "device index %d out of range [0, %d)", device, n);
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return fail(nullptr, STEREO_ERROR_INVALID_DEVICE, "hipGetDeviceProperties: %s", hipGetErrorString(e));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(nullptr, STEREO_ERROR_INVALID_DEVICE, "device %d is %s; this library is built for gfx950", device,
                    prop.gcnArchName);
    stereo_context* ctx = new stereo_context();
    ctx->device = device;
    e = hipSetDevice(device);
    if (e != hipSuccess) {
        delete ctx;
        return fail(nullptr, STEREO_ERROR_INVALID_DEVICE, "hipSetDevice: %s", hipGetErrorString(e));
    }
    e = stereo::noalloc_setup(device);
    if (e != hipSuccess) {
        delete ctx;
        return fail(nullptr, STEREO_ERROR_HIP, "scratch array of the no-allocation matcher: %s", hipGetErrorString(e));
    }
    *out = ctx;
    return STEREO_OK;
}

void stereo_destroy(stereo_context* ctx)
{
    if (ctx && ctx->stage) (void)hipHostFree(ctx->stage);
    delete ctx;
}

const char* stereo_last_error(const stereo_context* ctx)
{
    if (ctx) return ctx->last_error.c_str();
    std::lock_guard<std::mutex> lk(g_err_mutex);
    return g_create_error.c_str();
}

stereo_status stereo_set_timing(stereo_context* ctx, int enable)
{
    if (!ctx) return STEREO_ERROR_INVALID_ARGUMENT;
    ctx->timing = enable ? 1 : 0;
    return STEREO_OK;
}

stereo_status stereo_synchronize(stereo_context* ctx, stereo_stream stream)
{
    if (!ctx) return STEREO_ERROR_INVALID_ARGUMENT;
    HIP_TRY(ctx, hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
    return STEREO_OK;
}

size_t stereo_pixel_size(int32_t t)
{
    switch (t) {
    case STEREO_PIXEL_LUM_U8: return 1;
    case STEREO_PIXEL_LUM_U16: return 2;
    case STEREO_PIXEL_LUM_I16: return 2;
    case STEREO_PIXEL_LUM_U32: return 4;
    case STEREO_PIXEL_LUM_U64: return 8;
    case STEREO_PIXEL_LUM_F32: return 4;
    case STEREO_PIXEL_RGBA_U16: return 8;
    case STEREO_PIXEL_RGBA_U8: return 4;
    default: return 0;
    }
}

stereo_status stereo_image_alloc(stereo_context* ctx, stereo_image* img, int32_t w, int32_t h, int32_t type)
{
    if (!ctx || !img) return STEREO_ERROR_INVALID_ARGUMENT;
    stereo_image_format f{w, h, type};
    if (!format_valid(&f)) return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "invalid image format %dx%d type %d", w, h, type);
    stereo_status st = set_device(ctx);
    if (st) return st;
    GUARD(ctx, "stereo_image_alloc");
    const size_t row = (size_t)w * stereo_pixel_size(type);
    const size_t pitch = (row + 255) & ~(size_t)255;  // 256-byte rows: whole cache lines
    void* p = nullptr;
    HIP_TRY(ctx, hipMalloc(&p, pitch * (size_t)h), "hipMalloc");
    img->data = p;
    img->pitch = pitch;
    img->format = f;
    return STEREO_OK;
}

stereo_status stereo_image_free(stereo_context* ctx, stereo_image* img)
{
    if (!ctx || !img) return STEREO_ERROR_INVALID_ARGUMENT;
    if (img->data) HIP_TRY(ctx, hipFree(img->data), "hipFree");
    img->data = nullptr;
    return STEREO_OK;
}

namespace {

// Host rows <-> device rows through the context's pinned staging buffer: rows are
// packed on the host side and moved with one 2-D copy per chunk.  A 2-D copy straight
// from pageable memory goes row by row (about 25 us a row), which this avoids.
constexpr size_t kStageBytes = (size_t)8 << 20;

stereo_status staged_copy(stereo_context* ctx, bool to_device, uint8_t* dev, size_t dev_pitch, uint8_t* host,
                          size_t host_pitch, size_t row, int rows, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(ctx->stage_mu);
    const size_t need = row > kStageBytes ? row : kStageBytes;
    if (ctx->stage_bytes < need) {
        if (ctx->stage) (void)hipHostFree(ctx->stage);
        ctx->stage = nullptr;
        ctx->stage_bytes = 0;
        HIP_TRY(ctx, hipHostMalloc(&ctx->stage, need, hipHostMallocDefault), "hipHostMalloc");
        ctx->stage_bytes = need;
    }
    uint8_t* st = static_cast<uint8_t*>(ctx->stage);
    const int chunk = (int)(ctx->stage_bytes / row);
    for (int r0 = 0; r0 < rows; r0 += chunk) {
        const int n = rows - r0 < chunk ? rows - r0 : chunk;
        if (to_device) {
            for (int i = 0; i < n; ++i) std::memcpy(st + (size_t)i * row, host + (size_t)(r0 + i) * host_pitch, row);
            HIP_TRY(ctx, hipMemcpy2DAsync(dev + (size_t)r0 * dev_pitch, dev_pitch, st, row, row, n, hipMemcpyHostToDevice, s),
                    "hipMemcpy2DAsync");
            HIP_TRY(ctx, hipStreamSynchronize(s), "hipStreamSynchronize");
        } else {
            HIP_TRY(ctx, hipMemcpy2DAsync(st, row, dev + (size_t)r0 * dev_pitch, dev_pitch, row, n, hipMemcpyDeviceToHost, s),
                    "hipMemcpy2DAsync");
            HIP_TRY(ctx, hipStreamSynchronize(s), "hipStreamSynchronize");
            for (int i = 0; i < n; ++i) std::memcpy(host + (size_t)(r0 + i) * host_pitch, st + (size_t)i * row, row);
        }
    }
    return STEREO_OK;
}

}  // namespace

stereo_status stereo_image_upload(stereo_context* ctx, stereo_image* img, const void* host, size_t host_pitch,
                                  stereo_stream stream)
{
    if (!ctx || !img || !img->data || !host) return STEREO_ERROR_INVALID_ARGUMENT;
    const size_t row = (size_t)img->format.width * stereo_pixel_size(img->format.pixel_type);
    if (host_pitch < row) return fail(ctx, STEREO_ERROR_INVALID_ARGUMENT, "host pitch %zu < row bytes %zu", host_pitch, row);
    if (row == 0 || img->format.height < 1) return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "empty image");
    stereo_status st = set_device(ctx);
    if (st) return st;
    return staged_copy(ctx, true, static_cast<uint8_t*>(img->data), img->pitch,
                       const_cast<uint8_t*>(static_cast<const uint8_t*>(host)), host_pitch, row, img->format.height,
                       (hipStream_t)stream);
}

stereo_status stereo_image_download(stereo_context* ctx, const stereo_image* img, void* host, size_t host_pitch,
                                    stereo_stream stream)
{
    if (!ctx || !img || !img->data || !host) return STEREO_ERROR_INVALID_ARGUMENT;
    const size_t row = (size_t)img->format.width * stereo_pixel_size(img->format.pixel_type);
    if (host_pitch < row) return fail(ctx, STEREO_ERROR_INVALID_ARGUMENT, "host pitch %zu < row bytes %zu", host_pitch, row);
    if (row == 0 || img->format.height < 1) return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "empty image");
    stereo_status st = set_device(ctx);
    if (st) return st;
    return staged_copy(ctx, false, static_cast<uint8_t*>(img->data), img->pitch, static_cast<uint8_t*>(host), host_pitch,
                       row, img->format.height, (hipStream_t)stream);
}

stereo_status stereo_device_alloc(stereo_context* ctx, void** ptr, size_t bytes)
{
    if (!ctx || !ptr || bytes == 0) return STEREO_ERROR_INVALID_ARGUMENT;
    *ptr = nullptr;
    stereo_status st = set_device(ctx);
    if (st) return st;
    GUARD(ctx, "stereo_device_alloc");
    HIP_TRY(ctx, hipMalloc(ptr, bytes), "hipMalloc");
    return STEREO_OK;
}

stereo_status stereo_device_free(stereo_context* ctx, void* ptr)
{
    if (!ctx) return STEREO_ERROR_INVALID_ARGUMENT;
    if (ptr) HIP_TRY(ctx, hipFree(ptr), "hipFree");
    return STEREO_OK;
}

stereo_status stereo_memcpy(stereo_context* ctx, void* dst, const void* src, size_t bytes, int32_t kind)
{
    if (!ctx || !dst || !src || kind < 0 || kind > 2) return STEREO_ERROR_INVALID_ARGUMENT;
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    HIP_TRY(ctx, hipMemcpy(dst, src, bytes, k), "hipMemcpy");
    return STEREO_OK;
}

stereo_status stereo_driver_destroy(stereo_driver* d)
{
    if (!d) return STEREO_ERROR_INVALID_ARGUMENT;
    driver_release(d);
    delete d;
    return STEREO_OK;
}

const char* stereo_driver_name(const stereo_driver* d) { return d ? d->name : ""; }

stereo_status stereo_driver_last_time_ms(stereo_driver* d, double* ms)
{
    if (!d || !ms) return STEREO_ERROR_INVALID_ARGUMENT;
    if (!d->timed) return fail(d->ctx, STEREO_ERROR_NOT_READY, "%s: no timed process call (enable stereo_set_timing)", d->name);
    HIP_TRY(d->ctx, hipEventSynchronize(d->ev_stop), "hipEventSynchronize");
    float t = 0.f;
    HIP_TRY(d->ctx, hipEventElapsedTime(&t, d->ev_start, d->ev_stop), "hipEventElapsedTime");
    *ms = t;
    return STEREO_OK;
}

// ---- census -----------------------------------------------------------------------

stereo_status stereo_census_configure(stereo_context* ctx, stereo_driver** out, const stereo_image_format* fmt,
                                      const stereo_census_params* p)
{
    if (!ctx || !out || !fmt || !p) return STEREO_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!format_valid(fmt)) return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "Invalid image format");
    if (!stereo::census_supported(p->window_w, p->window_h, p->reference_3x3))
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "census window %dx%d (reference_3x3=%d) not supported", p->window_w,
                    p->window_h, p->reference_3x3);
    stereo_status st = set_device(ctx);
    if (st) return st;
    stereo_driver* d = new stereo_driver(stereo_driver::CENSUS);
    d->ctx = ctx;
    d->name = "Census Transform";
    d->fmt = *fmt;
    d->cen = *p;
    if ((st = driver_events(d))) { driver_release(d); delete d; return st; }
    *out = d;
    return STEREO_OK;
}

stereo_status stereo_census_process(stereo_driver* d, stereo_image* trg, const stereo_image* src, stereo_stream stream)
{
    if (!d || d->kind != stereo_driver::CENSUS || !trg || !src) return STEREO_ERROR_INVALID_ARGUMENT;
    stereo_context* ctx = d->ctx;
    if (trg->format.width != src->format.width || trg->format.height != src->format.height)
        return fail(ctx, STEREO_ERROR_FORMAT_MISMATCH, "Input and output images do not have the same format");
    stereo_status st;
    if ((st = check_image(ctx, src, d->fmt.width, d->fmt.height, STEREO_PIXEL_LUM_U8, "census source"))) return st;
    const int out_type = d->cen.reference_3x3 ? STEREO_PIXEL_LUM_U8 : STEREO_PIXEL_LUM_U64;
    if ((st = check_image(ctx, trg, d->fmt.width, d->fmt.height, out_type, "census target"))) return st;
    if ((st = set_device(ctx))) return st;
    GUARD(ctx, "stereo_census_process");
    hipStream_t s = (hipStream_t)stream;
    if ((st = begin(d, s))) return st;
    HIP_TRY(ctx, stereo::launch_census(view_of(src), view_of(trg), d->cen.window_w, d->cen.window_h, d->cen.reference_3x3, s),
            "census launch");
    return end(d, s);
}

// ---- matchers -----------------------------------------------------------------------

namespace {

// Work-list buffers of the band search around an initial map: one run of bins per
// tile, cut into items of kBandChunk disparities.
constexpr int kBandChunk = 32;

struct BandExt : DriverExt {
    stereo::BmFrame* frame_dev = nullptr;
    stereo::BmFrame frame{};  // last frame table written to the device
    stereo::BmTile* tiles = nullptr;
    stereo::BmItem* items = nullptr;
    int* nitems = nullptr;
    int* counter = nullptr;
    int* tile_evals = nullptr;
    uint32_t* tile_mask = nullptr;
    // key planes of the SAD / census band search (left, right; W x H), zero at every
    // matcher launch: cleared at allocation and by every merge; planes_dirty when a
    // matcher ran without it
    uint32_t* lplane = nullptr;
    uint32_t* rplane = nullptr;
    bool planes_dirty = false;
    int ntx = 0, nty = 0, max_items = 0;
    hipEvent_t ev_done = nullptr;  // recorded after every list finish launch
    bool done_recorded = false;
    ~BandExt() override
    {
        if (ev_done) (void)hipEventDestroy(ev_done);
        for (void* p : {(void*)frame_dev, (void*)tiles, (void*)items, (void*)nitems, (void*)counter, (void*)tile_evals,
                        (void*)tile_mask, (void*)lplane, (void*)rplane})
            if (p) (void)hipFree(p);
    }
};

stereo_status band_alloc(stereo_driver* d)
{
    BandExt* x = new BandExt();
    d->ext = x;
    const int W = d->fmt.width, H = d->fmt.height, D = d->mp.num_disparities;
    x->ntx = (W + stereo::BM_TILE_W - 1) / stereo::BM_TILE_W;
    x->nty = (H + stereo::BM_TILE_H - 1) / stereo::BM_TILE_H;
    const int nt = x->ntx * x->nty;
    if (d->mp.algorithm == STEREO_ALGO_NCC || d->mp.algorithm == STEREO_ALGO_AW1 ||
        d->mp.algorithm == STEREO_ALGO_AW2) {
        // the NCC and adaptive-window matchers mask their own tiles: only the per-tile bins
        hipError_t e;
        if ((e = hipMalloc((void**)&x->tile_evals, nt * sizeof(int))) ||
            (e = hipMalloc((void**)&x->tile_mask, nt * sizeof(uint32_t))))
            return hip_fail(d->ctx, e, "hipMalloc");
        return STEREO_OK;
    }
    x->max_items = nt * ((D + kBandChunk - 1) / kBandChunk);
    hipError_t e;
    if ((e = hipMalloc((void**)&x->frame_dev, sizeof(stereo::BmFrame))) ||
        (e = hipMalloc((void**)&x->tiles, nt * sizeof(stereo::BmTile))) ||
        (e = hipMalloc((void**)&x->items, (size_t)x->max_items * sizeof(stereo::BmItem))) ||
        (e = hipMalloc((void**)&x->nitems, sizeof(int))) || (e = hipMalloc((void**)&x->counter, sizeof(int))) ||
        (e = hipMalloc((void**)&x->tile_evals, nt * sizeof(int))) ||
        (e = hipMalloc((void**)&x->tile_mask, nt * sizeof(uint32_t))) ||
        (e = hipMalloc((void**)&x->lplane, (size_t)W * H * 4)) || (e = hipMalloc((void**)&x->rplane, (size_t)W * H * 4)) ||
        (e = hipMemset(x->lplane, 0, (size_t)W * H * 4)) || (e = hipMemset(x->rplane, 0, (size_t)W * H * 4)) ||
        (e = hipMemset(x->counter, 0, sizeof(int))) || (e = hipEventCreate(&x->ev_done)) ||
        (e = hipDeviceSynchronize()))
        return hip_fail(d->ctx, e, "hipMalloc");
    stereo::BmFrame& f = x->frame;
    std::memset(&f, 0, sizeof(f));
    f.L.pitch = f.R.pitch = d->pad_pitch;
    f.L.w = f.R.w = W;
    f.L.h = f.R.h = H;
    f.L.ptr = d->padL + (size_t)stereo::BM_PAD_LEFT * d->pad_elem;
    f.R.ptr = d->padR + (size_t)stereo::BM_PAD_LEFT * d->pad_elem;
    f.W = W;
    f.H = H;
    f.D = D;
    f.bx = d->bx;
    f.by = d->by;
    f.tol = d->mp.lr_tolerance;
    f.value_base = 0;
    f.ntx = x->ntx;
    f.nty = x->nty;
    f.tile_base = 0;
    f.row_base = 0;
    HIP_TRY(d->ctx, hipMemcpy(x->frame_dev, &f, sizeof(f), hipMemcpyHostToDevice), "hipMemcpy");
    return STEREO_OK;
}

}  // namespace


stereo_status stereo_matcher_validate(const stereo_image_format* fmt, const stereo_matcher_params* p, char* msg,
                                      size_t msg_len)
{
    auto say = [&](stereo_status st, const char* text) {
        if (msg && msg_len) std::snprintf(msg, msg_len, "%s", text);
        return st;
    };
    if (!fmt || !p) return say(STEREO_ERROR_INVALID_ARGUMENT, "null argument");
    if (!format_valid(fmt)) return say(STEREO_ERROR_INVALID_FORMAT, "Invalid image format");
    if (fmt->pixel_type != STEREO_PIXEL_LUM_U8)
        return say(STEREO_ERROR_UNSUPPORTED_PIXEL_TYPE, "matchers take LUM_U8 images");
    for (int i = 0; i < 4; ++i)
        if (p->reserved[i]) return say(STEREO_ERROR_INVALID_CONFIG, "reserved parameters must be 0");
    if (p->use_disp_input != 0 && p->use_disp_input != 1)
        return say(STEREO_ERROR_INVALID_CONFIG, "use_disp_input must be 0 or 1");
    if (p->use_disp_input && p->algorithm != STEREO_ALGO_SAD && p->algorithm != STEREO_ALGO_CENSUS &&
        p->algorithm != STEREO_ALGO_NCC && p->algorithm != STEREO_ALGO_AW1 && p->algorithm != STEREO_ALGO_AW2)
        return say(STEREO_ERROR_INVALID_CONFIG, "the band search around an initial disparity map is built for the SAD, "
                                                "census, NCC and adaptive-window matchers");
    if (p->lr_tolerance < 0) return say(STEREO_ERROR_INVALID_CONFIG, "lr_tolerance must be >= 0");
    if (p->num_disparities < 8 || p->num_disparities > 256 || p->num_disparities % 8)
        return say(STEREO_ERROR_INVALID_CONFIG, "num_disparities must be a multiple of 8 in [8, 256]");
    switch (p->algorithm) {
    case STEREO_ALGO_SAD:
        if (!stereo::bm_supported(stereo::BM_COST_SAD, p->window_w, p->window_h, p->num_disparities))
            return say(STEREO_ERROR_INVALID_CONFIG, "SAD window must be 5x5, 7x7 or 9x9");
        break;
    case STEREO_ALGO_CENSUS:
        if (!stereo::bm_supported(stereo::BM_COST_CENSUS, p->window_w, p->window_h, p->num_disparities))
            return say(STEREO_ERROR_INVALID_CONFIG, "census aggregation window must be 5x5, 7x7 or 9x9");
        if (!stereo::census_supported(p->census_w, p->census_h, 0))
            return say(STEREO_ERROR_INVALID_CONFIG, "census window must be 9x7, 7x9, 7x7, 5x5 or 3x3");
        break;
    case STEREO_ALGO_NCC:
        if (!stereo::ncc_supported(p->window_w, p->window_h, p->num_disparities))
            return say(STEREO_ERROR_INVALID_CONFIG, "NCC window must be 5x5, 7x7, 9x9 or 11x11");
        break;
    case STEREO_ALGO_AW1:
    case STEREO_ALGO_AW2:
        if (p->window_w != p->window_h || !stereo::aw_supported(p->window_w, p->num_disparities))
            return say(STEREO_ERROR_INVALID_CONFIG, "adaptive-window sub-patch must be 3x3 or 5x5 (Patch_9x9 / Patch_15x15)");
        break;
    default:
        return say(STEREO_ERROR_INVALID_CONFIG, "Invalid algorithm configuration.");
    }
    if (fmt->width > stereo::BM_MAX_WIDTH) return say(STEREO_ERROR_INVALID_FORMAT, "matcher images are limited to 16384 columns");
    const int bx = p->window_w / 2 + (p->algorithm == STEREO_ALGO_CENSUS ? p->census_w / 2 : 0);
    const int by = p->window_h / 2 + (p->algorithm == STEREO_ALGO_CENSUS ? p->census_h / 2 : 0);
    if (fmt->width <= 2 * bx || fmt->height <= 2 * by) return say(STEREO_ERROR_INVALID_FORMAT, "image smaller than the window");
    return say(STEREO_OK, "");
}

stereo_status stereo_matcher_configure(stereo_context* ctx, stereo_driver** out, const stereo_image_format* fmt,
                                       const stereo_matcher_params* p)
{
    if (!ctx || !out || !fmt || !p) return STEREO_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    char msg[256];
    stereo_status st = stereo_matcher_validate(fmt, p, msg, sizeof(msg));
    if (st) return fail(ctx, st, "%s", msg);
    if ((st = set_device(ctx))) return st;
    stereo_driver* d = new stereo_driver(stereo_driver::MATCHER);
    d->ctx = ctx;
    d->fmt = *fmt;
    d->mp = *p;
    const bool census = p->algorithm == STEREO_ALGO_CENSUS;
    const bool ncc = p->algorithm == STEREO_ALGO_NCC;
    const bool aw = p->algorithm == STEREO_ALGO_AW1 || p->algorithm == STEREO_ALGO_AW2;
    d->name = census ? "Block Matcher Census"
              : ncc  ? "Block Matcher NCC"
              : aw   ? (p->algorithm == STEREO_ALGO_AW1 ? "Block Matcher Adaptive Window 1" : "Block Matcher Adaptive Window 2")
                     : "Block Matcher SAD";
    d->bx = p->window_w / 2 + (census ? p->census_w / 2 : 0);
    d->by = p->window_h / 2 + (census ? p->census_h / 2 : 0);
    if (aw) d->bx = d->by = p->window_w + p->window_w / 2;
    if (ncc)
        stereo::ncc_grid(fmt->width, fmt->height, p->num_disparities, &d->ntx, &d->nty, &d->nseg, &d->seg_len);
    else if (aw)
        stereo::aw_grid(fmt->width, fmt->height, p->num_disparities, &d->ntx, &d->nty, &d->nseg, &d->seg_len);
    else
        stereo::bm_grid(fmt->width, fmt->height, p->num_disparities, &d->ntx, &d->nty, &d->nseg, &d->seg_len);
    d->pad_elem = census ? 8 : 1;
    const size_t cols = (size_t)stereo::BM_PAD_LEFT + fmt->width + stereo::BM_PAD_RIGHT;
    d->pad_pitch = (cols * d->pad_elem + 255) & ~(size_t)255;
    const size_t W = fmt->width, H = fmt->height;
    auto cleanup = [&](stereo_status s2) { driver_release(d); delete d; return s2; };
    if ((st = driver_events(d))) return cleanup(st);
    if (injected_failure()) return cleanup(fail(ctx, STEREO_ERROR_HIP, "stereo_matcher_configure: injected failure"));
    hipError_t e;
    if ((e = hipMalloc((void**)&d->padL, d->pad_pitch * H)) != hipSuccess) return cleanup(hip_fail(ctx, e, "hipMalloc"));
    if ((e = hipMalloc((void**)&d->padR, d->pad_pitch * H)) != hipSuccess) return cleanup(hip_fail(ctx, e, "hipMalloc"));
    if (ncc) {
        // per segment: stereo_matcher_scores reads the left keys after the frame
        if ((e = hipMalloc((void**)&d->lkey, (size_t)d->nseg * W * H * 4)) != hipSuccess)
            return cleanup(hip_fail(ctx, e, "hipMalloc"));
        if ((e = hipMalloc((void**)&d->rpart, (size_t)d->nseg * d->ntx * H * stereo::BM_NPART * 4)) != hipSuccess)
            return cleanup(hip_fail(ctx, e, "hipMalloc"));
    } else {
        // the key planes start at zero; every frame's merge leaves them so
        if ((e = hipMalloc((void**)&d->lkey, W * H * 4)) != hipSuccess) return cleanup(hip_fail(ctx, e, "hipMalloc"));
        if ((e = hipMalloc((void**)&d->rkey, W * H * 4)) != hipSuccess) return cleanup(hip_fail(ctx, e, "hipMalloc"));
        if ((e = hipMemset(d->lkey, 0, W * H * 4)) != hipSuccess) return cleanup(hip_fail(ctx, e, "hipMemset"));
        if ((e = hipMemset(d->rkey, 0, W * H * 4)) != hipSuccess) return cleanup(hip_fail(ctx, e, "hipMemset"));
    }
    if (ncc) {
        if ((e = hipMalloc((void**)&d->lstats, W * H * sizeof(int2))) != hipSuccess) return cleanup(hip_fail(ctx, e, "hipMalloc"));
        if ((e = hipMalloc((void**)&d->rstats, W * H * sizeof(int2))) != hipSuccess) return cleanup(hip_fail(ctx, e, "hipMalloc"));
    }
    // the pad columns only ever meet masked candidates; clear them once so that no
    // uninitialised memory is read
    if ((e = hipMemset(d->padL, 0, d->pad_pitch * H)) != hipSuccess) return cleanup(hip_fail(ctx, e, "hipMemset"));
    if ((e = hipMemset(d->padR, 0, d->pad_pitch * H)) != hipSuccess) return cleanup(hip_fail(ctx, e, "hipMemset"));
    if (p->use_disp_input && (st = band_alloc(d))) return cleanup(st);
    *out = d;
    return STEREO_OK;
}

// The launch sequence of one frame: inputs into the padded buffers (census or copies), the
// NCC statistics, the matcher and the merge / check.  bins: the NCC or adaptive-window band
// search's per-tile masks, or nullptr for the full range.  begin() has been called; this calls end().
static stereo_status matcher_run(stereo_driver* d, stereo_image* disp, bool dispid, stereo_image* disp_lr,
                                 stereo_image* disp_rl, const stereo_image* left, const stereo_image* right,
                                 const uint32_t* bins, int bins_nty, hipStream_t s)
{
    stereo_context* ctx = d->ctx;
    const int W = d->fmt.width, H = d->fmt.height;
    View pl, pr;
    pl.pitch = pr.pitch = d->pad_pitch;
    pl.w = pr.w = W;
    pl.h = pr.h = H;
    pl.ptr = d->padL + (size_t)stereo::BM_PAD_LEFT * d->pad_elem;
    pr.ptr = d->padR + (size_t)stereo::BM_PAD_LEFT * d->pad_elem;
    const bool census = d->mp.algorithm == STEREO_ALGO_CENSUS;
    const bool ncc = d->mp.algorithm == STEREO_ALGO_NCC;
    const bool aw = d->mp.algorithm == STEREO_ALGO_AW1 || d->mp.algorithm == STEREO_ALGO_AW2;
    if (census) {
        const View vl = view_of(left), vr = view_of(right);
        HIP_TRY(ctx, stereo::launch_census_pair(vl, pl, &vr, &pr, d->mp.census_w, d->mp.census_h, s), "census launch");
    } else {
        HIP_TRY(ctx, stereo::launch_pad_copy_pair(view_of(left), pl, view_of(right), pr, s), "copy launch");
    }
    if (ncc) {
        HIP_TRY(ctx, stereo::launch_ncc_stats_pair(view_of(left), d->lstats, 8388608.0f, view_of(right), d->rstats, 1.0f, W,
                                                   d->mp.window_w, d->mp.window_h, s),
                "ncc stats launch");
    }
    if (ctx->timing) HIP_TRY(ctx, hipEventRecord(d->ev_stage[0], s), "hipEventRecord");
    if (ncc) {
        stereo::NccParams np{};
        np.L = pl;
        np.R = pr;
        np.lstats = d->lstats;
        np.rstats = d->rstats;
        np.stats_pitch = W;
        np.lkey = d->lkey;
        np.rpart = d->rpart;
        np.W = W;
        np.H = H;
        np.D = d->mp.num_disparities;
        np.bx = d->bx;
        np.by = d->by;
        np.ntx = d->ntx;
        np.nseg = d->nseg;
        np.seg_len = d->seg_len;
        np.bins = bins;
        np.bins_nty = bins_nty;
        HIP_TRY(ctx, stereo::launch_ncc(d->mp.window_w, np, s), "ncc launch");
    } else if (aw) {
        if (d->planes_dirty) {
            // an earlier frame's merge did not run: clear what its matcher folded in
            HIP_TRY(ctx, stereo::launch_bm_clear(d->lkey, (size_t)W * H, s), "clear launch");
            HIP_TRY(ctx, stereo::launch_bm_clear(d->rkey, (size_t)W * H, s), "clear launch");
            d->planes_dirty = false;
        }
        stereo::AwParams ap;
        ap.L = pl;
        ap.R = pr;
        ap.lkey = d->lkey;
        ap.rkey = d->rkey;
        ap.W = W;
        ap.H = H;
        ap.D = d->mp.num_disparities;
        ap.bx = d->bx;
        ap.by = d->by;
        ap.ntx = d->ntx;
        ap.nseg = d->nseg;
        ap.seg_len = d->seg_len;
        ap.bins = bins;
        ap.bins_nty = bins_nty;
        d->planes_dirty = true;  // until the merge has run
        HIP_TRY(ctx, stereo::launch_aw(d->mp.window_w, ap, s), "adaptive-window launch");
    } else {
        if (d->planes_dirty) {
            // an earlier frame's merge did not run: clear what its matcher folded in
            HIP_TRY(ctx, stereo::launch_bm_clear(d->lkey, (size_t)W * H, s), "clear launch");
            HIP_TRY(ctx, stereo::launch_bm_clear(d->rkey, (size_t)W * H, s), "clear launch");
            d->planes_dirty = false;
        }
        stereo::BmParams bp;
        bp.L = pl;
        bp.R = pr;
        bp.lkey = d->lkey;
        bp.rkey = d->rkey;
        bp.W = W;
        bp.H = H;
        bp.D = d->mp.num_disparities;
        bp.bx = d->bx;
        bp.by = d->by;
        bp.ntx = d->ntx;
        bp.nseg = d->nseg;
        bp.seg_len = d->seg_len;
        // from the matcher launch on, the planes hold keys until the merge has run
        d->planes_dirty = true;
        HIP_TRY(ctx, stereo::launch_bm(census ? stereo::BM_COST_CENSUS : stereo::BM_COST_SAD, d->mp.window_w, bp, s),
                "matcher launch");
    }
    if (ctx->timing) HIP_TRY(ctx, hipEventRecord(d->ev_stage[1], s), "hipEventRecord");
    // a failure here leaves the key planes holding this frame's keys (the next frame
    // clears them first); STEREO_FAIL_AT can select it
    if (injected_failure())
        return fail(ctx, STEREO_ERROR_HIP, "%s: injected failure between the matcher and the merge", d->name);
    stereo::BmFinishParams f;
    f.lkey = d->lkey;
    f.rpart = d->rpart;
    if (d->rkey) {
        f.lplane = d->lkey;
        f.rplane = d->rkey;
    }
    f.disp = (disp && !dispid) ? view_of(disp) : null_view();
    f.dispid = (disp && dispid) ? view_of(disp) : null_view();
    f.disp_l = disp_lr ? view_of(disp_lr) : null_view();
    f.disp_r = disp_rl ? view_of(disp_rl) : null_view();
    f.W = W;
    f.H = H;
    f.D = d->mp.num_disparities;
    f.ntx = d->ntx;
    f.nseg = d->nseg;
    f.seg_len = d->seg_len;
    f.bx = d->bx;
    f.by = d->by;
    f.tol = d->mp.lr_tolerance;
    f.value_base = d->mp.algorithm == STEREO_ALGO_AW1 ? 255 * 3 * d->mp.window_w * d->mp.window_w : 0;
    if (!ncc && !aw) f.cnn = d->finish_cnn;
    HIP_TRY(ctx, stereo::launch_bm_finish(f, s), "finish launch");
    if (d->rkey) d->planes_dirty = false;
    return end(d, s);
}

stereo_status stereo_matcher_process(stereo_driver* d, stereo_image* disp, stereo_image* disp_lr, stereo_image* disp_rl,
                                     const stereo_image* left, const stereo_image* right, stereo_stream stream)
{
    if (!d || d->kind != stereo_driver::MATCHER || !left || !right) return STEREO_ERROR_INVALID_ARGUMENT;
    stereo_context* ctx = d->ctx;
    const int W = d->fmt.width, H = d->fmt.height;
    stereo_status st;
    if ((st = check_image(ctx, left, W, H, STEREO_PIXEL_LUM_U8, "left image"))) return st;
    if ((st = check_image(ctx, right, W, H, STEREO_PIXEL_LUM_U8, "right image"))) return st;
    bool dispid = false;
    if (disp) {
        if (disp->format.pixel_type == STEREO_PIXEL_RGBA_U16) dispid = true;
        if ((st = check_image(ctx, disp, W, H, dispid ? STEREO_PIXEL_RGBA_U16 : STEREO_PIXEL_LUM_I16, "disparity output")))
            return st;
    }
    if (disp_lr && (st = check_image(ctx, disp_lr, W, H, STEREO_PIXEL_LUM_I16, "left-to-right output"))) return st;
    if (disp_rl && (st = check_image(ctx, disp_rl, W, H, STEREO_PIXEL_LUM_I16, "right-to-left output"))) return st;
    if ((st = set_device(ctx))) return st;
    GUARD(ctx, "stereo_matcher_process");
    hipStream_t s = (hipStream_t)stream;
    if ((st = begin(d, s))) return st;

    return matcher_run(d, disp, dispid, disp_lr, disp_rl, left, right, nullptr, 0, s);
}

stereo_status stereo_matcher_process_init(stereo_driver* d, stereo_image* disp, stereo_image* disp_lr,
                                          stereo_image* disp_rl, const stereo_image* left, const stereo_image* right,
                                          const stereo_image* disp_init, stereo_stream stream)
{
    if (!d || d->kind != stereo_driver::MATCHER || !left || !right || !disp_init) return STEREO_ERROR_INVALID_ARGUMENT;
    stereo_context* ctx = d->ctx;
    BandExt* x = dynamic_cast<BandExt*>(d->ext);
    if (!x)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "%s: configured without use_disp_input", d->name);
    const int W = d->fmt.width, H = d->fmt.height;
    stereo_status st;
    if ((st = check_image(ctx, left, W, H, STEREO_PIXEL_LUM_U8, "left image"))) return st;
    if ((st = check_image(ctx, right, W, H, STEREO_PIXEL_LUM_U8, "right image"))) return st;
    bool dispid = false;
    if (disp) {
        if (disp->format.pixel_type == STEREO_PIXEL_RGBA_U16) dispid = true;
        if ((st = check_image(ctx, disp, W, H, dispid ? STEREO_PIXEL_RGBA_U16 : STEREO_PIXEL_LUM_I16, "disparity output")))
            return st;
    }
    if (disp_lr && (st = check_image(ctx, disp_lr, W, H, STEREO_PIXEL_LUM_I16, "left-to-right output"))) return st;
    if (disp_rl && (st = check_image(ctx, disp_rl, W, H, STEREO_PIXEL_LUM_I16, "right-to-left output"))) return st;
    const int iw = disp_init->format.width, ih = disp_init->format.height;
    const int ipt = disp_init->format.pixel_type;
    if (ipt != STEREO_PIXEL_LUM_I16 && ipt != STEREO_PIXEL_RGBA_U16)
        return fail(ctx, STEREO_ERROR_UNSUPPORTED_PIXEL_TYPE, "initial disparity: LUM_I16 or RGBA_U16");
    if (iw < 1 || ih < 1 || iw > W || ih > H)
        return fail(ctx, STEREO_ERROR_FORMAT_MISMATCH, "initial disparity %dx%d: at most the images' %dx%d", iw, ih, W, H);
    if ((st = check_image(ctx, disp_init, iw, ih, ipt, "initial disparity"))) return st;
    if ((st = set_device(ctx))) return st;
    GUARD(ctx, "stereo_matcher_process_init");
    const hipStream_t s = (hipStream_t)stream;
    stereo::InitBandArgs ib;
    ib.init = view_of(disp_init);
    ib.ref_format = ipt == STEREO_PIXEL_RGBA_U16 ? 1 : 0;
    ib.scale = W / iw;
    ib.W = W;
    ib.H = H;
    if (d->mp.algorithm == STEREO_ALGO_NCC || d->mp.algorithm == STEREO_ALGO_AW1 ||
        d->mp.algorithm == STEREO_ALGO_AW2) {
        // NCC, AW1, AW2: the per-tile bins, then the matcher's frame with every candidate
        // outside its pixel's bins masked and groups of disparities no pixel searches skipped
        if ((st = begin(d, s))) return st;
        stereo::CnnBandArgs ba{};
        ba.tile_evals = x->tile_evals;
        ba.tile_mask = x->tile_mask;
        ba.ntx = x->ntx;
        ba.nty = x->nty;
        ba.D = d->mp.num_disparities;
        HIP_TRY(ctx, stereo::launch_init_band_mask(ib, ba, s), "band launch");
        return matcher_run(d, disp, dispid, disp_lr, disp_rl, left, right, x->tile_mask, x->nty, s);
    }

    // outputs live in the frame table; rewrite it (synchronously) when they change
    stereo::BmFrame f = x->frame;
    f.disp = (disp && !dispid) ? view_of(disp) : null_view();
    f.dispid = (disp && dispid) ? view_of(disp) : null_view();
    f.disp_l = disp_lr ? view_of(disp_lr) : null_view();
    f.disp_r = disp_rl ? view_of(disp_rl) : null_view();
    if (std::memcmp(&f, &x->frame, sizeof(f)) != 0) {
        // the previous call's launches read the table and may still run on another stream
        if (x->done_recorded) HIP_TRY(ctx, hipEventSynchronize(x->ev_done), "hipEventSynchronize");
        HIP_TRY(ctx, hipStreamSynchronize(s), "hipStreamSynchronize");
        HIP_TRY(ctx, hipMemcpy(x->frame_dev, &f, sizeof(f), hipMemcpyHostToDevice), "hipMemcpy");
        x->frame = f;
    }
    if ((st = begin(d, s))) return st;
    const bool census = d->mp.algorithm == STEREO_ALGO_CENSUS;
    if (census) {
        const View vl = view_of(left), vr = view_of(right);
        HIP_TRY(ctx, stereo::launch_census_pair(vl, f.L, &vr, &f.R, d->mp.census_w, d->mp.census_h, s), "census launch");
    } else {
        HIP_TRY(ctx, stereo::launch_pad_copy_pair(view_of(left), f.L, view_of(right), f.R, s), "copy launch");
    }
    if (ctx->timing) HIP_TRY(ctx, hipEventRecord(d->ev_stage[0], s), "hipEventRecord");
    stereo::CnnBandArgs ba{};
    ba.tiles = x->tiles;
    ba.items = x->items;
    ba.nitems = x->nitems;
    ba.list_counter = x->counter;
    ba.tile_evals = x->tile_evals;
    ba.tile_mask = x->tile_mask;
    ba.ntx = x->ntx;
    ba.nty = x->nty;
    ba.D = d->mp.num_disparities;
    ba.chunk = kBandChunk;
    HIP_TRY(ctx, stereo::launch_init_band(ib, ba, s), "band launch");
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
    if (x->planes_dirty) {
        // an earlier call's matcher ran without its merge: clear its keys
        HIP_TRY(ctx, stereo::launch_bm_clear(x->lplane, (size_t)W * H, s), "clear launch");
        HIP_TRY(ctx, stereo::launch_bm_clear(x->rplane, (size_t)W * H, s), "clear launch");
    }
    x->planes_dirty = true;  // until the merge is enqueued behind the matcher
    HIP_TRY(ctx, stereo::launch_bm_list(census ? stereo::BM_COST_CENSUS : stereo::BM_COST_SAD, d->mp.window_w, lp, grid, s),
            "list matcher launch");
    if (ctx->timing) HIP_TRY(ctx, hipEventRecord(d->ev_stage[1], s), "hipEventRecord");
    HIP_TRY(ctx, stereo::launch_bm_list_finish(lp, W, s), "list finish launch");
    x->planes_dirty = false;
    HIP_TRY(ctx, hipEventRecord(x->ev_done, s), "hipEventRecord");
    x->done_recorded = true;
    return end(d, s);
}

stereo_status stereo_matcher_scores(stereo_driver* d, stereo_image* score, const stereo_image* left,
                                    const stereo_image* right, stereo_stream stream)
{
    if (!d || d->kind != stereo_driver::MATCHER || !score || !left || !right) return STEREO_ERROR_INVALID_ARGUMENT;
    stereo_context* ctx = d->ctx;
    if (d->mp.algorithm != STEREO_ALGO_NCC)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "%s: winning scores are produced by the NCC matcher", d->name);
    const int W = d->fmt.width, H = d->fmt.height;
    stereo_status st;
    if ((st = check_image(ctx, score, W, H, STEREO_PIXEL_LUM_F32, "score output"))) return st;
    if ((st = check_image(ctx, left, W, H, STEREO_PIXEL_LUM_U8, "left image"))) return st;
    if ((st = check_image(ctx, right, W, H, STEREO_PIXEL_LUM_U8, "right image"))) return st;
    if ((st = set_device(ctx))) return st;
    stereo::NccScoreParams f;
    f.L = view_of(left);
    f.R = view_of(right);
    f.lkey = d->lkey;
    f.score = view_of(score);
    f.lstats = d->lstats;
    f.rstats = d->rstats;
    f.stats_pitch = W;
    f.W = W;
    f.H = H;
    f.ax = d->mp.window_w / 2;
    f.ay = d->mp.window_h / 2;
    f.bx = d->bx;
    f.by = d->by;
    f.nseg = d->nseg;
    HIP_TRY(ctx, stereo::launch_ncc_score(f, (hipStream_t)stream), "ncc score launch");
    return STEREO_OK;
}

stereo_status stereo_matcher_stage_times(stereo_driver* d, double* ms, int32_t n)
{
    if (!d || d->kind != stereo_driver::MATCHER || !ms || n < 1) return STEREO_ERROR_INVALID_ARGUMENT;
    if (!d->timed) return fail(d->ctx, STEREO_ERROR_NOT_READY, "%s: no timed process call (enable stereo_set_timing)", d->name);
    HIP_TRY(d->ctx, hipEventSynchronize(d->ev_stop), "hipEventSynchronize");
    hipEvent_t marks[4] = {d->ev_start, d->ev_stage[0], d->ev_stage[1], d->ev_stop};
    for (int i = 0; i < 3 && i < n; ++i) {
        float t = 0.f;
        HIP_TRY(d->ctx, hipEventElapsedTime(&t, marks[i], marks[i + 1]), "hipEventElapsedTime");
        ms[i] = t;
    }
    return STEREO_OK;
}

// ---- reference-format cross check ----------------------------------------------------

stereo_status stereo_crosscheck_configure(stereo_context* ctx, stereo_driver** out, const stereo_image_format* fmt,
                                          const stereo_crosscheck_params* p)
{
    if (!ctx || !out || !fmt || !p) return STEREO_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!format_valid(fmt)) return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "Invalid image format");
    if (p->disp_delta_thresh < 0) return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "disp_delta_thresh must be >= 0");
    stereo_status st;
    if ((st = set_device(ctx))) return st;
    stereo_driver* d = new stereo_driver(stereo_driver::CROSSCHECK);
    d->ctx = ctx;
    d->name = "Cross Check Disparity Maps";
    d->fmt = *fmt;
    d->cc = *p;
    auto cleanup = [&](stereo_status s2) { driver_release(d); delete d; return s2; };
    if ((st = driver_events(d))) return cleanup(st);
    if (injected_failure()) return cleanup(fail(ctx, STEREO_ERROR_HIP, "stereo_crosscheck_configure: injected failure"));
    hipError_t e;
    if ((e = hipMalloc((void**)&d->range_dev, 2 * sizeof(uint32_t))) != hipSuccess) return cleanup(hip_fail(ctx, e, "hipMalloc"));
    if ((e = hipHostMalloc((void**)&d->range_host, 2 * sizeof(uint32_t), 0)) != hipSuccess)
        return cleanup(hip_fail(ctx, e, "hipHostMalloc"));
    if ((e = hipEventCreateWithFlags(&d->ev_range, hipEventDisableTiming)) != hipSuccess)
        return cleanup(hip_fail(ctx, e, "hipEventCreate"));
    d->range_host[0] = 0xFFF0 - 1;
    d->range_host[1] = 1;
    *out = d;
    return STEREO_OK;
}

stereo_status stereo_crosscheck_process(stereo_driver* d, stereo_image* result, const stereo_image* lr,
                                        const stereo_image* rl, stereo_stream stream)
{
    if (!d || d->kind != stereo_driver::CROSSCHECK || !result || !lr || !rl) return STEREO_ERROR_INVALID_ARGUMENT;
    stereo_context* ctx = d->ctx;
    if (lr->format.width != rl->format.width || lr->format.height != rl->format.height ||
        lr->format.pixel_type != rl->format.pixel_type)
        return fail(ctx, STEREO_ERROR_FORMAT_MISMATCH, "Input disparity images have different formats");
    if (lr->format.width != result->format.width || lr->format.height != result->format.height ||
        lr->format.pixel_type != result->format.pixel_type)
        return fail(ctx, STEREO_ERROR_FORMAT_MISMATCH, "Input and output disparity images have different formats");
    stereo_status st;
    const int W = d->fmt.width, H = d->fmt.height;
    if ((st = check_image(ctx, lr, W, H, STEREO_PIXEL_RGBA_U16, "left-to-right map"))) return st;
    if ((st = check_image(ctx, rl, W, H, STEREO_PIXEL_RGBA_U16, "right-to-left map"))) return st;
    if ((st = check_image(ctx, result, W, H, STEREO_PIXEL_RGBA_U16, "result map"))) return st;
    if ((st = set_device(ctx))) return st;
    GUARD(ctx, "stereo_crosscheck_process");
    hipStream_t s = (hipStream_t)stream;
    if ((st = begin(d, s))) return st;
    HIP_TRY(ctx,
            stereo::launch_crosscheck_ref(view_of(result), view_of(lr), view_of(rl), d->cc.disp_delta_thresh,
                                          d->cc.left_to_right, d->cc.disp_offset, d->range_dev, s),
            "crosscheck launch");
    HIP_TRY(ctx, hipMemcpyAsync(d->range_host, d->range_dev, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
            "hipMemcpyAsync");
    HIP_TRY(ctx, hipEventRecord(d->ev_range, s), "hipEventRecord");
    d->range_pending = true;
    return end(d, s);
}

stereo_status stereo_crosscheck_range(stereo_driver* d, uint16_t* mn, uint16_t* mx)
{
    if (!d || d->kind != stereo_driver::CROSSCHECK || !mn || !mx) return STEREO_ERROR_INVALID_ARGUMENT;
    if (d->range_pending) {
        HIP_TRY(d->ctx, hipEventSynchronize(d->ev_range), "hipEventSynchronize");
        d->range_pending = false;
    }
    *mn = (uint16_t)d->range_host[0];
    *mx = (uint16_t)d->range_host[1];
    return STEREO_OK;
}

}  // extern "C"
