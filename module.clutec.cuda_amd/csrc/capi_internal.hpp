// Internal plumbing shared by the C ABI translation units: context and driver
// objects, error reporting, fault injection and image checks.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "clutec_stereo.h"
#include "kernels.hpp"

using stereo::View;

// ------------------------------------------------------------------------------------
// error plumbing and fault injection
// ------------------------------------------------------------------------------------
namespace capi {

inline std::mutex g_err_mutex;
inline std::string g_create_error;

// STEREO_FAIL_AT=<n>: the n-th guarded operation (device allocation or kernel launch
// sequence) of the process fails with STEREO_ERROR_HIP, so error paths can be tested.
inline long g_fail_at = -2;
inline long g_guard_count = 0;

inline bool injected_failure()
{
    std::lock_guard<std::mutex> lk(g_err_mutex);
    if (g_fail_at == -2) {
        const char* e = std::getenv("STEREO_FAIL_AT");
        g_fail_at = e ? std::atol(e) : -1;
    }
    ++g_guard_count;
    return g_fail_at > 0 && g_guard_count == g_fail_at;
}

}  // namespace capi

struct stereo_context {
    int device = 0;
    int timing = 0;
    std::string last_error;
    // pinned staging buffer of the host image transfers (allocated on the first one)
    std::mutex stage_mu;
    void* stage = nullptr;
    size_t stage_bytes = 0;
};

namespace capi {

inline stereo_status fail(stereo_context* ctx, stereo_status st, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->last_error = buf;
    else {
        std::lock_guard<std::mutex> lk(g_err_mutex);
        g_create_error = buf;
    }
    return st;
}

inline stereo_status hip_fail(stereo_context* ctx, hipError_t e, const char* what)
{
    if (e == hipErrorOutOfMemory) return fail(ctx, STEREO_ERROR_OUT_OF_MEMORY, "%s: %s", what, hipGetErrorString(e));
    return fail(ctx, STEREO_ERROR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr, what)                          \
    do {                                                  \
        hipError_t _e = (expr);                           \
        if (_e != hipSuccess) return hip_fail(ctx, _e, what); \
    } while (0)

#define GUARD(ctx, what)                                                       \
    do {                                                                       \
        if (injected_failure())                                                \
            return fail(ctx, STEREO_ERROR_HIP, "%s: injected failure", what);  \
    } while (0)

inline View view_of(const stereo_image* img)
{
    View v;
    v.ptr = static_cast<uint8_t*>(img->data);
    v.pitch = img->pitch;
    v.w = img->format.width;
    v.h = img->format.height;
    return v;
}

inline View null_view()
{
    View v;
    v.ptr = nullptr;
    v.pitch = 0;
    v.w = 0;
    v.h = 0;
    return v;
}

inline bool format_valid(const stereo_image_format* f)
{
    return f && f->width > 0 && f->height > 0 && f->width <= 32768 && f->height <= 32768 &&
           stereo_pixel_size(f->pixel_type) > 0;
}

inline stereo_status check_image(stereo_context* ctx, const stereo_image* img, int w, int h, int type, const char* what)
{
    if (!img || !img->data) return fail(ctx, STEREO_ERROR_INVALID_ARGUMENT, "%s: null image", what);
    if (img->format.pixel_type != type)
        return fail(ctx, STEREO_ERROR_UNSUPPORTED_PIXEL_TYPE, "%s: pixel types of given images not supported", what);
    if (img->format.width != w || img->format.height != h)
        return fail(ctx, STEREO_ERROR_FORMAT_MISMATCH, "%s: image is %dx%d, driver configured for %dx%d", what,
                    img->format.width, img->format.height, w, h);
    if (img->pitch < (size_t)w * stereo_pixel_size(type))
        return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "%s: pitch %zu too small", what, img->pitch);
    return STEREO_OK;
}

}  // namespace capi

// ------------------------------------------------------------------------------------
// drivers
// ------------------------------------------------------------------------------------
// State of the drivers that live in their own translation unit (batched and learned
// matchers); released with the driver.
struct DriverExt {
    virtual ~DriverExt() = default;
};

struct stereo_driver {
    enum Kind { CENSUS, MATCHER, CROSSCHECK, BATCH, LEARNED, DENOISE, BM1, RECTIFY, COLOR, REFMATCH } kind;
    stereo_context* ctx = nullptr;
    const char* name = "";
    stereo_image_format fmt{};
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    hipEvent_t ev_stage[2] = {nullptr, nullptr};
    bool timed = false;

    // census
    stereo_census_params cen{};
    // matcher
    stereo_matcher_params mp{};
    int bx = 0, by = 0, ntx = 0, nty = 0, nseg = 1, seg_len = 0;
    uint8_t* padL = nullptr;  // padded matcher inputs (grey bytes or census words)
    uint8_t* padR = nullptr;
    size_t pad_pitch = 0;
    size_t pad_elem = 0;
    // keys: per disparity segment for NCC (lkey, rpart; stereo_matcher_scores reads the
    // left keys after the frame); one plane each for SAD, census and the adaptive windows
    // (lkey, rkey), cleared by every frame's merge -- planes_dirty when a frame's merge did
    // not run after its matcher
    uint32_t* lkey = nullptr;
    uint32_t* rpart = nullptr;
    uint32_t* rkey = nullptr;
    bool planes_dirty = false;
    // the learned proposal's coarse matcher: its merge also writes the network input
    stereo::CnnInputTarget finish_cnn{};
    int2* lstats = nullptr;  // NCC window statistics
    int2* rstats = nullptr;
    // crosscheck
    stereo_crosscheck_params cc{};
    uint32_t* range_dev = nullptr;
    uint32_t* range_host = nullptr;
    hipEvent_t ev_range = nullptr;
    bool range_pending = false;
    // batched / learned matchers
    DriverExt* ext = nullptr;

    explicit stereo_driver(Kind k) : kind(k) {}
};

namespace capi {

inline stereo_status driver_events(stereo_driver* d)
{
    HIP_TRY(d->ctx, hipEventCreate(&d->ev_start), "hipEventCreate");
    HIP_TRY(d->ctx, hipEventCreate(&d->ev_stop), "hipEventCreate");
    HIP_TRY(d->ctx, hipEventCreate(&d->ev_stage[0]), "hipEventCreate");
    HIP_TRY(d->ctx, hipEventCreate(&d->ev_stage[1]), "hipEventCreate");
    return STEREO_OK;
}

inline void driver_release(stereo_driver* d)
{
    if (d->ev_start) (void)hipEventDestroy(d->ev_start);
    if (d->ev_stop) (void)hipEventDestroy(d->ev_stop);
    for (hipEvent_t e : d->ev_stage)
        if (e) (void)hipEventDestroy(e);
    if (d->ev_range) (void)hipEventDestroy(d->ev_range);
    if (d->padL) (void)hipFree(d->padL);
    if (d->padR) (void)hipFree(d->padR);
    if (d->lkey) (void)hipFree(d->lkey);
    if (d->rpart) (void)hipFree(d->rpart);
    if (d->rkey) (void)hipFree(d->rkey);
    if (d->lstats) (void)hipFree(d->lstats);
    if (d->rstats) (void)hipFree(d->rstats);
    if (d->range_dev) (void)hipFree(d->range_dev);
    if (d->range_host) (void)hipHostFree(d->range_host);
    delete d->ext;
    d->ext = nullptr;
}

inline stereo_status begin(stereo_driver* d, hipStream_t s)
{
    d->timed = false;
    if (d->ctx->timing) HIP_TRY(d->ctx, hipEventRecord(d->ev_start, s), "hipEventRecord");
    return STEREO_OK;
}

inline stereo_status end(stereo_driver* d, hipStream_t s)
{
    if (d->ctx->timing) {
        HIP_TRY(d->ctx, hipEventRecord(d->ev_stop, s), "hipEventRecord");
        d->timed = true;
    }
    return STEREO_OK;
}

inline stereo_status set_device(stereo_context* ctx)
{
    HIP_TRY(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    return STEREO_OK;
}

}  // namespace capi
