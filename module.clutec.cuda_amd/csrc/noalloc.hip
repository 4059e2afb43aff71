// The fused SAD / census matcher for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// stereo_matcher_process runs on a configured driver whose workspace -- the padded
// inputs and the two key planes (the best left and right keys of every pixel) -- was
// allocated by stereo_matcher_configure.  This entry point takes the same parameters
// per call and places that workspace in `scratch`, a file-scope device array of
// STEREO_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the process
// ends.  stereo_create looks up the array's device address once per device
// (hipGetSymbolAddress with that device current); host code uses that address, never the
// array's name.  The key planes must start at zero, and the array keeps whatever the last
// call of any size left, so every call clears them with a kernel first; the pad columns
// of the inputs only meet masked candidates.  All calls on one GPU share the array, so
// calls that overlap on one GPU race; calls on one stream are ordered.
#include <hip/hip_runtime.h>

#include <mutex>

#include "capi_internal.hpp"

namespace {

static __device__ __align__(256) uint8_t scratch[STEREO_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

struct Layout {
    int ntx, nty, nseg, seg_len;
    size_t pad_elem, pad_pitch, pad_bytes, plane_bytes;
    size_t total() const { return 2 * pad_bytes + 2 * plane_bytes; }
};

// The driver's segmentation and workspace for a frame: two padded inputs, then the left
// and the right key plane.
Layout layout_for(int W, int H, int D, bool census)
{
    Layout l{};
    stereo::bm_grid(W, H, D, &l.ntx, &l.nty, &l.nseg, &l.seg_len);
    l.pad_elem = census ? 8 : 1;
    const size_t cols = (size_t)stereo::BM_PAD_LEFT + W + stereo::BM_PAD_RIGHT;
    l.pad_pitch = (cols * l.pad_elem + 255) & ~(size_t)255;
    l.pad_bytes = l.pad_pitch * H;
    l.plane_bytes = ((size_t)W * H * 4 + 255) & ~(size_t)255;
    return l;
}

}  // namespace

namespace stereo {

// Set-up of the no-allocation entry point, called by stereo_create with `device` current:
// the device address of `scratch` on that device, looked up once per device.
hipError_t noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = p;
    return hipSuccess;
}

}  // namespace stereo

using namespace capi;

extern "C" {

stereo_status stereo_matcher_process_noalloc(stereo_context* ctx, const stereo_matcher_params* p, stereo_image* disp,
                                             stereo_image* disp_lr, stereo_image* disp_rl, const stereo_image* left,
                                             const stereo_image* right, stereo_stream stream)
{
    if (!ctx || !p || !left || !right) return STEREO_ERROR_INVALID_ARGUMENT;
    if (p->algorithm != STEREO_ALGO_SAD && p->algorithm != STEREO_ALGO_CENSUS)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "stereo_matcher_process_noalloc: SAD or census matchers only");
    if (p->use_disp_input)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "stereo_matcher_process_noalloc: no band search (use_disp_input 0)");
    const stereo_image_format fmt = left->format;
    char msg[256];
    stereo_status st = stereo_matcher_validate(&fmt, p, msg, sizeof(msg));
    if (st) return fail(ctx, st, "%s", msg);
    const int W = fmt.width, H = fmt.height, D = p->num_disparities;
    if ((st = check_image(ctx, left, W, H, STEREO_PIXEL_LUM_U8, "left image"))) return st;
    if ((st = check_image(ctx, right, W, H, STEREO_PIXEL_LUM_U8, "right image"))) return st;
    bool dispid = false;
    if (disp) {
        dispid = disp->format.pixel_type == STEREO_PIXEL_RGBA_U16;
        if ((st = check_image(ctx, disp, W, H, dispid ? STEREO_PIXEL_RGBA_U16 : STEREO_PIXEL_LUM_I16, "disparity output")))
            return st;
    }
    if (disp_lr && (st = check_image(ctx, disp_lr, W, H, STEREO_PIXEL_LUM_I16, "left-to-right output"))) return st;
    if (disp_rl && (st = check_image(ctx, disp_rl, W, H, STEREO_PIXEL_LUM_I16, "right-to-left output"))) return st;

    const bool census = p->algorithm == STEREO_ALGO_CENSUS;
    const Layout lay = layout_for(W, H, D, census);
    if (lay.total() > STEREO_NOALLOC_SCRATCH_BYTES)
        return fail(ctx, STEREO_ERROR_INVALID_FORMAT,
                    "stereo_matcher_process_noalloc: a %dx%d %s frame needs %zu bytes of workspace, more than the %u "
                    "of the library's scratch array",
                    W, H, census ? "census" : "SAD", lay.total(), (unsigned)STEREO_NOALLOC_SCRATCH_BYTES);
    uint8_t* base;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = ctx->device >= 0 && ctx->device < kMaxDevices ? static_cast<uint8_t*>(g_scratch[ctx->device]) : nullptr;
    }
    if (!base)
        return fail(ctx, STEREO_ERROR_HIP, "stereo_matcher_process_noalloc: the scratch array of device %d is not set up",
                    ctx->device);
    if ((st = set_device(ctx))) return st;
    GUARD(ctx, "stereo_matcher_process_noalloc");
    const hipStream_t s = (hipStream_t)stream;

    uint8_t* padL = base;
    uint8_t* padR = padL + lay.pad_bytes;
    uint32_t* lkey = reinterpret_cast<uint32_t*>(padR + lay.pad_bytes);
    uint32_t* rkey = reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(lkey) + lay.plane_bytes);
    // both planes at once: they are contiguous
    HIP_TRY(ctx, stereo::launch_bm_clear(lkey, 2 * lay.plane_bytes / 4, s), "clear launch");
    View pl, pr;
    pl.pitch = pr.pitch = lay.pad_pitch;
    pl.w = pr.w = W;
    pl.h = pr.h = H;
    pl.ptr = padL + (size_t)stereo::BM_PAD_LEFT * lay.pad_elem;
    pr.ptr = padR + (size_t)stereo::BM_PAD_LEFT * lay.pad_elem;
    if (census) {
        const View vl = view_of(left), vr = view_of(right);
        HIP_TRY(ctx, stereo::launch_census_pair(vl, pl, &vr, &pr, p->census_w, p->census_h, s), "census launch");
    } else {
        HIP_TRY(ctx, stereo::launch_pad_copy_pair(view_of(left), pl, view_of(right), pr, s), "copy launch");
    }
    const int bx = p->window_w / 2 + (census ? p->census_w / 2 : 0);
    const int by = p->window_h / 2 + (census ? p->census_h / 2 : 0);
    stereo::BmParams bp;
    bp.L = pl;
    bp.R = pr;
    bp.lkey = lkey;
    bp.rkey = rkey;
    bp.W = W;
    bp.H = H;
    bp.D = D;
    bp.bx = bx;
    bp.by = by;
    bp.ntx = lay.ntx;
    bp.nseg = lay.nseg;
    bp.seg_len = lay.seg_len;
    HIP_TRY(ctx, stereo::launch_bm(census ? stereo::BM_COST_CENSUS : stereo::BM_COST_SAD, p->window_w, bp, s),
            "matcher launch");
    stereo::BmFinishParams f;
    f.lkey = nullptr;
    f.rpart = nullptr;
    f.lplane = lkey;
    f.rplane = rkey;
    f.disp = (disp && !dispid) ? view_of(disp) : null_view();
    f.dispid = (disp && dispid) ? view_of(disp) : null_view();
    f.disp_l = disp_lr ? view_of(disp_lr) : null_view();
    f.disp_r = disp_rl ? view_of(disp_rl) : null_view();
    f.W = W;
    f.H = H;
    f.D = D;
    f.ntx = lay.ntx;
    f.nseg = lay.nseg;
    f.seg_len = lay.seg_len;
    f.bx = bx;
    f.by = by;
    f.tol = p->lr_tolerance;
    f.value_base = 0;
    HIP_TRY(ctx, stereo::launch_bm_finish(f, s), "finish launch");
    return STEREO_OK;
}

}  // extern "C"
