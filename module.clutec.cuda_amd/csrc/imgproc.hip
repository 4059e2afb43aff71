// The steps on either side of the matchers: rectification of the input images
// (RectifyImage, CluTec.Cuda.ImgProc/RectifyImage.cu) and the visual / on-disk forms of the
// disparity ids (ColorDisparityImage, CluTec.Cuda.MultiView/ColorDisparityImage.cu).
//
// Rectification: target pixel (u, v) of the rectified camera (fx', fy', cx', cy') becomes
// the ray ((u - cx') / fx', (v - cy') / fy', 1), is rotated into the source camera (R),
// projected to (xn, yn), distorted with the OpenCV model (k1, k2, p1, p2, k3) and mapped
// through the source camera (fx, fy, cx, cy) (MapPixelTrgToSrc and DistortPX, :111-158).
// The pixel is sampled as CopyPixel does (:64-107): bilinear over the 2x2 block at the
// floor of the position when that block lies in the source (on values / 255), else the
// nearest pixel when it lies in the source, else 0.
//
// Colouring: per pixel, an id in [First, Last] maps to a grey level, a colour ramp, the id
// itself or the reversed id (DispRel); the other codes to a palette entry (Analysis) or 0.
//
// Every float step is one IEEE single operation rounded to nearest, contraction off,
// division correctly rounded, so the results equal a plain C evaluation of the formulas.
// Both kernels are one thread per output pixel, 64 x 4 pixels per workgroup.
#include <hip/hip_runtime.h>

#include <cmath>

#include "capi_internal.hpp"

namespace stereo {
namespace {

// NormFloatTo<T>: clamp to [0, 1] (NaN to 0), scale by the type's maximum, round half up
__device__ __forceinline__ uint32_t norm_to(float f, float maxv)
{
#pragma clang fp contract(off)
    f = fminf(fmaxf(f, 0.0f), 1.0f);
    return (uint32_t)(f * maxv + 0.5f);
}

struct RectArgs {
    View src, dst;
    float p[22];  // src fx, fy, cx, cy, k1, k2, p1, p2, k3, rect fx, fy, cx, cy, R[9]
};

__global__ __launch_bounds__(256) void rectify_kernel(RectArgs a)
{
#pragma clang fp contract(off)
    // the 256 values v / 255 of the bilinear tap, each the correctly rounded quotient
    __shared__ float norm[256];
    norm[threadIdx.x] = (float)threadIdx.x / 255.0f;
    __syncthreads();
    const int u = blockIdx.x * 64 + (threadIdx.x & 63);
    const int v = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (u >= a.dst.w || v >= a.dst.h) return;
    const float* p = a.p;
    const float x = ((float)u - p[11]) / p[9], y = ((float)v - p[12]) / p[10];
    const float* R = p + 13;
    const float X = R[0] * x + R[1] * y + R[2];
    const float Y = R[3] * x + R[4] * y + R[5];
    const float Z = R[6] * x + R[7] * y + R[8];
    const float xn = X / Z, yn = Y / Z;
    const float r2 = xn * xn + yn * yn;
    const float radial = 1.0f + r2 * (p[4] + r2 * (p[5] + r2 * p[8]));
    const float xy = xn * yn;
    const float xd = xn * radial + 2.0f * p[6] * xy + p[7] * (r2 + 2.0f * xn * xn);
    const float yd = yn * radial + p[6] * (r2 + 2.0f * yn * yn) + 2.0f * p[7] * xy;
    const float sx = p[0] * xd + p[2];
    const float sy = p[1] * yd + p[3];
    const float fx0 = floorf(sx), fy0 = floorf(sy);
    const float nx = floorf(sx + 0.5f), ny = floorf(sy + 0.5f);
    // positions beyond +-2^30 (or NaN) are outside
    const bool ok = fabsf(fx0) < 1073741824.0f && fabsf(fy0) < 1073741824.0f;
    const int ix = ok ? (int)fx0 : -2, iy = ok ? (int)fy0 : -2;
    const bool okn = fabsf(nx) < 1073741824.0f && fabsf(ny) < 1073741824.0f;
    const int jx = okn ? (int)nx : -1, jy = okn ? (int)ny : -1;
    const int sW = a.src.w, sH = a.src.h;
    uint32_t o = 0;
    if (ix >= 0 && iy >= 0 && ix + 2 <= sW && iy + 2 <= sH) {
        const float ax = sx - fx0, ay = sy - fy0;
        const uint8_t* s0 = row_ptr<const uint8_t>(a.src, iy) + ix;
        const uint8_t* s1 = row_ptr<const uint8_t>(a.src, iy + 1) + ix;
        float f = (1.0f - ax) * (1.0f - ay) * norm[s0[0]];
        f += ax * (1.0f - ay) * norm[s0[1]];
        f += (1.0f - ax) * ay * norm[s1[0]];
        f += ax * ay * norm[s1[1]];
        o = norm_to(f, 255.0f);
    } else if (jx >= 0 && jy >= 0 && jx < sW && jy < sH) {
        o = row_ptr<const uint8_t>(a.src, jy)[jx];
    }
    row_ptr<uint8_t>(a.dst, v)[u] = (uint8_t)o;
}

struct ColorArgs {
    View disp, dst;
    int dmin, dmax, inf_offset;
    uchar4 colors[7];  // Unknown, CannotEvaluate, Saturated, NotSpecific, NotFound, NotUnique, Inconsistent
};

// FMT 0: 4 x uint16 ids, 1: uint16 ids, 2: float ids.  STYLE 0 Analysis, 1 Gray, 2 Color,
// 3 DispAbs, 4 DispRel (ColorDisparityImage.h:50-57).
template <int FMT, int STYLE>
__global__ __launch_bounds__(256) void color_kernel(ColorArgs a)
{
#pragma clang fp contract(off)
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= a.disp.w || y >= a.disp.h) return;
    const float range = (float)(a.dmax - a.dmin);
    float v;
    int id = 0;
    bool valid;
    float fd;
    if (FMT == 2) {
        v = row_ptr<const float>(a.disp, y)[x];
        valid = v >= 1.0f && v <= (float)0xFFEF;
        fd = (v - (float)a.dmin) / range;
    } else {
        id = FMT == 0 ? (int)row_ptr<const ushort4>(a.disp, y)[x].x : (int)row_ptr<const uint16_t>(a.disp, y)[x];
        v = (float)id;
        valid = id >= 1 && id <= 0xFFEF;
        fd = (float)(id - a.dmin) / range;
    }
    if (STYLE == 0) {
        uchar4 o;
        if (valid) {
            const uint8_t c = (uint8_t)norm_to(fd, 255.0f);
            o = make_uchar4(c, c, c, 255);
        } else {
            const float codes[7] = {0.0f, 65535.0f, 65534.0f, 65532.0f, 65533.0f, 65531.0f, 65530.0f};
            o = make_uchar4(255, 0, 0, 255);
#pragma unroll
            for (int k = 6; k >= 0; --k)
                if (v == codes[k]) o = a.colors[k];
        }
        row_ptr<uchar4>(a.dst, y)[x] = o;
    } else if (STYLE == 2) {
        uchar4 o = make_uchar4(0, 0, 0, 255);
        if (valid) {
            // NormFloatToColor (CluTec.Cuda.Base/Conversion.h:56-71)
            const float fB = 1.0f - fminf(fd / 0.7f, 1.0f);
            const float fG = 1.0f - fminf(fabsf(fd - 0.5f) / 0.59f, 1.0f);
            const float fR = 1.0f - fminf(fabsf(fd - 0.8f) / 0.7f, 1.0f);
            const float fM = fmaxf(fmaxf(fR, fG), fB);
            o.x = (uint8_t)norm_to(fR / fM, 255.0f);
            o.y = (uint8_t)norm_to(fG / fM, 255.0f);
            o.z = (uint8_t)norm_to(fB / fM, 255.0f);
        }
        row_ptr<uchar4>(a.dst, y)[x] = o;
    } else {
        uint32_t g = 0;
        if (valid) {
            if (STYLE == 1) {
                g = norm_to(fd, 65535.0f);
            } else if (STYLE == 3) {
                g = (uint32_t)v;
            } else if (FMT == 2) {
                // float arithmetic, converted to 16 bits with truncation, saturating
                const float r = ((float)0x7FFF + (float)a.inf_offset) - (v - 1.0f);
                g = !(r > 0.0f) ? 0u : (r >= 65535.0f ? 65535u : (uint32_t)r);
            } else {
                g = (uint32_t)((0x7FFF + a.inf_offset) - (id - 1)) & 0xFFFFu;
            }
        }
        row_ptr<uint16_t>(a.dst, y)[x] = (uint16_t)g;
    }
}

struct RectExt : DriverExt {
    stereo_rectify_params p{};
};

struct ColorExt : DriverExt {
    stereo_color_params p{};
};

template <int FMT>
void color_launch(const ColorArgs& a, int style, dim3 grid, hipStream_t s)
{
    switch (style) {
    case 0: hipLaunchKernelGGL((color_kernel<FMT, 0>), grid, dim3(256), 0, s, a); break;
    case 1: hipLaunchKernelGGL((color_kernel<FMT, 1>), grid, dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((color_kernel<FMT, 2>), grid, dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL((color_kernel<FMT, 3>), grid, dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL((color_kernel<FMT, 4>), grid, dim3(256), 0, s, a); break;
    }
}

}  // namespace
}  // namespace stereo

using namespace capi;

extern "C" {

stereo_status stereo_rectify_configure(stereo_context* ctx, stereo_driver** out, const stereo_image_format* fmt,
                                       const stereo_rectify_params* p)
{
    if (!ctx || !out || !fmt || !p) return STEREO_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!format_valid(fmt)) return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "Invalid image format");
    if (fmt->pixel_type != STEREO_PIXEL_LUM_U8)
        return fail(ctx, STEREO_ERROR_UNSUPPORTED_PIXEL_TYPE, "Pixel types of given images not supported");
    for (int i = 0; i < 4; ++i)
        if (p->reserved[i]) return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "reserved parameters must be 0");
    const float* f = &p->src_fx;
    for (int i = 0; i < 22; ++i)
        if (!std::isfinite(f[i])) return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "rectify: parameter %d is not finite", i);
    if (p->src_fx == 0.0f || p->src_fy == 0.0f || p->rect_fx == 0.0f || p->rect_fy == 0.0f)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "rectify: focal lengths must be non-zero");
    stereo_status st;
    if ((st = set_device(ctx))) return st;
    stereo_driver* d = new stereo_driver(stereo_driver::RECTIFY);
    d->ctx = ctx;
    d->fmt = *fmt;
    d->name = "Rectify Image";
    stereo::RectExt* x = new stereo::RectExt();
    x->p = *p;
    d->ext = x;
    if ((st = driver_events(d))) {
        driver_release(d);
        delete d;
        return st;
    }
    *out = d;
    return STEREO_OK;
}

stereo_status stereo_rectify_process(stereo_driver* d, stereo_image* target, const stereo_image* source,
                                     stereo_stream stream)
{
    if (!d || d->kind != stereo_driver::RECTIFY || !target || !source) return STEREO_ERROR_INVALID_ARGUMENT;
    stereo_context* ctx = d->ctx;
    const stereo::RectExt* x = static_cast<const stereo::RectExt*>(d->ext);
    stereo_status st;
    if ((st = check_image(ctx, target, d->fmt.width, d->fmt.height, STEREO_PIXEL_LUM_U8, "rectified image"))) return st;
    if (!format_valid(&source->format))
        return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "rectify: invalid source image format");
    if ((st = check_image(ctx, source, source->format.width, source->format.height, STEREO_PIXEL_LUM_U8,
                          "source image")))
        return st;
    if ((st = set_device(ctx))) return st;
    GUARD(ctx, "stereo_rectify_process");
    const hipStream_t s = (hipStream_t)stream;
    if ((st = begin(d, s))) return st;
    stereo::RectArgs a;
    a.src = view_of(source);
    a.dst = view_of(target);
    std::memcpy(a.p, &x->p.src_fx, sizeof(a.p));
    const dim3 grid((d->fmt.width + 63) / 64, (d->fmt.height + 3) / 4);
    hipLaunchKernelGGL(stereo::rectify_kernel, grid, dim3(256), 0, s, a);
    HIP_TRY(ctx, hipGetLastError(), "rectify launch");
    return end(d, s);
}

stereo_status stereo_color_configure(stereo_context* ctx, stereo_driver** out, const stereo_image_format* fmt,
                                     const stereo_color_params* p)
{
    if (!ctx || !out || !fmt || !p) return STEREO_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!format_valid(fmt)) return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "Invalid image format");
    if (fmt->pixel_type != STEREO_PIXEL_RGBA_U16 && fmt->pixel_type != STEREO_PIXEL_LUM_U16 &&
        fmt->pixel_type != STEREO_PIXEL_LUM_F32)
        return fail(ctx, STEREO_ERROR_UNSUPPORTED_PIXEL_TYPE, "Disparity image pixel type not supported");
    for (int i = 0; i < 4; ++i)
        if (p->reserved[i]) return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "reserved parameters must be 0");
    if (p->disp_max <= p->disp_min)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "colour: disparity range [%u, %u] is empty", p->disp_min,
                    p->disp_max);
    stereo_status st;
    if ((st = set_device(ctx))) return st;
    stereo_driver* d = new stereo_driver(stereo_driver::COLOR);
    d->ctx = ctx;
    d->fmt = *fmt;
    d->name = "Color Disparity Image";
    stereo::ColorExt* x = new stereo::ColorExt();
    x->p = *p;
    d->ext = x;
    if ((st = driver_events(d))) {
        driver_release(d);
        delete d;
        return st;
    }
    *out = d;
    return STEREO_OK;
}

stereo_status stereo_color_process(stereo_driver* d, stereo_image* out, const stereo_image* disp, int32_t style,
                                   stereo_stream stream)
{
    if (!d || d->kind != stereo_driver::COLOR || !out || !disp) return STEREO_ERROR_INVALID_ARGUMENT;
    stereo_context* ctx = d->ctx;
    if (style < 0 || style > 4) return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "Given coloring style not supported");
    const stereo::ColorExt* x = static_cast<const stereo::ColorExt*>(d->ext);
    const int W = d->fmt.width, H = d->fmt.height;
    stereo_status st;
    if ((st = check_image(ctx, disp, W, H, d->fmt.pixel_type, "disparity image"))) return st;
    const int out_type = (style == 0 || style == 2) ? STEREO_PIXEL_RGBA_U8 : STEREO_PIXEL_LUM_U16;
    if ((st = check_image(ctx, out, W, H, out_type, "colour image"))) return st;
    if ((st = set_device(ctx))) return st;
    GUARD(ctx, "stereo_color_process");
    const hipStream_t s = (hipStream_t)stream;
    if ((st = begin(d, s))) return st;
    stereo::ColorArgs a;
    a.disp = view_of(disp);
    a.dst = view_of(out);
    a.dmin = x->p.disp_min;
    a.dmax = x->p.disp_max;
    a.inf_offset = x->p.disp_inf_offset;
    for (int k = 0; k < 7; ++k)
        a.colors[k] = make_uchar4(x->p.colors[k][0], x->p.colors[k][1], x->p.colors[k][2], x->p.colors[k][3]);
    const dim3 grid((W + 63) / 64, (H + 3) / 4);
    switch (d->fmt.pixel_type) {
    case STEREO_PIXEL_RGBA_U16: stereo::color_launch<0>(a, style, grid, s); break;
    case STEREO_PIXEL_LUM_U16: stereo::color_launch<1>(a, style, grid, s); break;
    default: stereo::color_launch<2>(a, style, grid, s); break;
    }
    HIP_TRY(ctx, hipGetLastError(), "colour launch");
    return end(d, s);
}

}  // extern "C"
