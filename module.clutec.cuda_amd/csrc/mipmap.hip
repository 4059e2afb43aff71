// Disparity mip maps (the role of CreateDispMipMaps, CluTec.Cuda.MultiView/
// CreateDispMipMaps.cu).  Level 0 is the disparity map as float: the reference format's
// channel-0 id when First <= id <= Last, else Unknown (0) (CopyDispChannel :66-94), or
// the plain map's d >= 0, else -1.  Level i+1, ((w-1)/2) x ((h-1)/2) of level i
// (DeviceSurfMipMap.h:78-86), is a [1 2 1] filter along x and then along y, each tap
// skipping the unknown marker and dividing by the weights it kept, the vertical pass
// halving the value (SmoothHorizontalLum :97-150, SmoothVerticalLum :153-218).  One
// thread computes one target pixel from its three horizontal taps, so a level is one
// launch without an intermediate buffer; every sum is formed in the reference's order
// (first value, twice the second, the third) with a correctly rounded division.
#include <hip/hip_runtime.h>

#include "capi_internal.hpp"

namespace stereo {
namespace {

__global__ __launch_bounds__(256) void mip_level0_kernel(View src, int ref_format, View dst)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= dst.w || y >= dst.h) return;
    float v;
    if (ref_format) {
        const uint32_t id = row_ptr<const ushort4>(src, y)[x].x;
        v = (id >= 1u && id <= 0xFFEFu) ? (float)id : 0.0f;
    } else {
        const int d = row_ptr<const int16_t>(src, y)[x];
        v = d >= 0 ? (float)d : -1.0f;
    }
    row_ptr<float>(dst, y)[x] = v;
}

__device__ __forceinline__ float mip_tap(float v0, float v1, float v2, float unknown)
{
    float cnt = 0.0f, sum = 0.0f;
    if (v0 != unknown) { cnt += 1.0f; sum += v0; }
    if (v1 != unknown) { cnt += 2.0f; sum += v1 * 2.0f; }
    if (v2 != unknown) { cnt += 1.0f; sum += v2; }
    return cnt > 0.0f ? __fdiv_rn(sum, cnt) : unknown;
}

__global__ __launch_bounds__(256) void mip_next_kernel(View src, float unknown, View dst)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= dst.w || y >= dst.h) return;
    float h[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const float* r = row_ptr<const float>(src, 2 * y + j) + 2 * x;
        h[j] = mip_tap(r[0], r[1], r[2], unknown);
    }
    const float v = mip_tap(h[0], h[1], h[2], unknown);
    row_ptr<float>(dst, y)[x] = v != unknown ? 0.5f * v : unknown;
}

}  // namespace
}  // namespace stereo

using namespace capi;

extern "C" stereo_status stereo_disp_mipmaps(stereo_context* ctx, const stereo_image* disp, stereo_image* levels,
                                             int32_t nlevels, stereo_stream stream)
{
    if (!ctx || !disp || !levels || nlevels < 1) return STEREO_ERROR_INVALID_ARGUMENT;
    const int32_t pt = disp->format.pixel_type;
    if (pt != STEREO_PIXEL_LUM_I16 && pt != STEREO_PIXEL_RGBA_U16)
        return fail(ctx, STEREO_ERROR_UNSUPPORTED_PIXEL_TYPE, "mip maps: disparity map must be LUM_I16 or RGBA_U16");
    if (!format_valid(&disp->format) || !disp->data)
        return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "mip maps: invalid disparity image");
    // "Mip-map level count larger than the maximal possible level" (CreateDispMipMaps.cu:262-265)
    int w = disp->format.width, h = disp->format.height;
    stereo_status st;
    for (int i = 0; i < nlevels; ++i) {
        if (i > 0) {
            if (w < 3 || h < 3)
                return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "mip maps: level %d is below the smallest level", i);
            w = (w - 1) / 2;
            h = (h - 1) / 2;
        }
        char what[32];
        std::snprintf(what, sizeof(what), "mip level %d", i);
        if ((st = check_image(ctx, &levels[i], w, h, STEREO_PIXEL_LUM_F32, what))) return st;
    }
    if ((st = set_device(ctx))) return st;
    GUARD(ctx, "stereo_disp_mipmaps");
    const hipStream_t s = (hipStream_t)stream;
    const float unknown = pt == STEREO_PIXEL_RGBA_U16 ? 0.0f : -1.0f;
    {
        const View dst = view_of(&levels[0]);
        dim3 grid((dst.w + 63) / 64, (dst.h + 3) / 4);
        hipLaunchKernelGGL(stereo::mip_level0_kernel, grid, dim3(256), 0, s, view_of(disp),
                           pt == STEREO_PIXEL_RGBA_U16 ? 1 : 0, dst);
        HIP_TRY(ctx, hipGetLastError(), "mip level 0 launch");
    }
    for (int i = 1; i < nlevels; ++i) {
        const View dst = view_of(&levels[i]);
        dim3 grid((dst.w + 63) / 64, (dst.h + 3) / 4);
        hipLaunchKernelGGL(stereo::mip_next_kernel, grid, dim3(256), 0, s, view_of(&levels[i - 1]), unknown, dst);
        HIP_TRY(ctx, hipGetLastError(), "mip level launch");
    }
    return STEREO_OK;
}
