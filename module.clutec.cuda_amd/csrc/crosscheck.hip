// Reference-format left-right cross check for gfx950.
//
// Restates CrossCheckDisparityMaps::Kernel::Transform (CrossCheckDisparityMaps.cu:62-109)
// per pixel on 4 x uint16 disparity-id maps: codes outside [First, Last] pass through,
// the right position is MapPixelPos(x, id - First, LtoR), positions outside the image
// and |id_R - id_L| > thresh become Inconsistent.  The reference accumulates the range of
// kept ids with two global atomics per pixel into a 128 x 2 buffer that the host then
// reduces (:86-94, :181-190); here each wave reduces its kept ids with DPP-free wave
// intrinsics and issues one atomicMin / atomicMax, so the range costs two atomics per
// 64 pixels.
#include "common.hpp"
#include "kernels.hpp"

namespace stereo {
namespace {

constexpr uint32_t ID_FIRST = 1, ID_LAST = 0xFFF0 - 1, ID_INCONSISTENT = 0xFFFF - 5;

__global__ void crosscheck_range_init(uint32_t* range)
{
    if (threadIdx.x == 0) { range[0] = ID_LAST; range[1] = ID_FIRST; }
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t t = (uint32_t)__shfl_xor((int)v, o, 64);
        v = t < v ? t : v;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t t = (uint32_t)__shfl_xor((int)v, o, 64);
        v = t > v ? t : v;
    }
    return v;
}

__global__ __launch_bounds__(256) void crosscheck_ref_kernel(View out, View lr, View rl, int thresh, int ltor,
                                                             int offset, uint32_t* range)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    uint32_t kmin = ID_LAST, kmax = ID_FIRST;
    if (x < lr.w && y < lr.h) {
        ushort4 pl = row_ptr<const ushort4>(lr, y)[x];
        const uint32_t id = pl.x;
        if (id >= ID_FIRST && id <= ID_LAST) {
            const int d = (int)id - (int)ID_FIRST;
            const int xr = ltor ? x - d - offset : x + d + offset;
            bool keep = false;
            if (xr >= 0 && xr < rl.w) {
                const int idr = row_ptr<const ushort4>(rl, y)[xr].x;
                keep = abs(idr - (int)id) <= thresh;
            }
            if (keep) { kmin = id; kmax = id; }
            else pl.x = (uint16_t)ID_INCONSISTENT;
        }
        row_ptr<ushort4>(out, y)[x] = pl;
    }
    kmin = wave_min(kmin);
    kmax = wave_max(kmax);
    if ((threadIdx.x & 63) == 0) {
        if (kmin != ID_LAST) atomicMin(&range[0], kmin);
        if (kmax != ID_FIRST) atomicMax(&range[1], kmax);
    }
}

}  // namespace

hipError_t launch_crosscheck_ref(const View& out, const View& lr, const View& rl, int thresh, int ltor, int offset,
                                 uint32_t* range, hipStream_t s)
{
    hipLaunchKernelGGL(crosscheck_range_init, dim3(1), dim3(64), 0, s, range);
    dim3 grid((lr.w + 63) / 64, (lr.h + 3) / 4);
    hipLaunchKernelGGL(crosscheck_ref_kernel, grid, dim3(256), 0, s, out, lr, rl, thresh, ltor, offset, range);
    return hipGetLastError();
}

}  // namespace stereo
