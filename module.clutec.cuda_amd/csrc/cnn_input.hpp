// Network input of the learned proposal for one quarter-resolution row, shared by
// cnn_input_kernel and the coarse census matcher's merge (bm_finish_planes_kernel),
// which writes the row's checked coarse map just before and so saves the launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"
#include "kernels.hpp"

namespace stereo {

// float to bf16, round to nearest, ties to even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Network input of row Y of frame n: ch0 = L/255, ch1 = R(x - dfill)/255 (bf16), ch2..15 =
// 0, where dfill is the coarse disparity with holes filled by the smaller of the nearest
// valid values to the left and right within 64 pixels (0 when none).  dfill is written as
// float for the band kernels.  256 threads, one row: each wave takes 64-pixel chunks and
// records which pixels are valid as a ballot in `valid` (LDS, CNN_MAX_W4 / 64 words); the
// nearest valid pixel within 64 then lies in the pixel's own chunk or in the neighbouring
// one, found with a leading / trailing bit count.
__device__ __forceinline__ void cnn_input_row(const View& L4, const View& R4, const View& dc, int n, uint16_t* x,
                                              float* dfill, uint64_t* valid, int Y)
{
    const int W = L4.w, H = L4.h;
    const int nch = (W + 63) >> 6;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int16_t* src = row_ptr<const int16_t>(dc, Y);
    for (int c = wv; c < nch; c += 4) {
        const int X = c * 64 + lane;
        const uint64_t m = __ballot(X < W && src[X] >= 0);
        if (lane == 0) valid[c] = m;
    }
    __syncthreads();
    const uint64_t le = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);  // lanes 0 .. lane
    for (int c = wv; c < nch; c += 4) {
        const int X = c * 64 + lane;
        if (X >= W) continue;
        int d = src[X];
        if (d < 0) {
            const uint64_t ml = valid[c] & le, mr = valid[c] & ~le;
            int l = -1, r = W;
            if (ml) l = c * 64 + 63 - __clzll((long long)ml);
            else if (c > 0 && valid[c - 1]) l = (c - 1) * 64 + 63 - __clzll((long long)valid[c - 1]);
            if (mr) r = c * 64 + __ffsll((unsigned long long)mr) - 1;
            else if (c + 1 < nch && valid[c + 1]) r = (c + 1) * 64 + __ffsll((unsigned long long)valid[c + 1]) - 1;
            const int a = (l >= 0 && X - l <= 64) ? src[l] : -1;
            const int b = (r < W && r - X <= 64) ? src[r] : -1;
            d = a < 0 ? (b < 0 ? 0 : b) : (b < 0 ? a : min(a, b));
        }
        const size_t p = ((size_t)n * H + Y) * W + X;
        dfill[p] = (float)d;
        const int xr = clampi(X - d, 0, W - 1);
        const float l = row_ptr<const uint8_t>(L4, Y)[X] * (1.0f / 255.0f);
        const float r = row_ptr<const uint8_t>(R4, Y)[xr] * (1.0f / 255.0f);
        uint4* o = reinterpret_cast<uint4*>(x + p * 16);
        o[0] = make_uint4((uint32_t)f2bf(l) | ((uint32_t)f2bf(r) << 16), 0, 0, 0);
        o[1] = make_uint4(0, 0, 0, 0);
    }
}

}  // namespace stereo
