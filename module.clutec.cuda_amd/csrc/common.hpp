// Shared device/host helpers for the gfx950 stereo engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace stereo {

// Non-owning pitched device image view (the role of _CDeviceImage /
// _CDeviceSurface, CluTec.Cuda.Base/DeviceImage.h:208-229): base pointer, row
// pitch in bytes, size in pixels.
struct View {
    uint8_t* ptr;
    size_t pitch;
    int w, h;
};

template <typename T>
__device__ __forceinline__ T* row_ptr(const View& v, int y)
{
    return reinterpret_cast<T*>(v.ptr + (size_t)y * v.pitch);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Winner-take-all key shared by every integer matcher.  A larger key is better:
//   key = ((65535 - cost) << 8) | (255 - d)
// so the maximum key has the smallest aggregated cost and, among equal costs,
// the smallest disparity.  Key 0 means "no valid candidate".  Requires
// cost <= 65280 and d <= 255.
__host__ __device__ __forceinline__ int key_disp(uint32_t key) { return key ? 255 - (int)(key & 255u) : -1; }
__host__ __device__ __forceinline__ int key_cost(uint32_t key) { return 65535 - (int)(key >> 8); }

// DPP controls (GFX9 encoding).
constexpr int DPP_ROW_SHL1 = 0x101;
constexpr int DPP_ROW_SHR1 = 0x111;
constexpr int DPP_WAVE_SHR1 = 0x138;

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, true);
}

// max(src shifted one lane up the wave (lane 0 reads 0), key): one v_max_u32_dpp.
__device__ __forceinline__ uint32_t max_shift_in(uint32_t acc, uint32_t key)
{
    const uint32_t s = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)acc, DPP_WAVE_SHR1, 0xf, 0xf, true);
    return s > key ? s : key;
}

}  // namespace stereo
