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

// One step of a running sum, s - out + in, where s is a sum of non-negative terms that
// include out (so s >= out), as one v_sad_u32 (|s - out| + in, the pattern below).  On
// packed u16 pairs it holds when each half of s is at least out's half: the 32-bit
// difference then borrows nothing between the halves.  Every operand stays below 2^31.
__device__ __forceinline__ uint32_t slide(uint32_t s, uint32_t out, uint32_t in)
{
    return (max(s, out) - min(s, out)) + in;
}

// Winner-take-all key shared by every integer matcher.  A larger key is better:
//   key = ((65535 - cost) << 8) | (255 - d)
// so the maximum key has the smallest aggregated cost and, among equal costs,
// the smallest disparity.  Key 0 means "no valid candidate".  Requires
// cost <= 65280 and d <= 255.
__host__ __device__ __forceinline__ int key_disp(uint32_t key) { return key ? 255 - (int)(key & 255u) : -1; }
__host__ __device__ __forceinline__ int key_cost(uint32_t key) { return 65535 - (int)(key >> 8); }

// Workgroups of a launch are dealt to the 8 XCDs of MI355X round-robin (workgroup b
// runs on XCD b % 8) and every XCD has its own L2.  xcd_remap renumbers them so that
// XCD k works through the contiguous range [k q, (k + 1) q) of the work items
// (q = n / 8), which keeps neighbouring tiles -- and the image rows they share --
// in one L2; the last n % 8 workgroups keep their number.  A bijection on [0, n).
__device__ __forceinline__ int xcd_remap(int b, int n)
{
    const int q = n >> 3;
    if (b >= 8 * q) return b;
    return (b & 7) * q + (b >> 3);
}

// DPP controls (GFX9 encoding).
constexpr int DPP_ROW_SHL1 = 0x101;
constexpr int DPP_ROW_SHR1 = 0x111;
constexpr int DPP_WAVE_SHR1 = 0x138;

// The empty asm keeps the move a plain v_mov_b32_dpp: without it the DPP combiner
// folds the move into the arithmetic that consumes it, and its v_subrev_u32_dpp
// form (D = src1 - dpp(src0)) produced wrong sums on gfx950 while the unfused
// sequence is exact.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t v)
{
    uint32_t r = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, true);
    asm volatile("" : "+v"(r));
    return r;
}

// max(src shifted one lane up the wave (lane 0 reads 0), key): one v_max_u32_dpp.
__device__ __forceinline__ uint32_t max_shift_in(uint32_t acc, uint32_t key)
{
    const uint32_t s = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)acc, DPP_WAVE_SHR1, 0xf, 0xf, true);
    return s > key ? s : key;
}

// N independent chains acc[i] = max(acc[i] shifted one lane up (lane 0 reads 0),
// key[i]) as one block of v_max_u32_dpp, so that the compiler can neither split the
// DPP move from the max nor schedule a VALU write between them.  The leading s_nop
// covers the two wait states a DPP read needs after a VALU write of its source;
// inside the block each source was written N instructions earlier (N >= 2).
__device__ __forceinline__ void max_shift_in14(uint32_t (&acc)[14], const uint32_t (&key)[14])
{
    asm volatile("s_nop 1\n"
                 "v_max_u32_dpp %0, %0, %14 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %1, %1, %15 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %2, %2, %16 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %3, %3, %17 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %4, %4, %18 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %5, %5, %19 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %6, %6, %20 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %7, %7, %21 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %8, %8, %22 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %9, %9, %23 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %10, %10, %24 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %11, %11, %25 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %12, %12, %26 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %13, %13, %27 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]), "+v"(acc[4]), "+v"(acc[5]), "+v"(acc[6]), "+v"(acc[7]), "+v"(acc[8]), "+v"(acc[9]), "+v"(acc[10]), "+v"(acc[11]), "+v"(acc[12]), "+v"(acc[13])
                 : "v"(key[0]), "v"(key[1]), "v"(key[2]), "v"(key[3]), "v"(key[4]), "v"(key[5]), "v"(key[6]), "v"(key[7]), "v"(key[8]), "v"(key[9]), "v"(key[10]), "v"(key[11]), "v"(key[12]), "v"(key[13]));
}

// The same chains with a signed minimum, for keys that are all negative as int32: the 0
// that lane 0 reads is then the identity.
__device__ __forceinline__ void min_shift_in14(uint32_t (&acc)[14], const uint32_t (&key)[14])
{
    asm volatile("s_nop 1\n"
                 "v_min_i32_dpp %0, %0, %14 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_min_i32_dpp %1, %1, %15 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_min_i32_dpp %2, %2, %16 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_min_i32_dpp %3, %3, %17 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_min_i32_dpp %4, %4, %18 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_min_i32_dpp %5, %5, %19 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_min_i32_dpp %6, %6, %20 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_min_i32_dpp %7, %7, %21 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_min_i32_dpp %8, %8, %22 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_min_i32_dpp %9, %9, %23 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_min_i32_dpp %10, %10, %24 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_min_i32_dpp %11, %11, %25 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_min_i32_dpp %12, %12, %26 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_min_i32_dpp %13, %13, %27 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]), "+v"(acc[4]), "+v"(acc[5]), "+v"(acc[6]), "+v"(acc[7]), "+v"(acc[8]), "+v"(acc[9]), "+v"(acc[10]), "+v"(acc[11]), "+v"(acc[12]), "+v"(acc[13])
                 : "v"(key[0]), "v"(key[1]), "v"(key[2]), "v"(key[3]), "v"(key[4]), "v"(key[5]), "v"(key[6]), "v"(key[7]), "v"(key[8]), "v"(key[9]), "v"(key[10]), "v"(key[11]), "v"(key[12]), "v"(key[13]));
}

__device__ __forceinline__ void max_shift_in13(uint32_t (&acc)[13], const uint32_t (&key)[13])
{
    asm volatile("s_nop 1\n"
                 "v_max_u32_dpp %0, %0, %13 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %1, %1, %14 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %2, %2, %15 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %3, %3, %16 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %4, %4, %17 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %5, %5, %18 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %6, %6, %19 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %7, %7, %20 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %8, %8, %21 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %9, %9, %22 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %10, %10, %23 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %11, %11, %24 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %12, %12, %25 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]), "+v"(acc[4]), "+v"(acc[5]), "+v"(acc[6]), "+v"(acc[7]), "+v"(acc[8]), "+v"(acc[9]), "+v"(acc[10]), "+v"(acc[11]), "+v"(acc[12])
                 : "v"(key[0]), "v"(key[1]), "v"(key[2]), "v"(key[3]), "v"(key[4]), "v"(key[5]), "v"(key[6]), "v"(key[7]), "v"(key[8]), "v"(key[9]), "v"(key[10]), "v"(key[11]), "v"(key[12]));
}

__device__ __forceinline__ void max_shift_in12(uint32_t (&acc)[12], const uint32_t (&key)[12])
{
    asm volatile("s_nop 1\n"
                 "v_max_u32_dpp %0, %0, %12 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %1, %1, %13 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %2, %2, %14 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %3, %3, %15 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %4, %4, %16 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %5, %5, %17 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %6, %6, %18 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %7, %7, %19 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %8, %8, %20 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %9, %9, %21 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %10, %10, %22 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 "v_max_u32_dpp %11, %11, %23 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
                 : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]), "+v"(acc[4]), "+v"(acc[5]), "+v"(acc[6]), "+v"(acc[7]), "+v"(acc[8]), "+v"(acc[9]), "+v"(acc[10]), "+v"(acc[11])
                 : "v"(key[0]), "v"(key[1]), "v"(key[2]), "v"(key[3]), "v"(key[4]), "v"(key[5]), "v"(key[6]), "v"(key[7]), "v"(key[8]), "v"(key[9]), "v"(key[10]), "v"(key[11]));
}

}  // namespace stereo
