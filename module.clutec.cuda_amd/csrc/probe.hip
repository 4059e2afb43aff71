// Device calibration for the roofline: the INT32 VALU issue rate measured on the
// running device.  Every lane runs CHAINS independent chains of one instruction
// (v_add_u32, or v_bcnt_u32_b32, which is a population count plus an accumulate),
// written as inline assembly so that the compiler can neither fold nor reorder them;
// the grid fills every SIMD several times over.  lane-ops/s = lanes x iterations x
// CHAINS x UNROLL / kernel time.  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include "capi_internal.hpp"

namespace stereo {
namespace {

constexpr int CHAINS = 8;
constexpr int UNROLL = 16;
constexpr int PROBE_BLOCKS = 8192;

template <int OP>
__global__ __launch_bounds__(256) void valu_peak_kernel(uint32_t* out, int iters)
{
    uint32_t a[CHAINS];
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) a[c] = threadIdx.x * 7u + c;
    const uint32_t b = blockIdx.x | 1u;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
#pragma unroll
            for (int c = 0; c < CHAINS; ++c) {
                if (OP == 0)
                    asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[c]) : "v"(b));
                else
                    asm volatile("v_bcnt_u32_b32 %0, %1, %0" : "+v"(a[c]) : "v"(b));
            }
        }
    }
    uint32_t s = 0;
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) s ^= a[c];
    out[blockIdx.x * 256 + threadIdx.x] = s;  // keeps the chains live
}

typedef __bf16 probe_bf2 __attribute__((ext_vector_type(2)));
typedef float probe_f2 __attribute__((ext_vector_type(2)));

// Counts the float bit patterns, NaNs excepted, whose bf16 from v_cvt_pk_bf16_f32 (the
// one-launch network's epilogue) differs from the integer round-to-nearest-even of the
// network's other kernels and of its CPU reference; all 2^32 patterns, grid-stride.
__global__ __launch_bounds__(256) void bf16_rounding_kernel(unsigned long long* bad)
{
    unsigned long long cnt = 0;
    for (unsigned long long i = (unsigned long long)blockIdx.x * 256 + threadIdx.x; i < (1ull << 32);
         i += (unsigned long long)gridDim.x * 256) {
        const uint32_t u = (uint32_t)i;
        if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu)) continue;
        const float f = __uint_as_float(u);
        const probe_f2 v = {f, f};
        const uint32_t hw = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, probe_bf2)) & 0xFFFFu;
        const uint32_t sw = (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
        cnt += hw != sw;
    }
    if (cnt) atomicAdd(bad, cnt);
}

}  // namespace
}  // namespace stereo

using namespace capi;

extern "C" stereo_status stereo_valu_peak(stereo_context* ctx, int32_t op, double* tera_ops)
{
    if (!ctx || !tera_ops || (op != 0 && op != 1)) return STEREO_ERROR_INVALID_ARGUMENT;
    stereo_status st;
    if ((st = set_device(ctx))) return st;
    GUARD(ctx, "stereo_valu_peak");
    uint32_t* out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const int iters = 64;
    hipError_t e = hipMalloc((void**)&out, (size_t)stereo::PROBE_BLOCKS * 256 * 4);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float best = 0.f;
    for (int rep = 0; rep < 4 && e == hipSuccess; ++rep) {
        // the first launch is a warm-up; the fastest of the rest counts
        e = hipEventRecord(e0, nullptr);
        if (e == hipSuccess) {
            if (op == 0)
                hipLaunchKernelGGL(stereo::valu_peak_kernel<0>, dim3(stereo::PROBE_BLOCKS), dim3(256), 0, nullptr, out, iters);
            else
                hipLaunchKernelGGL(stereo::valu_peak_kernel<1>, dim3(stereo::PROBE_BLOCKS), dim3(256), 0, nullptr, out, iters);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && rep > 0 && (best == 0.f || ms < best)) best = ms;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (out) (void)hipFree(out);
    HIP_TRY(ctx, e, "valu probe");
    const double ops = (double)stereo::PROBE_BLOCKS * 256 * iters * stereo::UNROLL * stereo::CHAINS;
    *tera_ops = best > 0.f ? ops / (best * 1e-3) / 1e12 : 0.0;
    return STEREO_OK;
}

extern "C" stereo_status stereo_bf16_rounding_check(stereo_context* ctx, uint64_t* mismatches)
{
    if (!ctx || !mismatches) return STEREO_ERROR_INVALID_ARGUMENT;
    stereo_status st;
    if ((st = set_device(ctx))) return st;
    GUARD(ctx, "stereo_bf16_rounding_check");
    unsigned long long* d = nullptr;
    unsigned long long h = 0;
    hipError_t e = hipMalloc((void**)&d, sizeof(h));
    if (e == hipSuccess) e = hipMemset(d, 0, sizeof(h));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(stereo::bf16_rounding_kernel, dim3(4096), dim3(256), 0, nullptr, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost);
    if (d) (void)hipFree(d);
    HIP_TRY(ctx, e, "bf16 rounding check");
    *mismatches = h;
    return STEREO_OK;
}
