// Census transform kernels for gfx950.
//
// Reference: CensusTransform::Kernel::Transform (CluTec.Cuda.ImgProc/CensusTransform.cu:59-120)
// compares the 8 neighbours of a 3x3 window against the centre with strict '>' and
// packs them raster-ordered, most significant bit first, into a uint8.  The wide
// variant keeps those rules for a CW x CH window packed into uint64: raster position
// k maps to bit CW*CH-1-k, the centre bit is always 0.  Pixels whose window leaves
// the image are written as 0 (the reference leaves them untouched).
//
// Layout: a 64 x 32 output tile per 256-thread workgroup; the input tile plus halo
// is staged in LDS once, one pixel per dword so that every read is a conflict-free
// ds_read_b32.  Each thread produces an 8-row column strip and walks the strip's
// 8 + CH - 1 input rows once, feeding every output row whose window covers the row
// (the rows arrive in window order, so the raster bit order is kept): about 16 LDS
// reads per pixel instead of CW x CH.  A descriptor bit costs two VALU operations:
// t = centre - neighbour (negative exactly when neighbour > centre), then
// acc = v_alignbit(acc, t, 31), which shifts the sign bit of t into the accumulator.
// Both images of a pair go through one launch (blockIdx.z).  On request the same
// launch writes the 4 x 4 box means (sum + 8) >> 4 of both images (the quarter-
// resolution pair of the learned proposal), read from the staged tile.
//
// The matchers take their descriptors from census_tile_pk instead: the same comparisons
// at one VALU operation per bit, in a bit order of its own (see there).  The operator
// (launch_census) keeps the raster order above.
#include "common.hpp"
#include "kernels.hpp"

namespace stereo {

namespace {

constexpr int CT_W = 64;
constexpr int CT_H = 32;
constexpr int CT_STRIP = 8;

template <int CW, int CH>
struct CensusShape {
    static constexpr int RX = CW / 2, RY = CH / 2;
    static constexpr int LW = CT_W + 2 * RX, LH = CT_H + 2 * RY;
};

// One 64 x 32 tile of one image (origin x0, y0); q: its 4 x 4 box means, or q.ptr == nullptr.
template <int CW, int CH>
__device__ __forceinline__ void census_tile(const View& src, const View& dst, const View& q, int x0, int y0,
                                            uint32_t (*tile)[CensusShape<CW, CH>::LW + 1])
{
    constexpr int RX = CW / 2, RY = CH / 2, NB = CW * CH;
    constexpr int LW = CT_W + 2 * RX;
    constexpr int LH = CT_H + 2 * RY;
    constexpr int NHI = NB > 32 ? NB - 32 : 0;  // bits landing in the high word
    const int tid = threadIdx.x;
    {
        // every load of the thread issued before the first store: one load latency for
        // the staging instead of one per position (a position past the tile repeats the
        // last one and is not stored)
        constexpr int NS = LH * LW, NIT = (NS + 255) / 256;
        uint32_t a[NIT];
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
            const int i = min(tid + 256 * k, NS - 1);
            const int r = i / LW, c = i - r * LW;
            a[k] = row_ptr<const uint8_t>(src, clampi(y0 - RY + r, 0, src.h - 1))[clampi(x0 - RX + c, 0, src.w - 1)];
        }
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
            const int i = tid + 256 * k;
            if (i < NS) tile[i / LW][i % LW] = a[k];
        }
    }
    __syncthreads();

    const int cx = tid & 63;
    const int strip = tid >> 6;
    const int x = x0 + cx;
    const uint32_t* col = &tile[strip * CT_STRIP][cx];
    uint32_t hi[CT_STRIP], lo[CT_STRIP], cen[CT_STRIP];
#pragma unroll
    for (int k = 0; k < CT_STRIP; ++k) {
        cen[k] = col[(k + RY) * (LW + 1) + RX];
        hi[k] = 0;
        lo[k] = 0;
    }
    // one input row in flight ahead of the one being used; the compiler barrier at
    // the end of each row keeps the LDS loads of later rows from being hoisted
    // (which would hold all 14 rows in registers and cost most of the occupancy)
    constexpr int NR = CT_STRIP + CH - 1;
    uint32_t v[CW], vn[CW];
#pragma unroll
    for (int dx = 0; dx < CW; ++dx) v[dx] = col[dx];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if (r + 1 < NR) {
#pragma unroll
            for (int dx = 0; dx < CW; ++dx) vn[dx] = col[(r + 1) * (LW + 1) + dx];
        }
#pragma unroll
        for (int k = 0; k < CT_STRIP; ++k) {
            const int dy = r - k;
            if (dy < 0 || dy >= CH) continue;
#pragma unroll
            for (int dx = 0; dx < CW; ++dx) {
                const int bit = dy * CW + dx;
                const uint32_t t = cen[k] - v[dx];
                if (bit < NHI)
                    hi[k] = __builtin_amdgcn_alignbit(hi[k], t, 31);
                else
                    lo[k] = __builtin_amdgcn_alignbit(lo[k], t, 31);
            }
        }
#pragma unroll
        for (int dx = 0; dx < CW; ++dx) v[dx] = vn[dx];
        asm volatile("" ::: "memory");
    }
    // the descriptors are complete here: without this the compiler sinks each
    // pixel's bit chain into its bounds-checked store and keeps every row live
#pragma unroll
    for (int k = 0; k < CT_STRIP; ++k) asm volatile("" : "+v"(hi[k]), "+v"(lo[k]));
#pragma unroll
    for (int k = 0; k < CT_STRIP; ++k) {
        const int y = y0 + strip * CT_STRIP + k;
        const bool inside = x >= RX && x < src.w - RX && y >= RY && y < src.h - RY;
        if (x < dst.w && y < dst.h) {
            const uint64_t v = inside ? (((uint64_t)hi[k] << 32) | lo[k]) : 0ull;
            row_ptr<uint64_t>(dst, y)[x] = v;
        }
    }
    // quarter resolution: 16 x 8 box means of the tile (the image size is a
    // multiple of 4, so every 4 x 4 block lies inside the image)
    if (q.ptr && tid < (CT_W / 4) * (CT_H / 4)) {
        const int qx = tid & 15, qy = tid >> 4;
        const int X = x0 / 4 + qx, Y = y0 / 4 + qy;
        if (X < q.w && Y < q.h) {
            uint32_t sum = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int i = 0; i < 4; ++i) sum += tile[RY + 4 * qy + j][RX + 4 * qx + i];
            row_ptr<uint8_t>(q, Y)[X] = (uint8_t)((sum + 8) >> 4);
        }
    }
}

// The matchers' own descriptors.  They compare every neighbour of the window with the
// centre by the same rule as census_tile, but lay the bits out in an order made for
// packed arithmetic; the matchers only ever take the Hamming distance of two
// descriptors of one layout, which does not depend on the order.  The staged tile
// holds, per word, pixels c and c + 1 of a row as two 16-bit halves, so one LDS read
// brings two neighbours and one v_pk_sub_i16 compares both with the centre
// (replicated in both halves): a half is negative exactly when its neighbour is
// greater.  A v_perm_b32 with the sign-replicating selectors turns the signs of two
// such results into four 0x00 / 0xFF bytes, and one v_and_or_b32 keeps bit s of each
// byte: four descriptor bits for four VALU operations, against eight in census_tile.
// A row's neighbour pairs go dx = (0, 1), (2, 3), ...; with an odd width the last
// neighbour of two consecutive rows forms a pair, and the last row's alone is paired
// with 0 (never greater than the centre, so that bit is always 0).  The results are
// taken in window-row order, two at a time, into bit s of the bytes of the low word
// (first 8 pairs of results), then of the high word.  The centre's own comparison
// gives 0, as in census_tile.
typedef short census_s2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t census_pk_gt(uint32_t cen2, uint32_t pair)
{
    const census_s2 c = __builtin_bit_cast(census_s2, cen2);
    const census_s2 n = __builtin_bit_cast(census_s2, pair);
    return __builtin_bit_cast(uint32_t, (census_s2)(c - n));
}

// (s & m) | w as one v_and_or_b32 (the compiler otherwise pairs the ORs of two groups into
// a v_or3_b32 after two v_and_b32: three operations for two groups instead of two)
__device__ __forceinline__ uint32_t census_and_or(uint32_t s, uint32_t m, uint32_t w)
{
    uint32_t r;
    asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(s), "s"(m), "v"(w));
    return r;
}

template <int CW, int CH>
__device__ __forceinline__ void census_tile_pk(const View& src, const View& dst, const View& q, int x0, int y0,
                                               uint32_t (*tile)[CensusShape<CW, CH>::LW + 4])
{
    constexpr int RX = CW / 2, RY = CH / 2;
    constexpr int LW = CT_W + 2 * RX;
    constexpr int LH = CT_H + 2 * RY;
    constexpr int NP = CW / 2;        // neighbour pairs of a window row
    constexpr bool ODD = (CW & 1) != 0;  // a last neighbour at dx = CW - 1
    static_assert((CW * CH + 3) / 4 <= 16, "descriptor longer than 64 bits");
    constexpr int LS = LW + 4;  // LDS row stride: a multiple of 4 words for 16-byte stores
    const int tid = threadIdx.x;
    // Interior tiles of 4-byte-aligned images with a 9-wide window (the staged columns
    // x0 - 4 .. x0 + 71 all inside the row and the first of them on a dword boundary): a
    // thread loads dwords D, D' of a row and builds the four words of D's pixels with
    // v_perm_b32, pixel k in the low half and pixel k + 1 (the next dword's first for
    // k = 3) in the high half.  Elsewhere every word takes two clamped byte loads.
    const bool dwords = RX == 4 && x0 >= RX && x0 - RX + 4 * (LW / 4 + 1) <= src.w &&
                        ((reinterpret_cast<uintptr_t>(src.ptr) | src.pitch) & 3) == 0;
    if (dwords) {
        constexpr int NG = LW / 4, NS = LH * NG, NIT = (NS + 255) / 256;
        uint32_t d0[NIT], d1[NIT];
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
            const int i = min(tid + 256 * k, NS - 1);
            const int r = i / NG, g = i - r * NG;
            const uint32_t* row = reinterpret_cast<const uint32_t*>(
                row_ptr<const uint8_t>(src, clampi(y0 - RY + r, 0, src.h - 1)) + x0 - RX) + g;
            d0[k] = row[0];
            d1[k] = row[1];
        }
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
            const int i = tid + 256 * k;
            if (i < NS) {
                const int r = i / NG, g = i - r * NG;
                uint4 w;
                w.x = __builtin_amdgcn_perm(d1[k], d0[k], 0x0C010C00u);
                w.y = __builtin_amdgcn_perm(d1[k], d0[k], 0x0C020C01u);
                w.z = __builtin_amdgcn_perm(d1[k], d0[k], 0x0C030C02u);
                w.w = __builtin_amdgcn_perm(d1[k], d0[k], 0x0C040C03u);
                *reinterpret_cast<uint4*>(&tile[r][4 * g]) = w;
            }
        }
    } else {
        // every load of the thread issued before the first store, as in census_tile
        constexpr int NS = LH * LW, NIT = (NS + 255) / 256;
        uint32_t a[NIT], b[NIT];
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
            const int i = min(tid + 256 * k, NS - 1);
            const int r = i / LW, c = i - r * LW;
            const uint8_t* row = row_ptr<const uint8_t>(src, clampi(y0 - RY + r, 0, src.h - 1));
            const int gx = x0 - RX + c;
            a[k] = row[clampi(gx, 0, src.w - 1)];
            b[k] = row[clampi(gx + 1, 0, src.w - 1)];
        }
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
            const int i = tid + 256 * k;
            if (i < NS) tile[i / LW][i % LW] = a[k] | (b[k] << 16);
        }
    }
    __syncthreads();

    const int cx = tid & 63;
    const int strip = tid >> 6;
    const int x = x0 + cx;
    const uint32_t* col = &tile[strip * CT_STRIP][cx];
    uint32_t acc[CT_STRIP][2], pend[CT_STRIP], cen2[CT_STRIP];
    bool has[CT_STRIP];
    int slot[CT_STRIP];
#pragma unroll
    for (int k = 0; k < CT_STRIP; ++k) {
        cen2[k] = __builtin_amdgcn_perm(0u, col[(k + RY) * LS + RX], 0x01000100u);
        acc[k][0] = acc[k][1] = 0u;
        pend[k] = 0u;
        has[k] = false;
        slot[k] = 0;
    }
    // one result into pixel k's descriptor: every second one completes a group of four
    auto push = [&](int k, uint32_t t) {
        if (!has[k]) {
            pend[k] = t;
            has[k] = true;
            return;
        }
        const uint32_t s = __builtin_amdgcn_perm(pend[k], t, 0x0B0A0908u);
        uint32_t& w = acc[k][slot[k] >> 3];
        w = census_and_or(s, 0x01010101u << (slot[k] & 7), w);
        ++slot[k];
        has[k] = false;
    };
    constexpr int NR = CT_STRIP + CH - 1;
    constexpr int NV = NP + (ODD ? 1 : 0);
    uint32_t v[NV], vn[NV], sprev = 0u;
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = col[2 * j];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if (r + 1 < NR) {
#pragma unroll
            for (int j = 0; j < NV; ++j) vn[j] = col[(r + 1) * LS + 2 * j];
        }
        // the last neighbours of this row and the one before, and this row's alone
        const uint32_t qpair = ODD ? __builtin_amdgcn_perm(v[NV - 1], sprev, 0x05040100u) : 0u;
        const uint32_t qlast = ODD ? (v[NV - 1] & 0xFFFFu) : 0u;
#pragma unroll
        for (int k = 0; k < CT_STRIP; ++k) {
            const int dy = r - k;
            if (dy < 0 || dy >= CH) continue;
#pragma unroll
            for (int j = 0; j < NP; ++j) push(k, census_pk_gt(cen2[k], v[j]));
            if (ODD) {
                if (dy & 1) push(k, census_pk_gt(cen2[k], qpair));
                else if (dy == CH - 1) push(k, census_pk_gt(cen2[k], qlast));
            }
        }
        if (ODD) sprev = v[NV - 1];
#pragma unroll
        for (int j = 0; j < NV; ++j) v[j] = vn[j];
        asm volatile("" ::: "memory");
    }
#pragma unroll
    for (int k = 0; k < CT_STRIP; ++k)
        if (has[k]) push(k, 0u);
#pragma unroll
    for (int k = 0; k < CT_STRIP; ++k) asm volatile("" : "+v"(acc[k][0]), "+v"(acc[k][1]));
#pragma unroll
    for (int k = 0; k < CT_STRIP; ++k) {
        const int y = y0 + strip * CT_STRIP + k;
        const bool inside = x >= RX && x < src.w - RX && y >= RY && y < src.h - RY;
        if (x < dst.w && y < dst.h) {
            const uint64_t d = inside ? (((uint64_t)acc[k][1] << 32) | acc[k][0]) : 0ull;
            row_ptr<uint64_t>(dst, y)[x] = d;
        }
    }
    if (q.ptr && tid < (CT_W / 4) * (CT_H / 4)) {
        const int qx = tid & 15, qy = tid >> 4;
        const int X = x0 / 4 + qx, Y = y0 / 4 + qy;
        if (X < q.w && Y < q.h) {
            uint32_t sum = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int i = 0; i < 4; ++i) sum += tile[RY + 4 * qy + j][RX + 4 * qx + i] & 0xFFFFu;
            row_ptr<uint8_t>(q, Y)[X] = (uint8_t)((sum + 8) >> 4);
        }
    }
}

// Reference 3x3 census into uint8 (CensusTransform.cu:103-119 bit layout).
__global__ __launch_bounds__(256) void census3x3_ref_kernel(View src, View dst)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= dst.w || y >= dst.h) return;
    uint32_t v = 0;
    if (x >= 1 && x < src.w - 1 && y >= 1 && y < src.h - 1) {
        const uint8_t* r0 = row_ptr<const uint8_t>(src, y - 1);
        const uint8_t* r1 = row_ptr<const uint8_t>(src, y);
        const uint8_t* r2 = row_ptr<const uint8_t>(src, y + 1);
        const uint32_t c = r1[x];
        v |= (r0[x - 1] > c) ? 0x80u : 0u;
        v |= (r0[x] > c) ? 0x40u : 0u;
        v |= (r0[x + 1] > c) ? 0x20u : 0u;
        v |= (r1[x - 1] > c) ? 0x10u : 0u;
        v |= (r1[x + 1] > c) ? 0x08u : 0u;
        v |= (r2[x - 1] > c) ? 0x04u : 0u;
        v |= (r2[x] > c) ? 0x02u : 0u;
        v |= (r2[x + 1] > c) ? 0x01u : 0u;
    }
    row_ptr<uint8_t>(dst, y)[x] = (uint8_t)v;
}

// PK: the matchers' bit order (census_tile_pk), else the raster order of the operator.
template <int CW, int CH, bool PK>
__global__ __launch_bounds__(256) void census_wide_kernel(View src0, View dst0, View src1, View dst1, View q0, View q1)
{
    using S = CensusShape<CW, CH>;
    if (PK) {
        __shared__ __align__(16) uint32_t tile[S::LH][S::LW + 4];
        census_tile_pk<CW, CH>(blockIdx.z ? src1 : src0, blockIdx.z ? dst1 : dst0, blockIdx.z ? q1 : q0,
                               blockIdx.x * CT_W, blockIdx.y * CT_H, tile);
    } else {
        __shared__ uint32_t tile[S::LH][S::LW + 1];
        census_tile<CW, CH>(blockIdx.z ? src1 : src0, blockIdx.z ? dst1 : dst0, blockIdx.z ? q1 : q0,
                            blockIdx.x * CT_W, blockIdx.y * CT_H, tile);
    }
}

// Every image of a batch in one launch: workgroup b takes tile b - jobs[i].base of the
// image i whose tiles hold it (the jobs in order of base).
__device__ __forceinline__ int job_of(const ImageJob* jobs, int njobs, int b)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].base <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int CW, int CH>
__global__ __launch_bounds__(256) void census_batch_kernel(const ImageJob* jobs, int njobs)
{
    using S = CensusShape<CW, CH>;
    __shared__ __align__(16) uint32_t tile[S::LH][S::LW + 4];
    const int i = job_of(jobs, njobs, blockIdx.x);
    const ImageJob j = jobs[i];
    const int t = blockIdx.x - j.base;
    const int ntx = (j.src.w + CT_W - 1) / CT_W;
    const View none{nullptr, 0, 0, 0};
    census_tile_pk<CW, CH>(j.src, j.dst, none, (t % ntx) * CT_W, (t / ntx) * CT_H, tile);
}

// Rows of every image of a batch into the padded matcher inputs: workgroup b copies row
// b - jobs[i].base of image i.
__global__ __launch_bounds__(256) void pad_copy_batch_kernel(const ImageJob* jobs, int njobs)
{
    const ImageJob j = jobs[job_of(jobs, njobs, blockIdx.x)];
    const int y = blockIdx.x - j.base;
    const uint8_t* sp = row_ptr<const uint8_t>(j.src, y);
    uint8_t* dp = row_ptr<uint8_t>(j.dst, y);
    const int W = j.src.w;
    const bool aligned = ((reinterpret_cast<uintptr_t>(sp) | reinterpret_cast<uintptr_t>(dp)) & 3) == 0;
    for (int x = threadIdx.x * 4; x < W; x += 256 * 4) {
        if (aligned && x + 4 <= W) {
            *reinterpret_cast<uint32_t*>(dp + x) = *reinterpret_cast<const uint32_t*>(sp + x);
        } else {
            for (int k = 0; k < 4 && x + k < W; ++k) dp[x + k] = sp[x + k];
        }
    }
}

// The bytes of two W x H grey images into the padded matcher inputs in one launch: one
// workgroup row per image row, 4 bytes per lane when both rows are 4-byte aligned.
__global__ __launch_bounds__(256) void pad_copy_kernel(View s0, View d0, View s1, View d1)
{
    const View& src = blockIdx.z ? s1 : s0;
    const View& dst = blockIdx.z ? d1 : d0;
    const int y = blockIdx.y;
    const uint8_t* sp = row_ptr<const uint8_t>(src, y);
    uint8_t* dp = row_ptr<uint8_t>(dst, y);
    const int W = src.w;
    const bool aligned = ((reinterpret_cast<uintptr_t>(sp) | reinterpret_cast<uintptr_t>(dp)) & 3) == 0;
    for (int x = (blockIdx.x * 256 + threadIdx.x) * 4; x < W; x += gridDim.x * 256 * 4) {
        if (aligned && x + 4 <= W) {
            *reinterpret_cast<uint32_t*>(dp + x) = *reinterpret_cast<const uint32_t*>(sp + x);
        } else {
            for (int k = 0; k < 4 && x + k < W; ++k) dp[x + k] = sp[x + k];
        }
    }
}

template <int CW, int CH, bool PK>
hipError_t launch_wide(const View& src0, const View& dst0, const View* src1, const View* dst1, const View* q0,
                       const View* q1, hipStream_t s)
{
    dim3 grid((src0.w + CT_W - 1) / CT_W, (src0.h + CT_H - 1) / CT_H, src1 ? 2 : 1);
    const View none{nullptr, 0, 0, 0};
    hipLaunchKernelGGL((census_wide_kernel<CW, CH, PK>), grid, dim3(256), 0, s, src0, dst0, src1 ? *src1 : src0,
                       dst1 ? *dst1 : dst0, q0 ? *q0 : none, q1 ? *q1 : none);
    return hipGetLastError();
}

template <bool PK>
hipError_t launch_wide_any(const View& src0, const View& dst0, const View* src1, const View* dst1, int cw, int ch,
                           hipStream_t s, const View* q0, const View* q1)
{
    if (src1 && (src1->w != src0.w || src1->h != src0.h)) return hipErrorInvalidValue;
    if (q0 || q1) {
        // quarter-resolution outputs: sizes a multiple of 4, one per image
        if ((src0.w & 3) || (src0.h & 3) || (q1 && !src1)) return hipErrorInvalidValue;
        if (q0 && (q0->w != src0.w / 4 || q0->h != src0.h / 4)) return hipErrorInvalidValue;
        if (q1 && (q1->w != src0.w / 4 || q1->h != src0.h / 4)) return hipErrorInvalidValue;
    }
    if (cw == 9 && ch == 7) return launch_wide<9, 7, PK>(src0, dst0, src1, dst1, q0, q1, s);
    if (cw == 7 && ch == 9) return launch_wide<7, 9, PK>(src0, dst0, src1, dst1, q0, q1, s);
    if (cw == 5 && ch == 5) return launch_wide<5, 5, PK>(src0, dst0, src1, dst1, q0, q1, s);
    if (cw == 7 && ch == 7) return launch_wide<7, 7, PK>(src0, dst0, src1, dst1, q0, q1, s);
    if (cw == 3 && ch == 3) return launch_wide<3, 3, PK>(src0, dst0, src1, dst1, q0, q1, s);
    return hipErrorInvalidValue;
}

}  // namespace

bool census_supported(int cw, int ch, int ref3x3)
{
    if (ref3x3) return cw == 3 && ch == 3;
    return (cw == 9 && ch == 7) || (cw == 7 && ch == 9) || (cw == 5 && ch == 5) || (cw == 7 && ch == 7) ||
           (cw == 3 && ch == 3);
}

hipError_t launch_census(const View& src, const View& dst, int cw, int ch, int ref3x3, hipStream_t s)
{
    if (ref3x3) {
        dim3 grid((src.w + 63) / 64, (src.h + 3) / 4);
        hipLaunchKernelGGL(census3x3_ref_kernel, grid, dim3(256), 0, s, src, dst);
        return hipGetLastError();
    }
    return launch_wide_any<false>(src, dst, nullptr, nullptr, cw, ch, s, nullptr, nullptr);
}

hipError_t launch_pad_copy_pair(const View& src0, const View& dst0, const View& src1, const View& dst1, hipStream_t s)
{
    if (src0.w != src1.w || src0.h != src1.h || dst0.w < src0.w || dst1.w < src1.w || dst0.h < src0.h ||
        dst1.h < src1.h || src0.w < 1 || src0.h < 1)
        return hipErrorInvalidValue;
    dim3 grid((src0.w + 1023) / 1024, src0.h, 2);
    hipLaunchKernelGGL(pad_copy_kernel, grid, dim3(256), 0, s, src0, dst0, src1, dst1);
    return hipGetLastError();
}

int census_tiles(int w, int h) { return ((w + CT_W - 1) / CT_W) * ((h + CT_H - 1) / CT_H); }

hipError_t launch_census_batch(const ImageJob* jobs, int njobs, int total_tiles, int cw, int ch, hipStream_t s)
{
    if (njobs < 1 || total_tiles < 1) return hipErrorInvalidValue;
    const dim3 grid(total_tiles), block(256);
    if (cw == 9 && ch == 7) hipLaunchKernelGGL((census_batch_kernel<9, 7>), grid, block, 0, s, jobs, njobs);
    else if (cw == 7 && ch == 9) hipLaunchKernelGGL((census_batch_kernel<7, 9>), grid, block, 0, s, jobs, njobs);
    else if (cw == 5 && ch == 5) hipLaunchKernelGGL((census_batch_kernel<5, 5>), grid, block, 0, s, jobs, njobs);
    else if (cw == 7 && ch == 7) hipLaunchKernelGGL((census_batch_kernel<7, 7>), grid, block, 0, s, jobs, njobs);
    else if (cw == 3 && ch == 3) hipLaunchKernelGGL((census_batch_kernel<3, 3>), grid, block, 0, s, jobs, njobs);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_pad_copy_batch(const ImageJob* jobs, int njobs, int total_rows, hipStream_t s)
{
    if (njobs < 1 || total_rows < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pad_copy_batch_kernel, dim3(total_rows), dim3(256), 0, s, jobs, njobs);
    return hipGetLastError();
}

hipError_t launch_census_pair(const View& src0, const View& dst0, const View* src1, const View* dst1, int cw, int ch,
                              hipStream_t s, const View* q0, const View* q1)
{
    return launch_wide_any<true>(src0, dst0, src1, dst1, cw, ch, s, q0, q1);
}

}  // namespace stereo
