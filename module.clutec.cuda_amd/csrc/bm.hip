// Fused SAD / census-Hamming block matcher for gfx950: cost, separable box
// aggregation, left-to-right and right-to-left winner-take-all in one pass over
// the cost volume, followed by a small kernel that merges the right-view
// partials and applies the left-right check.
//
// Reference operators restated here (plain specification, see the oracle):
//   per-pixel costs      AbsoluteDifference  BlockMatcherNCC.cu:147-151
//                        CensusDifference    BlockMatcherNCC.cu:153-158
//   box aggregation      AlgoArraySum::SumCache (Kernel.Algo.ArraySum.h:149-189):
//                        separable running sums, slide-in / slide-out
//   winner-take-all      strict comparison scanning d upward (BlockMatcherNCC.cu:298)
//   left-right check     CrossCheckDisparityMaps::Kernel::Transform (:62-109)
//
// Inputs are padded views: columns [-BM_PAD_LEFT, W + BM_PAD_RIGHT) of every row
// are readable memory (values there only meet masked candidates), so the loads
// need no clamping.  Rows are clamped.
//
// Workgroup = 256 lanes = 4 waves; tile = 64 output columns x 56 output rows; all D
// disparities are streamed through the tile two at a time:
//   row phase     lane = (H row 0..63, run of 18 columns).  Costs of the pair
//                 (d, d+1) are packed as two uint16 halves of one register; the
//                 right-image window of the lane lives in registers and slides by
//                 register renaming across 8 disparities.  The horizontal running
//                 sum over 2*AX+1 columns takes its halo from the neighbouring runs
//                 with DPP row shifts, and the sums H go to LDS (double buffered).
//   column phase  lane = output column, wave = 14-row group.  Vertical running
//                 sums of H give the aggregated cost A of 14 pixels; the key
//                 (0xFF << 24 | A << 8 | d, see min_i32) is built with one
//                 v_perm_b32.  The left winner is a signed minimum per pixel.  The
//                 right winner for right pixel x-d is carried along the wave: each
//                 step shifts the running minima one lane up with a DPP wave shift
//                 fused into v_min_i32, so lane x always holds right pixel x-d.
//                 Running sums slide with one v_sad_u32 each (slide).  The keys
//                 leave the tile complemented, as (65535-A) << 8 | 255-d.  What leaves lane 63 is a
//                 partial result for the tile to the right; lane 63 parks it in
//                 LDS and the wave stores 64 steps at a time, coalesced.
#include "common.hpp"
#include "cnn_input.hpp"
#include "kernels.hpp"

namespace stereo {
namespace bm {

constexpr int TW = 64;
constexpr int TH = 56;
constexpr int HROWS = 64;
constexpr int RUN = 18;
constexpr int HSTRIDE = 72;  // == 8 (mod 32): the b64 H stores of a 16-lane group hit 32 distinct banks
constexpr int RG = TH / 4;

struct CostCensus {
    static constexpr int DC = 4;  // disparities per right-window load
    // workgroups per CU asked of the work-list kernel: three fit its LDS, and the
    // register cap (168) costs it only a few spills outside the disparity loop
    static constexpr int LIST_WG = 3;
    using E = uint64_t;           // element type in memory
    struct T { uint32_t lo, hi; };
    static __device__ __forceinline__ T make(E q) { return T{(uint32_t)q, (uint32_t)(q >> 32)}; }
    // two Hamming costs packed as (c(b0) | c(b1) << 16)
    static __device__ __forceinline__ uint32_t pair(T a, T b0, T b1)
    {
        uint32_t c1 = __builtin_popcount(a.lo ^ b1.lo) + __builtin_popcount(a.hi ^ b1.hi);
        return __builtin_popcount(a.lo ^ b0.lo) + __builtin_popcount(a.hi ^ b0.hi) + (c1 << 16);
    }
};

struct CostSad {
    static constexpr int DC = 8;
    static constexpr int LIST_WG = 2;  // three would spill inside the disparity loop
    using E = uint8_t;
    using T = uint32_t;
    static __device__ __forceinline__ T make(E v) { return v; }
    static __device__ __forceinline__ uint32_t pair(T a, T b0, T b1)
    {
        const uint32_t c1 = __builtin_amdgcn_sad_u8(a, b1, 0u);
        return __builtin_amdgcn_sad_u8(a, b0, c1 << 16);
    }
};

// One workgroup's work: a 64 x 56 tile of one frame and a disparity range [ds, de).
// Keys go out in one of two ways:
//   partials  right pixel xr's partial key of tile row r at rp[r * rp_stride + xr - x0 +
//             255 - rp_kofs], the left key of tile pixel (c, r) at lk[r * lk_stride + c];
//             a later merge takes the maximum over the items and segments;
//   planes    (ATOM) one W x H plane of right keys (rp = its row y0, rp_stride = W) and one
//             of left keys (lk = pixel (x0, y0), lk_stride = W), merged in place by
//             atomicMax: every segment of a tile folds its keys into the same planes,
//             which the merge kernel reads and clears for the next frame.
struct Job {
    View L, R;
    int W, H, bx;
    int x0, y0, ds, de;
    uint32_t* rp;
    int rp_stride, rp_kofs;
    uint32_t* lk;
    int lk_stride;
};

__device__ __forceinline__ bool job_is_edge(const Job& j)
{
    return (j.x0 < j.bx + j.de - 1) || (j.x0 + TW > j.W - j.bx);
}

// Whether any pixel of the tile has a candidate in [ds, de): left pixel x takes d when
// bx <= x < W - bx and d <= x - bx.  Near the left edge a whole tile's segment can fall
// short (1080p, D = 256: the last segment on the first three tile columns); such a job
// would only fold zero keys, which the key planes skip.
__device__ __forceinline__ bool job_has_candidates(const Job& j)
{
    const int xmax = min(j.x0 + TW - 1, j.W - j.bx - 1);
    return xmax >= j.x0 && xmax - j.bx >= j.ds;
}

// Keys inside the tile: 0xFF000000 | A << 8 | d for aggregated cost A and disparity d,
// negative as int32, so that the signed minimum picks the least cost and, among equal
// costs, the least d, and 0 stands for no candidate.  Their complement is the plane key
// (65535 - A) << 8 | (255 - d), folded with atomicMax where 0 is again no candidate.
__device__ __forceinline__ uint32_t min_i32(uint32_t a, uint32_t b) { return (int32_t)a < (int32_t)b ? a : b; }
__device__ __forceinline__ uint32_t plane_key(uint32_t k) { return k ? ~k : 0u; }

// tid: the thread index (threadIdx.x, or an opaque copy of it, see bm_list_kernel)
template <class Cost, int AX, int AY, bool EDGE, bool ATOM>
__device__ __forceinline__ void bm_tile(const Job& p, uint32_t (*hbuf)[HROWS * HSTRIDE], uint32_t (*ebuf)[RG][64],
                                        int tid)
{
    static_assert(TH + 2 * AY <= HROWS, "tile too tall for the aggregation window");
    static_assert(AX <= 4, "horizontal radius limited by the run halo");
    constexpr int DC = Cost::DC;
    const int x0 = p.x0, y0 = p.y0;
    const int W = p.W, H = p.H;
    const int ds = p.ds, de = p.de;

    // row-phase identity
    const int run = tid & 3;
    const int hr = tid >> 2;
    const int gyr = clampi(y0 - AY + hr, 0, H - 1);
    const int xs = x0 - 4 + RUN * run;
    using E = typename Cost::E;
    const E* const rowL = row_ptr<const E>(p.L, gyr) + xs;
    const E* rowR = row_ptr<const E>(p.R, gyr) + xs - ds - (DC - 1);  // window start for d0 = ds
    typename Cost::T cl[RUN];
#pragma unroll
    for (int i = 0; i < RUN; ++i) cl[i] = Cost::make(rowL[i]);

    // column-phase identity (gw is the wave index, uniform across the wave)
    const int lane = tid & 63;
    const int gw = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int x = x0 + lane;
    const int rows_here = H - (y0 + gw * RG);  // output rows of this wave inside the image
    uint32_t* const rp = ATOM ? p.rp + (size_t)gw * RG * p.rp_stride : p.rp + (size_t)gw * RG * p.rp_stride - p.rp_kofs;
    uint32_t* const eb = &ebuf[gw][0][0];
    const bool lane_ok = x >= p.bx && x < W - p.bx;
    const int dlim = x - p.bx;
    uint32_t lacc[RG], racc[RG];
#pragma unroll
    for (int i = 0; i < RG; ++i) { lacc[i] = 0; racc[i] = 0; }

    int s = 0;
    for (int d0 = ds; d0 < de; d0 += DC, rowR -= DC) {
        typename Cost::T cr[RUN + DC - 1];
#pragma unroll
        for (int k = 0; k < RUN + DC - 1; ++k) cr[k] = Cost::make(rowR[k]);

#pragma unroll
        for (int q = 0; q < DC / 2; ++q, ++s) {
            const int d = d0 + 2 * q;
            uint32_t* hb = hbuf[s & 1];

            // ---- row phase: packed costs of (d, d+1) and horizontal sums
            uint32_t e[RUN + 2 * AX];
#pragma unroll
            for (int i = 0; i < RUN; ++i)
                e[AX + i] = Cost::pair(cl[i], cr[i - 2 * q + DC - 1], cr[i - 2 * q + DC - 2]);
#pragma unroll
            for (int i = 0; i < AX; ++i) {
                e[i] = dpp_mov<DPP_ROW_SHR1>(e[AX + RUN - AX + i]);
                e[AX + RUN + i] = dpp_mov<DPP_ROW_SHL1>(e[AX + i]);
            }
            uint32_t h = 0;
#pragma unroll
            for (int i = 0; i <= 2 * AX; ++i) h += e[i];
            uint2* hrow = reinterpret_cast<uint2*>(hb + hr * HSTRIDE + RUN * run);
#pragma unroll
            for (int i = 0; i < RUN; i += 2) {
                if (i > 0) h = slide(h, e[i - 1], e[i + 2 * AX]);
                const uint32_t h0 = h;
                h = slide(h, e[i], e[i + 1 + 2 * AX]);
                hrow[i / 2] = make_uint2(h0, h);
            }
            __syncthreads();

            // ---- column phase: vertical sums, keys, both winner-take-alls
            uint32_t hv[RG + 2 * AY];
            const uint32_t* hcol = hb + gw * RG * HSTRIDE + lane + 4;
#pragma unroll
            for (int i = 0; i < RG + 2 * AY; ++i) hv[i] = hcol[i * HSTRIDE];
            uint32_t a = 0;
#pragma unroll
            for (int i = 0; i <= 2 * AY; ++i) a += hv[i];
            const uint32_t ddv = (uint32_t)d | ((uint32_t)(d + 1) << 8);
            bool ok0 = true, ok1 = true;
            if (EDGE) { ok0 = lane_ok && d <= dlim; ok1 = lane_ok && d + 1 <= dlim; }
            uint32_t k0[RG], k1[RG];
#pragma unroll
            for (int i = 0; i < RG; ++i) {
                if (i > 0) a = slide(a, hv[i - 1], hv[i + 2 * AY]);
                k0[i] = __builtin_amdgcn_perm(ddv, a, 0x0D010004u);
                k1[i] = __builtin_amdgcn_perm(ddv, a, 0x0D030205u);
                if (EDGE) { k0[i] = ok0 ? k0[i] : 0u; k1[i] = ok1 ? k1[i] : 0u; }
                lacc[i] = min_i32(lacc[i], min_i32(k0[i], k1[i]));
            }
            // right-view chain: one lane per step; all rows first so that no DPP read
            // waits on the VALU write just before it, then lane 63 parks what leaves
            static_assert(RG == 14, "the right-view chain block is written for 14 rows");
            min_shift_in14(racc, k0);
            if (lane == 63) {
#pragma unroll
                for (int i = 0; i < RG; ++i) eb[i * 64 + (d & 63)] = racc[i];
            }
            min_shift_in14(racc, k1);
            if (lane == 63) {
#pragma unroll
                for (int i = 0; i < RG; ++i) eb[i * 64 + ((d + 1) & 63)] = racc[i];
            }
            // flush the lane-63 emissions every 64 disparities (and at the end)
            const int dl = d + 1;
            if ((dl & 63) == 63 || dl == de - 1) {
                const int base = dl & ~63;
                if (lane <= (dl & 63) && base + lane >= ds) {
                    const int k = 318 - (base + lane);
                    if (ATOM) {
                        // right pixel x0 + k - 255, inside the image
                        const int xr = x0 + k - 255;
                        if (xr >= 0 && xr < W) {
#pragma unroll
                            for (int i = 0; i < RG; ++i) {
                                const uint32_t v = eb[i * 64 + lane];
                                if (i < rows_here && v) atomicMax(&rp[(size_t)i * p.rp_stride + xr], ~v);
                            }
                        }
                    } else {
#pragma unroll
                        for (int i = 0; i < RG; ++i)
                            if (i < rows_here) rp[i * p.rp_stride + k] = plane_key(eb[i * 64 + lane]);
                    }
                }
            }
        }
    }

    // what is still in flight belongs to right pixels x0+lane-(de-1)
    if (ATOM) {
        const int xr = x0 + lane - (de - 1);
        if (xr >= 0 && xr < W) {
#pragma unroll
            for (int i = 0; i < RG; ++i)
                if (i < rows_here && racc[i]) atomicMax(&rp[(size_t)i * p.rp_stride + xr], ~racc[i]);
        }
    } else {
        const int k = lane + 255 - (de - 1);
#pragma unroll
        for (int i = 0; i < RG; ++i)
            if (i < rows_here) rp[i * p.rp_stride + k] = plane_key(racc[i]);
    }
    if (x < W) {
        uint32_t* lk = p.lk + (size_t)gw * RG * p.lk_stride + lane;
#pragma unroll
        for (int i = 0; i < RG; ++i) {
            if (i >= rows_here) continue;
            if (ATOM) {
                if (lacc[i]) atomicMax(&lk[(size_t)i * p.lk_stride], ~lacc[i]);
            } else {
                lk[(size_t)i * p.lk_stride] = plane_key(lacc[i]);
            }
        }
    }
}

template <class Cost, int AX, int AY, bool ATOM>
__device__ __forceinline__ void run_job(const Job& j, uint32_t (*hbuf)[HROWS * HSTRIDE], uint32_t (*ebuf)[RG][64],
                                        int tid)
{
    if (job_is_edge(j))
        bm_tile<Cost, AX, AY, true, ATOM>(j, hbuf, ebuf, tid);
    else
        bm_tile<Cost, AX, AY, false, ATOM>(j, hbuf, ebuf, tid);
}

// Grid launch over one frame: one workgroup per (tile x, tile y, disparity segment),
// renumbered so that every XCD works on a contiguous band of tiles.
template <class Cost, int AX, int AY>
__global__ __launch_bounds__(256, 2) void bm_fused_kernel(BmParams p)
{
    __shared__ uint32_t hbuf[2][HROWS * HSTRIDE];
    __shared__ uint32_t ebuf[4][RG][64];
    // one-dimensional grid of ntx * nty * nseg workgroups
    const int w = xcd_remap(blockIdx.x, gridDim.x);
    // segment fastest: the segments of a tile run together on one XCD, so their atomics
    // on the same key-plane lines meet in its L2, as do the census rows they read
    const int seg = w % p.nseg, rest = w / p.nseg;
    const int tx = rest % p.ntx, ty = rest / p.ntx;
    Job j;
    j.L = p.L;
    j.R = p.R;
    j.W = p.W;
    j.H = p.H;
    j.bx = p.bx;
    j.x0 = tx * TW;
    j.y0 = ty * TH;
    j.ds = seg * p.seg_len;
    j.de = min(p.D, j.ds + p.seg_len);
    // the key planes: every segment of the tile folds into the same rows
    j.rp = p.rkey + (size_t)j.y0 * p.W;
    j.rp_stride = p.W;
    j.rp_kofs = 0;
    j.lk = p.lkey + (size_t)j.y0 * p.W + j.x0;
    j.lk_stride = p.W;
    if (!job_has_candidates(j)) return;
    run_job<Cost, AX, AY, true>(j, hbuf, ebuf, threadIdx.x);
}

// Persistent launch over a work list: each workgroup takes items blockIdx.x,
// blockIdx.x + gridDim.x, ... up to the device-side count.
template <class Cost, int AX, int AY>
__global__ __launch_bounds__(256, Cost::LIST_WG) void bm_list_kernel(BmListParams p)
{
    __shared__ uint32_t hbuf[2][HROWS * HSTRIDE];
    __shared__ uint32_t ebuf[4][RG][64];
    __shared__ int next;
    const int n = *p.nitems;
    // Items go out in schedule order: workgroup b takes item b first, with no counter
    // traffic at the launch; after that a device counter (reset to 0 by the finish
    // kernel) hands item gridDim.x + c to the c-th workgroup that comes back free.
    for (int k = blockIdx.x; k < n;) {
        const int i = p.sched ? p.sched[k] : k;
        const BmItem it = p.items[i];
        const BmFrame& f = p.frames[it.frame];
        Job j;
        j.L = f.L;
        j.R = f.R;
        j.W = f.W;
        j.H = f.H;
        j.bx = f.bx;
        j.x0 = it.tx * TW;
        j.y0 = it.ty * TH;
        j.ds = it.ds;
        j.de = it.de;
        // the frame's key planes
        j.rp = p.rplane + f.plane_off + (size_t)j.y0 * f.W;
        j.rp_stride = f.W;
        j.rp_kofs = 0;
        j.lk = p.lplane + f.plane_off + (size_t)j.y0 * f.W + j.x0;
        j.lk_stride = f.W;
        if (job_has_candidates(j)) {
            // an opaque copy of the thread index per item: the addresses derived from it
            // are recomputed for each item instead of being hoisted out of this loop into
            // registers that stay live through the whole item
            int tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            run_job<Cost, AX, AY, true>(j, hbuf, ebuf, tid);
        }
        __syncthreads();  // the item is done with the LDS buffers, and `next` is free
        if (threadIdx.x == 0) next = (int)gridDim.x + atomicAdd(p.counter, 1);
        __syncthreads();
        k = next;
    }
}

// TPS: the most tiles a segment's left pixels xr + ds .. xr + de - 1 can span,
// (seg_len - 1) / 64 + 2, at most 5 (seg_len <= 256).
template <int TPS>
__device__ __forceinline__ uint32_t rkey_at(const BmFinishParams& f, int xr, int y)
{
    // right pixel xr meets left pixels xr+d; segment s covers d in [ds, de), and the
    // tile holding left pixel x keeps the partial under index xr - x0 + 255.  A
    // segment spans at most TPS tiles.  Two segments' loads are issued
    // together and unconditionally: past the segment's last tile a load repeats that
    // tile (the max is unchanged), and a segment with no tile reads a fixed address
    // whose value is discarded.
    uint32_t best = 0;
    for (int s0 = 0; s0 < f.nseg; s0 += 2) {
        uint32_t v[2][TPS];
        bool any[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int s = min(s0 + q, f.nseg - 1);
            const int ds = s * f.seg_len;
            const int de = min(f.D, ds + f.seg_len);
            const int tlo = (xr + ds) / 64;
            const int thi = min((xr + de - 1) / 64, f.ntx - 1);
            any[q] = tlo <= thi;
            const uint32_t* base = f.rpart + ((size_t)s * f.ntx * f.H + y) * BM_NPART + (xr + 255);
#pragma unroll
            for (int j = 0; j < TPS; ++j) {
                const int t = min(tlo + j, thi);
                const uint32_t* a = any[q] ? base + ((size_t)t * f.H * BM_NPART - 64 * t) : f.rpart;
                v[q][j] = *a;
            }
        }
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int j = 0; j < TPS; ++j) best = (any[q] && v[q][j] > best) ? v[q][j] : best;
    }
    return best;
}

// Output of one left pixel: the checked disparity and the maps, in every requested
// format.  lk = merged left key, rrow = merged right keys of the row.
__device__ __forceinline__ void write_pixel(const View& disp, const View& disp_l, const View& disp_r,
                                            const View& dispid, int x, int y, bool inside, uint32_t lk,
                                            const uint32_t* rrow, int tol, int value_base)
{
    int dl = -1, dr = -1, outd = -1;
    if (inside) {
        dl = key_disp(lk);
        if (dl >= 0) {
            const int dxr = key_disp(rrow[x - dl]);
            outd = (dxr >= 0 && abs(dxr - dl) <= tol) ? dl : -1;
        }
        dr = key_disp(rrow[x]);
    }
    if (disp.ptr) row_ptr<int16_t>(disp, y)[x] = (int16_t)outd;
    if (disp_l.ptr) row_ptr<int16_t>(disp_l, y)[x] = (int16_t)dl;
    if (disp_r.ptr) row_ptr<int16_t>(disp_r, y)[x] = (int16_t)dr;
    if (dispid.ptr) {
        uint16_t id;
        if (!inside) id = 0;                                // Unknown
        else if (outd >= 0) id = (uint16_t)(outd + 1);     // First + d
        else if (dl >= 0) id = 0xFFFFu - 5u;                // Inconsistent
        else id = 0xFFFFu - 2u;                             // NotFound
        int cost = lk ? key_cost(lk) : 0;
        if (lk && value_base) cost = value_base - cost;
        ushort4 v;
        v.x = id;
        v.y = (uint16_t)(cost > 65535 ? 65535 : cost);
        v.z = 0;
        v.w = 0;
        row_ptr<ushort4>(dispid, y)[x] = v;
    }
}

// One workgroup per image row: first every right pixel of the row merges its partial
// keys (contiguous reads, one pass) into LDS, then every left pixel looks up the right
// winner at x - dL for the check.  The left keys of a thread's first FIN_PF pixels are
// loaded before the barrier, so their latency overlaps the right-key merge.
constexpr int FIN_PF = 8;

__device__ __forceinline__ uint32_t lkey_at(const BmFinishParams& f, int x, int y)
{
    uint32_t lk = 0;
    for (int s = 0; s < f.nseg; ++s) {
        const uint32_t v = f.lkey[(size_t)s * f.W * f.H + (size_t)y * f.W + x];
        lk = v > lk ? v : lk;
    }
    return lk;
}

template <int TPS>
__global__ __launch_bounds__(256) void bm_finish_kernel(BmFinishParams f)
{
    extern __shared__ uint32_t rrow[];
    const int y = blockIdx.x;
    const int W = f.W;
    const bool row_ok = y >= f.by && y < f.H - f.by;
    // the left keys of the thread's first FIN_PF pixels, two segments per batch of loads
    // (a position past the row reads the row's last pixel; the value is discarded;
    // slots wholly past the row are skipped)
    uint32_t lpre[FIN_PF];
    {
        const size_t plane = (size_t)W * f.H;
#pragma unroll
        for (int k = 0; k < FIN_PF; ++k) lpre[k] = 0u;
        for (int s0 = 0; s0 < f.nseg; s0 += 2) {
            const size_t sa = (size_t)s0 * plane, sb = (size_t)min(s0 + 1, f.nseg - 1) * plane;
            uint32_t la[FIN_PF], lb[FIN_PF];
#pragma unroll
            for (int k = 0; k < FIN_PF; ++k) {
                la[k] = lb[k] = 0u;
                if (256 * k < W) {
                    const size_t i = (size_t)y * W + min((int)threadIdx.x + 256 * k, W - 1);
                    la[k] = f.lkey[sa + i];
                    lb[k] = f.lkey[sb + i];
                }
            }
#pragma unroll
            for (int k = 0; k < FIN_PF; ++k) {
                const uint32_t m = la[k] > lb[k] ? la[k] : lb[k];
                lpre[k] = m > lpre[k] ? m : lpre[k];
            }
        }
#pragma unroll
        for (int k = 0; k < FIN_PF; ++k) {
            const int x = threadIdx.x + 256 * k;
            if (!(row_ok && x >= f.bx && x < W - f.bx)) lpre[k] = 0u;
        }
    }
#pragma unroll 2
    for (int x = threadIdx.x; x < W; x += 256) {
        const bool col_ok = x >= f.bx && x < W - f.bx;
        const uint32_t r = rkey_at<TPS>(f, x, y);
        rrow[x] = (row_ok && col_ok) ? r : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < FIN_PF; ++k) {
        const int x = threadIdx.x + 256 * k;
        if (x < W)
            write_pixel(f.disp, f.disp_l, f.disp_r, f.dispid, x, y, row_ok && x >= f.bx && x < W - f.bx, lpre[k], rrow,
                        f.tol, f.value_base);
    }
    for (int x = threadIdx.x + 256 * FIN_PF; x < W; x += 256) {
        const bool inside = row_ok && x >= f.bx && x < W - f.bx;
        write_pixel(f.disp, f.disp_l, f.disp_r, f.dispid, x, y, inside, inside ? lkey_at(f, x, y) : 0u, rrow, f.tol,
                    f.value_base);
    }
}

// Check of one image row from the key planes that the matchers fold their keys into
// with atomicMax: the workgroup reads the row's right keys into LDS and every left pixel
// looks up the right winner at x - dL.  Each key is cleared by the thread that read it,
// so the planes start the next frame at zero.
__device__ __forceinline__ void planes_row(const View& disp, const View& disp_l, const View& disp_r, const View& dispid,
                                           int W, int y, bool row_ok, int bx, int tol, int value_base, uint32_t* lrow,
                                           uint32_t* rrowg, uint32_t* rrow)
{
    uint32_t lpre[FIN_PF];
#pragma unroll
    for (int k = 0; k < FIN_PF; ++k) {
        const int x = threadIdx.x + 256 * k;
        lpre[k] = 0u;
        if (x < W) {
            lpre[k] = lrow[x];
            lrow[x] = 0u;
        }
    }
    for (int x = threadIdx.x; x < W; x += 256) {
        const bool col_ok = x >= bx && x < W - bx;
        const uint32_t r = rrowg[x];
        rrowg[x] = 0u;
        rrow[x] = (row_ok && col_ok) ? r : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < FIN_PF; ++k) {
        const int x = threadIdx.x + 256 * k;
        if (x < W) {
            const bool inside = row_ok && x >= bx && x < W - bx;
            write_pixel(disp, disp_l, disp_r, dispid, x, y, inside, inside ? lpre[k] : 0u, rrow, tol, value_base);
        }
    }
    for (int x = threadIdx.x + 256 * FIN_PF; x < W; x += 256) {
        const bool inside = row_ok && x >= bx && x < W - bx;
        const uint32_t lk = lrow[x];
        lrow[x] = 0u;
        write_pixel(disp, disp_l, disp_r, dispid, x, y, inside, inside ? lk : 0u, rrow, tol, value_base);
    }
}

// SAD / census grid launch: one workgroup per image row.
__global__ __launch_bounds__(256) void bm_finish_planes_kernel(BmFinishParams f)
{
    extern __shared__ uint32_t rrow[];
    const int y = blockIdx.x;
    const bool row_ok = y >= f.by && y < f.H - f.by;
    planes_row(f.disp, f.disp_l, f.disp_r, f.dispid, f.W, y, row_ok, f.bx, f.tol, f.value_base,
               f.lplane + (size_t)y * f.W, f.rplane + (size_t)y * f.W, rrow);
    if (f.cnn.x) {
        // the learned proposal's network input from the row just written
        __shared__ uint64_t valid[CNN_MAX_W4 / 64];
        __syncthreads();
        cnn_input_row(f.cnn.L4, f.cnn.R4, f.disp, f.cnn.n, f.cnn.x, f.cnn.dfill, valid, y);
    }
}

// Clears n words (the key planes, where a caller cannot rely on the last frame's merge
// having cleared them).
__global__ __launch_bounds__(256) void bm_clear_kernel(uint32_t* p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0u;
}

// Merge and check for the work-list matcher: one workgroup per row of every frame, from
// the frame's key planes (row r of the launch is row r - row_base of the frame whose rows
// it falls in).  Also returns the item counter to 0 for the next matcher launch.
__global__ __launch_bounds__(256) void bm_list_finish_kernel(BmListParams p)
{
    extern __shared__ uint32_t rrow[];
    const int r = blockIdx.x;
    if (r == 0 && threadIdx.x == 0) *p.counter = 0;  // the matcher launch before this one has ended
    // the frame whose rows hold r: the last one with row_base <= r (frames in row order),
    // by binary search, so the dependent loads before the row's work do not grow with the
    // batch (64 frames: 6 instead of up to 63)
    int fi = 0, hi = p.nframes - 1;
    while (fi < hi) {
        const int mid = (fi + hi + 1) >> 1;
        if (p.frames[mid].row_base <= r) fi = mid;
        else hi = mid - 1;
    }
    const BmFrame& f = p.frames[fi];
    const int y = r - f.row_base;
    const bool row_ok = y >= f.by && y < f.H - f.by;
    const size_t row = f.plane_off + (size_t)y * f.W;
    planes_row(f.disp, f.disp_l, f.disp_r, f.dispid, f.W, y, row_ok, f.bx, f.tol, f.value_base, p.lplane + row,
               p.rplane + row, rrow);
}

}  // namespace bm


template <class Cost>
static hipError_t launch_list_cost(const BmListParams& p, int ww, int grid, hipStream_t s)
{
    if (ww == 9) hipLaunchKernelGGL((bm::bm_list_kernel<Cost, 4, 4>), dim3(grid), dim3(256), 0, s, p);
    else if (ww == 7) hipLaunchKernelGGL((bm::bm_list_kernel<Cost, 3, 3>), dim3(grid), dim3(256), 0, s, p);
    else hipLaunchKernelGGL((bm::bm_list_kernel<Cost, 2, 2>), dim3(grid), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_bm_list(int kind, int ww, const BmListParams& p, int grid, hipStream_t s)
{
    if (grid < 1 || !p.counter || !p.lplane || !p.rplane) return hipErrorInvalidValue;
    if (kind == BM_COST_CENSUS) return launch_list_cost<bm::CostCensus>(p, ww, grid, s);
    return launch_list_cost<bm::CostSad>(p, ww, grid, s);
}

hipError_t launch_bm_list_finish(const BmListParams& p, int max_width, hipStream_t s)
{
    if (max_width < 1 || max_width > BM_LIST_MAX_WIDTH || p.total_rows < 1 || !p.counter || !p.lplane || !p.rplane)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(bm::bm_list_finish_kernel, dim3(p.total_rows), dim3(256), (size_t)max_width * 4, s, p);
    return hipGetLastError();
}

bool bm_supported(int kind, int ww, int wh, int D)
{
    if (D < 8 || D > 256 || (D % 8) != 0) return false;
    if (kind == BM_COST_CENSUS) return (ww == 9 && wh == 9) || (ww == 7 && wh == 7) || (ww == 5 && wh == 5);
    if (kind == BM_COST_SAD) return (ww == 9 && wh == 9) || (ww == 7 && wh == 7) || (ww == 5 && wh == 5);
    return false;
}

void bm_grid(int W, int H, int D, int* ntx, int* nty, int* nseg, int* seg_len)
{
    *ntx = (W + bm::TW - 1) / bm::TW;
    *nty = (H + bm::TH - 1) / bm::TH;
    // Split the disparity range so that the launch holds enough workgroups for the 256
    // CUs (3 resident each) to finish together.  Every segment folds into the same key
    // planes, so the merge does not grow with the count; what a segment costs is its
    // set-up and its atomics.  Measured best: whole 64-disparity blocks (the kernel parks
    // its right-view emissions in blocks of 64) when they give at least 1200 workgroups
    // (1080p D=256: 4 x 64, 2160p), else segments of 16 (640x480 D=64, 720p D=128), where
    // 8 already costs more in set-up than it gains in balance.
    const int tiles = *ntx * *nty;
    int len = (long)tiles * ((D + 63) / 64) >= 1200 ? 64 : 16;
    if (len > D) len = D;
    *seg_len = len;
    *nseg = (D + len - 1) / len;
}

template <class Cost>
static hipError_t launch_cost(const BmParams& p, int ww, hipStream_t s)
{
    dim3 grid(p.ntx * ((p.H + bm::TH - 1) / bm::TH) * p.nseg);
    if (ww == 9) hipLaunchKernelGGL((bm::bm_fused_kernel<Cost, 4, 4>), grid, dim3(256), 0, s, p);
    else if (ww == 7) hipLaunchKernelGGL((bm::bm_fused_kernel<Cost, 3, 3>), grid, dim3(256), 0, s, p);
    else hipLaunchKernelGGL((bm::bm_fused_kernel<Cost, 2, 2>), grid, dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_bm(int kind, int ww, const BmParams& p, hipStream_t s)
{
    if (kind == BM_COST_CENSUS) return launch_cost<bm::CostCensus>(p, ww, s);
    return launch_cost<bm::CostSad>(p, ww, s);
}

hipError_t launch_bm_clear(uint32_t* p, size_t n, hipStream_t s)
{
    if (!n) return hipSuccess;
    const size_t blocks = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
    hipLaunchKernelGGL(bm::bm_clear_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_bm_finish(const BmFinishParams& f, hipStream_t s)
{
    if (f.W > BM_MAX_WIDTH || f.seg_len < 1 || f.seg_len > 256) return hipErrorInvalidValue;
    const size_t lds = (size_t)f.W * sizeof(uint32_t);
    if (f.lplane) {
        if (!f.rplane) return hipErrorInvalidValue;
        if (f.cnn.x && (!f.disp.ptr || !f.cnn.dfill || f.W > CNN_MAX_W4 || f.cnn.L4.w != f.W || f.cnn.L4.h != f.H ||
                        f.cnn.R4.w != f.W || f.cnn.R4.h != f.H))
            return hipErrorInvalidValue;
        hipLaunchKernelGGL(bm::bm_finish_planes_kernel, dim3(f.H), dim3(256), lds, s, f);
        return hipGetLastError();
    }
    const int tps = (f.seg_len - 1) / bm::TW + 2;
    if (tps <= 2) hipLaunchKernelGGL(bm::bm_finish_kernel<2>, dim3(f.H), dim3(256), lds, s, f);
    else if (tps == 3) hipLaunchKernelGGL(bm::bm_finish_kernel<3>, dim3(f.H), dim3(256), lds, s, f);
    else if (tps == 4) hipLaunchKernelGGL(bm::bm_finish_kernel<4>, dim3(f.H), dim3(256), lds, s, f);
    else hipLaunchKernelGGL(bm::bm_finish_kernel<5>, dim3(f.H), dim3(256), lds, s, f);
    return hipGetLastError();
}

}  // namespace stereo
