// Adaptive-window block matchers (AW1 / AW2) for gfx950.
//
// Reference: AlgoAdaptWnd_W32H16::SNeighborhood<Two_OutOf_Four>::Eval
// (CluTec.Cuda.ImgProc/Kernel.Algo.AdaptWnd_W32H16.h:240-285) and
// AlgoAdaptWnd3D Two_OutOf_Four (Kernel.Algo.AdaptWnd3D.h:244-300): the window of a pixel
// is its centre s x s sub-patch plus the two best of the four edge neighbours at
// distance s (above, left, below, right).  BlockMatcherAW1 scores similarity
// 255 - |a - b| (BlockMatcherAW1.cu:191-195) and keeps the maximum with strict '>'
// (:294); BlockMatcherAW2 scores |a - b| (BlockMatcherAW2.cu:129-133) and keeps the
// minimum with strict '<' (:167).  Summed over three sub-patches of s*s pixels the AW1
// similarity is 255*3*s*s minus the AW2 SAD of the same three sub-patches, and the two
// most similar neighbours are the two with the smallest SAD, so both matchers pick the
// same disparity with the same tie rule; this kernel computes the AW2 integer value
// and the AW1 value is derived from it.
//
// Layout (pairs (d, d+1) packed in uint16 halves, as in bm.hip): tile = 64 columns x
// 48 rows, 4 waves.  Row phase: lane = (H row, run of 20 columns), |a-b| and horizontal
// s-sums into LDS.  Column phase: lane = column, wave = 12 rows; vertical s-sums give
// the sub-patch sums S at the pixel's column (for rows o-s, o, o+s) and at columns
// x-s and x+s; the two smallest neighbours come from packed uint16 min/max; the winner
// key, both winner-take-alls and the right-view chain are those of the SAD matcher.
#include "common.hpp"
#include "kernels.hpp"

namespace stereo {
namespace aw {

constexpr int TW = 64;
constexpr int RG = 12;
constexpr int TH = 4 * RG;   // 48
constexpr int HROWS = 64;
constexpr int RUN = 20;      // 4 runs cover cost columns [-8, 72)
constexpr int HSTRIDE = 88;
constexpr int DC = 8;

__device__ __forceinline__ uint32_t pk_min(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned short, a),
                                                                 __builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned short, b)));
}

__device__ __forceinline__ uint32_t pk_max(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned short, a),
                                                                 __builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned short, b)));
}

// BAND: the band search around an initial map (MatcherUseDisparity of BlockMatcherAW1.cu:1336
// and BlockMatcherAW2.cu:858): only the 8-disparity bins set in the tile's masks p.bins
// are searched (the same per-tile bins as the SAD, census and NCC matchers).
template <int SP, bool EDGE, bool BAND>
__device__ __forceinline__ void aw_tile(const AwParams& p, uint32_t (*hbuf)[HROWS * HSTRIDE], uint32_t (*ebuf)[RG][64], int tx, int ty, int seg)
{
    constexpr int R = SP / 2;        // sub-patch radius
    constexpr int B = SP + R;        // window radius
    static_assert(TH + 2 * B <= HROWS, "tile too tall");
    static_assert(B + 8 <= 16, "runs cover the window's columns");
    constexpr int NC = RG + 2 * SP;  // centre-column sub-patch sums per wave (rows o-s .. o+s)
    const int tid = threadIdx.x;
    const int x0 = tx * TW, y0 = ty * TH;
    const int W = p.W, H = p.H;
    const int ds = seg * p.seg_len;
    const int de = min(p.D, ds + p.seg_len);

    // band search: the tile's rows lie in at most two 64 x BM_TILE_H bin tiles; rows before
    // ysplit take m0, the others m1
    uint32_t m0 = ~0u, m1 = ~0u;
    int ysplit = 1 << 30;
    if (BAND) {
        const int tb0 = y0 / BM_TILE_H, tb1 = min((y0 + TH - 1) / BM_TILE_H, p.bins_nty - 1);
        m0 = p.bins[tb0 * p.ntx + tx];
        m1 = p.bins[tb1 * p.ntx + tx];
        ysplit = (tb0 + 1) * BM_TILE_H;
        const int b0 = ds >> 3, b1 = (de - 1) >> 3;
        // nothing to search in this segment: its keys would all be empty, which the key
        // planes already hold
        if (((m0 | m1) & (uint32_t)((2ull << b1) - (1ull << b0))) == 0u) return;
    }

    // row phase identity: H row hr is image row y0 - B + hr
    const int run = tid & 3;
    const int hr = tid >> 2;
    const int gyr = clampi(y0 - B + hr, 0, H - 1);
    const int xs = x0 - 8 + RUN * run;
    const uint8_t* const rowL = row_ptr<const uint8_t>(p.L, gyr) + xs;
    const uint8_t* rowR = row_ptr<const uint8_t>(p.R, gyr) + xs - ds - (DC - 1);
    uint32_t cl[RUN];
#pragma unroll
    for (int i = 0; i < RUN; ++i) cl[i] = rowL[i];

    // column phase identity
    const int lane = tid & 63;
    const int gw = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int x = x0 + lane;
    const int rows_here = H - (y0 + gw * RG);
    // the wave's first row of the right-key plane; every segment folds into it
    uint32_t* const rp = p.rkey + (size_t)(y0 + gw * RG) * W;
    uint32_t* const eb = &ebuf[gw][0][0];
    const bool lane_ok = x >= p.bx && x < W - p.bx;
    const int dlim = x - p.bx;
    uint32_t lacc[RG], racc[RG];
#pragma unroll
    for (int i = 0; i < RG; ++i) { lacc[i] = 0; racc[i] = 0; }

    // after the pair ending at disparity dl: at the end of a block of 64 disparities (or of
    // the segment) the right-view values parked in the wave's LDS rows go to the partials
    auto flush = [&](int dl) {
        if ((dl & 63) == 63 || dl == de - 1) {
            const int base = dl & ~63;
            // right pixel x0 + 63 - (base + lane), inside the image
            const int xr = x0 + 63 - (base + lane);
            if (lane <= (dl & 63) && base + lane >= ds && xr >= 0 && xr < W) {
#pragma unroll
                for (int i = 0; i < RG; ++i) {
                    const uint32_t v = eb[i * 64 + lane];
                    if (i < rows_here && v) atomicMax(&rp[(size_t)i * W + xr], v);
                }
            }
        }
    };

    int s = 0;
    for (int d0 = ds; d0 < de; d0 += DC, rowR -= DC) {
        // the band search: rows of this wave that search the group (bit o: row o), wave-uniform
        uint32_t keep = (1u << RG) - 1u;
        if (BAND) {
            const uint32_t bit = 1u << (d0 >> 3);  // the group's bin (d0 is a multiple of 8)
            const uint32_t lo_rows = (1u << clampi(ysplit - (y0 + gw * RG), 0, RG)) - 1u;
            keep = ((m0 & bit) ? lo_rows : 0u) | ((m1 & bit) ? (((1u << RG) - 1u) & ~lo_rows) : 0u);
            if (((m0 | m1) & bit) == 0u) {
                // no pixel of the tile searches these 8 disparities: the right-view chain
                // moves on by 8 lanes with no new candidates.  The values that leave the
                // wave at lane 63 during those 8 steps sit in lanes 62 .. 55 now.
                if (lane >= 63 - DC && lane < 63) {
                    const int slot = (d0 + 62 - lane) & 63;
#pragma unroll
                    for (int i = 0; i < RG; ++i) eb[i * 64 + slot] = racc[i];
                }
#pragma unroll
                for (int i = 0; i < RG; ++i) {
                    const uint32_t v = __shfl_up(racc[i], DC);
                    racc[i] = lane >= DC ? v : 0u;
                }
                flush(d0 + DC - 1);
                continue;
            }
        }
        uint32_t cr[RUN + DC - 1];
#pragma unroll
        for (int k = 0; k < RUN + DC - 1; ++k) cr[k] = rowR[k];
#pragma unroll
        for (int q = 0; q < DC / 2; ++q, ++s) {
            const int d = d0 + 2 * q;
            uint32_t* hb = hbuf[s & 1];
            // ---- row phase: packed |a-b| of (d, d+1) and horizontal s-sums
            uint32_t e[RUN + 2 * R];
#pragma unroll
            for (int i = 0; i < RUN; ++i) {
                const uint32_t c1 = __builtin_amdgcn_sad_u8(cl[i], cr[i - 2 * q + DC - 2], 0u);
                e[R + i] = __builtin_amdgcn_sad_u8(cl[i], cr[i - 2 * q + DC - 1], c1 << 16);
            }
#pragma unroll
            for (int i = 0; i < R; ++i) {
                e[i] = dpp_mov<DPP_ROW_SHR1>(e[R + RUN - R + i]);
                e[R + RUN + i] = dpp_mov<DPP_ROW_SHL1>(e[R + i]);
            }
            uint32_t h = 0;
#pragma unroll
            for (int i = 0; i <= 2 * R; ++i) h += e[i];
            uint2* hrow = reinterpret_cast<uint2*>(hb + hr * HSTRIDE + RUN * run);
#pragma unroll
            for (int i = 0; i < RUN; i += 2) {
                if (i > 0) h = slide(h, e[i - 1], e[i + 2 * R]);
                const uint32_t h0 = h;
                h = slide(h, e[i], e[i + 1 + 2 * R]);
                hrow[i / 2] = make_uint2(h0, h);
            }
            __syncthreads();

            // ---- column phase
            // centre column: sub-patch sums for rows gw*RG - SP .. gw*RG + RG - 1 + SP
            const uint32_t* hc = hb + (gw * RG) * HSTRIDE + lane + 8;
            uint32_t sc[NC];
            {
                uint32_t v = 0;
#pragma unroll
                for (int j = 0; j <= 2 * R; ++j) v += hc[j * HSTRIDE];
                sc[0] = v;
#pragma unroll
                for (int k = 1; k < NC; ++k) {
                    v = slide(v, hc[(k - 1) * HSTRIDE], hc[(k + 2 * R) * HSTRIDE]);
                    sc[k] = v;
                }
            }
            // left / right columns at distance SP: output rows only
            uint32_t sl[RG], sr[RG];
            {
                const uint32_t* hl = hc + SP * HSTRIDE - SP;
                const uint32_t* hq = hc + SP * HSTRIDE + SP;
                uint32_t vl = 0, vr = 0;
#pragma unroll
                for (int j = 0; j <= 2 * R; ++j) { vl += hl[j * HSTRIDE]; vr += hq[j * HSTRIDE]; }
                sl[0] = vl;
                sr[0] = vr;
#pragma unroll
                for (int k = 1; k < RG; ++k) {
                    vl = slide(vl, hl[(k - 1) * HSTRIDE], hl[(k + 2 * R) * HSTRIDE]);
                    vr = slide(vr, hq[(k - 1) * HSTRIDE], hq[(k + 2 * R) * HSTRIDE]);
                    sl[k] = vl;
                    sr[k] = vr;
                }
            }
            const uint32_t ddv = (uint32_t)(255 - d) | ((uint32_t)(254 - d) << 8);
            bool ok0 = true, ok1 = true;
            if (EDGE) { ok0 = lane_ok && d <= dlim; ok1 = lane_ok && d + 1 <= dlim; }
            uint32_t k0[RG], k1[RG];
#pragma unroll
            for (int o = 0; o < RG; ++o) {
                const uint32_t top = sc[o], ctr = sc[o + SP], bot = sc[o + 2 * SP];
                const uint32_t lo1 = pk_min(top, sl[o]), hi1 = pk_max(top, sl[o]);
                const uint32_t lo2 = pk_min(bot, sr[o]), hi2 = pk_max(bot, sr[o]);
                const uint32_t m1 = pk_min(lo1, lo2);
                const uint32_t m2 = pk_min(pk_max(lo1, lo2), pk_min(hi1, hi2));
                const uint32_t abar = 0xFFFFFFFFu - (ctr + m1 + m2);  // 65535 - AW2 in each half
                k0[o] = __builtin_amdgcn_perm(ddv, abar, 0x0C010004u);
                k1[o] = __builtin_amdgcn_perm(ddv, abar, 0x0C030205u);
                if (EDGE) { k0[o] = ok0 ? k0[o] : 0u; k1[o] = ok1 ? k1[o] : 0u; }
                if (BAND) {
                    const bool kp = (keep >> o) & 1u;
                    k0[o] = kp ? k0[o] : 0u;
                    k1[o] = kp ? k1[o] : 0u;
                }
                const uint32_t kk = k0[o] > k1[o] ? k0[o] : k1[o];
                lacc[o] = lacc[o] > kk ? lacc[o] : kk;
            }
            static_assert(RG == 12, "right-view chain written for 12 rows");
            max_shift_in12(racc, k0);
            if (lane == 63) {
#pragma unroll
                for (int i = 0; i < RG; ++i) eb[i * 64 + (d & 63)] = racc[i];
            }
            max_shift_in12(racc, k1);
            if (lane == 63) {
#pragma unroll
                for (int i = 0; i < RG; ++i) eb[i * 64 + ((d + 1) & 63)] = racc[i];
            }
            flush(d + 1);
        }
    }
    {
        // what is still in flight belongs to right pixels x0 + lane - (de - 1)
        const int xr = x0 + lane - (de - 1);
        if (xr >= 0 && xr < W) {
#pragma unroll
            for (int i = 0; i < RG; ++i)
                if (i < rows_here && racc[i]) atomicMax(&rp[(size_t)i * W + xr], racc[i]);
        }
    }
    if (x < W) {
        uint32_t* lk = p.lkey + (size_t)(y0 + gw * RG) * W + x;
#pragma unroll
        for (int i = 0; i < RG; ++i)
            if (i < rows_here && lacc[i]) atomicMax(&lk[(size_t)i * W], lacc[i]);
    }
}

template <int SP, bool BAND>
__global__ __launch_bounds__(256, 2) void aw_fused_kernel(AwParams p)
{
    __shared__ uint32_t hbuf[2][HROWS * HSTRIDE];
    __shared__ uint32_t ebuf[4][RG][64];
    // one-dimensional grid of ntx * nty * nseg workgroups, renumbered per XCD; the band
    // search keeps the hardware's round-robin over the XCDs, since its work sits in a few
    // segments and image regions
    const int nty = (p.H + TH - 1) / TH;
    const int w = BAND ? (int)blockIdx.x : xcd_remap(blockIdx.x, gridDim.x);
    const int tx = w % p.ntx, rest = w / p.ntx;
    const int ty = rest % nty, seg = rest / nty;
    const int x0 = tx * TW;
    const int de = min(p.D, (seg + 1) * p.seg_len);
    // no pixel of the tile has a candidate in the segment (left pixel x takes d when
    // bx <= x < W - bx and d <= x - bx): the job would only fold zero keys
    const int xmax = min(x0 + TW - 1, p.W - p.bx - 1);
    if (xmax < x0 || xmax - p.bx < seg * p.seg_len) return;
    const bool edge = (x0 < p.bx + de - 1) || (x0 + TW > p.W - p.bx);
    if (edge)
        aw_tile<SP, true, BAND>(p, hbuf, ebuf, tx, ty, seg);
    else
        aw_tile<SP, false, BAND>(p, hbuf, ebuf, tx, ty, seg);
}

}  // namespace aw

bool aw_supported(int sub, int D) { return (sub == 5 || sub == 3) && D >= 8 && D <= 256 && D % 8 == 0; }

void aw_grid(int W, int H, int D, int* ntx, int* nty, int* nseg, int* seg_len)
{
    *ntx = (W + aw::TW - 1) / aw::TW;
    *nty = (H + aw::TH - 1) / aw::TH;
    const int tiles = *ntx * *nty;
    int n = (1200 + tiles - 1) / tiles;
    if (n > D / 8) n = D / 8;
    if (n > BM_MAX_SEG) n = BM_MAX_SEG;
    if (n < 1) n = 1;
    int len = (D + n - 1) / n;
    len = (len + 7) & ~7;
    *seg_len = len;
    *nseg = (D + len - 1) / len;
}

hipError_t launch_aw(int sub, const AwParams& p, hipStream_t s)
{
    dim3 grid(p.ntx * ((p.H + aw::TH - 1) / aw::TH) * p.nseg);
    if (sub == 5) {
        if (p.bins)
            hipLaunchKernelGGL((aw::aw_fused_kernel<5, true>), grid, dim3(256), 0, s, p);
        else
            hipLaunchKernelGGL((aw::aw_fused_kernel<5, false>), grid, dim3(256), 0, s, p);
    } else {
        if (p.bins)
            hipLaunchKernelGGL((aw::aw_fused_kernel<3, true>), grid, dim3(256), 0, s, p);
        else
            hipLaunchKernelGGL((aw::aw_fused_kernel<3, false>), grid, dim3(256), 0, s, p);
    }
    return hipGetLastError();
}

}  // namespace stereo
