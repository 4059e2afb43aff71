// Normalised cross-correlation block matcher for gfx950.
//
// Reference: BlockMatcherNCC (CluTec.Cuda.MultiView/BlockMatcherNCC.cu): EvalMsd :338-373
// (base mean / std-dev), EvalSim :180-336 (ncc = (E[IJ] - muI muJ) / (sdI sdJ), zero when
// the norm is small, clamped at 0 before quantisation), strict '>' winner-take-all over d
// (:298).  The reference forms E[x^2] - mu^2 in float, which cancels catastrophically;
// here every box sum is an exact integer:
//     cov = N*SIJ - SI*SJ,  var = N*SS - S^2,  score = cov / sqrt(varI * varJ)
// and only the last step is FP32: score = float(cov) * rI * rJ with rI = 1/sqrt(varI)
// rounded once from FP64.  A window whose standard deviation is below one grey level
// (var < N^2) scores 0.  The winner key is ((floor(max(score,0) * 2^23)) << 8 | 255-d):
// the best score, lowest d on ties, the same key layout as the SAD / census matcher so
// the right-view merge and the check are shared.
//
// Kernel layout (one disparity per step): workgroup = 4 waves, tile = 64 columns x 52
// rows.  Row phase: lane = (H row 0..63, run of 19 columns), products I*J(x-d) and
// their horizontal sums over 2*AX+1 columns (halo from the neighbouring runs by DPP)
// into LDS (one buffer, two barriers per disparity).  Column phase: lane = column,
// wave = 13 rows: vertical running sums give SIJ, the right window's (SJ, rJ) come
// from an LDS tile of the right statistics that is refreshed every 16 disparities,
// the left window's are registers.  Left winner in
// a register, right winner carried along the wave by a DPP shift (see bm.hip).
#include "common.hpp"
#include "kernels.hpp"

namespace stereo {
namespace ncc {

constexpr int TW = 64;
constexpr int RG = 13;
constexpr int TH = 4 * RG;  // 52
constexpr int HROWS = 64;
constexpr int RUN = 19;
constexpr int COLS = 4 * RUN;  // 76 cost columns: tile columns [-6, 70)
constexpr int HSTRIDE = 76;
constexpr int DCH = 8;         // disparities per group (one band-search bin)
constexpr int SDC = 16;        // disparities per right-statistics tile
constexpr int EMIT = 32;       // right-view values parked per wave between flushes
constexpr int SCOLS = TW + SDC - 1;

// Per-pixel window statistics: .x = window sum S (int), .y = float bits of the
// reciprocal standard-deviation factor scale / sqrt(N SS - S^2) (0 when the
// variance is below N^2, and 0 outside the matchable area).  A workgroup does a
// 64 x 32 block: the input rows it needs go to LDS, horizontal sums of v and v^2
// over 2 ax + 1 columns go to LDS, and each lane slides the vertical sums down 8
// rows of one column.  Both images of a pair go through one launch (blockIdx.z).
constexpr int ST_W = 64, ST_H = 32, ST_MAXA = 6;

__global__ __launch_bounds__(256) void ncc_stats_kernel(View img0, int2* out0, float scale0, View img1, int2* out1,
                                                        float scale1, int pitch_px, int ax, int ay)
{
    const View img = blockIdx.z ? img1 : img0;
    int2* const out = blockIdx.z ? out1 : out0;
    const float scale = blockIdx.z ? scale1 : scale0;
    __shared__ uint8_t pix[ST_H + 2 * ST_MAXA][ST_W + 2 * ST_MAXA];
    __shared__ int hs[ST_H + 2 * ST_MAXA][ST_W];
    __shared__ int hss[ST_H + 2 * ST_MAXA][ST_W];
    const int x0 = blockIdx.x * ST_W, y0 = blockIdx.y * ST_H;
    const int rows = ST_H + 2 * ay, cols = ST_W + 2 * ax;
    {
        // lane = column (and column + 64 for the halo), wave = every fourth row; all
        // loads (clamped to the image) are issued before the first store
        constexpr int NR = (ST_H + 2 * ST_MAXA + 3) / 4;
        const int c = threadIdx.x & 63, r0 = threadIdx.x >> 6;
        const int xa = clampi(x0 - ax + c, 0, img.w - 1), xb = clampi(x0 - ax + 64 + c, 0, img.w - 1);
        uint8_t va[NR], vb[NR];
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const uint8_t* row = row_ptr<const uint8_t>(img, clampi(y0 - ay + r0 + 4 * i, 0, img.h - 1));
            va[i] = row[xa];
            vb[i] = row[xb];
        }
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const int r = r0 + 4 * i;
            if (r < rows) {
                pix[r][c] = va[i];
                if (64 + c < cols) pix[r][64 + c] = vb[i];
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < rows * ST_W; k += 256) {
        const int r = k / ST_W, c = k - r * ST_W;
        int a = 0, b = 0;
        for (int i = 0; i <= 2 * ax; ++i) {
            const int v = pix[r][c + i];
            a += v;
            b += v * v;
        }
        hs[r][c] = a;
        hss[r][c] = b;
    }
    __syncthreads();
    const int c = threadIdx.x & 63, rg = threadIdx.x >> 6;  // 4 groups of 8 rows
    const int x = x0 + c;
    const long long n = (long long)(2 * ax + 1) * (2 * ay + 1);
    int sv = 0, ssv = 0;
    for (int j = 0; j < 2 * ay; ++j) {
        sv += hs[rg * 8 + j][c];
        ssv += hss[rg * 8 + j][c];
    }
    for (int i = 0; i < 8; ++i) {
        const int r = rg * 8 + i;  // output row y0 + r uses LDS rows r .. r + 2 ay
        sv += hs[r + 2 * ay][c];
        ssv += hss[r + 2 * ay][c];
        const int y = y0 + r;
        if (x < img.w && y < img.h) {
            int2 o = make_int2(0, 0);
            if (x >= ax && x < img.w - ax && y >= ay && y < img.h - ay) {
                const long long var = n * (long long)ssv - (long long)sv * sv;
                float rf = 0.0f;
                if (var >= n * n) rf = (float)((double)scale / sqrt((double)var));
                o = make_int2(sv, __float_as_int(rf));
            }
            out[(size_t)y * pitch_px + x] = o;
        }
        sv -= hs[r][c];
        ssv -= hss[r][c];
    }
}

template <int AX, int AY, bool EDGE, bool BAND>
__device__ __forceinline__ void ncc_tile(const NccParams& p, uint32_t* hb, int16_t* ssum, float* srinv,
                                         uint32_t (*ebuf)[RG][EMIT], int tx, int ty, int seg)
{
    static_assert(TH + 2 * AY <= HROWS, "tile too tall for the window");
    static_assert(AX <= 6 && 2 * AX + 64 <= COLS, "window wider than the run layout");
    const int tid = threadIdx.x;
    const int x0 = tx * TW, y0 = ty * TH;
    const int W = p.W, H = p.H;
    const int ds = seg * p.seg_len;
    const int de = min(p.D, ds + p.seg_len);
    constexpr int N = (2 * AX + 1) * (2 * AY + 1);

    // row phase identity
    const int run = tid & 3;
    const int hr = tid >> 2;
    const int gyr = clampi(y0 - AY + hr, 0, H - 1);
    const int xs = x0 - 6 + RUN * run;
    const uint8_t* const rowL = row_ptr<const uint8_t>(p.L, gyr) + xs;
    const uint8_t* const rowR = row_ptr<const uint8_t>(p.R, gyr) + xs;
    uint32_t il[RUN];
#pragma unroll
    for (int i = 0; i < RUN; ++i) il[i] = rowL[i];

    // column phase identity
    const int lane = tid & 63;
    const int gw = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int x = x0 + lane;
    const int rows_here = H - (y0 + gw * RG);
    uint32_t* const rp = p.rpart + (((size_t)seg * p.ntx + tx) * H + y0 + gw * RG) * BM_NPART;
    uint32_t* const eb = &ebuf[gw][0][0];  // [RG][EMIT]
    const bool lane_ok = x >= p.bx && x < W - p.bx;
    const int dlim = x - p.bx;
    int nsi[RG];     // -SI of the left window
    float rfi[RG];   // 2^23 / sqrt(varI), or 0
    uint32_t lacc[RG], racc[RG];
#pragma unroll
    for (int i = 0; i < RG; ++i) {
        const int y = min(y0 + gw * RG + i, H - 1);
        const int2 st = p.lstats[(size_t)y * p.stats_pitch + min(x, W - 1)];
        nsi[i] = -st.x;
        rfi[i] = __int_as_float(st.y);
        lacc[i] = 0;
        racc[i] = 0;
    }
    // band search: the tile's rows lie in at most two 64 x BM_TILE_H bin tiles; rows before
    // ysplit take m0, the others m1 (uniform per wave and row); a group of disparities
    // whose bins neither mask holds is skipped
    uint32_t m0 = ~0u, m1 = ~0u;
    int ysplit = 1 << 30;
    if (BAND) {
        const int tb0 = y0 / BM_TILE_H, tb1 = min((y0 + TH - 1) / BM_TILE_H, p.bins_nty - 1);
        m0 = p.bins[tb0 * p.ntx + tx];
        m1 = p.bins[tb1 * p.ntx + tx];
        ysplit = (tb0 + 1) * BM_TILE_H;
        const int b0 = ds >> 3, b1 = (de - 1) >> 3;
        if (((m0 | m1) & (uint32_t)((2ull << b1) - (1ull << b0))) == 0u) {
            // nothing to search in this segment: every right-view partial the loop below
            // would write (slots 256 - de .. 318 - ds) and every left key is the empty key
            const int lane = tid & 63, gw = tid >> 6;
            const int rows_here = min(RG, H - (y0 + gw * RG));
            uint32_t* const rp = p.rpart + (((size_t)seg * p.ntx + tx) * H + y0 + gw * RG) * BM_NPART;
            for (int kk = 256 - de + lane; kk <= 318 - ds; kk += 64)
                for (int i = 0; i < rows_here; ++i) rp[i * BM_NPART + kk] = 0u;
            const int x = x0 + lane;
            if (x < W) {
                uint32_t* lk = p.lkey + (size_t)seg * W * H + (size_t)(y0 + gw * RG) * W + x;
                for (int i = 0; i < rows_here; ++i) lk[(size_t)i * W] = 0u;
            }
            return;
        }
    }

    // right pixels of the run for the current disparity d, byte i of rw = R(xs + i - d)
    // (little-endian over RWW words), and the one that enters next, R(xs - d - 1): one load
    // per disparity instead of RUN (every offset down to xs - de - 1 lies in the matcher's
    // left padding)
    constexpr int RWW = (RUN + 3) / 4;
    // after disparity d: at the end of a block of EMIT disparities (or of the segment) the
    // right-view values parked in the wave's LDS rows go to the partials
    auto flush = [&](int d) {
        if ((d & (EMIT - 1)) == EMIT - 1 || d == de - 1) {
            const int base = d & ~(EMIT - 1);
            if (lane <= (d & (EMIT - 1)) && base + lane >= ds) {
                const int kk = 318 - (base + lane);
#pragma unroll
                for (int i = 0; i < RG; ++i)
                    if (i < rows_here) rp[i * BM_NPART + kk] = eb[i * EMIT + lane];
            }
        }
    };
    uint32_t rw[RWW], rnext = 0;
    int wd = -1;  // the disparity rw holds
    int sdc = -SDC;  // the first disparity of the right-statistics tile in LDS (none yet)
    for (int dc = ds; dc < de; dc += DCH) {
        const int dce = min(de, dc + DCH);
        // the band search: rows of this wave that search the group (bit r: row r), wave-uniform
        uint32_t keep = (1u << RG) - 1u;
        if (BAND) {
            const uint32_t bit = 1u << (dc >> 3);  // the group's bin (dc is a multiple of 8)
            const uint32_t lo_rows = (1u << clampi(ysplit - (y0 + gw * RG), 0, RG)) - 1u;
            keep = ((m0 & bit) ? lo_rows : 0u) | ((m1 & bit) ? (((1u << RG) - 1u) & ~lo_rows) : 0u);
            if (((m0 | m1) & bit) == 0u) {
                // no pixel of the tile searches these 8 disparities: the right-view chain
                // moves on by 8 lanes with no new candidates.  The values that leave the
                // wave at lane 63 during those 8 steps sit in lanes 62 .. 55 now.
                if (lane >= 63 - DCH && lane < 63) {
                    const int slot = (dc + 62 - lane) & (EMIT - 1);
#pragma unroll
                    for (int i = 0; i < RG; ++i) eb[i * EMIT + slot] = racc[i];
                }
#pragma unroll
                for (int i = 0; i < RG; ++i) {
                    const uint32_t v = __shfl_up(racc[i], DCH);
                    racc[i] = lane >= DCH ? v : 0u;
                }
                flush(dce - 1);
                continue;
            }
        }
        // right statistics for disparities sdc .. sdc+SDC-1 (right columns x0-sdc-(SDC-1) ..
        // x0+63-sdc), rows of the tile, loaded when the group leaves the tile in LDS: the
        // thread's loads (the last index repeated past the end) are all issued before the
        // barrier, the LDS stores after it
        if (dc - sdc >= SDC) {
            sdc = dc;
            constexpr int SN = TH * SCOLS, SIT = (SN + 255) / 256, SB = SIT;
            static_assert(SIT % SB == 0, "whole batches");
            // an opaque copy of the thread index: the addresses are recomputed here rather
            // than hoisted out of the disparity loop into live registers
            int t0 = tid;
            asm volatile("" : "+v"(t0));
            auto sload = [&](int j) {
                const int k = min(t0 + 256 * j, SN - 1);
                const int r = k / SCOLS, c = k - r * SCOLS;
                const int y = min(y0 + r, H - 1);
                const int xc = clampi(x0 - sdc - (SDC - 1) + c, 0, W - 1);
                return p.rstats[(size_t)y * p.stats_pitch + xc];
            };
#pragma unroll
            for (int j0 = 0; j0 < SIT; j0 += SB) {
                int2 v[SB];
#pragma unroll
                for (int j = 0; j < SB; ++j) v[j] = sload(j0 + j);
                if (j0 == 0) __syncthreads();
#pragma unroll
                for (int j = 0; j < SB; ++j) {
                    const int k = tid + 256 * (j0 + j);
                    if (k < SN) {
                        ssum[k] = (int16_t)v[j].x;
                        srinv[k] = __int_as_float(v[j].y);
                    }
                }
            }
        }
        if (wd != dc) {
            uint32_t b[RUN];
#pragma unroll
            for (int i = 0; i < RUN; ++i) b[i] = rowR[i - dc];
            rnext = rowR[-dc - 1];
#pragma unroll
            for (int k = 0; k < RWW; ++k) {
                uint32_t w = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (4 * k + j < RUN) w |= b[4 * k + j] << (8 * j);
                rw[k] = w;
            }
        }
        for (int d = dc; d < dce; ++d) {
            // ---- row phase: products and horizontal sums for disparity d
            uint32_t e[RUN + 2 * AX];
#pragma unroll
            for (int i = 0; i < RUN; ++i) e[AX + i] = il[i] * ((rw[i >> 2] >> (8 * (i & 3))) & 0xffu);
#pragma unroll
            for (int k = RWW - 1; k > 0; --k) rw[k] = __builtin_amdgcn_alignbyte(rw[k], rw[k - 1], 3);
            rw[0] = (rw[0] << 8) | rnext;
            rnext = rowR[-d - 2];
            wd = d + 1;
            __syncthreads();  // the previous column phase has read the H buffer
#pragma unroll
            for (int i = 0; i < AX; ++i) {
                e[i] = dpp_mov<DPP_ROW_SHR1>(e[AX + RUN - AX + i]);
                e[AX + RUN + i] = dpp_mov<DPP_ROW_SHL1>(e[AX + i]);
            }
            uint32_t h = 0;
#pragma unroll
            for (int i = 0; i <= 2 * AX; ++i) h += e[i];
            uint32_t* hrow = hb + hr * HSTRIDE + RUN * run;
#pragma unroll
            for (int i = 0; i < RUN; ++i) {
                if (i > 0) h = slide(h, e[i - 1], e[i + 2 * AX]);
                hrow[i] = h;
            }
            __syncthreads();

            // ---- column phase
            uint32_t hv[RG + 2 * AY];
            const uint32_t* hcol = hb + gw * RG * HSTRIDE + lane + 6;
#pragma unroll
            for (int i = 0; i < RG + 2 * AY; ++i) hv[i] = hcol[i * HSTRIDE];
            uint32_t sij = 0;
#pragma unroll
            for (int i = 0; i <= 2 * AY; ++i) sij += hv[i];
            const int soff = (gw * RG) * SCOLS + lane + (SDC - 1) - (d - sdc);
            const int16_t* srow_s = ssum + soff;
            const float* srow_r = srinv + soff;
            const uint32_t dd = 255u - (uint32_t)d;
            bool ok = true;
            if (EDGE) ok = lane_ok && d <= dlim;
            uint32_t key[RG];
#pragma unroll
            for (int i = 0; i < RG; ++i) {
                if (i > 0) sij = slide(sij, hv[i - 1], hv[i + 2 * AY]);
                const int sj = srow_s[i * SCOLS];
                const int cov = __mul24(nsi[i], sj) + (int)__umul24((uint32_t)N, sij);
                const float sc = (float)cov * rfi[i] * srow_r[i * SCOLS];
                // v_cvt_u32_f32 truncates and clamps negative values to 0: the key's
                // floor(max(score, 0) * 2^23) in one instruction
                uint32_t q;
                asm("v_cvt_u32_f32 %0, %1" : "=v"(q) : "v"(sc));
                uint32_t k = (q << 8) | dd;
                if (EDGE) k = ok ? k : 0u;
                if (BAND) k = ((keep >> i) & 1u) ? k : 0u;
                key[i] = k;
                lacc[i] = lacc[i] > k ? lacc[i] : k;
            }
            static_assert(RG == 13, "right-view chain written for 13 rows");
            max_shift_in13(racc, key);
            if (lane == 63) {
#pragma unroll
                for (int i = 0; i < RG; ++i) eb[i * EMIT + (d & (EMIT - 1))] = racc[i];
            }
            flush(d);
        }
    }
    {
        const int kk = lane + 255 - (de - 1);
#pragma unroll
        for (int i = 0; i < RG; ++i)
            if (i < rows_here) rp[i * BM_NPART + kk] = racc[i];
    }
    if (x < W) {
        uint32_t* lk = p.lkey + (size_t)seg * W * H + (size_t)(y0 + gw * RG) * W + x;
#pragma unroll
        for (int i = 0; i < RG; ++i)
            if (i < rows_here) lk[(size_t)i * W] = lacc[i];
    }
}

// BAND: the band search's instantiation, a kernel of its own so that the full-range
// kernel keeps its register count (and three workgroups per CU)
template <int AX, int AY, bool BAND>
__global__ __launch_bounds__(256, 3) void ncc_fused_kernel(NccParams p)
{
    __shared__ uint32_t hbuf[HROWS * HSTRIDE];
    // right statistics of the tile: window sums (at most 121 x 255, an int16) and 1/sigma
    __shared__ int16_t ssum[TH * SCOLS];
    __shared__ float srinv[TH * SCOLS];
    __shared__ uint32_t ebuf[4][RG][EMIT];
    // one-dimensional grid of ntx * nty * nseg workgroups, renumbered per XCD
    const int nty = (p.H + TH - 1) / TH;
    // the band search keeps the hardware's round-robin over the XCDs: its work sits in a
    // few segments and image regions, and a contiguous range per XCD would load some
    // XCDs far more than others
    const int w = BAND ? (int)blockIdx.x : xcd_remap(blockIdx.x, gridDim.x);
    const int tx = w % p.ntx, rest = w / p.ntx;
    const int ty = rest % nty, seg = rest / nty;
    const int x0 = tx * TW;
    const int de = min(p.D, (seg + 1) * p.seg_len);
    const bool edge = (x0 < p.bx + de - 1) || (x0 + TW > p.W - p.bx);
    if (edge)
        ncc_tile<AX, AY, true, BAND>(p, hbuf, ssum, srinv, ebuf, tx, ty, seg);
    else
        ncc_tile<AX, AY, false, BAND>(p, hbuf, ssum, srinv, ebuf, tx, ty, seg);
}

// Winning score of every left pixel: the same FP32 arithmetic as the matcher, on a
// window sum computed directly for the winning disparity.
__global__ __launch_bounds__(256) void ncc_score_kernel(NccScoreParams f)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= f.W || y >= f.H) return;
    float out = __int_as_float(0x7fc00000);  // NaN: no match
    uint32_t lk = 0;
    if (x >= f.bx && x < f.W - f.bx && y >= f.by && y < f.H - f.by)
        for (int s = 0; s < f.nseg; ++s) {
            const uint32_t v = f.lkey[(size_t)s * f.W * f.H + (size_t)y * f.W + x];
            lk = v > lk ? v : lk;
        }
    const int d = key_disp(lk);
    if (d >= 0) {
        int sij = 0;
        for (int j = -f.ay; j <= f.ay; ++j) {
            const uint8_t* rl = row_ptr<const uint8_t>(f.L, y + j);
            const uint8_t* rr = row_ptr<const uint8_t>(f.R, y + j);
            for (int i = -f.ax; i <= f.ax; ++i) sij += (int)rl[x + i] * (int)rr[x + i - d];
        }
        const int2 a = f.lstats[(size_t)y * f.stats_pitch + x];
        const int2 b = f.rstats[(size_t)y * f.stats_pitch + x - d];
        const int n = (2 * f.ax + 1) * (2 * f.ay + 1);
        const int cov = n * sij - a.x * b.x;
        // lstats carry the 2^23 key scale; undo it exactly (power of two)
        out = (float)cov * (__int_as_float(a.y) * (1.0f / 8388608.0f)) * __int_as_float(b.y);
    }
    row_ptr<float>(f.score, y)[x] = out;
}

}  // namespace ncc

bool ncc_supported(int ww, int wh, int D)
{
    if (D < 8 || D > 256 || (D % 8) != 0) return false;
    return (ww == 11 && wh == 11) || (ww == 9 && wh == 9) || (ww == 5 && wh == 5) || (ww == 7 && wh == 7);
}

void ncc_grid(int W, int H, int D, int* ntx, int* nty, int* nseg, int* seg_len)
{
    *ntx = (W + ncc::TW - 1) / ncc::TW;
    *nty = (H + ncc::TH - 1) / ncc::TH;
    const int tiles = *ntx * *nty;
    int n = (1200 + tiles - 1) / tiles;
    if (n > D / 16) n = D / 16;
    if (n > BM_MAX_SEG) n = BM_MAX_SEG;
    if (n < 1) n = 1;
    int len = (D + n - 1) / n;
    len = (len + 15) & ~15;
    *seg_len = len;
    *nseg = (D + len - 1) / len;
}

hipError_t launch_ncc_stats(const View& img, int2* out, int pitch_px, int ww, int wh, float scale, hipStream_t s)
{
    if (ww / 2 > ncc::ST_MAXA || wh / 2 > ncc::ST_MAXA) return hipErrorInvalidValue;
    dim3 grid((img.w + ncc::ST_W - 1) / ncc::ST_W, (img.h + ncc::ST_H - 1) / ncc::ST_H, 1);
    hipLaunchKernelGGL(ncc::ncc_stats_kernel, grid, dim3(256), 0, s, img, out, scale, img, out, scale, pitch_px, ww / 2,
                       wh / 2);
    return hipGetLastError();
}

hipError_t launch_ncc_stats_pair(const View& img0, int2* out0, float scale0, const View& img1, int2* out1, float scale1,
                                 int pitch_px, int ww, int wh, hipStream_t s)
{
    if (ww / 2 > ncc::ST_MAXA || wh / 2 > ncc::ST_MAXA || img0.w != img1.w || img0.h != img1.h)
        return hipErrorInvalidValue;
    dim3 grid((img0.w + ncc::ST_W - 1) / ncc::ST_W, (img0.h + ncc::ST_H - 1) / ncc::ST_H, 2);
    hipLaunchKernelGGL(ncc::ncc_stats_kernel, grid, dim3(256), 0, s, img0, out0, scale0, img1, out1, scale1, pitch_px,
                       ww / 2, wh / 2);
    return hipGetLastError();
}

hipError_t launch_ncc(int ww, const NccParams& p, hipStream_t s)
{
    dim3 grid(p.ntx * ((p.H + ncc::TH - 1) / ncc::TH) * p.nseg);
#define NCC_LAUNCH(A)                                                                          \
    if (p.bins)                                                                                   \
        hipLaunchKernelGGL((ncc::ncc_fused_kernel<A, A, true>), grid, dim3(256), 0, s, p);         \
    else                                                                                          \
        hipLaunchKernelGGL((ncc::ncc_fused_kernel<A, A, false>), grid, dim3(256), 0, s, p);        \
    break
    switch (ww) {
    case 11: NCC_LAUNCH(5);
    case 9: NCC_LAUNCH(4);
    case 7: NCC_LAUNCH(3);
    default: NCC_LAUNCH(2);
    }
#undef NCC_LAUNCH
    return hipGetLastError();
}

hipError_t launch_ncc_score(const NccScoreParams& f, hipStream_t s)
{
    dim3 grid((f.W + 63) / 64, (f.H + 3) / 4);
    hipLaunchKernelGGL(ncc::ncc_score_kernel, grid, dim3(256), 0, s, f);
    return hipGetLastError();
}

}  // namespace stereo
