// The reference's BlockMatcherNCC and BlockMatcherAW1 operators, with their block-tiled
// results: the same disparities, test patches, decisions and four output channels as
// the reference's 32-thread blocks (CluTec.Cuda.MultiView/BlockMatcherNCC.cu:614-1255,
// BlockMatcherAW1.cu:657-1263).
//
// Both operators launch one 32-thread block per tile of result origins, full blocks only
// (EvalThreadConfig_ReadDepBlockSize, KernelDriverBase.cpp:166-197).  A block decides
// which disparities its results see: groups of 16 from the range start while the group
// start is below min(Max, MaxDisparity(X0)); a group whose test area lies inside the
// right image evaluates 16; otherwise, when the block's rows fit, a narrower group at the
// reference's placement (:789-876).  Each result follows the maximum of its similarity
// with the values around it, and is written with a code and three channels.
//   NCC: tile (33 - P) / 2 * 2 x 16 from a 33 x (15 + P) read area.  The similarity is
//        the reference's float formula over exact integer window sums -- (SIJ / N -
//        mean_I mean_J) / (sd_I sd_J), 0 where sd_I sd_J < 1 -- clamped at 0 and
//        truncated to a byte (EvalSim :180-336, EvalMsd :338-373); the gate is the mean
//        |L(x) - L(x + 1)| (EvalGrad :388-410).
//   AW1: tile 18 x 6 (sub-patch 5) or 24 x 10 (sub-patch 3) from a 32 x 20 / 32 x 18 read
//        area; the similarity is the centre sub-patch plus the two largest edge
//        neighbours of 255 - |a - b| or the signed product (EvalSimAW :246-316); the gate is
//        the window's standard deviation, truncated to an int, in % (EvalGrad :331-417).
// Floating point is evaluated as written: contraction off, correctly rounded division and
// square root, so the byte scores and 16-bit channels equal the oracle's bit for bit.
//
// Layout on gfx950: one wave per workgroup.  NCC: a wave takes two vertically adjacent
// reference blocks of one block column (they share their groups), and lane 32 b + t plays
// the reference thread t of block b, holding the same CG results.  Per group of 16
// disparities the right rows sit in LDS with the right windows' mean and deviation; per
// disparity each lane forms one column's running 16-row sums of L R into LDS, then
// its results' running row sums and the score.  AW1: a wave takes one reference block;
// per disparity the column sums, then the sub-patch sums go through LDS.  The band search
// around an initial map (MatcherUseDisparity) runs one block per wave.
#include <hip/hip_runtime.h>

#include "capi_internal.hpp"

namespace stereo {
namespace {

constexpr int R_SDC = 16;
constexpr int R_FIRST = 1, R_LAST = 0xFFEF, R_CANNOT = 0xFFFF, R_SATURATED = 0xFFFE, R_NOT_FOUND = 0xFFFD,
              R_NOT_SPECIFIC = 0xFFFC;

struct RefArgs {
    View L, R, out, init;  // init.ptr == nullptr: the plain matcher
    int dmin, dmax, offset, ltor;
    int imin_delta;  // NCC iMinDeltaThresh
    float min_delta, sim_thresh, grad_thresh;
    int nby;  // reference block rows
};

// A group of disparities: labels D0 .. D0 + n - 1; the test origin of result column 0
// for label D0 + r is base - r (left to right) or base + r; the right columns used are
// [lo, lo + n + TBX - 1).  X-only: a block whose rows do not fit evaluates nothing.
struct RefGroup {
    int n, base, lo;
};

template <int TBX>
__device__ __forceinline__ RefGroup ref_group(int X0, int D0, int W, int offset, int ltor, int cap)
{
    RefGroup g;
    const int tx = (ltor ? X0 - D0 - offset : X0 + D0 + offset) - ltor * (R_SDC - 1);
    if (tx >= 0 && tx + TBX + R_SDC - 1 <= W) {
        g.n = cap;
        g.base = tx + ltor * (R_SDC - 1);
    } else {
        int dx, sx;
        if (tx >= 0) {
            dx = W - tx - TBX;
            sx = tx;
        } else {
            dx = tx + (R_SDC - 1);
            sx = 0;
        }
        g.n = dx >= 0 ? dx + 1 : 0;
        g.base = sx + ltor * (g.n - 1);
    }
    g.lo = ltor ? g.base - (g.n - 1) : g.base;
    return g;
}

// The reference thread's initial value of one written result in the band search
// (:960-1001): the initial id at the result's centre over the scale (outside the map:
// Unknown), scaled in 16 bits when it is an id, as it is when it is a code.
__device__ __forceinline__ int ref_init_value(const View& init, int scale, int cx, int cy)
{
    const int px = cx / scale, py = cy / scale;
    const int id = (px < init.w && py < init.h) ? row_ptr<const ushort4>(init, py)[px].x : 0;
    return id <= R_LAST ? (scale * id) & 0xFFFF : id;
}

// MinDisp / MaxDisp over the 32 lanes of a half-wave (:418-484): each lane's plain min / max
// from (Last, First) counts only when valid; lo > hi or none valid: no range.
__device__ __forceinline__ bool ref_block_range(int tlo, int thi, int& lo, int& hi)
{
    int l = (tlo >= R_FIRST && tlo <= R_LAST) ? tlo : 0x7FFFFFFF;
    int h = (thi >= R_FIRST && thi <= R_LAST) ? thi : -1;
#pragma unroll
    for (int m = 16; m >= 1; m >>= 1) {
        l = min(l, __shfl_xor(l, m, 32));
        h = max(h, __shfl_xor(h, m, 32));
    }
    lo = l;
    hi = h;
    return l != 0x7FFFFFFF && h >= 0 && l <= h;
}

struct Tracker {
    int disp, mx, prev, prev2, nxt, nxt2, last;
};

// Copies the ROWS x COLS bytes of img at (x0, y0) into dst[r][c] (row pitch DP), zero
// outside the image or at columns c >= cmax, with NT threads: a thread issues its loads
// FB at a time before storing them, so that FB are in flight together (a load-then-store
// loop keeps one).
template <int ROWS, int COLS, int DP, int NT>
__device__ __forceinline__ void fill_window(uint8_t* dst, const View& img, int x0, int y0, int cmax, int tid_in)
{
    constexpr int N = ROWS * COLS, K = (N + NT - 1) / NT, FB = 8;
    // an opaque copy of the thread index: the element positions derived from it are
    // recomputed per call instead of being hoisted out of the caller's loops into
    // registers that stay live through them
    int tid = tid_in;
    asm volatile("" : "+v"(tid));
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += FB) {
        uint32_t v[FB];
#pragma unroll
        for (int k = 0; k < FB; ++k) {
            const int i = tid + NT * (k0 + k);
            const int r = i / COLS, c = i - r * COLS;
            const int x = x0 + c, y = y0 + r;
            v[k] = (k0 + k < K && i < N && c < cmax && x >= 0 && x < img.w && y >= 0 && y < img.h)
                       ? row_ptr<const uint8_t>(img, y)[x]
                       : 0u;
        }
#pragma unroll
        for (int k = 0; k < FB; ++k) {
            const int i = tid + NT * (k0 + k);
            if (k0 + k < K && i < N) {
                const int r = i / COLS, c = i - r * COLS;
                dst[r * DP + c] = (uint8_t)v[k];
            }
        }
    }
}

// The peak update (NCC EvalSim :280-333, AW1 EvalSimAW :262-315), branch-free.
__device__ __forceinline__ void track(Tracker& s, int v, int d, bool upd)
{
    const bool act = upd && s.disp <= R_LAST;
    const bool above = act && v > s.mx;
    const bool b1 = act && !above && s.disp + 1 == d + R_FIRST;
    const bool b2 = act && !above && s.disp + 2 == d + R_FIRST;
    s.prev2 = above ? s.prev : s.prev2;
    s.prev = above ? s.last : s.prev;
    s.mx = above ? v : s.mx;
    s.disp = above ? d + R_FIRST : s.disp;
    s.nxt = (above || b1) ? v : s.nxt;
    s.nxt2 = (above || b2) ? v : s.nxt2;
    s.last = act ? v : s.last;
}

// ------------------------------------------------------------------------------------
// NCC
// ------------------------------------------------------------------------------------

// x / N correctly rounded, for an integer 0 <= x <= N * 65025 (a window sum of bytes or
// of products of two bytes over N = P * P pixels): the product with the rounded reciprocal
// and one residual step, x - q0 N being exact in the fused multiply-add.  Equal to the
// IEEE division for every such x and P = 5, 7, 9, 11, 15 (checked exhaustively on the CPU).
template <int N>
__device__ __forceinline__ float div_window(int x)
{
    constexpr float R = 1.0f / (float)N;
    const float xf = (float)x;
    const float q0 = xf * R;
    const float r = __builtin_fmaf(-q0, (float)N, xf);
    return __builtin_fmaf(r, R, q0);
}

template <int P>
struct NccShape {
    static constexpr int CG = (32 - P + 1) / 2;
    static constexpr int PCX = 2 * CG;
    static constexpr int TBX = 33;
    static constexpr int TBY = 16 + P - 1;
    static constexpr int TW = TBX + R_SDC - 1;     // 48 test columns
    static constexpr int NC = R_SDC + PCX - 1;     // right window starts a group can use
};

// NB reference blocks per wave: 2 for the matcher, 1 for the band search.
template <int P, int NB>
__global__ __launch_bounds__(64) void ncc_ref_kernel(RefArgs a)
{
#pragma clang fp contract(off)
    using S = NccShape<P>;
    constexpr int CG = S::CG, PCX = S::PCX, TW = S::TW, NC = S::NC;
    constexpr int ROWS = 16 * NB, DR = ROWS + P - 1;

    // The left pixels are needed only until the left statistics and the lanes' column
    // registers are formed; the right windows' statistics take their place.  Column sums:
    // of L R per disparity, of the right windows' sums and squares per group, and (at the
    // start) of the left windows and gradients (columns 32 .. 63), one plane used in turn;
    // rows of 66 words (2 mod 32), so that the 16 rows and two column groups of a
    // half-wave's row-sum reads fall in 32 different banks.
    constexpr int SC = 66;
    struct Lds {
        union {
            uint8_t sl[DR][S::TBX];
            struct {
                float tmean[ROWS][NC + 1], tsd[ROWS][NC + 1];
            } st;
        } u;
        uint8_t sr[DR][TW];
        int scol[ROWS][SC];
    };
    __shared__ Lds lds;
    auto& sl = lds.u.sl;
    auto& sr = lds.sr;
    auto& tmean = lds.u.st.tmean;
    auto& tsd = lds.u.st.tsd;
    auto& scol = lds.scol;

    const int l = threadIdx.x;
    const int W = a.L.w, H = a.L.h;
    const int X0 = blockIdx.x * PCX;
    const int b = l >> 5, t = l & 31;  // reference block and thread
    const int by = blockIdx.y * NB + b;
    const int Y0 = by * 16;
    const int Y0s = blockIdx.y * NB * 16;
    const bool lane_ok = b < NB;
    const bool block_ok = lane_ok && by < a.nby && X0 + P <= W && Y0 + P <= H;
    const bool rows_fit = Y0 + S::TBY <= H;
    // this lane's results: row iy, columns ix0 .. ix0 + CG - 1 (AlgoArraySum_W32H16)
    const int iy = t & 15, ix0 = (t >> 4) * CG;
    const int row = 16 * b + iy;  // in the strip

    fill_window<DR, S::TBX, S::TBX, 64>(&sl[0][0], a.L, X0, Y0s, S::TBX, l);
    __syncthreads();

    bool written[CG];
    Tracker tr[CG];
    float bmean[CG], bsd[CG];
#pragma unroll
    for (int k = 0; k < CG; ++k) {
        written[k] = block_ok && X0 + ix0 + k + P < W && Y0 + iy + P < H;
        tr[k] = Tracker{0, 0, 0, 0, 0, 0, 0};
    }

    // left window sums (EvalMsd) and, for the matcher, the gradient sums (EvalGrad): the
    // sums and gradients first, then the sums of squares in the same plane
    bool cannot[CG];
    {
        const int c = l & 31, r0 = 16 * (lane_ok ? b : 0);
        int s1 = 0, sg = 0;
        for (int j = 0; j < P; ++j) {
            const int v = sl[r0 + j][c], w = sl[r0 + j][c + 1];
            s1 += v;
            sg += abs(v - w);
        }
        if (lane_ok) {
            scol[r0][c] = s1;
            scol[r0][c + 32] = sg;
        }
        for (int j = 1; j < 16; ++j) {
            const int vo = sl[r0 + j - 1][c], wo = sl[r0 + j - 1][c + 1];
            const int vn = sl[r0 + j + P - 1][c], wn = sl[r0 + j + P - 1][c + 1];
            s1 += vn - vo;
            sg += abs(vn - wn) - abs(vo - wo);
            if (lane_ok) {
                scol[r0 + j][c] = s1;
                scol[r0 + j][c + 32] = sg;
            }
        }
    }
    __syncthreads();
    {
        int s1 = 0, sg = 0;
        const int rr = lane_ok ? row : iy;
#pragma unroll
        for (int i = 0; i < P; ++i) {
            s1 += scol[rr][ix0 + i];
            sg += scol[rr][ix0 + i + 32];
        }
#pragma unroll
        for (int k = 0; k < CG; ++k) {
            if (k > 0) {
                s1 += scol[rr][ix0 + k + P - 1] - scol[rr][ix0 + k - 1];
                sg += scol[rr][ix0 + k + P - 1 + 32] - scol[rr][ix0 + k - 1 + 32];
            }
            bmean[k] = div_window<P * P>(s1);
            cannot[k] = div_window<P * P>(sg) < a.grad_thresh;
        }
    }
    __syncthreads();
    {
        const int c = l & 31, r0 = 16 * (lane_ok ? b : 0);
        int s2 = 0;
        for (int j = 0; j < P; ++j) {
            const int v = sl[r0 + j][c];
            s2 += v * v;
        }
        if (lane_ok) scol[r0][c] = s2;
        for (int j = 1; j < 16; ++j) {
            const int vo = sl[r0 + j - 1][c], vn = sl[r0 + j + P - 1][c];
            s2 += vn * vn - vo * vo;
            if (lane_ok) scol[r0 + j][c] = s2;
        }
    }
    __syncthreads();
    {
        int s2 = 0;
        const int rr = lane_ok ? row : iy;
#pragma unroll
        for (int i = 0; i < P; ++i) s2 += scol[rr][ix0 + i];
#pragma unroll
        for (int k = 0; k < CG; ++k) {
            if (k > 0) s2 += scol[rr][ix0 + k + P - 1] - scol[rr][ix0 + k - 1];
            const float m = bmean[k], m2 = div_window<P * P>(s2);
            const float sq = m * m;
            bsd[k] = sqrtf(m2 - sq);
        }
    }

    const bool band = NB == 1 && a.init.ptr != nullptr;
    int d_start, d_stop, cap;
    bool eval_block;  // the block evaluates disparities
    if (!band) {
        // the early-out of a block whose every result is CannotEvaluate (:741-775)
        bool any = false;
#pragma unroll
        for (int k = 0; k < CG; ++k) {
            tr[k].disp = cannot[k] ? R_CANNOT : 0;
            any = any || !cannot[k];
        }
        const uint64_t bal = __ballot(any);
        eval_block = ((bal >> (32 * b)) & 0xFFFFFFFFull) != 0;
        const int maxd = a.ltor ? X0 - a.offset + 1 : W - X0 - a.offset;  // MaxDisparity
        d_start = a.dmin;
        d_stop = a.dmax < maxd ? a.dmax : maxd;
        cap = R_SDC;
    } else {
        const int scale = W / a.init.w;
        int tlo = R_LAST, thi = R_FIRST;
#pragma unroll
        for (int k = 0; k < CG; ++k) {
            if (!written[k]) continue;
            const int v = ref_init_value(a.init, scale, X0 + ix0 + k + P / 2, Y0 + iy + P / 2);
            tr[k].disp = v;
            if (v <= R_LAST) {
                tlo = min(tlo, v);
                thi = max(thi, v);
            }
        }
        int lo, hi;
        // lanes 0 .. 31 hold the block's reference threads; every lane takes their range,
        // so that the whole wave goes on (or stops) together and shares the group loads
        const bool valid = ref_block_range(tlo, thi, lo, hi);
        lo = __builtin_amdgcn_readfirstlane(lo);
        hi = __builtin_amdgcn_readfirstlane(hi);
        if (!__builtin_amdgcn_readfirstlane((int)valid)) {
            // no valid range: every written result keeps its initial value (:1003-1027)
#pragma unroll
            for (int k = 0; k < CG; ++k)
                if (written[k])
                    row_ptr<ushort4>(a.out, Y0 + iy + P / 2)[X0 + ix0 + k + P / 2] =
                        make_ushort4((uint16_t)tr[k].disp, 0, 0, 0);
            return;
        }
        d_start = lo < 4 ? 0 : lo - 4;
        const int d_end = hi + 4;  // inclusive
        eval_block = d_end <= 32767;  // the reference's 16-bit loop variables wrap beyond
        d_stop = d_end + 1;
        cap = min(R_SDC, d_end - d_start + 1);
    }
    const bool upd = block_ok && rows_fit && eval_block;

    // this lane's column for the per-disparity sums: the left pixels of its 16 rows
    uint32_t lreg[16 + P - 1];
    {
        const int c = l & 31, r0 = 16 * (lane_ok ? b : 0);
#pragma unroll
        for (int j = 0; j < 16 + P - 1; ++j) lreg[j] = sl[r0 + j][c];
    }

    const bool wave_evaluates = __ballot(upd) != 0;
    for (int D0 = d_start; wave_evaluates && D0 < d_stop; D0 += R_SDC) {
        const RefGroup g = ref_group<S::TBX>(X0, D0, W, a.offset, a.ltor, cap);
        if (g.n <= 0) continue;
        __syncthreads();  // the previous group's readers of sr, scol and the statistics are done
        fill_window<DR, TW, TW, 64>(&sr[0][0], a.R, g.lo, Y0s, g.n + S::TBX - 1, l);
        __syncthreads();
        // the right windows' statistics at every start column the group uses: the means
        // from the column sums, then the deviations from the column sums of squares (a
        // thread finishes the same columns in both passes, so it reads its own means)
        for (int pass = 0; pass < 2; ++pass) {
            if (pass) __syncthreads();  // the means' readers of the plane are done
            for (int i = l; i < NB * TW; i += 64) {
                const int c = i % TW, r0 = 16 * (i / TW);
                int s = 0;
                for (int j = 0; j < P; ++j) {
                    const int v = sr[r0 + j][c];
                    s += pass ? v * v : v;
                }
                scol[r0][c] = s;
                for (int j = 1; j < 16; ++j) {
                    const int vo = sr[r0 + j - 1][c], vn = sr[r0 + j + P - 1][c];
                    s += pass ? vn * vn - vo * vo : vn - vo;
                    scol[r0 + j][c] = s;
                }
            }
            __syncthreads();
            for (int i = l; i < ROWS * 2; i += 64) {
                const int rr = i >> 1, c0 = (i & 1) * ((NC + 1) / 2);
                const int c1 = min(NC, c0 + (NC + 1) / 2);
                int s = 0;
                for (int k = 0; k < P; ++k) s += scol[rr][c0 + k];
                for (int c = c0; c < c1; ++c) {
                    if (c > c0) s += scol[rr][c + P - 1] - scol[rr][c - 1];
                    if (!pass) {
                        tmean[rr][c] = div_window<P * P>(s);
                    } else {
                        const float m = tmean[rr][c], m2 = div_window<P * P>(s);
                        const float sq = m * m;
                        tsd[rr][c] = sqrtf(m2 - sq);
                    }
                }
            }
        }
        for (int r = 0; r < g.n; ++r) {
            const int d = D0 + r;
            const int sh = (a.ltor ? g.base - r : g.base + r) - g.lo;  // test origin column in sr
            __syncthreads();  // the previous disparity's row sums have read scol
            // column sums of L R over this lane's 16 rows
            {
                const int c = l & 31, r0 = 16 * (lane_ok ? b : 0);
                int q[16 + P - 1];
#pragma unroll
                for (int j = 0; j < 16 + P - 1; ++j) q[j] = (int)(lreg[j] * (uint32_t)sr[r0 + j][c + sh]);
                int s = 0;
#pragma unroll
                for (int j = 0; j < P; ++j) s += q[j];
                if (lane_ok) scol[r0][c] = s;
#pragma unroll
                for (int j = 1; j < 16; ++j) {
                    s = (int)slide((uint32_t)s, (uint32_t)q[j - 1], (uint32_t)q[j + P - 1]);
                    if (lane_ok) scol[r0 + j][c] = s;
                }
            }
            __syncthreads();
            // row sums and the score of each result (EvalSim's result lambda)
            {
                const int rr = lane_ok ? row : iy;
                const int* const cr = scol[rr];
                int s3 = 0;
#pragma unroll
                for (int i = 0; i < P; ++i) s3 += cr[ix0 + i];
#pragma unroll
                for (int k = 0; k < CG; ++k) {
                    if (k > 0) s3 = (int)slide((uint32_t)s3, (uint32_t)cr[ix0 + k - 1], (uint32_t)cr[ix0 + k + P - 1]);
                    const float cross = div_window<P * P>(s3);
                    const float tm = tmean[rr][sh + ix0 + k], ts = tsd[rr][sh + ix0 + k];
                    const float norm = bsd[k] * ts;
                    float ncc;
                    if (norm < 1.0f) {
                        ncc = 0.0f;
                    } else {
                        const float mm = bmean[k] * tm;
                        ncc = (cross - mm) / norm;
                    }
                    ncc = fmaxf(0.0f, ncc);
                    const int u = (int)fminf(255.0f, ncc * 255.0f);
                    track(tr[k], u, d, upd);
                }
            }
        }
    }

    // WriteDisparity (:491-592)
#pragma unroll
    for (int k = 0; k < CG; ++k) {
        if (!written[k]) continue;
        const Tracker& s = tr[k];
        ushort4 o;
        const int dv = s.disp;
        if (dv > R_LAST || dv == 0) {
            o = make_ushort4((uint16_t)dv, 0, 0, 0);
        } else {
            const float fmax = (float)s.mx / 255.0f;
            const int diff1 = max(abs(s.prev - s.mx), abs(s.prev2 - s.mx));
            const int diff2 = max(abs(s.nxt - s.mx), abs(s.nxt2 - s.mx));
            int code;
            if (fmax >= a.sim_thresh) {
                const bool sat = dv <= R_FIRST + a.dmin + 1 || dv >= R_FIRST + a.dmax - 1;
                code = sat ? R_SATURATED : ((diff1 >= a.imin_delta && diff2 >= a.imin_delta) ? dv : R_NOT_SPECIFIC);
            } else {
                code = R_NOT_FOUND;
            }
            o = make_ushort4((uint16_t)code, (uint16_t)s.mx, (uint16_t)diff1, (uint16_t)diff2);
        }
        row_ptr<ushort4>(a.out, Y0 + iy + P / 2)[X0 + ix0 + k + P / 2] = o;
    }
}

// ------------------------------------------------------------------------------------
// AW1
// ------------------------------------------------------------------------------------

template <int SP>
struct AwShape {
    static constexpr int CG = (32 - SP + 1) / 2;
    static constexpr int SPCX = 2 * CG;              // sub-patch columns
    static constexpr int P3 = 3 * SP;                // window
    static constexpr int PCX = SPCX - 2 * SP;        // result columns
    static constexpr int PCY = 16 - 2 * SP;          // result rows
    static constexpr int TBX = 32;
    static constexpr int TBY = PCY + P3 - 1;
    static constexpr int TW = TBX + R_SDC - 1;       // 47 test columns
    static constexpr int NR = PCX * PCY;
    static constexpr int RPL = (NR + 63) / 64;       // results per lane
    static constexpr int NE = 3 * SP * SP;           // EffectivePatchElementCount
};

// SimValueToFloat (:155-159) and SimFloatToInteger (:172-176)
__device__ __forceinline__ float aw1_sim_float(int v, int ne)
{
#pragma clang fp contract(off)
    return ((float)v / (float)ne) * (100.0f / 255.0f);
}

__device__ __forceinline__ uint32_t aw1_sim_u16(float f)
{
#pragma clang fp contract(off)
    float t = (f / 100.0f + 1.0f) / 2.0f;
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return (uint32_t)(t * 65535.0f);
}

template <int SP, int SIM>
__global__ __launch_bounds__(64) void aw1_ref_kernel(RefArgs a)
{
#pragma clang fp contract(off)
    using S = AwShape<SP>;
    constexpr int PCX = S::PCX, PCY = S::PCY, TBY = S::TBY, TW = S::TW, NR = S::NR, RPL = S::RPL;

    __shared__ uint8_t sl[TBY][32];
    __shared__ uint8_t sr[TBY][TW + 1];
    __shared__ int scol[16][33];
    __shared__ int ssp[16][33];
    __shared__ float stab[256];

    const int l = threadIdx.x;
    const int W = a.L.w, H = a.L.h;
    const int X0 = blockIdx.x * PCX, Y0 = blockIdx.y * PCY;
    const bool block_ok = (int)blockIdx.y < a.nby && X0 + S::P3 <= W && Y0 + S::P3 <= H;
    if (!block_ok) return;  // uniform over the wave
    const bool rows_fit = Y0 + TBY <= H;

    if (SIM == 1)
        for (int i = l; i < 256; i += 64) stab[i] = 2.0f * ((float)i / 255.0f - 0.5f);
    for (int i = l; i < TBY * 32; i += 64) {
        const int r = i / 32, c = i % 32;
        const int x = X0 + c, y = Y0 + r;
        sl[r][c] = (x < W && y < H) ? row_ptr<const uint8_t>(a.L, y)[x] : 0;
    }
    __syncthreads();

    // this lane's results: row-major positions l + 64 q of the block's PCX x PCY
    int rx[RPL], ry[RPL];
    bool has[RPL], written[RPL];
    Tracker tr[RPL];
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
        const int pos = l + 64 * q;
        has[q] = pos < NR;
        rx[q] = has[q] ? pos % PCX : 0;
        ry[q] = has[q] ? pos / PCX : 0;
        written[q] = has[q] && X0 + rx[q] + S::P3 < W && Y0 + ry[q] + S::P3 < H;
        tr[q] = Tracker{0, -32767, 0, 0, 0, 0, 0};
    }

    const bool band = a.init.ptr != nullptr;
    int d_start, d_stop, cap;
    bool eval_block;
    if (!band) {
        // EvalGrad: the P3 x P3 window's standard deviation, truncated, in % of 255
        if (l < 32) {
            int s1 = 0, s2 = 0;
            for (int j = 0; j < S::P3; ++j) {
                const int v = sl[j][l];
                s1 += v;
                s2 += v * v;
            }
            scol[0][l] = s1;
            ssp[0][l] = s2;
            for (int j = 1; j < PCY; ++j) {
                const int vo = sl[j - 1][l], vn = sl[j + S::P3 - 1][l];
                s1 += vn - vo;
                s2 += vn * vn - vo * vo;
                scol[j][l] = s1;
                ssp[j][l] = s2;
            }
        }
        __syncthreads();
        bool any = false;
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            if (!has[q]) continue;
            int s1 = 0, s2 = 0;
            for (int i = 0; i < S::P3; ++i) {
                s1 += scol[ry[q]][rx[q] + i];
                s2 += ssp[ry[q]][rx[q] + i];
            }
            const float m = (float)s1 / (float)(S::P3 * S::P3), m2 = (float)s2 / (float)(S::P3 * S::P3);
            const float sq = m * m;
            const int isd = (int)sqrtf(fabsf(m2 - sq));
            const float sd = (float)isd / 255.0f * 100.0f;
            if (sd < a.grad_thresh) tr[q].disp = R_CANNOT;
            else any = true;
        }
        eval_block = __ballot(any) != 0;
        const int maxd = a.ltor ? X0 - a.offset + 1 : W - X0 - a.offset;
        d_start = a.dmin;
        d_stop = a.dmax < maxd ? a.dmax : maxd;
        cap = R_SDC;
    } else {
        // reference thread t holds the positions 32 r + t: lanes t and t + 32 together
        const int scale = W / a.init.w;
        int tlo = R_LAST, thi = R_FIRST;
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            if (!written[q]) continue;
            const int v = ref_init_value(a.init, scale, X0 + rx[q] + S::P3 / 2, Y0 + ry[q] + S::P3 / 2);
            tr[q].disp = v;
            if (v <= R_LAST) {
                tlo = min(tlo, v);
                thi = max(thi, v);
            }
        }
        tlo = min(tlo, __shfl_xor(tlo, 32));
        thi = max(thi, __shfl_xor(thi, 32));
        int lo, hi;
        if (!ref_block_range(tlo, thi, lo, hi)) {
#pragma unroll
            for (int q = 0; q < RPL; ++q)
                if (written[q])
                    row_ptr<ushort4>(a.out, Y0 + ry[q] + S::P3 / 2)[X0 + rx[q] + S::P3 / 2] =
                        make_ushort4((uint16_t)tr[q].disp, 0, 0, 0);
            return;
        }
        d_start = lo < 4 ? 0 : lo - 4;
        const int d_end = hi + 4;
        eval_block = d_end <= 32767;
        d_stop = d_end + 1;
        cap = min(R_SDC, d_end - d_start + 1);
    }
    const bool upd = rows_fit && eval_block;  // uniform

    for (int D0 = d_start; upd && D0 < d_stop; D0 += R_SDC) {
        const RefGroup g = ref_group<S::TBX>(X0, D0, W, a.offset, a.ltor, cap);
        if (g.n <= 0) continue;
        __syncthreads();
        for (int i = l; i < TBY * TW; i += 64) {
            const int r = i / TW, c = i % TW;
            const int x = g.lo + c, y = Y0 + r;
            sr[r][c] = (c < g.n + S::TBX - 1 && x >= 0 && x < W && y < H) ? row_ptr<const uint8_t>(a.R, y)[x] : 0;
        }
        for (int r = 0; r < g.n; ++r) {
            const int d = D0 + r;
            const int sh = (a.ltor ? g.base - r : g.base + r) - g.lo;
            __syncthreads();
            // column sums of SP rows for the 16 sub-patch top rows and 32 columns
            {
                const int c = l & 31, r0 = (l >> 5) * 8;
                int q[8 + SP - 1];
#pragma unroll
                for (int j = 0; j < 8 + SP - 1; ++j) {
                    const int u = sl[r0 + j][c], v = sr[r0 + j][c + sh];
                    if (SIM == 0) {
                        q[j] = 255 - abs(u - v);
                    } else {
                        const float f = stab[u] * stab[v];
                        q[j] = (int)(f * 255.0f);
                    }
                }
                int s = 0;
#pragma unroll
                for (int j = 0; j < SP; ++j) s += q[j];
                scol[r0][c] = s;
#pragma unroll
                for (int j = 1; j < 8; ++j) {
                    // AbsoluteDifference terms are non-negative: one v_sad_u32 per step
                    s = SIM == 0 ? (int)slide((uint32_t)s, (uint32_t)q[j - 1], (uint32_t)q[j + SP - 1])
                                 : s + q[j + SP - 1] - q[j - 1];
                    scol[r0 + j][c] = s;
                }
            }
            __syncthreads();
            // sub-patch sums: row l % 16, columns 8 (l / 16) .. + 8 of the 2 CG
            {
                const int rr = l & 15, c0 = (l >> 4) * 8;
                int s = 0;
#pragma unroll
                for (int i = 0; i < SP; ++i) s += scol[rr][c0 + i];
                ssp[rr][c0] = s;
#pragma unroll
                for (int i = 1; i < 8; ++i) {
                    if (c0 + i + SP - 1 < 32) {
                        s = SIM == 0 ? (int)slide((uint32_t)s, (uint32_t)scol[rr][c0 + i - 1], (uint32_t)scol[rr][c0 + i + SP - 1])
                                     : s + scol[rr][c0 + i + SP - 1] - scol[rr][c0 + i - 1];
                        ssp[rr][c0 + i] = s;
                    }
                }
            }
            __syncthreads();
            // Two_OutOf_Four (Kernel.Algo.AdaptWnd_W32H16.h:240-285) and the peak update
#pragma unroll
            for (int q = 0; q < RPL; ++q) {
                const int x = rx[q], y = ry[q];
                int m1 = ssp[y][x + SP], m2 = -32767, v;
                v = ssp[y + SP][x];
                m2 = max(m2, min(m1, v));
                m1 = max(m1, v);
                v = ssp[y + 2 * SP][x + SP];
                m2 = max(m2, min(m1, v));
                m1 = max(m1, v);
                v = ssp[y + SP][x + 2 * SP];
                m2 = max(m2, min(m1, v));
                m1 = max(m1, v);
                const int aw = (int)(int16_t)(m1 + m2 + ssp[y + SP][x + SP]);
                track(tr[q], aw, d, has[q]);
            }
        }
    }

    // WriteDisparity (:509-635)
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
        if (!written[q]) continue;
        const Tracker& s = tr[q];
        ushort4 o;
        const int dv = s.disp;
        if (dv > R_LAST || dv == 0) {
            o = make_ushort4((uint16_t)dv, 0, 0, 0);
        } else {
            const float fmax = aw1_sim_float(s.mx, S::NE);
            const float dp = fmaxf(fabsf(aw1_sim_float(s.prev, S::NE) - fmax), fabsf(aw1_sim_float(s.prev2, S::NE) - fmax));
            const float dn = fmaxf(fabsf(aw1_sim_float(s.nxt, S::NE) - fmax), fabsf(aw1_sim_float(s.nxt2, S::NE) - fmax));
            const bool specific = dp >= a.min_delta && dn >= a.min_delta;
            int code;
            if (fmax >= a.sim_thresh) {
                const bool sat = dv <= R_FIRST + a.dmin + 1 || dv >= R_FIRST + a.dmax - 1;
                code = sat ? R_SATURATED : (specific ? dv : R_NOT_SPECIFIC);
            } else {
                code = R_NOT_FOUND;
            }
            o = make_ushort4((uint16_t)code, (uint16_t)aw1_sim_u16(fmax), (uint16_t)aw1_sim_u16(dp),
                             (uint16_t)aw1_sim_u16(dn));
        }
        row_ptr<ushort4>(a.out, Y0 + ry[q] + S::P3 / 2)[X0 + rx[q] + S::P3 / 2] = o;
    }
}

// AW1 over NB vertically stacked reference blocks of one block column, 256 lanes: the
// blocks share their disparity groups (they depend on the block column only), and the
// similarity column sums and sub-patch sums are functions of the position and the
// disparity, so the stack forms them once per disparity for all its blocks' results
// (the one-block kernel forms 16 sub-patch rows for 6 or 10 result rows).  Decisions,
// tracker and write-out per result as in aw1_ref_kernel; the full-range search only.
template <int SP, int SIM, int NB>
__global__ __launch_bounds__(256) void aw1_ref_stack_kernel(RefArgs a)
{
#pragma clang fp contract(off)
    using S = AwShape<SP>;
    constexpr int PCX = S::PCX, PCY = S::PCY, TW = S::TW, P3 = S::P3, SPCX = S::SPCX;
    constexpr int RROWS = PCY * NB;        // result rows of the stack
    constexpr int SROWS = RROWS + 2 * SP;  // sub-patch top rows its results read
    constexpr int IROWS = SROWS + SP - 1;  // image rows
    constexpr int NRES = PCX * RROWS;
    constexpr int RPT = (NRES + 255) / 256;       // results per thread
    constexpr int CR = (SROWS + 7) / 8;           // column-sum rows per thread
    constexpr int GR = (RROWS + 7) / 8;           // gate column-sum rows per thread
    constexpr int CPG = (SPCX + 3) / 4;           // sub-patch columns per thread
    static_assert(IROWS == RROWS + P3 - 1, "the gate and the sub-patches read the same rows");
    static_assert(4 * SROWS <= 256, "one thread per sub-patch row and column group");

    // row stride of the column and sub-patch sums: 50 (18 mod 32) puts the two rows of 18
    // results a half-wave reads in different banks (sub-patch 5); 33 otherwise
    constexpr int SS = SP == 5 ? 50 : 33;
    __shared__ uint8_t sl[IROWS][32];
    __shared__ uint8_t sr[IROWS][TW + 1];
    __shared__ int scol[SROWS][SS];
    __shared__ int ssp[SROWS][SS];
    __shared__ float stab[256];
    __shared__ int sflag[NB];
    __shared__ int sany;

    const int l = threadIdx.x;
    const int W = a.L.w, H = a.L.h;
    const int X0 = blockIdx.x * PCX;
    const int by0 = blockIdx.y * NB;
    const int Y0 = by0 * PCY;

    if (SIM == 1) stab[l] = 2.0f * ((float)l / 255.0f - 0.5f);
    for (int i = l; i < IROWS * 32; i += 256) {
        const int r = i / 32, c = i % 32;
        const int x = X0 + c, y = Y0 + r;
        sl[r][c] = (x < W && y < H) ? row_ptr<const uint8_t>(a.L, y)[x] : 0;
    }
    if (l < NB) sflag[l] = 0;
    if (l == 0) sany = 0;

    // this thread's results: positions l + 256 q of the stack's PCX x RROWS, row-major
    int rx[RPT], ry[RPT], rb[RPT];
    bool has[RPT], written[RPT], upd[RPT];
    Tracker tr[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int pos = l + 256 * q;
        rx[q] = pos < NRES ? pos % PCX : 0;
        ry[q] = pos < NRES ? pos / PCX : 0;
        rb[q] = ry[q] / PCY;
        const int yb = Y0 + rb[q] * PCY;  // the block's origin row
        const bool block_ok = by0 + rb[q] < a.nby && X0 + P3 <= W && yb + P3 <= H;
        has[q] = pos < NRES && block_ok;
        written[q] = has[q] && X0 + rx[q] + P3 < W && Y0 + ry[q] + P3 < H;
        tr[q] = Tracker{0, -32767, 0, 0, 0, 0, 0};
    }
    __syncthreads();

    // EvalGrad: each result's P3 x P3 window standard deviation, truncated, in % of 255
    {
        const int c = l & 31, r0 = (l >> 5) * GR;
        int s1 = 0, s2 = 0;
        for (int j = 0; j < P3; ++j) {
            const int v = r0 + j < IROWS ? sl[r0 + j][c] : 0;
            s1 += v;
            s2 += v * v;
        }
        if (r0 < RROWS) {
            scol[r0][c] = s1;
            ssp[r0][c] = s2;
        }
        for (int j = 1; j < GR; ++j) {
            if (r0 + j >= RROWS) break;
            const int vo = sl[r0 + j - 1][c], vn = sl[r0 + j + P3 - 1][c];
            s1 += vn - vo;
            s2 += vn * vn - vo * vo;
            scol[r0 + j][c] = s1;
            ssp[r0 + j][c] = s2;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        if (!has[q]) continue;
        int s1 = 0, s2 = 0;
        for (int i = 0; i < P3; ++i) {
            s1 += scol[ry[q]][rx[q] + i];
            s2 += ssp[ry[q]][rx[q] + i];
        }
        const float m = (float)s1 / (float)(P3 * P3), m2 = (float)s2 / (float)(P3 * P3);
        const float sq = m * m;
        const int isd = (int)sqrtf(fabsf(m2 - sq));
        const float sd = (float)isd / 255.0f * 100.0f;
        if (sd < a.grad_thresh) tr[q].disp = R_CANNOT;
        else sflag[rb[q]] = 1;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        // the block's rows fit and one of its results evaluates (:741-775)
        const bool rows_fit = Y0 + rb[q] * PCY + S::TBY <= H;
        upd[q] = has[q] && rows_fit && sflag[rb[q]] != 0;
        if (upd[q]) sany = 1;
    }
    __syncthreads();
    const bool evaluate = sany != 0;

    const int maxd = a.ltor ? X0 - a.offset + 1 : W - X0 - a.offset;
    const int d_stop = a.dmax < maxd ? a.dmax : maxd;
    for (int D0 = a.dmin; evaluate && D0 < d_stop; D0 += R_SDC) {
        const RefGroup g = ref_group<S::TBX>(X0, D0, W, a.offset, a.ltor, R_SDC);
        if (g.n <= 0) continue;
        __syncthreads();
        for (int i = l; i < IROWS * TW; i += 256) {
            const int r = i / TW, c = i % TW;
            const int x = g.lo + c, y = Y0 + r;
            sr[r][c] = (c < g.n + S::TBX - 1 && x >= 0 && x < W && y < H) ? row_ptr<const uint8_t>(a.R, y)[x] : 0;
        }
        for (int r = 0; r < g.n; ++r) {
            const int d = D0 + r;
            const int sh = (a.ltor ? g.base - r : g.base + r) - g.lo;
            __syncthreads();
            // column sums of SP rows: the SROWS sub-patch top rows x 32 columns
            {
                const int c = l & 31, r0 = (l >> 5) * CR;
                int qv[CR + SP - 1];
#pragma unroll
                for (int j = 0; j < CR + SP - 1; ++j) {
                    const int rr = r0 + j < IROWS ? r0 + j : IROWS - 1;
                    const int u = sl[rr][c], v = sr[rr][c + sh];
                    if (SIM == 0) {
                        qv[j] = 255 - abs(u - v);
                    } else {
                        const float f = stab[u] * stab[v];
                        qv[j] = (int)(f * 255.0f);
                    }
                }
                int sum = 0;
#pragma unroll
                for (int j = 0; j < SP; ++j) sum += qv[j];
                if (r0 < SROWS) scol[r0][c] = sum;
#pragma unroll
                for (int j = 1; j < CR; ++j) {
                    sum = SIM == 0 ? (int)slide((uint32_t)sum, (uint32_t)qv[j - 1], (uint32_t)qv[j + SP - 1])
                                   : sum + qv[j + SP - 1] - qv[j - 1];
                    if (r0 + j < SROWS) scol[r0 + j][c] = sum;
                }
            }
            __syncthreads();
            // sub-patch sums: row l / 4, columns CPG (l % 4) .. + CPG
            if (l < 4 * SROWS) {
                const int rr = l >> 2, c0 = (l & 3) * CPG;
                int sum = 0;
#pragma unroll
                for (int i = 0; i < SP; ++i) sum += scol[rr][c0 + i];
                if (c0 < SPCX) ssp[rr][c0] = sum;
#pragma unroll
                for (int i = 1; i < CPG; ++i) {
                    if (c0 + i < SPCX) {
                        sum = SIM == 0 ? (int)slide((uint32_t)sum, (uint32_t)scol[rr][c0 + i - 1], (uint32_t)scol[rr][c0 + i + SP - 1])
                                       : sum + scol[rr][c0 + i + SP - 1] - scol[rr][c0 + i - 1];
                        ssp[rr][c0 + i] = sum;
                    }
                }
            }
            __syncthreads();
            // Two_OutOf_Four (Kernel.Algo.AdaptWnd_W32H16.h:240-285) and the peak update
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int x = rx[q], y = ry[q];
                int m1 = ssp[y][x + SP], m2 = -32767, v;
                v = ssp[y + SP][x];
                m2 = max(m2, min(m1, v));
                m1 = max(m1, v);
                v = ssp[y + 2 * SP][x + SP];
                m2 = max(m2, min(m1, v));
                m1 = max(m1, v);
                v = ssp[y + SP][x + 2 * SP];
                m2 = max(m2, min(m1, v));
                m1 = max(m1, v);
                const int aw = (int)(int16_t)(m1 + m2 + ssp[y + SP][x + SP]);
                track(tr[q], aw, d, upd[q]);
            }
        }
    }

    // WriteDisparity (:509-635)
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        if (!written[q]) continue;
        const Tracker& t = tr[q];
        ushort4 o;
        const int dv = t.disp;
        if (dv > R_LAST || dv == 0) {
            o = make_ushort4((uint16_t)dv, 0, 0, 0);
        } else {
            const float fmax = aw1_sim_float(t.mx, S::NE);
            const float dp = fmaxf(fabsf(aw1_sim_float(t.prev, S::NE) - fmax), fabsf(aw1_sim_float(t.prev2, S::NE) - fmax));
            const float dn = fmaxf(fabsf(aw1_sim_float(t.nxt, S::NE) - fmax), fabsf(aw1_sim_float(t.nxt2, S::NE) - fmax));
            const bool specific = dp >= a.min_delta && dn >= a.min_delta;
            int code;
            if (fmax >= a.sim_thresh) {
                const bool sat = dv <= R_FIRST + a.dmin + 1 || dv >= R_FIRST + a.dmax - 1;
                code = sat ? R_SATURATED : (specific ? dv : R_NOT_SPECIFIC);
            } else {
                code = R_NOT_FOUND;
            }
            o = make_ushort4((uint16_t)code, (uint16_t)aw1_sim_u16(fmax), (uint16_t)aw1_sim_u16(dp),
                             (uint16_t)aw1_sim_u16(dn));
        }
        row_ptr<ushort4>(a.out, Y0 + ry[q] + P3 / 2)[X0 + rx[q] + P3 / 2] = o;
    }
}

struct RefExt : DriverExt {
    int algorithm = 0;  // STEREO_ALGO_NCC or STEREO_ALGO_AW1
    stereo_ncc_params ncc{};
    stereo_aw1_params aw{};
};

template <int P>
hipError_t ncc_ref_launch(const RefArgs& a, int nbx, hipStream_t s)
{
    if (a.init.ptr)
        hipLaunchKernelGGL((ncc_ref_kernel<P, 1>), dim3(nbx, a.nby), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL((ncc_ref_kernel<P, 2>), dim3(nbx, (a.nby + 1) / 2), dim3(64), 0, s, a);
    return hipGetLastError();
}

// blocks stacked per workgroup in the full-range search: 8 x 6 result rows (sub-patch 5),
// 5 x 10 (sub-patch 3)
template <int SP>
constexpr int aw1_stack() { return SP == 5 ? 8 : 5; }

template <int SP>
hipError_t aw1_ref_launch(const RefArgs& a, int sim, int nbx, hipStream_t s)
{
    constexpr int NB = aw1_stack<SP>();
    const dim3 sgrid(nbx, (a.nby + NB - 1) / NB);
    if (a.init.ptr) {
        if (sim == 0) hipLaunchKernelGGL((aw1_ref_kernel<SP, 0>), dim3(nbx, a.nby), dim3(64), 0, s, a);
        else hipLaunchKernelGGL((aw1_ref_kernel<SP, 1>), dim3(nbx, a.nby), dim3(64), 0, s, a);
    } else {
        if (sim == 0) hipLaunchKernelGGL((aw1_ref_stack_kernel<SP, 0, NB>), sgrid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((aw1_ref_stack_kernel<SP, 1, NB>), sgrid, dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

}  // namespace
}  // namespace stereo

using namespace capi;

namespace {

stereo_status ref_common_checks(stereo_context* ctx, const stereo_image_format* fmt, int dmin, int dmax, int offset,
                                const int32_t* reserved, int min_w, int min_h, const char* name)
{
    if (!format_valid(fmt)) return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "Invalid image format");
    if (fmt->pixel_type != STEREO_PIXEL_LUM_U8)
        return fail(ctx, STEREO_ERROR_UNSUPPORTED_PIXEL_TYPE, "Pixel types of given images not supported");
    for (int i = 0; i < 4; ++i)
        if (reserved[i]) return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "reserved parameters must be 0");
    if (dmin < 0 || dmax <= dmin || dmax > 4096 || offset < -4096 || offset > 4096)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "disparity range [%d, %d) offset %d: need 0 <= min < max <= 4096",
                    dmin, dmax, offset);
    if (fmt->width < min_w || fmt->height < min_h || fmt->width > 16384 || fmt->height > 16384)
        return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "%s: image %dx%d, need %d <= width <= 16384 and %d <= height <= 16384",
                    name, fmt->width, fmt->height, min_w, min_h);
    return STEREO_OK;
}

stereo_status ref_make_driver(stereo_context* ctx, stereo_driver** out, const stereo_image_format* fmt,
                              stereo::RefExt* x, const char* name)
{
    stereo_status st;
    if ((st = set_device(ctx))) {
        delete x;
        return st;
    }
    stereo_driver* d = new stereo_driver(stereo_driver::REFMATCH);
    d->ctx = ctx;
    d->fmt = *fmt;
    d->name = name;
    d->ext = x;
    if ((st = driver_events(d))) {
        driver_release(d);
        delete d;
        return st;
    }
    *out = d;
    return STEREO_OK;
}

stereo_status ref_check_io(stereo_driver* d, stereo_image* disp, const stereo_image* left, const stereo_image* right,
                           const stereo_image* disp_init, int use_init, const char* name)
{
    stereo_context* ctx = d->ctx;
    const int W = d->fmt.width, H = d->fmt.height;
    stereo_status st;
    if ((st = check_image(ctx, left, W, H, STEREO_PIXEL_LUM_U8, "left image"))) return st;
    if ((st = check_image(ctx, right, W, H, STEREO_PIXEL_LUM_U8, "right image"))) return st;
    if ((st = check_image(ctx, disp, W, H, STEREO_PIXEL_RGBA_U16, "disparity image"))) return st;
    if (use_init) {
        if (!disp_init || !disp_init->data)
            return fail(ctx, STEREO_ERROR_INVALID_ARGUMENT, "%s: use_disp_input needs an initial disparity map", name);
        const int iw = disp_init->format.width, ih = disp_init->format.height;
        if (iw < 1 || ih < 1 || iw > W || ih > H)
            return fail(ctx, STEREO_ERROR_FORMAT_MISMATCH, "initial disparity map %dx%d for a %dx%d image", iw, ih, W, H);
        if ((st = check_image(ctx, disp_init, iw, ih, STEREO_PIXEL_RGBA_U16, "initial disparity map"))) return st;
    } else if (disp_init) {
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "%s: configured without use_disp_input", name);
    }
    return STEREO_OK;
}

}  // namespace

extern "C" {

stereo_status stereo_ncc_configure(stereo_context* ctx, stereo_driver** out, const stereo_image_format* fmt,
                                   const stereo_ncc_params* p)
{
    if (!ctx || !out || !fmt || !p) return STEREO_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    const int P = p->patch;
    if (P != 11 && P != 9 && P != 5) return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "Invalid algorithm configuration.");
    stereo_status st = ref_common_checks(ctx, fmt, p->disp_min, p->disp_max, p->disp_offset, p->reserved, 48, 15 + P,
                                         "BlockMatcherNCC");
    if (st) return st;
    stereo::RefExt* x = new stereo::RefExt();
    x->algorithm = STEREO_ALGO_NCC;
    x->ncc = *p;
    return ref_make_driver(ctx, out, fmt, x, "Block Matcher NCC");
}

stereo_status stereo_aw1_configure(stereo_context* ctx, stereo_driver** out, const stereo_image_format* fmt,
                                   const stereo_aw1_params* p)
{
    if (!ctx || !out || !fmt || !p) return STEREO_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (p->patch != 15 && p->patch != 9) return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "Invalid algorithm configuration.");
    if (p->sim_type != 0 && p->sim_type != 1)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "Given similarity type not supported.");
    const int sp = p->patch / 3;
    const int tby = (16 - 2 * sp) + 3 * sp - 1;
    stereo_status st =
        ref_common_checks(ctx, fmt, p->disp_min, p->disp_max, p->disp_offset, p->reserved, 48, tby, "BlockMatcherAW1");
    if (st) return st;
    stereo::RefExt* x = new stereo::RefExt();
    x->algorithm = STEREO_ALGO_AW1;
    x->aw = *p;
    return ref_make_driver(ctx, out, fmt, x, "Block Matcher Adaptive Window 1");
}

static stereo_status ref_process(stereo_driver* d, stereo_image* disp, const stereo_image* left,
                                 const stereo_image* right, const stereo_image* disp_init, stereo_stream stream,
                                 int algorithm)
{
    if (!d || d->kind != stereo_driver::REFMATCH || !disp || !left || !right) return STEREO_ERROR_INVALID_ARGUMENT;
    const stereo::RefExt* x = static_cast<const stereo::RefExt*>(d->ext);
    if (x->algorithm != algorithm) return STEREO_ERROR_INVALID_ARGUMENT;
    stereo_context* ctx = d->ctx;
    const bool ncc = algorithm == STEREO_ALGO_NCC;
    const int use_init = ncc ? x->ncc.use_disp_input : x->aw.use_disp_input;
    stereo_status st;
    if ((st = ref_check_io(d, disp, left, right, disp_init, use_init, d->name))) return st;
    if ((st = set_device(ctx))) return st;
    GUARD(ctx, ncc ? "stereo_ncc_process" : "stereo_aw1_process");
    const hipStream_t s = (hipStream_t)stream;
    if ((st = begin(d, s))) return st;
    stereo::RefArgs a;
    a.L = view_of(left);
    a.R = view_of(right);
    a.out = view_of(disp);
    a.init = use_init ? view_of(disp_init) : null_view();
    const int W = d->fmt.width, H = d->fmt.height;
    hipError_t e;
    if (ncc) {
        const stereo_ncc_params& p = x->ncc;
        a.dmin = p.disp_min;
        a.dmax = p.disp_max;
        a.offset = p.disp_offset;
        a.ltor = p.left_to_right ? 1 : 0;
        a.imin_delta = p.min_delta_thresh;
        a.min_delta = 0.0f;
        a.sim_thresh = p.ncc_thresh;
        a.grad_thresh = p.grad_thresh;
        // EvalThreadConfig_ReadDepBlockSize over a 33 x (15 + P) read area, full blocks only
        const int P = p.patch, PCX = 2 * ((32 - P + 1) / 2);
        const int nbx = (W - (33 - PCX)) / PCX;
        a.nby = (H - (P - 1)) / 16;
        switch (P) {
        case 11: e = stereo::ncc_ref_launch<11>(a, nbx, s); break;
        case 9: e = stereo::ncc_ref_launch<9>(a, nbx, s); break;
        default: e = stereo::ncc_ref_launch<5>(a, nbx, s); break;
        }
    } else {
        const stereo_aw1_params& p = x->aw;
        a.dmin = p.disp_min;
        a.dmax = p.disp_max;
        a.offset = p.disp_offset;
        a.ltor = p.left_to_right ? 1 : 0;
        a.imin_delta = 0;
        a.min_delta = p.min_delta_thresh;
        a.sim_thresh = p.sim_thresh;
        a.grad_thresh = p.grad_thresh;
        const int sp = p.patch / 3, CG = (32 - sp + 1) / 2;
        const int PCX = 2 * CG - 2 * sp, PCY = 16 - 2 * sp, TBY = PCY + 3 * sp - 1;
        const int nbx = (W - (32 - PCX)) / PCX;
        a.nby = (H - (TBY - PCY)) / PCY;
        e = sp == 5 ? stereo::aw1_ref_launch<5>(a, p.sim_type, nbx, s) : stereo::aw1_ref_launch<3>(a, p.sim_type, nbx, s);
    }
    HIP_TRY(ctx, e, ncc ? "BlockMatcherNCC launch" : "BlockMatcherAW1 launch");
    return end(d, s);
}

stereo_status stereo_ncc_process(stereo_driver* d, stereo_image* disp, const stereo_image* left,
                                 const stereo_image* right, const stereo_image* disp_init, stereo_stream stream)
{
    return ref_process(d, disp, left, right, disp_init, stream, STEREO_ALGO_NCC);
}

stereo_status stereo_aw1_process(stereo_driver* d, stereo_image* disp, const stereo_image* left,
                                 const stereo_image* right, const stereo_image* disp_init, stereo_stream stream)
{
    return ref_process(d, disp, left, right, disp_init, stream, STEREO_ALGO_AW1);
}

}  // extern "C"
