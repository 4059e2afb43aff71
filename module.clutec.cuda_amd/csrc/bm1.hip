// BlockMatcher1 (CluTec.Cuda.MultiView/BlockMatcher1.cu), with the reference's block-tiled
// results: the same disparities, test patches, codes and similarity channels as the
// reference's 32-thread blocks.
//
// The reference tiles the result origins into blocks of PCX x 4 (PCX = 8 * ((32 - P + 1) / 8),
// AlgoArraySum<32, 2, P, P>), counted from a 33 x (P + 3) read area
// (KernelDriverBase.cpp:166-197).  A block decides which disparities its results see:
// groups of 16 from the range start while the group start is below min(Max,
// MaxDisparity(X0)); a group whose 48-column test area lies inside the right image
// evaluates 16 disparities (past Max for the last group); otherwise, when the block's
// P + 3 rows fit, a partial group evaluates the disparities of a narrower area, with the
// reference's placement of the test patches (BlockMatcher1.cu:747-838).  Each result
// tracks the similarity maximum with the values around it (EvalSim :244-302) and is
// written with its code and three 16-bit similarity channels (WriteDisparity :447-575).
//
// Layout on gfx950: one 256-thread workgroup takes a strip of NB reference blocks of one
// block column, which share their evaluation list (bottom blocks whose rows do not fit
// take none).  The strip's left pixels, and per group of 16 disparities the right pixels,
// sit in LDS as bytes.  Per disparity, each thread forms running P-row column sums of the
// similarity over a segment of rows into LDS (double-buffered), then a running P-column
// row sum over its results and the peak update.  With 16 blocks per strip, each wave's
// column sums are its own 16 rows, so the waves synchronise only per group of 16
// disparities, not per disparity.  The band search around an initial map
// (MatcherUseDisparity :857-1203) gives every block its own range, and runs one block per
// workgroup.
#include <hip/hip_runtime.h>

#include "capi_internal.hpp"

namespace stereo {
namespace {

constexpr int B1_SDC = 16, B1_TBX = 33, B1_TEST_W = 48;
constexpr int B1_FIRST = 1, B1_LAST = 0xFFEF, B1_CANNOT = 0xFFFF, B1_SATURATED = 0xFFFE, B1_NOT_FOUND = 0xFFFD,
              B1_NOT_SPECIFIC = 0xFFFC;

struct Bm1Args {
    View L, R, out, init;  // init.ptr == nullptr: the plain matcher
    int dmin, dmax, offset, ltor;
    float min_delta, sim_thresh, grad_thresh;
    int nby;  // reference block rows
};

template <int P>
struct Bm1Shape {
    static constexpr int CG = (32 - P + 1) / 8;
    static constexpr int PCX = 8 * CG;
    static constexpr int TBY = P + 3;
};

// SimValueToFloat (:155-159) and SimFloatToInteger (:172-176)
__device__ __forceinline__ float bm1_sim_float(int v, int P)
{
#pragma clang fp contract(off)
    return ((float)v / (float)(P * P)) * (100.0f / 127.0f);
}

__device__ __forceinline__ uint32_t bm1_sim_u16(float f)
{
#pragma clang fp contract(off)
    float t = (f / 100.0f + 1.0f) / 2.0f;
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return (uint32_t)(t * 65535.0f);
}

// One group of disparities of a block: labels D0 .. D0 + n - 1; the test origin of result
// column 0 for label D0 + r is base - r (left to right) or base + r; the right pixels used
// are the columns [lo, lo + n + 31).
struct Bm1Group {
    int n, base, lo;
};

// The group at D0 for a block at X0 whose P + 3 rows fit (:747-838); n = 0 when nothing is
// evaluated.  cap is 16, or the band search's count (:1088).
__device__ __forceinline__ Bm1Group bm1_group(int X0, int D0, int W, int offset, int ltor, int cap)
{
    Bm1Group g;
    const int tx = (ltor ? X0 - D0 - offset : X0 + D0 + offset) - ltor * (B1_SDC - 1);
    if (tx >= 0 && tx + B1_TEST_W <= W) {
        g.n = cap;
        g.base = tx + ltor * (B1_SDC - 1);
    } else {
        int dx, sx;
        if (tx >= 0) {
            dx = W - tx - B1_TBX;
            sx = tx;
        } else {
            dx = tx + (B1_SDC - 1);
            sx = 0;
        }
        g.n = dx >= 0 ? dx + 1 : 0;
        g.base = sx + ltor * (g.n - 1);
    }
    g.lo = ltor ? g.base - (g.n - 1) : g.base;
    return g;
}

template <int P, int SIM, int NB>
__global__ __launch_bounds__(256) void bm1_kernel(Bm1Args a)
{
#pragma clang fp contract(off)
    using S = Bm1Shape<P>;
    constexpr int PCX = S::PCX, CG = S::CG;
    constexpr int ROWS = 4 * NB, DR = ROWS + P - 1;
    // vertical phase: 32 columns x SEG rows per thread
    constexpr int SEG = ROWS * 32 >= 256 ? ROWS * 32 / 256 : 1;
    constexpr int VT = ROWS * 32 / SEG;  // threads with a segment
    // row phase: RPT consecutive results of one row per thread
    constexpr int RPT = ROWS * PCX >= 256 ? ROWS * PCX / 256 : 1;
    constexpr int TPR = PCX / RPT;  // threads per result row
    static_assert(PCX % RPT == 0 && SEG * VT == ROWS * 32, "tile shape");

    __shared__ uint8_t sl[DR][32];
    __shared__ uint8_t sr[DR][B1_TEST_W];
    __shared__ int sv[2][ROWS][32];
    __shared__ float stab[256];
    __shared__ int sinit[4 * PCX];
    __shared__ int sred[2][32];

    const int t = threadIdx.x;
    const int W = a.L.w, H = a.L.h;
    const int X0 = blockIdx.x * PCX;
    const int by0 = blockIdx.y * NB;
    const int Y0s = by0 * 4;

    if (SIM == 1) stab[t] = 2.0f * ((float)t / 255.0f - 0.5f);
    for (int i = t; i < DR * 32; i += 256) {
        const int r = i / 32, c = i % 32;
        const int x = X0 + c, y = Y0s + r;
        sl[r][c] = (x < W && y < H) ? row_ptr<const uint8_t>(a.L, y)[x] : 0;
    }

    // this thread's results: row rr, columns ix0 .. ix0 + RPT - 1
    const int rr = t / TPR, ix0 = (t % TPR) * RPT;
    const bool has_row = rr < ROWS && t < TPR * ROWS;
    const int bro = rr / 4;                 // block of the strip
    const int Y0 = Y0s + 4 * bro;           // its origin row
    const bool block_ok = has_row && by0 + bro < a.nby && X0 + P <= W && Y0 + P <= H;
    const bool rows_fit = Y0 + S::TBY <= H;  // the block's test rows lie in the image
    const bool upd = block_ok && rows_fit;   // the block evaluates disparities
    int disp[RPT], mx[RPT], prev[RPT], prev2[RPT], nxt[RPT], nxt2[RPT], last[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        disp[k] = 0;
        mx[k] = -32767;
        prev[k] = prev2[k] = nxt[k] = nxt2[k] = last[k] = 0;
    }
    __syncthreads();

    const bool band = NB == 1 && a.init.ptr != nullptr;  // the band search runs one block per workgroup
    int d_start, d_stop, cap;  // groups D0 = d_start + 16 k while D0 < d_stop
    bool any = true;
    if (!band) {
        // EvalGrad (:325-361): the left patch's standard deviation in % of 255
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (!block_ok) continue;
            int s = 0, s2 = 0;
            for (int j = 0; j < P; ++j)
                for (int i = 0; i < P; ++i) {
                    const int v = sl[rr + j][ix0 + k + i];
                    s += v;
                    s2 += v * v;
                }
            const float mean = (float)s / (float)(P * P);
            const float mean2 = (float)s2 / (float)(P * P);
            const float sq = mean * mean;
            const float sd = sqrtf(fabsf(mean2 - sq)) / 255.0f * 100.0f;
            if (sd < a.grad_thresh) disp[k] = B1_CANNOT;
        }
        const int maxd = a.ltor ? X0 - a.offset + 1 : W - X0 - a.offset;  // MaxDisparity
        d_start = a.dmin;
        d_stop = a.dmax < maxd ? a.dmax : maxd;
        cap = B1_SDC;
    } else {
        // MatcherUseDisparity (:912-963), one block per workgroup: reference thread u holds
        // the results ((u / 4) * CG + r, u % 4); a written result takes the initial id at its
        // centre over the scale (outside the initial map: Unknown), scaled in 16 bits when it
        // is an id, as it is when it is a code.  Each reference thread's plain min / max from
        // (Last, First) is then reduced keeping only valid operands.
        const int scale = W / a.init.w;
        if (t < 32) {
            int tlo = B1_LAST, thi = B1_FIRST;
            for (int r = 0; r < CG; ++r) {
                const int ix = (t / 4) * CG + r, iy = t % 4;
                int code = 0;
                if (X0 + ix + P < W && Y0s + iy + P < H) {
                    const int px = (X0 + ix + P / 2) / scale, py = (Y0s + iy + P / 2) / scale;
                    const int id = (px < a.init.w && py < a.init.h) ? row_ptr<const ushort4>(a.init, py)[px].x : 0;
                    if (id <= B1_LAST) {
                        code = (scale * id) & 0xFFFF;
                        tlo = min(tlo, code);
                        thi = max(thi, code);
                    } else {
                        code = id;
                    }
                }
                sinit[iy * PCX + ix] = code;
            }
            sred[0][t] = (tlo >= B1_FIRST && tlo <= B1_LAST) ? tlo : -1;
            sred[1][t] = (thi >= B1_FIRST && thi <= B1_LAST) ? thi : -1;
        }
        __syncthreads();
        int lo = -1, hi = -1;
        for (int u = 0; u < 32; ++u) {
            const int l = sred[0][u], h = sred[1][u];
            if (l >= 0) lo = lo < 0 ? l : min(lo, l);
            if (h >= 0) hi = hi < 0 ? h : max(hi, h);
        }
#pragma unroll
        for (int k = 0; k < RPT; ++k)
            if (has_row) disp[k] = sinit[rr * PCX + ix0 + k];
        if (lo < 0 || hi < 0 || lo > hi) {
            // no valid range: every written result keeps its initial code (:967-999)
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int ix = ix0 + k, iy = rr;
                if (block_ok && X0 + ix + P < W && Y0 + iy % 4 + P < H)
                    row_ptr<ushort4>(a.out, Y0 + iy % 4 + P / 2)[X0 + ix + P / 2] =
                        make_ushort4((uint16_t)disp[k], 0, 0, 0);
            }
            return;
        }
        d_start = lo < 4 ? 0 : lo - 4;
        const int d_end = hi + 4;  // inclusive (:1080-1082)
        // beyond 16 bits the reference's loop variables wrap: nothing is evaluated
        any = d_end <= 32767;
        d_stop = d_end + 1;
        cap = min(B1_SDC, d_end - d_start + 1);
    }

    // the left pixels of this thread's column segment do not change with the disparity:
    // they are read from LDS once, into registers
    uint32_t lreg[SEG + P - 1];
    {
        const int c = t % 32, r0 = (t / 32) * SEG;
#pragma unroll
        for (int j = 0; j < SEG + P - 1; ++j) lreg[j] = t < VT ? sl[r0 + j][c] : 0u;
    }

    int buf = 0;
    for (int D0 = d_start; any && D0 < d_stop; D0 += B1_SDC) {
        const Bm1Group g = bm1_group(X0, D0, W, a.offset, a.ltor, cap);
        if (g.n <= 0) continue;
        __syncthreads();  // the previous group's column sums have read sr
        for (int i = t; i < DR * B1_TEST_W; i += 256) {
            const int r = i / B1_TEST_W, c = i % B1_TEST_W;
            const int x = g.lo + c, y = Y0s + r;
            sr[r][c] = (c < g.n + 31 && x >= 0 && x < W && y < H) ? row_ptr<const uint8_t>(a.R, y)[x] : 0;
        }
        __syncthreads();
        for (int r = 0; r < g.n; ++r) {
            const int d = D0 + r;
            const int sh = (a.ltor ? g.base - r : g.base + r) - g.lo;  // test origin column in sr
            // column sums of P rows over this thread's segment
            if (t < VT) {
                const int c = t % 32, r0 = (t / 32) * SEG;
                // AbsoluteDifference: (255 - x) >> 1 == 127 - (x >> 1) for 0 <= x <= 255, and
                // v_sad_u8(l, q, 0) gives |l - q|, so a column sum is 127 P minus a sum of
                // t = |l - q| >> 1.  SignedProduct: the similarity, negated.
                int t[SEG + P - 1];
#pragma unroll
                for (int j = 0; j < SEG + P - 1; ++j) {
                    const uint32_t l = lreg[j], q = sr[r0 + j][c + sh];
                    if (SIM == 0) {
                        t[j] = (int)(__builtin_amdgcn_sad_u8(l, q, 0u) >> 1);
                    } else {
                        const float f = stab[l] * stab[q];
                        t[j] = -(int)(f * 127.0f);
                    }
                }
                constexpr int base = SIM == 0 ? 127 * P : 0;
                int s = 0;
#pragma unroll
                for (int j = 0; j < P; ++j) s += t[j];
                sv[buf][r0][c] = base - s;
#pragma unroll
                for (int j = 1; j < SEG; ++j) {
                    // AbsoluteDifference terms are non-negative: one v_sad_u32 per step
                    s = SIM == 0 ? (int)slide((uint32_t)s, (uint32_t)t[j - 1], (uint32_t)t[j + P - 1])
                                 : s + t[j + P - 1] - t[j - 1];
                    sv[buf][r0 + j][c] = base - s;
                }
            }
            if (NB == 16) {
                // wave w writes and reads only the column-sum rows 16 w .. 16 w + 15, and a
                // wave's LDS operations execute in order: ordering the compiler suffices
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            } else {
                __syncthreads();
            }
            // row sums of P columns and the peak update (:262-300); every thread reads
            // its row (in LDS for all threads), results outside an evaluating block
            // leave their state unchanged
            {
                const int (&vrow)[32] = sv[buf][has_row ? rr : 0];
                int s = 0;
#pragma unroll
                for (int i = 0; i < P; ++i) s += vrow[ix0 + i];
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    if (k > 0)
                        s = SIM == 0 ? (int)slide((uint32_t)s, (uint32_t)vrow[ix0 + k - 1], (uint32_t)vrow[ix0 + k + P - 1])
                                     : s + vrow[ix0 + k + P - 1] - vrow[ix0 + k - 1];
                    // branch-free: every state value is a select
                    const int v = (int)(int16_t)s;
                    const bool act = upd && disp[k] <= B1_LAST;
                    const bool above = act && v > mx[k];
                    const bool b1 = act && !above && disp[k] + 1 == d + B1_FIRST;
                    const bool b2 = act && !above && disp[k] + 2 == d + B1_FIRST;
                    prev2[k] = above ? prev[k] : prev2[k];
                    prev[k] = above ? last[k] : prev[k];
                    mx[k] = above ? v : mx[k];
                    disp[k] = above ? d + B1_FIRST : disp[k];
                    nxt[k] = (above || b1) ? v : nxt[k];
                    nxt2[k] = (above || b2) ? v : nxt2[k];
                    last[k] = act ? v : last[k];
                }
            }
            buf ^= 1;
        }
    }

    // WriteDisparity (:447-575)
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int ix = ix0 + k, iy = rr % 4;
        if (!block_ok || !(X0 + ix + P < W && Y0 + iy + P < H)) continue;
        ushort4 o;
        const int dv = disp[k];
        if (dv > B1_LAST || dv == 0) {
            o = make_ushort4((uint16_t)dv, 0, 0, 0);
        } else {
            const float fmax = bm1_sim_float(mx[k], P);
            const float dp = fmaxf(fabsf(bm1_sim_float(prev[k], P) - fmax), fabsf(bm1_sim_float(prev2[k], P) - fmax));
            const float dn = fmaxf(fabsf(bm1_sim_float(nxt[k], P) - fmax), fabsf(bm1_sim_float(nxt2[k], P) - fmax));
            const bool specific = dp >= a.min_delta && dn >= a.min_delta;
            int code;
            if (fmax >= a.sim_thresh) {
                const bool sat = dv <= B1_FIRST + a.dmin + 1 || dv >= B1_FIRST + a.dmax - 1;
                code = sat ? B1_SATURATED : (specific ? dv : B1_NOT_SPECIFIC);
            } else {
                code = B1_NOT_FOUND;
            }
            o = make_ushort4((uint16_t)code, (uint16_t)bm1_sim_u16(fmax), (uint16_t)bm1_sim_u16(dp),
                             (uint16_t)bm1_sim_u16(dn));
        }
        row_ptr<ushort4>(a.out, Y0 + iy + P / 2)[X0 + ix + P / 2] = o;
    }
}

struct Bm1Ext : DriverExt {
    stereo_bm1_params p{};
};

template <int P, int SIM>
hipError_t bm1_launch(const Bm1Args& a, int nbx, hipStream_t s)
{
    if (a.init.ptr) {
        hipLaunchKernelGGL((bm1_kernel<P, SIM, 1>), dim3(nbx, a.nby), dim3(256), 0, s, a);
    } else {
        hipLaunchKernelGGL((bm1_kernel<P, SIM, 16>), dim3(nbx, (a.nby + 15) / 16), dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

template <int P>
hipError_t bm1_dispatch_sim(const Bm1Args& a, int sim, int nbx, hipStream_t s)
{
    return sim == 0 ? bm1_launch<P, 0>(a, nbx, s) : bm1_launch<P, 1>(a, nbx, s);
}

}  // namespace
}  // namespace stereo

using namespace capi;

extern "C" {

stereo_status stereo_bm1_configure(stereo_context* ctx, stereo_driver** out, const stereo_image_format* fmt,
                                   const stereo_bm1_params* p)
{
    if (!ctx || !out || !fmt || !p) return STEREO_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!format_valid(fmt)) return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "Invalid image format");
    if (fmt->pixel_type != STEREO_PIXEL_LUM_U8)
        return fail(ctx, STEREO_ERROR_UNSUPPORTED_PIXEL_TYPE, "Pixel types of given images not supported");
    const int P = p->patch;
    if (P != 5 && P != 9 && P != 11 && P != 15)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "Invalid algorithm configuration.");
    if (p->sim_type != 0 && p->sim_type != 1)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "Given similarity type not supported.");
    for (int i = 0; i < 4; ++i)
        if (p->reserved[i]) return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "reserved parameters must be 0");
    if (p->disp_min < 0 || p->disp_max <= p->disp_min || p->disp_max > 4096 || p->disp_offset < -4096 ||
        p->disp_offset > 4096)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "disparity range [%d, %d) offset %d: need 0 <= min < max <= 4096",
                    p->disp_min, p->disp_max, p->disp_offset);
    if (fmt->width < 48 || fmt->height < P + 3 || fmt->width > 16384 || fmt->height > 16384)
        return fail(ctx, STEREO_ERROR_INVALID_FORMAT,
                    "BlockMatcher1: image %dx%d, need 48 <= width <= 16384 and %d <= height <= 16384", fmt->width,
                    fmt->height, P + 3);
    stereo_status st;
    if ((st = set_device(ctx))) return st;
    stereo_driver* d = new stereo_driver(stereo_driver::BM1);
    d->ctx = ctx;
    d->fmt = *fmt;
    d->name = "Block Matcher Adaptive Window 1";
    stereo::Bm1Ext* x = new stereo::Bm1Ext();
    x->p = *p;
    d->ext = x;
    if ((st = driver_events(d))) {
        driver_release(d);
        delete d;
        return st;
    }
    *out = d;
    return STEREO_OK;
}

stereo_status stereo_bm1_process(stereo_driver* d, stereo_image* disp, const stereo_image* left,
                                 const stereo_image* right, const stereo_image* disp_init, stereo_stream stream)
{
    if (!d || d->kind != stereo_driver::BM1 || !disp || !left || !right) return STEREO_ERROR_INVALID_ARGUMENT;
    stereo_context* ctx = d->ctx;
    const stereo::Bm1Ext* x = static_cast<const stereo::Bm1Ext*>(d->ext);
    const int W = d->fmt.width, H = d->fmt.height;
    stereo_status st;
    if ((st = check_image(ctx, left, W, H, STEREO_PIXEL_LUM_U8, "left image"))) return st;
    if ((st = check_image(ctx, right, W, H, STEREO_PIXEL_LUM_U8, "right image"))) return st;
    if ((st = check_image(ctx, disp, W, H, STEREO_PIXEL_RGBA_U16, "disparity image"))) return st;
    if (x->p.use_disp_input) {
        if (!disp_init || !disp_init->data)
            return fail(ctx, STEREO_ERROR_INVALID_ARGUMENT, "BlockMatcher1: use_disp_input needs an initial disparity map");
        const int iw = disp_init->format.width, ih = disp_init->format.height;
        if (iw < 1 || ih < 1 || iw > W || ih > H)
            return fail(ctx, STEREO_ERROR_FORMAT_MISMATCH, "initial disparity map %dx%d for a %dx%d image", iw, ih, W, H);
        if ((st = check_image(ctx, disp_init, iw, ih, STEREO_PIXEL_RGBA_U16, "initial disparity map"))) return st;
    } else if (disp_init) {
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "BlockMatcher1: configured without use_disp_input");
    }
    if ((st = set_device(ctx))) return st;
    GUARD(ctx, "stereo_bm1_process");
    const hipStream_t s = (hipStream_t)stream;
    if ((st = begin(d, s))) return st;
    const int P = x->p.patch;
    stereo::Bm1Args a;
    a.L = view_of(left);
    a.R = view_of(right);
    a.out = view_of(disp);
    a.init = x->p.use_disp_input ? view_of(disp_init) : stereo::View{};
    a.dmin = x->p.disp_min;
    a.dmax = x->p.disp_max;
    a.offset = x->p.disp_offset;
    a.ltor = x->p.left_to_right ? 1 : 0;
    a.min_delta = x->p.min_delta_thresh;
    a.sim_thresh = x->p.sim_thresh;
    a.grad_thresh = x->p.grad_thresh;
    // EvalThreadConfig_ReadDepBlockSize: blocks over the area less the read border
    const int PCX = 8 * ((32 - P + 1) / 8);
    const int effW = W - (33 - PCX), effH = H - (P - 1);
    // full blocks only (bOnlyFullBlocks defaults to true, KernelDriverBase.h:117-121)
    const int nbx = effW / PCX;
    a.nby = effH / 4;
    hipError_t e;
    switch (P) {
    case 5: e = stereo::bm1_dispatch_sim<5>(a, x->p.sim_type, nbx, s); break;
    case 9: e = stereo::bm1_dispatch_sim<9>(a, x->p.sim_type, nbx, s); break;
    case 11: e = stereo::bm1_dispatch_sim<11>(a, x->p.sim_type, nbx, s); break;
    default: e = stereo::bm1_dispatch_sim<15>(a, x->p.sim_type, nbx, s); break;
    }
    HIP_TRY(ctx, e, "BlockMatcher1 launch");
    return end(d, s);
}

}  // extern "C"
