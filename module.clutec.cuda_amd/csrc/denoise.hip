// Disparity de-noise (the role of Statistics::DispDenoise, CluTec.Cuda.MultiView/
// Statistics.DispDenoise.cu:110-330).  For every pixel, a P x P window (offsets
// -P/2 .. P-1-P/2; positions outside the image moved by P into it, :160-165) gives the
// count, sum and sum of squares of the valid values.  With too little support (count <
// P*P * support ratio) the pixel becomes unknown; a valid pixel is replaced by the
// rounded mean of the other valid values when it lies within the cut-off of it, else
// unknown; an invalid pixel is filled with the rounded window mean when the window's
// standard deviation is below the cut-off (and filling is on).  Every float step is one
// IEEE single operation rounded to nearest, with contraction into FMA switched off, so
// the result matches a plain C evaluation of the same formulas.  The square root is not
// evaluated: for a float c > 0 and x >= 0, RN(sqrt(x)) < c holds exactly when x < m*m,
// m the midpoint of c and the float below it (m*m is exact in double, and no float x
// equals it), so the test is one double comparison against a limit set by the host.
//
// Layout: one wave per 64 x 16 output tile.  Each lane loads its column of the tile and
// its P-1 halo rows (positions moved by P at the borders; lanes below P-1 also the
// column 64 further right), every load issued before the first is used, and stores the
// values in LDS: a valid value as itself, an invalid one as a negative number from which
// the input value is recovered (itself for int16 maps, -(id+1) for ids).  Then the lane
// walks down its column: per row the P values to its right give the row's count, sum and
// sum of squares, a ring of the last P rows' sums in registers keeps the window's totals
// by adding the new row and dropping the oldest, and from the P-th row on each step
// decides one output pixel.
#include <hip/hip_runtime.h>

#include <cmath>

#include "capi_internal.hpp"

namespace stereo {
namespace {

constexpr int DN_TW = 64, DN_TH = 16;

struct DenoiseArgs {
    View in, out;
    float need;      // (float)(P*P) * support ratio
    float cutoff;
    double sd_limit2;  // sqrt(var) < cutoff  <=>  var < sd_limit2
    int fill;
};

// The limit of the fill test: the square of the midpoint between the cut-off and the float
// below it; 0 when no root can be below the cut-off (c <= 0 or NaN).
inline double sd_limit2(float c)
{
    if (!(c > 0.0f)) return 0.0;
    if (std::isinf(c)) return INFINITY;
    const double m = 0.5 * ((double)std::nextafter(c, 0.0f) + (double)c);
    return m * m;
}

// A position of the window moved into the image: one shift by P suffices for images of at
// least P x P; positions that stay outside only feed outputs outside the image.
__device__ __forceinline__ int dn_reflect(int v, int n, int P)
{
    if (v < 0) v += P;
    if (v >= n) v -= P;
    return v;
}

// LDS code of an input value: >= 0 valid, < 0 invalid (see above)
template <bool REF>
__device__ __forceinline__ int dn_code(int raw)
{
    if (REF) return (raw >= 1 && raw <= 0xFFEF) ? raw : -raw - 1;
    return raw;
}

template <bool REF>
__device__ __forceinline__ int dn_load(const View& in, int x, int y)
{
    return REF ? (int)row_ptr<const ushort4>(in, y)[x].x : (int)row_ptr<const int16_t>(in, y)[x];
}

template <int P, bool REF>
__global__ __launch_bounds__(64) void denoise_kernel(DenoiseArgs a)
{
#pragma clang fp contract(off)
    constexpr int ROWS = DN_TH + P - 1, COLS = DN_TW + P - 1;
    __shared__ int val[ROWS][COLS];
    const int lane = threadIdx.x;
    const int W = a.in.w, H = a.in.h;
    const int x0 = blockIdx.x * DN_TW, y0 = blockIdx.y * DN_TH;
    // positions that stay outside the image after the move are read at a clamped position
    // and coded invalid; they only feed outputs outside the image
    const int xa = dn_reflect(x0 - P / 2 + lane, W, P);
    const int xb = dn_reflect(x0 - P / 2 + lane + DN_TW, W, P);
    const int xac = min(max(xa, 0), W - 1), xbc = min(max(xb, 0), W - 1);
    int va[ROWS], vb[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int y = min(max(dn_reflect(y0 - P / 2 + r, H, P), 0), H - 1);
        va[r] = dn_load<REF>(a.in, xac, y);
        vb[r] = dn_load<REF>(a.in, xbc, y);
    }
    const bool ina = xa >= 0 && xa < W, inb = xb >= 0 && xb < W;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int y = dn_reflect(y0 - P / 2 + r, H, P);
        const bool iny = y >= 0 && y < H;
        val[r][lane] = (ina && iny) ? dn_code<REF>(va[r]) : -1;
        if (lane < P - 1) val[r][lane + DN_TW] = (inb && iny) ? dn_code<REF>(vb[r]) : -1;
    }
    __syncthreads();
    const int x = x0 + lane;
    const int unknown = REF ? 0 : -1;
    int rc[P], rs[P];
    unsigned long long rq[P];
    int cnt = 0, sum = 0;
    unsigned long long sum2 = 0;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        int hc = 0, hs = 0;
        unsigned long long hq = 0;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            // branch-free: invalid values add nothing; values are below 2^16, so the
            // 24-bit product is the exact square
            const int v = val[r][lane + k];
            const unsigned u = v >= 0 ? (unsigned)v : 0u;
            hc += v >= 0 ? 1 : 0;
            hs += (int)u;
            hq += (unsigned)__umul24(u, u);
        }
        cnt += hc;
        sum += hs;
        sum2 += hq;
        rc[r % P] = hc;
        rs[r % P] = hs;
        rq[r % P] = hq;
        if (r < P - 1) continue;
        const int y = y0 + r - (P - 1);
        if (x < W && y < H) {
            const int v = val[r - (P - 1) + P / 2][lane + P / 2];
            const int raw = (v >= 0 || !REF) ? v : -v - 1;
            // one quotient serves both cases: the mean of the other values for a valid
            // pixel, (sum - v) / (count - 1), and the window mean for an invalid one (a
            // wave holds both kinds, so the operands are selected, not branched on).  The
            // sums are below 2^24, so (float)(sum - v) equals (float)sum - (float)v, and
            // the sum of squares (below 2^53) converts through double with one rounding.
            // A zero count gives NaN, which fails every test below.
            const bool valid = v >= 0;
            const float fd = (float)v;
            const float mean = (float)(valid ? sum - v : sum) / (float)(valid ? cnt - 1 : cnt);
            const float mean2 = (float)(double)sum2 / (float)cnt;
            const float sq = mean * mean;
            const float var = fabsf(mean2 - sq);
            const int rounded = (int)floorf(mean + 0.5f);
            const int keep = fabsf(fd - mean) < a.cutoff ? rounded : unknown;
            const int filled = (a.fill && (double)var < a.sd_limit2) ? rounded : raw;
            const int res = !((float)cnt >= a.need) ? unknown : (valid ? keep : filled);
            if (REF) row_ptr<uint16_t>(a.out, y)[x] = (uint16_t)res;
            else row_ptr<int16_t>(a.out, y)[x] = (int16_t)res;
        }
        cnt -= rc[(r + 1) % P];
        sum -= rs[(r + 1) % P];
        sum2 -= rq[(r + 1) % P];
    }
}

struct DenoiseExt : DriverExt {
    stereo_denoise_params p{};
};

}  // namespace
}  // namespace stereo

using namespace capi;

extern "C" {

stereo_status stereo_denoise_configure(stereo_context* ctx, stereo_driver** out, const stereo_image_format* fmt,
                                       const stereo_denoise_params* p)
{
    if (!ctx || !out || !fmt || !p) return STEREO_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!format_valid(fmt)) return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "Invalid image format");
    if (fmt->pixel_type != STEREO_PIXEL_RGBA_U16 && fmt->pixel_type != STEREO_PIXEL_LUM_I16)
        return fail(ctx, STEREO_ERROR_UNSUPPORTED_PIXEL_TYPE, "de-noise: disparity maps are RGBA_U16 or LUM_I16");
    const int P = p->patch;
    if (P != 3 && P != 5 && P != 7 && P != 9 && P != 11 && P != 16)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "Invalid algorithm configuration.");
    for (int i = 0; i < 4; ++i)
        if (p->reserved[i]) return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "reserved parameters must be 0");
    if (fmt->width < P || fmt->height < P)
        return fail(ctx, STEREO_ERROR_INVALID_FORMAT, "de-noise: image %dx%d smaller than the %dx%d window", fmt->width,
                    fmt->height, P, P);
    stereo_status st;
    if ((st = set_device(ctx))) return st;
    stereo_driver* d = new stereo_driver(stereo_driver::DENOISE);
    d->ctx = ctx;
    d->fmt = *fmt;
    d->name = "Statistics disparity de-noise Filter";
    stereo::DenoiseExt* x = new stereo::DenoiseExt();
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

stereo_status stereo_denoise_process(stereo_driver* d, stereo_image* out, const stereo_image* in, stereo_stream stream)
{
    if (!d || d->kind != stereo_driver::DENOISE || !out || !in) return STEREO_ERROR_INVALID_ARGUMENT;
    stereo_context* ctx = d->ctx;
    const stereo::DenoiseExt* x = static_cast<const stereo::DenoiseExt*>(d->ext);
    const int W = d->fmt.width, H = d->fmt.height;
    const bool ref = d->fmt.pixel_type == STEREO_PIXEL_RGBA_U16;
    stereo_status st;
    if ((st = check_image(ctx, in, W, H, d->fmt.pixel_type, "disparity input"))) return st;
    if ((st = check_image(ctx, out, W, H, ref ? STEREO_PIXEL_LUM_U16 : STEREO_PIXEL_LUM_I16, "de-noised output")))
        return st;
    if ((st = set_device(ctx))) return st;
    GUARD(ctx, "stereo_denoise_process");
    const hipStream_t s = (hipStream_t)stream;
    if ((st = begin(d, s))) return st;
    stereo::DenoiseArgs a;
    a.in = view_of(in);
    a.out = view_of(out);
    const int P = x->p.patch;
    a.need = (float)(P * P) * x->p.support_ratio;
    a.cutoff = x->p.mean_cutoff;
    a.sd_limit2 = stereo::sd_limit2(a.cutoff);
    a.fill = x->p.fill ? 1 : 0;
    const dim3 grid((W + stereo::DN_TW - 1) / stereo::DN_TW, (H + stereo::DN_TH - 1) / stereo::DN_TH);
    // 4 x uint16 EDisparityId pixels in and uint16 ids out, or int16 d in and out
#define DN_LAUNCH(PP)                                                                                   \
    case PP:                                                                                            \
        if (ref) hipLaunchKernelGGL((stereo::denoise_kernel<PP, true>), grid, dim3(64), 0, s, a);      \
        else hipLaunchKernelGGL((stereo::denoise_kernel<PP, false>), grid, dim3(64), 0, s, a);         \
        break;
    switch (P) {
        DN_LAUNCH(3)
        DN_LAUNCH(5)
        DN_LAUNCH(7)
        DN_LAUNCH(9)
        DN_LAUNCH(11)
        DN_LAUNCH(16)
    default: return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "Invalid algorithm configuration.");
    }
#undef DN_LAUNCH
    HIP_TRY(ctx, hipGetLastError(), "de-noise launch");
    return end(d, s);
}

}  // extern "C"
