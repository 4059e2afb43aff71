// Learned disparity proposal: a small convolutional network on BF16 MFMA (FP32
// accumulation) with hand-written forward, backward and Adam kernels, and the
// kernels that turn its output into per-tile disparity ranges for the work-list
// matcher.
//
// The network has no counterpart in the reference.  It stands where the
// reference's MatcherUseDisparity takes an initial disparity map and searches a
// band around it (BlockMatcherNCC.cu:894-1255: the band is the warp's min/max of
// the scaled initial disparities, widened by 4).  Here the initial map is the
// census matcher at 1/4 resolution; the network sees the quarter-resolution left
// image and the right image warped by that map, and predicts a residual for the
// centre and a confidence logit.  The bands are centred on the coarse map (the
// residual is an option); tiles whose confidence is low search the full range.
//
// Convolutions: 3x3, zero padding, dilation r, NHWC bf16 activations with 16 or 32
// channels, weights [Cout][9][Cin] bf16 (tap t = 3*ky + kx at offset
// ((ky-1) r, (kx-1) r)), r <= CNN_MAX_DIL.  A workgroup computes a tile of
// CNN_TILE_ROWS rows x 64 pixels: it stages the input tile with its halo of r pixels
// in LDS once (each input pixel is then read by 9 taps from LDS, not from L2), and
// wave w computes row w of the tile, four 16-pixel blocks sharing every weight
// operand, with v_mfma_f32_16x16x16_bf16:
//   D[co][px] += A[co][ci] * B[ci][px]
//   A (weights): lane (col, grp) holds co = col, ci = 4 grp .. 4 grp + 3 -> one 8-byte
//                load per tap and channel block, all kept in VGPRs
//   B (pixels):  lane (col, grp) holds px = col, ci = 4 grp .. +3 -> one ds_read_b64
//                of the staged NHWC pixel (pixel stride 2 CIN + 8 bytes: the 16
//                columns of a read fall on distinct bank pairs)
//   D:           lane (col, grp) holds px = col, co = 4 grp .. +3 -> one 8-byte store
// The data gradient is the same kernel with the transposed weights [Cin][9][Cout]
// and the tap offsets negated.  The weight gradient uses the pixels as the K
// dimension: D[co][ci] += dY[px][co] * X[px + off][ci].
#include "common.hpp"
#include "cnn_input.hpp"
#include "kernels.hpp"

namespace stereo {
namespace cnn {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float((h & 0xFFFFu) << 16); }
__device__ __forceinline__ bool bf_positive(uint32_t h) { return (h & 0x8000u) == 0 && (h & 0x7FFFu) != 0; }

template <int CIN, int COUT, int EPI>
__global__ __launch_bounds__(256) void conv3x3_kernel(CnnConvArgs a)
{
    constexpr int KC = CIN / 16, MB = COUT / 16;
    constexpr int PS = 2 * CIN + 8;  // staged pixel stride in bytes
    extern __shared__ __align__(16) uint8_t tile[];
    const int lane = threadIdx.x & 63;
    const int col = lane & 15, grp = lane >> 4;
    const int wv = threadIdx.x >> 6;
    const int ntx = (a.W + 63) >> 6, nty = (a.H + CNN_TILE_ROWS - 1) / CNN_TILE_ROWS;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int tx = b % ntx, r = b / ntx;
    const int ty = r % nty, n = r / nty;
    const int x0 = tx * 64, y0 = ty * CNN_TILE_ROWS;
    const int d = a.dil;
    const int TWp = 64 + 2 * d, THp = CNN_TILE_ROWS + 2 * d;

    // stage the input tile (zero outside the image).  A staged row holds TWp * QP
    // 16-byte chunks (at most 288); thread i owns chunks i and i + 256 of every row,
    // so the column arithmetic is done once and each row only adds its offset.  The
    // rows go in groups of four, all loads of a group in flight together; the
    // chunks are written to LDS as 8-byte halves.
    constexpr int QP = CIN / 8;  // 16-byte chunks per pixel
    const int cpr = TWp * QP;
    const int c0 = threadIdx.x, c1 = threadIdx.x + 256;
    const int rx0 = c0 / QP, q0 = c0 % QP, rx1 = c1 / QP, q1 = c1 % QP;
    const int gx0 = x0 - d + rx0, gx1 = x0 - d + rx1;
    const bool h0 = c0 < cpr, h1 = c1 < cpr;
    const bool ok0 = h0 && gx0 >= 0 && gx0 < a.W, ok1 = h1 && gx1 >= 0 && gx1 < a.W;
    const size_t col0 = (size_t)gx0 * CIN + q0 * 8, col1 = (size_t)gx1 * CIN + q1 * 8;
    const uint16_t* img = a.x + (size_t)n * a.H * a.W * CIN;
    for (int r0 = 0; r0 < THp; r0 += 4) {
        uint4 v0[4], v1[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gy = y0 - d + r0 + j;
            const bool rok = r0 + j < THp && gy >= 0 && gy < a.H;
            const uint16_t* row = img + (rok ? (size_t)gy * a.W * CIN : 0);
            v0[j] = (rok && ok0) ? *reinterpret_cast<const uint4*>(row + col0) : make_uint4(0u, 0u, 0u, 0u);
            v1[j] = (rok && ok1) ? *reinterpret_cast<const uint4*>(row + col1) : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (r0 + j >= THp) break;
            uint8_t* trow = tile + (size_t)(r0 + j) * TWp * PS;
            if (h0) {
                uint2* dst = reinterpret_cast<uint2*>(trow + rx0 * PS + q0 * 16);
                dst[0] = make_uint2(v0[j].x, v0[j].y);
                dst[1] = make_uint2(v0[j].z, v0[j].w);
            }
            if (h1) {
                uint2* dst = reinterpret_cast<uint2*>(trow + rx1 * PS + q1 * 16);
                dst[0] = make_uint2(v1[j].x, v1[j].y);
                dst[1] = make_uint2(v1[j].z, v1[j].w);
            }
        }
    }
    v4s wr[MB][9][KC];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int t = 0; t < 9; ++t)
#pragma unroll
            for (int kc = 0; kc < KC; ++kc)
                wr[mb][t][kc] = *reinterpret_cast<const v4s*>(a.w + ((size_t)(mb * 16 + col) * 9 + t) * CIN + kc * 16 + 4 * grp);
    __syncthreads();

    const int yy = y0 + wv;
    if (yy >= a.H) return;
    const int step = d * a.sign;
    v4f acc[4][MB];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) acc[k][mb] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int ry = wv + d + (t / 3 - 1) * step;
        const int rx = col + d + (t % 3 - 1) * step;
        const uint8_t* src = tile + (ry * TWp + rx) * PS + 8 * grp;
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) {
            v4s bv[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) bv[k] = *reinterpret_cast<const v4s*>(src + k * 16 * PS + kc * 32);
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int mb = 0; mb < MB; ++mb)
                    acc[k][mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wr[mb][t][kc], bv[k], acc[k][mb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int px = x0 + k * 16 + col;
        if (px >= a.W) continue;
        const size_t o = (((size_t)n * a.H + yy) * a.W + px) * COUT;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            const int c0 = mb * 16 + 4 * grp;
            float v[4] = {acc[k][mb][0], acc[k][mb][1], acc[k][mb][2], acc[k][mb][3]};
            if (a.bias) {
#pragma unroll
                for (int i = 0; i < 4; ++i) v[i] += a.bias[c0 + i];
            }
            if (EPI == CNN_EPI_F32) {
                *reinterpret_cast<float4*>(static_cast<float*>(a.y) + o + c0) = make_float4(v[0], v[1], v[2], v[3]);
            } else {
                if (EPI == CNN_EPI_RELU_BF16) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) v[i] = v[i] > 0.f ? v[i] : 0.f;
                } else {  // CNN_EPI_MASK_BF16: gradient through the ReLU of the input activation
                    const uint2 m = *reinterpret_cast<const uint2*>(a.mask + o + c0);
                    if (!bf_positive(m.x)) v[0] = 0.f;
                    if (!bf_positive(m.x >> 16)) v[1] = 0.f;
                    if (!bf_positive(m.y)) v[2] = 0.f;
                    if (!bf_positive(m.y >> 16)) v[3] = 0.f;
                }
                uint2 pk;
                pk.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
                pk.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
                *reinterpret_cast<uint2*>(static_cast<uint16_t*>(a.y) + o + c0) = pk;
            }
        }
    }
}

// Weight and bias gradients, one partial sum per workgroup:
//   part[blk][(co*9 + t)*CIN + ci] = sum over the workgroup's pixels of dY[p][co] X[p+off_t][ci]
//   part[blk][COUT*9*CIN + co]     = sum of dY[p][co]
// Wave w of the workgroup owns the 16x16 block (co block w / KC, ci block w % KC).
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void conv3x3_dw_kernel(CnnDwArgs a)
{
    constexpr int KC = CIN / 16, MB = COUT / 16;
    const int lane = threadIdx.x & 63;
    const int col = lane & 15, grp = lane >> 4;
    const int w = threadIdx.x >> 6;
    if (w >= MB * KC) return;
    const int mb = w / KC, kb = w % KC;
    const int co = mb * 16 + col;  // A row of this lane
    const int ci = kb * 16 + col;  // B column of this lane
    v4f acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    const int nbx = (a.W + 15) >> 4;
    const int total = a.N * a.H * nbx;
    for (int b = blockIdx.x; b < total; b += gridDim.x) {
        const int bxi = b % nbx;
        const int r = b / nbx;
        const int yy = r % a.H;
        const int n = r / a.H;
        const int p0 = bxi * 16 + 4 * grp;
        v4s av;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int px = p0 + i;
            const bool ok = px < a.W;
            const size_t idx = ok ? (((size_t)n * a.H + yy) * a.W + px) * COUT + co : 0;  // in-bounds address
            const uint16_t g = ok ? a.dy[idx] : (uint16_t)0;
            av[i] = (short)g;
            bsum += bf2f(g);
        }
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int sy = yy + (t / 3 - 1) * a.dil;
            const bool rok = sy >= 0 && sy < a.H;
            v4s bv;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int sx = p0 + i + (t % 3 - 1) * a.dil;
                const bool ok = rok && sx >= 0 && sx < a.W && p0 + i < a.W;
                const size_t idx = ok ? (((size_t)n * a.H + sy) * a.W + sx) * CIN + ci : 0;
                const short v = (short)a.x[idx];
                bv[i] = ok ? v : (short)0;
            }
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
        }
    }
    float* part = a.part + (size_t)blockIdx.x * (COUT * 9 * CIN + COUT);
    // D[m][n]: lane (col, grp) holds m = 4 grp + i (co), n = col (ci)
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) part[((mb * 16 + 4 * grp + i) * 9 + t) * CIN + kb * 16 + col] = acc[t][i];
    if (kb == 0) {
        bsum += __shfl_xor(bsum, 16);
        bsum += __shfl_xor(bsum, 32);
        if (grp == 0) part[COUT * 9 * CIN + co] = bsum;
    }
}

__global__ __launch_bounds__(256) void reduce_parts_kernel(const float* part, int nparts, int n, float* out)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    float s = 0.f;
    for (int b = 0; b < nparts; ++b) s += part[(size_t)b * n + j];
    out[j] = s;
}

// Loss over the network output (FP32, 16 channels; channel 0 = centre residual in
// quarter-resolution pixels, channel 1 = confidence logit) against the label
// (quarter-resolution disparity, NaN = none) relative to the filled coarse disparity
// and the confidence label (0 or 1, NaN = none):
//   Huber(e), e = out0 - (label - dfill)       where the disparity label is set
//   BCE(out1, c), c = the confidence label     where it is set, else c = (|e| <= tau)
// with scale = 1 / P.  Writes the bf16 gradient of the output and per-workgroup
// sums (huber, bce, labelled pixels, confident targets).
__global__ __launch_bounds__(256) void loss_kernel(const float* out, const float* label, const float* clabel,
                                                   const float* dfill, uint16_t* g, float* part, int P, float scale,
                                                   float lambda, float tau)
{
    __shared__ float red[4][256];
    const int p = blockIdx.x * 256 + threadIdx.x;
    float hub = 0.f, bce = 0.f, cnt = 0.f, good = 0.f;
    if (p < P) {
        const float o0 = out[(size_t)p * 16], o1 = out[(size_t)p * 16 + 1];
        const float t = label[p] - dfill[p];
        const float cl = clabel[p];
        float g0 = 0.f, g1 = 0.f;
        float c = -1.f;  // confidence target: the given label, else |e| <= tau, else none
        if (t == t) {
            const float e = o0 - t;
            const float ae = fabsf(e);
            hub = ae <= 1.f ? 0.5f * e * e : ae - 0.5f;
            g0 = fminf(fmaxf(e, -1.f), 1.f) * scale;
            c = ae <= tau ? 1.f : 0.f;
            cnt = 1.f;
        }
        if (cl == cl) c = cl;
        if (c >= 0.f) {
            const float sg = 1.f / (1.f + expf(-o1));
            bce = fmaxf(o1, 0.f) - o1 * c + log1pf(expf(-fabsf(o1)));
            g1 = lambda * (sg - c) * scale;
            good = c;
            if (!(t == t)) cnt = 1.f;
        }
        uint4 z = make_uint4(0, 0, 0, 0);
        z.x = (uint32_t)f2bf(g0) | ((uint32_t)f2bf(g1) << 16);
        uint4* gp = reinterpret_cast<uint4*>(g + (size_t)p * 16);
        gp[0] = z;
        gp[1] = make_uint4(0, 0, 0, 0);
    }
    red[0][threadIdx.x] = hub;
    red[1][threadIdx.x] = bce;
    red[2][threadIdx.x] = cnt;
    red[3][threadIdx.x] = good;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 4) part[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

// Adam on the FP32 master parameters of one layer (weights [COUT][9][CIN] then
// bias [COUT]); refreshes the bf16 forward weights [COUT][9][CIN] and the
// transposed data-gradient weights [CIN][9][COUT].
__global__ __launch_bounds__(256) void adam_kernel(float* param, const float* grad, float* m, float* v, int n, int cin,
                                                   int cout, float lr, float b1, float b2, float eps, float c1, float c2,
                                                   uint16_t* wf, uint16_t* wb)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    float pj = param[j];
    if (lr != 0.f) {
        const float gj = grad[j];
        const float mj = b1 * m[j] + (1.f - b1) * gj;
        const float vj = b2 * v[j] + (1.f - b2) * gj * gj;
        m[j] = mj;
        v[j] = vj;
        pj -= lr * (mj * c1) / (sqrtf(vj * c2) + eps);
        param[j] = pj;
    }
    const int nw = cout * 9 * cin;
    if (j < nw) {
        const int co = j / (9 * cin), rem = j - co * 9 * cin, t = rem / cin, ci = rem - t * cin;
        const uint16_t h = f2bf(pj);
        wf[j] = h;
        wb[((size_t)ci * 9 + t) * cout + co] = h;
    }
}

// ---- data preparation -----------------------------------------------------------

// Network input of frame n, one workgroup per row (cnn_input_row).
__global__ __launch_bounds__(256) void cnn_input_kernel(View L4, View R4, View dc, int n, uint16_t* x, float* dfill)
{
    __shared__ uint64_t valid[CNN_MAX_W4 / 64];
    cnn_input_row(L4, R4, dc, n, x, dfill, valid, blockIdx.x);
}

// Training label of quarter pixel (X, Y) of frame n from the full-resolution checked
// map: the mean of the valid disparities of its 4x4 block divided by 4, when at
// least 8 are valid and they span at most 4; NaN otherwise.
__global__ __launch_bounds__(256) void labels_kernel(View disp, int n, int W4, int H4, float* labels)
{
    const int X = blockIdx.x * 64 + (threadIdx.x & 63);
    const int Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= W4 || Y >= H4) return;
    int cnt = 0, sum = 0, mn = 1 << 30, mx = -1;
    for (int j = 0; j < 4; ++j) {
        const int16_t* row = row_ptr<const int16_t>(disp, 4 * Y + j) + 4 * X;
        for (int i = 0; i < 4; ++i) {
            const int d = row[i];
            if (d >= 0) {
                ++cnt;
                sum += d;
                mn = min(mn, d);
                mx = max(mx, d);
            }
        }
    }
    const float v = (cnt >= 8 && mx - mn <= 4) ? (float)sum / (float)(4 * cnt) : __int_as_float(0x7fc00000);
    labels[((size_t)n * H4 + Y) * W4 + X] = v;
}

// Confidence targets of frame n from a quarter-resolution LUM_U8 map: 0 -> 0, 1 -> 1,
// anything else -> NaN (no target).
__global__ __launch_bounds__(256) void conf_labels_kernel(View conf, int n, float* out)
{
    const int X = blockIdx.x * 64 + (threadIdx.x & 63);
    const int Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= conf.w || Y >= conf.h) return;
    const uint8_t v = row_ptr<const uint8_t>(conf, Y)[X];
    out[((size_t)n * conf.h + Y) * conf.w + X] = v <= 1 ? (float)v : __int_as_float(0x7fc00000);
}

// Per-tile disparity sets, one wave per tile.  Tile (tx, ty) covers quarter pixels
// [16 tx, 16 tx + 15] x [14 ty, 14 ty + 13] (64 x 56 full-resolution pixels); with a
// margin of one quarter pixel, every quarter pixel proposes the band
// [4 c - band/2 - 4, 4 c + band/2 + 4] around its centre c = dfill (+ the network's
// residual when centre_residual is set).  The tile keeps the union as a mask of
// 8-disparity bins (bit b: [8b, 8b + 8)), or all bins below D when its smallest
// confidence logit is below the threshold.
__global__ __launch_bounds__(256) void band_mask_kernel(CnnBandArgs a)
{
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int ntiles = a.ntx * a.nty;
    if (t < ntiles) {
        const int tx = t % a.ntx, ty = t / a.ntx;
        const int qx0 = max(0, 16 * tx - 1), qx1 = min(a.W4 - 1, 16 * tx + 16);
        const int qy0 = max(0, 14 * ty - 1), qy1 = min(a.H4 - 1, 14 * ty + 14);
        const int nx = qx1 - qx0 + 1, ny = qy1 - qy0 + 1;
        const int nbins = a.D >> 3;
        uint64_t mask = 0;
        float lmin = 1e30f;
        for (int k = lane; k < nx * ny && nx > 0 && ny > 0; k += 64) {
            const int qy = qy0 + k / nx, qx = qx0 + k % nx;
            const size_t p = (size_t)qy * a.W4 + qx;
            const float c = 4.f * (a.centre_residual ? a.dfill[p] + a.out[p * 16] : a.dfill[p]);
            lmin = fminf(lmin, a.out[p * 16 + 1]);
            const int lo = max(0, (int)floorf(c) - a.band / 2 - 4);
            const int hi = min(a.D - 1, (int)ceilf(c) + a.band / 2 + 4);
            if (lo <= hi) {
                const int b0 = lo >> 3, b1 = hi >> 3;
                mask |= ((2ull << b1) - 1ull) & ~((1ull << b0) - 1ull);
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            mask |= (uint64_t)__shfl_xor((unsigned long long)mask, o);
            lmin = fminf(lmin, __shfl_xor(lmin, o));
        }
        const uint64_t all = (nbins >= 64) ? ~0ull : ((1ull << nbins) - 1ull);
        const bool fallback = nx <= 0 || ny <= 0 || !(lmin >= a.logit_thresh) || mask == 0;
        if (lane == 0) {
            a.tile_mask[t] = (uint32_t)(fallback ? all : (mask & all));
            a.tile_evals[t] = fallback ? (1 << 16) : 0;  // the count is added with the items
        }
    }

}

// Tile bins from a caller's initial disparity map, one wave per tile (the role of
// MatcherUseDisparity, BlockMatcherNCC.cu:894-1255: scale = W / init width, the values
// scaled, min and max over the block, searched from min - 4 to max + 4).  Every pixel
// of the tile reads the initial map at (x / scale, y / scale), clamped; a valid value d
// (>= 0, or an id in [First, Last] as id - First) proposes scale x d.  The tile keeps
// the bins of [max(0, min - 4), min(D - 1, max + 4)], or all bins when none is valid.
__global__ __launch_bounds__(256) void init_band_mask_kernel(InitBandArgs ib, CnnBandArgs a)
{
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int ntiles = a.ntx * a.nty;
    if (t >= ntiles) return;
    const int tx = t % a.ntx, ty = t / a.ntx;
    const int x0 = tx * BM_TILE_W, y0 = ty * BM_TILE_H;
    const int nx = min(BM_TILE_W, ib.W - x0), ny = min(BM_TILE_H, ib.H - y0);
    int lo = 1 << 30, hi = -1;
    for (int k = lane; k < nx * ny; k += 64) {
        const int y = y0 + k / nx, x = x0 + k % nx;
        const int X = min(x / ib.scale, ib.init.w - 1), Y = min(y / ib.scale, ib.init.h - 1);
        int d;
        if (ib.ref_format) {
            const int id = row_ptr<const ushort4>(ib.init, Y)[X].x;
            d = (id >= 1 && id <= 0xFFEF) ? id - 1 : -1;
        } else {
            d = row_ptr<const int16_t>(ib.init, Y)[X];
        }
        if (d >= 0) {
            lo = min(lo, ib.scale * d);
            hi = max(hi, ib.scale * d);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o));
        hi = max(hi, __shfl_xor(hi, o));
    }
    const int nbins = a.D >> 3;
    const uint32_t all = nbins >= 32 ? 0xFFFFFFFFu : ((1u << nbins) - 1u);
    uint32_t mask = all;
    bool fallback = true;
    if (hi >= 0) {
        const int blo = max(0, lo - 4), bhi = min(a.D - 1, hi + 4);
        if (blo <= bhi) {
            const int b0 = blo >> 3, b1 = bhi >> 3;
            mask = (uint32_t)((((2ull << b1) - 1ull) & ~((1ull << b0) - 1ull)) & all);
            fallback = false;
        }
    }
    if (lane == 0) {
        a.tile_mask[t] = mask;
        a.tile_evals[t] = fallback ? (1 << 16) : 0;
    }
}

// The work list of the tiles' masks (one workgroup): every run of set bins is cut
// into items of at most `chunk` disparities; a block-wide scan places each tile's
// items and the tile table records them.
__global__ __launch_bounds__(1024) void band_items_kernel(CnnBandArgs a)
{
    const int ntiles = a.ntx * a.nty;
    __shared__ int scan[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < ntiles; base += 1024) {
        const int tt = base + threadIdx.x;
        uint32_t m = 0;
        int cnt = 0;
        if (tt < ntiles) {
            m = a.tile_mask[tt];
            for (uint32_t r = m; r;) {
                const int r0 = __builtin_ctz(r);
                const uint32_t sh = r >> r0;
                const int len = (~sh == 0u) ? 32 - r0 : __builtin_ctz(~sh);
                cnt += (8 * len + a.chunk - 1) / a.chunk;
                r &= ~((len >= 32 ? 0xFFFFFFFFu : ((1u << len) - 1u)) << r0);
            }
        }
        scan[threadIdx.x] = cnt;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int v = threadIdx.x >= off ? scan[threadIdx.x - off] : 0;
            __syncthreads();
            scan[threadIdx.x] += v;
            __syncthreads();
        }
        const int first = carry + scan[threadIdx.x] - cnt;
        if (tt < ntiles) {
            int k = first, evals = 0, lo = 1 << 30, hi = 0;
            for (uint32_t r = m; r;) {
                const int r0 = __builtin_ctz(r);
                const uint32_t sh = r >> r0;
                const int len = (~sh == 0u) ? 32 - r0 : __builtin_ctz(~sh);
                const int ds0 = 8 * r0, de0 = 8 * (r0 + len);
                for (int ds = ds0; ds < de0; ds += a.chunk, ++k) {
                    BmItem it;
                    it.frame = 0;
                    it.tx = tt % a.ntx;
                    it.ty = tt / a.ntx;
                    it.ds = ds;
                    it.de = min(de0, ds + a.chunk);
                    a.items[k] = it;
                }
                evals += de0 - ds0;
                lo = min(lo, ds0);
                hi = max(hi, de0);
                r &= ~((len >= 32 ? 0xFFFFFFFFu : ((1u << len) - 1u)) << r0);
            }
            BmTile bt;
            bt.lo = cnt ? lo : 0;
            bt.hi = cnt ? hi : 0;
            bt.first_item = first;
            bt.nitems = cnt;
            a.tiles[tt] = bt;
            a.tile_evals[tt] = evals | (a.tile_evals[tt] & (1 << 16));
        }
        __syncthreads();
        if (threadIdx.x == 1023) carry += scan[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.nitems[0] = carry;
        a.list_counter[0] = 0;
    }
}


// Inference through the five 16-channel layers in one launch.  A workgroup of 16 waves
// owns an output tile of FZ_TW x FZ_TH quarter pixels and keeps every intermediate
// activation of its tile in LDS: the input with the network's whole halo (the dilations
// add up to 10), then each layer's output over the region the remaining layers read
// (halo 9, 7, 3, 1, 0), ping-ponging between two staged buffers, so that nothing but
// the input and the final FP32 output crosses HBM.  Each layer is computed exactly as
// conv3x3_kernel computes it -- the same MFMA operands and tap order, the same epilogue
// (bias, ReLU, bf16 rounding) -- and activations outside the image are set to zero, the
// padding the next layer would see; the output therefore equals the layer-by-layer
// launches bit for bit.
constexpr int FZ_TW = 32, FZ_TH = 16, FZ_PS = 40;  // staged pixel stride: 16 bf16 channels + 8 bytes
constexpr int FZ_AW = FZ_TW + 20, FZ_AH = FZ_TH + 20;  // buffer A: the input region (halo 10)
constexpr int FZ_BW = FZ_TW + 18, FZ_BH = FZ_TH + 18;  // buffer B: layer 0's output region (halo 9)
constexpr int FZ_IL = 4;  // blocks of 16 output pixels a wave computes at once
constexpr int FZ_NW = 16;  // waves per workgroup: four per SIMD, the one workgroup a CU holds

// The halo of each layer's input (grow[5]: the tile itself).
constexpr int fz_grow(int l) { return l == 0 ? 10 : l == 1 ? 9 : l == 2 ? 7 : l == 3 ? 3 : l == 4 ? 1 : 0; }

typedef __bf16 fz_bf2 __attribute__((ext_vector_type(2)));
typedef float fz_f2 __attribute__((ext_vector_type(2)));

// Two floats to packed bf16, rounded to nearest even with v_cvt_pk_bf16_f32; equal to
// f2bf for every float but NaN (all 2^32 patterns compared on the device).
__device__ __forceinline__ uint32_t fz_pack_bf16(float lo, float hi)
{
    const fz_f2 v = {lo, hi};
    const fz_bf2 b = __builtin_convertvector(v, fz_bf2);
    return __builtin_bit_cast(uint32_t, b);
}

struct FzWeights {
    v4s w[9];
    float bias[4];
};

__device__ __forceinline__ void fz_load_weights(const CnnFusedArgs& a, int l, int col, int grp, FzWeights& wt)
{
#pragma unroll
    for (int t = 0; t < 9; ++t) wt.w[t] = *reinterpret_cast<const v4s*>(a.w[l] + ((size_t)col * 9 + t) * 16 + 4 * grp);
#pragma unroll
    for (int i = 0; i < 4; ++i) wt.bias[i] = a.bias[l][4 * grp + i];
}

// Layer L of the tile: every output region size and the dilation are constants.
template <int L>
__device__ __forceinline__ void fz_layer(const CnnFusedArgs& a, const uint8_t* in, uint8_t* out, const FzWeights& wt,
                                         int col, int grp, int wv, int x0, int y0, int n)
{
    constexpr int gi = fz_grow(L), go = fz_grow(L + 1), dil = gi - go;
    constexpr int rwi = FZ_TW + 2 * gi;  // input region row length
    constexpr int rwo = FZ_TW + 2 * go, rho = FZ_TH + 2 * go;
    constexpr int nbx = (rwo + 15) >> 4;
    constexpr int nitems = nbx * rho;
    // FZ_IL blocks of 16 output pixels per wave at once, each with its own accumulator,
    // their taps interleaved (a block past the end repeats the first one's reads and
    // writes nothing)
    for (int item0 = wv; item0 < nitems; item0 += FZ_NW * FZ_IL) {
        int oy[FZ_IL], ox[FZ_IL], base[FZ_IL];
        bool live[FZ_IL];
        v4f acc[FZ_IL];
#pragma unroll
        for (int u = 0; u < FZ_IL; ++u) {
            const int item = item0 + u * FZ_NW;
            live[u] = item < nitems;
            const int it = live[u] ? item : item0;
            oy[u] = it / nbx;
            ox[u] = (it % nbx) * 16 + col;
            const int oxr = min(ox[u], rwo - 1);  // lanes past the region read its last pixel
            base[u] = ((oy[u] + dil) * rwi + oxr + dil) * FZ_PS + 8 * grp;
            acc[u] = v4f{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int toff = ((t / 3 - 1) * dil * rwi + (t % 3 - 1) * dil) * FZ_PS;
            v4s bv[FZ_IL];
#pragma unroll
            for (int u = 0; u < FZ_IL; ++u) bv[u] = *reinterpret_cast<const v4s*>(in + base[u] + toff);
#pragma unroll
            for (int u = 0; u < FZ_IL; ++u)
                acc[u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wt.w[t], bv[u], acc[u], 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < FZ_IL; ++u) {
            if (!live[u] || ox[u] >= rwo) continue;
            // D: lane (col, grp) holds pixel ox, channels 4 grp .. 4 grp + 3
            const int gx = x0 - go + ox[u], gy = y0 - go + oy[u];
            const bool inside = gx >= 0 && gx < a.W && gy >= 0 && gy < a.H;
            float v[4] = {acc[u][0], acc[u][1], acc[u][2], acc[u][3]};
#pragma unroll
            for (int i = 0; i < 4; ++i) v[i] += wt.bias[i];
            if (L < 4) {
#pragma unroll
                for (int i = 0; i < 4; ++i) v[i] = (inside && v[i] > 0.f) ? v[i] : 0.f;
                uint2 pk;
                pk.x = fz_pack_bf16(v[0], v[1]);
                pk.y = fz_pack_bf16(v[2], v[3]);
                *reinterpret_cast<uint2*>(out + (oy[u] * rwo + ox[u]) * FZ_PS + 8 * grp) = pk;
            } else if (inside) {
                *reinterpret_cast<float4*>(a.out + (((size_t)n * a.H + gy) * a.W + gx) * 16 + 4 * grp) =
                    make_float4(v[0], v[1], v[2], v[3]);
            }
        }
    }
}

__global__ __launch_bounds__(64 * FZ_NW) void cnn_fused_forward_kernel(CnnFusedArgs a)
{
    __shared__ __align__(16) uint8_t bufA[FZ_AW * FZ_AH * FZ_PS];
    __shared__ __align__(16) uint8_t bufB[FZ_BW * FZ_BH * FZ_PS];
    const int lane = threadIdx.x & 63;
    const int col = lane & 15, grp = lane >> 4;
    const int wv = threadIdx.x >> 6;
    const int ntx = (a.W + FZ_TW - 1) / FZ_TW, nty = (a.H + FZ_TH - 1) / FZ_TH;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int tx = b % ntx, r = b / ntx;
    const int ty = r % nty, n = r / nty;
    const int x0 = tx * FZ_TW, y0 = ty * FZ_TH;
    const uint16_t* img = a.x + (size_t)n * a.H * a.W * 16;

    // a layer's weights are loaded while the layer before it computes
    FzWeights w0, w1;
    fz_load_weights(a, 0, col, grp, w0);
    // stage the input region (halo 10), zero outside the image: 4 x 8-byte chunks per pixel
    for (int i = threadIdx.x; i < FZ_AW * FZ_AH * 4; i += 64 * FZ_NW) {
        const int px = i >> 2, q = i & 3;
        const int gx = x0 - 10 + px % FZ_AW, gy = y0 - 10 + px / FZ_AW;
        uint2 v = make_uint2(0u, 0u);
        if (gx >= 0 && gx < a.W && gy >= 0 && gy < a.H)
            v = *reinterpret_cast<const uint2*>(img + ((size_t)gy * a.W + gx) * 16 + 4 * q);
        *reinterpret_cast<uint2*>(bufA + px * FZ_PS + 8 * q) = v;
    }
    __syncthreads();
    fz_load_weights(a, 1, col, grp, w1);
    fz_layer<0>(a, bufA, bufB, w0, col, grp, wv, x0, y0, n);
    __syncthreads();
    fz_load_weights(a, 2, col, grp, w0);
    fz_layer<1>(a, bufB, bufA, w1, col, grp, wv, x0, y0, n);
    __syncthreads();
    fz_load_weights(a, 3, col, grp, w1);
    fz_layer<2>(a, bufA, bufB, w0, col, grp, wv, x0, y0, n);
    __syncthreads();
    fz_load_weights(a, 4, col, grp, w0);
    fz_layer<3>(a, bufB, bufA, w1, col, grp, wv, x0, y0, n);
    __syncthreads();
    fz_layer<4>(a, bufA, bufB, w0, col, grp, wv, x0, y0, n);
}

}  // namespace cnn

hipError_t launch_cnn_conv(int cin, int cout, int epi, const CnnConvArgs& a, hipStream_t s)
{
    // one workgroup per tile of CNN_TILE_ROWS x 64 output pixels
    if (a.N < 1 || a.H < 1 || a.W < 1 || a.dil < 1 || a.dil > CNN_MAX_DIL) return hipErrorInvalidValue;
    const long ntiles = (long)a.N * ((a.H + CNN_TILE_ROWS - 1) / CNN_TILE_ROWS) * ((a.W + 63) / 64);
    if (ntiles > 0x7fffffffL) return hipErrorInvalidValue;
    const int grid = (int)ntiles;
    const size_t lds = (size_t)(64 + 2 * a.dil) * (CNN_TILE_ROWS + 2 * a.dil) * (2 * cin + 8);
#define CNN_CONV(CI, CO, E)                                                                             \
    if (cin == CI && cout == CO && epi == E) {                                                          \
        hipLaunchKernelGGL((cnn::conv3x3_kernel<CI, CO, E>), dim3(grid), dim3(256), lds, s, a);       \
        return hipGetLastError();                                                                       \
    }
    CNN_CONV(16, 16, CNN_EPI_RELU_BF16)
    CNN_CONV(16, 16, CNN_EPI_F32)
    CNN_CONV(16, 16, CNN_EPI_MASK_BF16)
    CNN_CONV(16, 32, CNN_EPI_RELU_BF16)
    CNN_CONV(32, 32, CNN_EPI_RELU_BF16)
    CNN_CONV(32, 16, CNN_EPI_F32)
    CNN_CONV(16, 32, CNN_EPI_F32)
    CNN_CONV(32, 32, CNN_EPI_F32)
    CNN_CONV(32, 32, CNN_EPI_MASK_BF16)
    CNN_CONV(16, 32, CNN_EPI_MASK_BF16)
    CNN_CONV(32, 16, CNN_EPI_RELU_BF16)
#undef CNN_CONV
    return hipErrorInvalidValue;
}

int cnn_dw_parts(const CnnDwArgs& a)
{
    const long total = (long)a.N * a.H * ((a.W + 15) / 16);
    long g = (total + 7) / 8;
    if (g > 512) g = 512;
    if (g < 1) g = 1;
    return (int)g;
}

hipError_t launch_cnn_dw(int cin, int cout, const CnnDwArgs& a, float* out, hipStream_t s)
{
    const int grid = cnn_dw_parts(a);
    if (cin == 16 && cout == 16) hipLaunchKernelGGL((cnn::conv3x3_dw_kernel<16, 16>), dim3(grid), dim3(256), 0, s, a);
    else if (cin == 16 && cout == 32) hipLaunchKernelGGL((cnn::conv3x3_dw_kernel<16, 32>), dim3(grid), dim3(256), 0, s, a);
    else if (cin == 32 && cout == 32) hipLaunchKernelGGL((cnn::conv3x3_dw_kernel<32, 32>), dim3(grid), dim3(256), 0, s, a);
    else if (cin == 32 && cout == 16) hipLaunchKernelGGL((cnn::conv3x3_dw_kernel<32, 16>), dim3(grid), dim3(256), 0, s, a);
    else return hipErrorInvalidValue;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int n = cout * 9 * cin + cout;
    hipLaunchKernelGGL(cnn::reduce_parts_kernel, dim3((n + 255) / 256), dim3(256), 0, s, (const float*)a.part, grid, n, out);
    return hipGetLastError();
}

hipError_t launch_cnn_loss(const float* out, const float* label, const float* clabel, const float* dfill, uint16_t* g,
                           float* part, int P, float scale, float lambda, float tau, hipStream_t s)
{
    hipLaunchKernelGGL(cnn::loss_kernel, dim3((P + 255) / 256), dim3(256), 0, s, out, label, clabel, dfill, g, part, P,
                       scale, lambda, tau);
    return hipGetLastError();
}

hipError_t launch_cnn_conf_labels(const View& conf, int n, float* out, hipStream_t s)
{
    dim3 grid((conf.w + 63) / 64, (conf.h + 3) / 4);
    hipLaunchKernelGGL(cnn::conf_labels_kernel, grid, dim3(256), 0, s, conf, n, out);
    return hipGetLastError();
}

hipError_t launch_cnn_labels(const View& disp, int n, int W4, int H4, float* labels, hipStream_t s)
{
    dim3 grid((W4 + 63) / 64, (H4 + 3) / 4);
    hipLaunchKernelGGL(cnn::labels_kernel, grid, dim3(256), 0, s, disp, n, W4, H4, labels);
    return hipGetLastError();
}

hipError_t launch_cnn_adam(float* param, const float* grad, float* m, float* v, int n, int cin, int cout, float lr,
                           float b1, float b2, float eps, float c1, float c2, uint16_t* wf, uint16_t* wb, hipStream_t s)
{
    hipLaunchKernelGGL(cnn::adam_kernel, dim3((n + 255) / 256), dim3(256), 0, s, param, grad, m, v, n, cin, cout, lr, b1,
                       b2, eps, c1, c2, wf, wb);
    return hipGetLastError();
}

hipError_t launch_cnn_fused_forward(const CnnFusedArgs& a, hipStream_t s)
{
    if (a.N < 1 || a.H < 1 || a.W < 1) return hipErrorInvalidValue;
    const long ntiles = (long)a.N * ((a.H + cnn::FZ_TH - 1) / cnn::FZ_TH) * ((a.W + cnn::FZ_TW - 1) / cnn::FZ_TW);
    if (ntiles > 0x7fffffffL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cnn::cnn_fused_forward_kernel, dim3((int)ntiles), dim3(64 * cnn::FZ_NW), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_cnn_input(const View& L4, const View& R4, const View& dc, int n, uint16_t* x, float* dfill,
                            hipStream_t s)
{
    if (L4.w > CNN_MAX_W4) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cnn::cnn_input_kernel, dim3(L4.h), dim3(256), 0, s, L4, R4, dc, n, x, dfill);
    return hipGetLastError();
}

hipError_t launch_init_band(const InitBandArgs& ib, const CnnBandArgs& a, hipStream_t s)
{
    if (a.D > 256 || a.D % 8 || a.chunk < 8 || a.chunk % 8 || ib.scale < 1 || !a.list_counter) return hipErrorInvalidValue;
    const int ntiles = a.ntx * a.nty;
    if (ntiles < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cnn::init_band_mask_kernel, dim3((ntiles + 3) / 4), dim3(256), 0, s, ib, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(cnn::band_items_kernel, dim3(1), dim3(1024), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_init_band_mask(const InitBandArgs& ib, const CnnBandArgs& a, hipStream_t s)
{
    if (a.D > 256 || a.D % 8 || ib.scale < 1) return hipErrorInvalidValue;
    const int ntiles = a.ntx * a.nty;
    if (ntiles < 1 || !a.tile_mask || !a.tile_evals) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cnn::init_band_mask_kernel, dim3((ntiles + 3) / 4), dim3(256), 0, s, ib, a);
    return hipGetLastError();
}

hipError_t launch_band_items(const CnnBandArgs& a, hipStream_t s)
{
    if (a.D > 256 || a.D % 8 || a.chunk < 8 || a.chunk % 8) return hipErrorInvalidValue;
    const int ntiles = a.ntx * a.nty;
    if (ntiles < 1 || !a.list_counter) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cnn::band_mask_kernel, dim3((ntiles + 3) / 4), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(cnn::band_items_kernel, dim3(1), dim3(1024), 0, s, a);
    return hipGetLastError();
}

}  // namespace stereo
