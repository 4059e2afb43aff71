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
// centre and a confidence logit.  Tiles whose confidence is low search the full
// range.
//
// Convolutions: 3x3, zero padding, dilation r, NHWC bf16 activations with 16 or 32
// channels, weights [Cout][9][Cin] bf16 (tap t = 3*ky + kx at offset
// ((ky-1) r, (kx-1) r)).  One wave computes 16 consecutive pixels of a row for all
// output channels with v_mfma_f32_16x16x16_bf16:
//   D[co][px] += A[co][ci] * B[ci][px]
//   A (weights): lane (col, grp) holds co = col, ci = 4 grp .. 4 grp + 3 -> one 8-byte
//                load per tap and channel block, all kept in VGPRs
//   B (pixels):  lane (col, grp) holds px = col, ci = 4 grp .. +3 -> one 8-byte load
//                of the NHWC pixel
//   D:           lane (col, grp) holds px = col, co = 4 grp .. +3 -> one 8-byte store
// The data gradient is the same kernel with the transposed weights [Cin][9][Cout]
// and the tap offsets negated.  The weight gradient uses the pixels as the K
// dimension: D[co][ci] += dY[px][co] * X[px + off][ci].
#include "common.hpp"
#include "kernels.hpp"

namespace stereo {
namespace cnn {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest, ties to even (finite inputs)
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float((h & 0xFFFFu) << 16); }
__device__ __forceinline__ bool bf_positive(uint32_t h) { return (h & 0x8000u) == 0 && (h & 0x7FFFu) != 0; }

template <int CIN, int COUT, int EPI>
__global__ __launch_bounds__(256) void conv3x3_kernel(CnnConvArgs a)
{
    constexpr int KC = CIN / 16, MB = COUT / 16;
    const int lane = threadIdx.x & 63;
    const int col = lane & 15, grp = lane >> 4;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * 4;
    v4s wr[MB][9][KC];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int t = 0; t < 9; ++t)
#pragma unroll
            for (int kc = 0; kc < KC; ++kc)
                wr[mb][t][kc] = *reinterpret_cast<const v4s*>(a.w + ((size_t)(mb * 16 + col) * 9 + t) * CIN + kc * 16 + 4 * grp);
    const int nbx = (a.W + 15) >> 4;
    const int total = a.N * a.H * nbx;
    const int step = a.dil * a.sign;
    // the input vectors of block b (9 taps x KC channel blocks), zero outside the image
    auto load = [&](int b, v4s (&in)[9][KC]) {
        const int px = (b % nbx) * 16 + col;
        const int r = b / nbx;
        const int yy = r % a.H, n = r / a.H;
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int sy = yy + (t / 3 - 1) * step;
            const int sx = px + (t % 3 - 1) * step;
            const bool ok = sy >= 0 && sy < a.H && sx >= 0 && sx < a.W;
            const uint16_t* src = ok ? a.x + (((size_t)n * a.H + sy) * a.W + sx) * CIN + 4 * grp : a.x;
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) {
                v4s bv = *reinterpret_cast<const v4s*>(src + kc * 16);
                in[t][kc] = ok ? bv : v4s{0, 0, 0, 0};
            }
        }
    };
    // software pipeline: the loads of the wave's next block are in flight while
    // the MFMAs of the current one run
    v4s cur[9][KC], nxt[9][KC];
    if (wave < total) load(wave, cur);
    for (int b = wave; b < total; b += nwaves) {
        if (b + nwaves < total) load(b + nwaves, nxt);
        const int bxi = b % nbx;
        const int r = b / nbx;
        const int yy = r % a.H;
        const int n = r / a.H;
        const int px = bxi * 16 + col;
        v4f acc[MB];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) acc[mb] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < 9; ++t)
#pragma unroll
            for (int kc = 0; kc < KC; ++kc)
#pragma unroll
                for (int mb = 0; mb < MB; ++mb)
                    acc[mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wr[mb][t][kc], cur[t][kc], acc[mb], 0, 0, 0);
#pragma unroll
        for (int t = 0; t < 9; ++t)
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) cur[t][kc] = nxt[t][kc];
        if (px < a.W) {
            const size_t o = (((size_t)n * a.H + yy) * a.W + px) * COUT;
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                const int c0 = mb * 16 + 4 * grp;
                float v[4] = {acc[mb][0], acc[mb][1], acc[mb][2], acc[mb][3]};
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
// (quarter-resolution disparity, NaN = none) relative to the filled coarse disparity:
//   Huber(e), e = out0 - (label - dfill);   BCE(out1, c) with c = (|e| <= tau)
// averaged over labelled pixels (scale = 1 / count).  Writes the bf16 gradient of
// the output and per-workgroup sums (huber, bce, count, correct).
__global__ __launch_bounds__(256) void loss_kernel(const float* out, const float* label, const float* dfill, uint16_t* g,
                                                   float* part, int P, float scale, float lambda, float tau)
{
    __shared__ float red[4][256];
    const int p = blockIdx.x * 256 + threadIdx.x;
    float hub = 0.f, bce = 0.f, cnt = 0.f, good = 0.f;
    if (p < P) {
        const float o0 = out[(size_t)p * 16], o1 = out[(size_t)p * 16 + 1];
        const float t = label[p] - dfill[p];
        float g0 = 0.f, g1 = 0.f;
        if (t == t) {
            const float e = o0 - t;
            const float ae = fabsf(e);
            hub = ae <= 1.f ? 0.5f * e * e : ae - 0.5f;
            const float c = ae <= tau ? 1.f : 0.f;
            const float sg = 1.f / (1.f + expf(-o1));
            bce = fmaxf(o1, 0.f) - o1 * c + log1pf(expf(-fabsf(o1)));
            g0 = fminf(fmaxf(e, -1.f), 1.f) * scale;
            g1 = lambda * (sg - c) * scale;
            cnt = 1.f;
            good = c;
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

// 4x4 box average with rounding: out(x, y) = (sum of in(4x.., 4y..) + 8) >> 4.
__global__ __launch_bounds__(256) void downsample4_kernel(View in, View out)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= out.w || y >= out.h) return;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t v = *reinterpret_cast<const uint32_t*>(row_ptr<const uint8_t>(in, 4 * y + j) + 4 * x);
        s += __builtin_amdgcn_sad_u8(v, 0u, 0u);
    }
    row_ptr<uint8_t>(out, y)[x] = (uint8_t)((s + 8) >> 4);
}

// Network input of frame n: ch0 = L/255, ch1 = R(x - dfill)/255 (bf16), ch2..15 = 0,
// where dfill is the coarse disparity with holes filled by the smaller of the
// nearest valid values to the left and right within 64 pixels (0 when none).
// dfill is written as float for the range kernel.  One workgroup per row; the
// row of coarse disparities is searched in LDS.
constexpr int CNN_MAX_W4 = 4096;

__global__ __launch_bounds__(256) void cnn_input_kernel(View L4, View R4, View dc, int n, uint16_t* x, float* dfill)
{
    // nearest valid index to the left (max-scan) and to the right (min-scan) of
    // every column, one row per workgroup
    __shared__ int16_t drow[CNN_MAX_W4];
    __shared__ int16_t lft[CNN_MAX_W4], rgt[CNN_MAX_W4];
    const int Y = blockIdx.x;
    const int W = L4.w, H = L4.h;
    const int16_t* src = row_ptr<const int16_t>(dc, Y);
    for (int X = threadIdx.x; X < W; X += 256) {
        const int16_t v = src[X];
        drow[X] = v;
        lft[X] = v >= 0 ? (int16_t)X : (int16_t)-1;
        rgt[X] = v >= 0 ? (int16_t)X : (int16_t)CNN_MAX_W4;
    }
    __syncthreads();
    for (int off = 1; off < W; off <<= 1) {
        int16_t a[CNN_MAX_W4 / 256], b[CNN_MAX_W4 / 256];
        int k = 0;
        for (int X = threadIdx.x; X < W; X += 256, ++k) {
            a[k] = X >= off ? max(lft[X], lft[X - off]) : lft[X];
            b[k] = X + off < W ? min(rgt[X], rgt[X + off]) : rgt[X];
        }
        __syncthreads();
        k = 0;
        for (int X = threadIdx.x; X < W; X += 256, ++k) {
            lft[X] = a[k];
            rgt[X] = b[k];
        }
        __syncthreads();
    }
    for (int X = threadIdx.x; X < W; X += 256) {
        int d = drow[X];
        if (d < 0) {
            const int l = lft[X], r = rgt[X];
            const int a = (l >= 0 && X - l <= 64) ? drow[l] : -1;
            const int b = (r < W && r - X <= 64) ? drow[r] : -1;
            d = a < 0 ? (b < 0 ? 0 : b) : (b < 0 ? a : min(a, b));
        }
        const size_t p = ((size_t)n * H + Y) * W + X;
        dfill[p] = (float)d;
        const int xr = clampi(X - d, 0, W - 1);
        const float l = row_ptr<const uint8_t>(L4, Y)[X] * (1.0f / 255.0f);
        const float r = row_ptr<const uint8_t>(R4, Y)[xr] * (1.0f / 255.0f);
        uint4* o = reinterpret_cast<uint4*>(x + p * 16);
        o[0] = make_uint4((uint32_t)f2bf(l) | ((uint32_t)f2bf(r) << 16), 0, 0, 0);
        o[1] = make_uint4(0, 0, 0, 0);
    }
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

// Per-tile disparity sets, one wave per tile.  Tile (tx, ty) covers quarter pixels
// [16 tx, 16 tx + 15] x [14 ty, 14 ty + 13] (64 x 56 full-resolution pixels); with a
// margin of one quarter pixel, every quarter pixel proposes the band
// [4 c - band/2 - 4, 4 c + band/2 + 4] around its centre c = dfill + out0.  The
// tile keeps the union as a mask of 8-disparity bins (bit b: [8b, 8b + 8)), or all
// bins below D when its smallest confidence logit is below the threshold.
__global__ __launch_bounds__(256) void band_mask_kernel(CnnBandArgs a)
{
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int ntiles = a.ntx * a.nty;
    if (t >= ntiles) return;
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
        const float c = 4.f * (a.dfill[p] + a.out[p * 16]);
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
        a.tile_evals[t] = fallback ? (1 << 16) : 0;  // the count is filled in by the item kernel
    }
}

// Work items of the bin masks (one workgroup of 1024 lanes): every run of set bins
// becomes items of at most `chunk` disparities; a block-wide scan places them.
__global__ __launch_bounds__(1024) void band_items_kernel(CnnBandArgs a)
{
    __shared__ int scan[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int ntiles = a.ntx * a.nty;
    for (int base = 0; base < ntiles; base += 1024) {
        const int t = base + threadIdx.x;
        uint32_t m = 0;
        int cnt = 0;
        if (t < ntiles) {
            m = a.tile_mask[t];
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
        if (t < ntiles) {
            int k = first, evals = 0, lo = 1 << 30, hi = 0;
            for (uint32_t r = m; r;) {
                const int r0 = __builtin_ctz(r);
                const uint32_t sh = r >> r0;
                const int len = (~sh == 0u) ? 32 - r0 : __builtin_ctz(~sh);
                const int ds0 = 8 * r0, de0 = 8 * (r0 + len);
                for (int ds = ds0; ds < de0; ds += a.chunk, ++k) {
                    BmItem it;
                    it.frame = 0;
                    it.tx = t % a.ntx;
                    it.ty = t / a.ntx;
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
            a.tiles[t] = bt;
            a.tile_evals[t] = evals | (a.tile_evals[t] & (1 << 16));
        }
        __syncthreads();
        if (threadIdx.x == 1023) carry += scan[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.nitems[0] = carry;
}

}  // namespace cnn

hipError_t launch_cnn_conv(int cin, int cout, int epi, const CnnConvArgs& a, hipStream_t s)
{
    // each wave loads the layer's weights into VGPRs once and then walks about four
    // 16-pixel blocks, so that the weight loads are amortised
    const int nbx = (a.W + 15) / 16;
    const long total = (long)a.N * a.H * nbx;
    long g = (total + 15) / 16;
    if (g > 1024) g = 1024;
    if (g < 1) g = 1;
    const int grid = (int)g;
    if (total < 1) return hipErrorInvalidValue;
#define CNN_CONV(CI, CO, E)                                                                             \
    if (cin == CI && cout == CO && epi == E) {                                                          \
        hipLaunchKernelGGL((cnn::conv3x3_kernel<CI, CO, E>), dim3(grid), dim3(256), 0, s, a);         \
        return hipGetLastError();                                                                       \
    }
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
    if (cin == 16 && cout == 32) hipLaunchKernelGGL((cnn::conv3x3_dw_kernel<16, 32>), dim3(grid), dim3(256), 0, s, a);
    else if (cin == 32 && cout == 32) hipLaunchKernelGGL((cnn::conv3x3_dw_kernel<32, 32>), dim3(grid), dim3(256), 0, s, a);
    else if (cin == 32 && cout == 16) hipLaunchKernelGGL((cnn::conv3x3_dw_kernel<32, 16>), dim3(grid), dim3(256), 0, s, a);
    else return hipErrorInvalidValue;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int n = cout * 9 * cin + cout;
    hipLaunchKernelGGL(cnn::reduce_parts_kernel, dim3((n + 255) / 256), dim3(256), 0, s, (const float*)a.part, grid, n, out);
    return hipGetLastError();
}

hipError_t launch_cnn_loss(const float* out, const float* label, const float* dfill, uint16_t* g, float* part, int P,
                           float scale, float lambda, float tau, hipStream_t s)
{
    hipLaunchKernelGGL(cnn::loss_kernel, dim3((P + 255) / 256), dim3(256), 0, s, out, label, dfill, g, part, P, scale,
                       lambda, tau);
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

hipError_t launch_downsample4(const View& in, const View& out, hipStream_t s)
{
    dim3 grid((out.w + 63) / 64, (out.h + 3) / 4);
    hipLaunchKernelGGL(cnn::downsample4_kernel, grid, dim3(256), 0, s, in, out);
    return hipGetLastError();
}

hipError_t launch_cnn_input(const View& L4, const View& R4, const View& dc, int n, uint16_t* x, float* dfill,
                            hipStream_t s)
{
    if (L4.w > cnn::CNN_MAX_W4) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cnn::cnn_input_kernel, dim3(L4.h), dim3(256), 0, s, L4, R4, dc, n, x, dfill);
    return hipGetLastError();
}

hipError_t launch_band_items(const CnnBandArgs& a, hipStream_t s)
{
    if (a.D > 256 || a.D % 8 || a.chunk < 8 || a.chunk % 8) return hipErrorInvalidValue;
    const int ntiles = a.ntx * a.nty;
    hipLaunchKernelGGL(cnn::band_mask_kernel, dim3((ntiles + 3) / 4), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(cnn::band_items_kernel, dim3(1), dim3(1024), 0, s, a);
    return hipGetLastError();
}

}  // namespace stereo
