/*
 * Seeded synthetic stereo pairs for the block-matcher workloads.
 *
 * The reference ships no test images (its only demo loads a Middlebury file from a
 * hard-coded Windows path, CluTec.Cuda.UserTest1/UsertTest1.cpp:59-63), so every
 * workload here is generated from a seed.  The generator is plain C so that the
 * host, the tests and the CPU baseline consume byte-identical inputs.
 *
 * Recipe (all steps in double precision, raster order, one PRNG stream):
 *   1. PRNG: xoshiro256** whose four state words come from SplitMix64(seed).
 *   2. Left image: uniform bytes (top 8 bits of each draw), Gaussian blur
 *      sigma 1.5 / radius 4 (separable, normalised, edge-clamped), then a
 *      min/max rescale to 0..255 rounded with llround.
 *   3. Disparity field: background value, then `nregions` axis-aligned
 *      rectangles painted in order.  Size U[W/20, W/4] x U[H/20, H/4], position
 *      uniform inside the image.  A region is constant (d in U[0, dmax]) or a
 *      slanted plane d0 + g.(p - centre) with |g| <= 0.05 px/px, clamped to
 *      [0, dmax] and rounded per pixel.
 *   4. Right image: forward warp R(x - d, y) = L(x, y); on collisions the larger
 *      disparity wins.  Never-written (disoccluded) pixels get fresh uniform
 *      bytes.  Finally every right pixel gets N(0, 2^2) noise (Box-Muller, one
 *      normal per two uniforms), rounded and clamped to 0..255.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "synth.h"

typedef struct { uint64_t s[4]; } xoshiro_t;

static uint64_t splitmix64(uint64_t* x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t xo_next(xoshiro_t* r)
{
    const uint64_t result = rotl64(r->s[1] * 5, 7) * 9;
    const uint64_t t = r->s[1] << 17;
    r->s[2] ^= r->s[0];
    r->s[3] ^= r->s[1];
    r->s[1] ^= r->s[2];
    r->s[0] ^= r->s[3];
    r->s[2] ^= t;
    r->s[3] = rotl64(r->s[3], 45);
    return result;
}

static void xo_seed(xoshiro_t* r, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) r->s[i] = splitmix64(&x);
}

/* uniform in [0, 1) with 53 random bits */
static double xo_u01(xoshiro_t* r) { return (double)(xo_next(r) >> 11) * (1.0 / 9007199254740992.0); }

/* uniform in (0, 1] (safe for log) */
static double xo_u01_open0(xoshiro_t* r) { return (double)((xo_next(r) >> 11) + 1) * (1.0 / 9007199254740992.0); }

/* integer uniform in [a, b] (b >= a) */
static int xo_range(xoshiro_t* r, int a, int b)
{
    uint64_t span = (uint64_t)(b - a) + 1u;
    return a + (int)(xo_next(r) % span);
}

static uint8_t xo_byte(xoshiro_t* r) { return (uint8_t)(xo_next(r) >> 56); }

static double xo_normal(xoshiro_t* r)
{
    double u1 = xo_u01_open0(r);
    double u2 = xo_u01(r);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

int stereo_synth_version(void) { return 1; }

int stereo_synth_pair(uint64_t seed, int W, int H, int kind, int nregions, int dmax, int dbg,
                      uint8_t* L, uint8_t* R, int16_t* gt)
{
    if (W < 20 || H < 20 || nregions < 0 || dmax < 0 || dmax > 32767 || dbg < 0 || dbg > dmax)
        return -1;
    if (kind < SYNTH_CONSTANT || kind > SYNTH_MIXED) return -1;
    if (!L || !R) return -1;

    const size_t N = (size_t)W * (size_t)H;
    double* a = (double*)malloc(N * sizeof(double));
    double* b = (double*)malloc(N * sizeof(double));
    int16_t* d = (int16_t*)malloc(N * sizeof(int16_t));
    int16_t* z = (int16_t*)malloc(N * sizeof(int16_t));
    if (!a || !b || !d || !z) { free(a); free(b); free(d); free(z); return -2; }

    xoshiro_t rng;
    xo_seed(&rng, seed);

    /* 1. noise */
    for (size_t i = 0; i < N; ++i) a[i] = (double)xo_byte(&rng);

    /* 2. separable Gaussian blur, sigma 1.5, radius 4, clamped edges */
    double k[9];
    double ks = 0.0;
    for (int i = -4; i <= 4; ++i) { k[i + 4] = exp(-(double)(i * i) / (2.0 * 1.5 * 1.5)); ks += k[i + 4]; }
    for (int i = 0; i < 9; ++i) k[i] /= ks;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double s = 0.0;
            for (int i = -4; i <= 4; ++i) s += k[i + 4] * a[(size_t)y * W + clampi(x + i, 0, W - 1)];
            b[(size_t)y * W + x] = s;
        }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double s = 0.0;
            for (int i = -4; i <= 4; ++i) s += k[i + 4] * b[(size_t)clampi(y + i, 0, H - 1) * W + x];
            a[(size_t)y * W + x] = s;
        }
    double mn = a[0], mx = a[0];
    for (size_t i = 1; i < N; ++i) { if (a[i] < mn) mn = a[i]; if (a[i] > mx) mx = a[i]; }
    const double scale = (mx > mn) ? 255.0 / (mx - mn) : 0.0;
    for (size_t i = 0; i < N; ++i) L[i] = (uint8_t)clampi((int)llround((a[i] - mn) * scale), 0, 255);

    /* 3. disparity field */
    for (size_t i = 0; i < N; ++i) d[i] = (int16_t)dbg;
    for (int r = 0; r < nregions; ++r) {
        const int w = xo_range(&rng, W / 20, W / 4);
        const int h = xo_range(&rng, H / 20, H / 4);
        const int x0 = xo_range(&rng, 0, W - w);
        const int y0 = xo_range(&rng, 0, H - h);
        int slanted = (kind == SYNTH_SLANTED);
        if (kind == SYNTH_MIXED) slanted = xo_u01(&rng) < 0.5;
        const int d0 = xo_range(&rng, 0, dmax);
        if (!slanted) {
            for (int y = y0; y < y0 + h; ++y)
                for (int x = x0; x < x0 + w; ++x) d[(size_t)y * W + x] = (int16_t)d0;
        } else {
            const double th = xo_u01(&rng) * 6.283185307179586;
            const double m = xo_u01(&rng) * 0.05;
            const double gx = m * cos(th), gy = m * sin(th);
            const double cx = x0 + 0.5 * (w - 1), cy = y0 + 0.5 * (h - 1);
            for (int y = y0; y < y0 + h; ++y)
                for (int x = x0; x < x0 + w; ++x) {
                    double v = (double)d0 + gx * ((double)x - cx) + gy * ((double)y - cy);
                    if (v < 0.0) v = 0.0;
                    if (v > (double)dmax) v = (double)dmax;
                    d[(size_t)y * W + x] = (int16_t)llround(v);
                }
        }
    }

    /* 4. forward warp with z-buffer (larger disparity wins) */
    for (size_t i = 0; i < N; ++i) { z[i] = -1; R[i] = 0; }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const int dv = d[(size_t)y * W + x];
            const int xr = x - dv;
            if (xr < 0) continue;
            const size_t j = (size_t)y * W + xr;
            if (dv > z[j]) { z[j] = (int16_t)dv; R[j] = L[(size_t)y * W + x]; }
        }
    for (size_t i = 0; i < N; ++i)
        if (z[i] < 0) R[i] = xo_byte(&rng);
    for (size_t i = 0; i < N; ++i) {
        const double v = (double)R[i] + 2.0 * xo_normal(&rng);
        R[i] = (uint8_t)clampi((int)llround(v), 0, 255);
    }

    if (gt) memcpy(gt, d, N * sizeof(int16_t));
    free(a); free(b); free(d); free(z);
    return 0;
}
