/* Runs every oracle entry point on small and edge-case inputs, built with
 * -fsanitize=address,undefined by tests/test_oracle_sanitize.py -- TEST
 * INFRASTRUCTURE ONLY.  Exit 0 when every call returned its expected status. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cnn_oracle.h"
#include "stereo_oracle.h"

static uint32_t rng = 12345u;
static uint32_t next(void)
{
    rng ^= rng << 13;
    rng ^= rng >> 17;
    rng ^= rng << 5;
    return rng;
}

static int fails = 0;
#define EXPECT(cond)                                                  \
    do {                                                              \
        if (!(cond)) {                                                \
            fprintf(stderr, "failed: %s (line %d)\n", #cond, __LINE__); \
            ++fails;                                                  \
        }                                                             \
    } while (0)

static void run_matchers(int W, int H, int D)
{
    const size_t N = (size_t)W * H;
    uint8_t* L = malloc(N);
    uint8_t* R = malloc(N);
    for (size_t i = 0; i < N; ++i) { L[i] = (uint8_t)next(); R[i] = (uint8_t)next(); }
    uint64_t* cl = malloc(N * 8);
    uint64_t* cr = malloc(N * 8);
    int16_t* dL = malloc(N * 2);
    int16_t* dR = malloc(N * 2);
    int16_t* out = malloc(N * 2);
    int32_t* cost = malloc(N * 4);
    int32_t* vol = malloc(N * (size_t)D * 4);
    double* sc = malloc(N * 8);
    double* vd = malloc(N * (size_t)D * 8);
    EXPECT(orc_census(L, W, H, 9, 7, cl) == 0);
    EXPECT(orc_census(R, W, H, 9, 7, cr) == 0);
    uint8_t* c3 = malloc(N);
    EXPECT(orc_census3x3_ref(L, W, H, c3) == 0);
    EXPECT(orc_bm(ORC_COST_SAD, L, R, W, H, D, 9, 9, 1, 1, 1, dL, dR, out, cost, 3) == 0);
    EXPECT(orc_bm(ORC_COST_CENSUS, cl, cr, W, H, D, 5, 5, 9, 7, 1, dL, dR, out, NULL, 2) == 0);
    uint32_t* bins = malloc(N * 4);
    for (size_t i = 0; i < N; ++i) bins[i] = next();
    EXPECT(orc_bm_bins(ORC_COST_SAD, L, R, W, H, D, 5, 5, 1, 1, 1, bins, dL, dR, out, 2) == 0);
    free(bins);
    EXPECT(orc_bm_brute(ORC_COST_SAD, L, R, W, H, D, 7, 7, 1, 1, 1, dL, dR, out, vol) == 0);
    EXPECT(orc_ncc(L, R, W, H, D, 11, 11, 1, dL, dR, out, sc, vd, 2) == 0);
    EXPECT(orc_aw(2, L, R, W, H, D, 5, 1, dL, dR, out, cost, vol, 2) == 0);
    EXPECT(orc_aw(1, L, R, W, H, D, 3, 1, dL, dR, out, cost, NULL, 1) == 0);
    EXPECT(orc_crosscheck_plain(dL, dR, W, H, 1, out) == 0);
    uint16_t* a = malloc(N * 8);
    uint16_t* b = malloc(N * 8);
    uint16_t* o = malloc(N * 8);
    for (size_t i = 0; i < N * 4; ++i) { a[i] = (uint16_t)next(); b[i] = (uint16_t)next(); }
    uint16_t mn = 0, mx = 0;
    EXPECT(orc_crosscheck_ref(a, b, W, H, 1, 1, 0, o, &mn, &mx) == 0);
    EXPECT(orc_crosscheck_ref(a, b, W, H, 2, 0, 3, o, &mn, &mx) == 0);
    free(L); free(R); free(cl); free(cr); free(dL); free(dR); free(out); free(cost);
    free(vol); free(sc); free(vd); free(c3); free(a); free(b); free(o);
}

static void run_cnn(void)
{
    const int N = 2, H = 5, W = 7, ci = 3, co = 4;
    float* x = malloc(sizeof(float) * N * H * W * ci);
    float* w = malloc(sizeof(float) * co * 9 * ci);
    float* b = malloc(sizeof(float) * co);
    float* y = malloc(sizeof(float) * N * H * W * co);
    float* dx = malloc(sizeof(float) * N * H * W * ci);
    double* dw = malloc(sizeof(double) * co * 9 * ci);
    double* db = malloc(sizeof(double) * co);
    for (int i = 0; i < N * H * W * ci; ++i) x[i] = (float)(next() % 100) / 50.0f - 1.0f;
    for (int i = 0; i < co * 9 * ci; ++i) w[i] = (float)(next() % 100) / 50.0f - 1.0f;
    for (int i = 0; i < co; ++i) b[i] = 0.5f;
    for (int dil = 1; dil <= 6; dil += 5) {
        EXPECT(orc_conv3x3_fwd(x, N, H, W, ci, w, b, co, dil, y) == 0);
        EXPECT(orc_conv3x3_bwd(x, N, H, W, ci, w, y, co, dil, dx, dw, db) == 0);
        EXPECT(orc_conv3x3_bwd(x, N, H, W, ci, w, y, co, dil, NULL, dw, NULL) == 0);
    }
    EXPECT(orc_conv3x3_fwd(x, N, H, W, ci, w, b, co, 0, y) != 0);
    free(x); free(w); free(b); free(y); free(dx); free(dw); free(db);
}

/* The operators around the matchers: mip maps, de-noise, BlockMatcher1 (plain and band
 * search), rectification and colouring, at edge and ragged sizes. */
static void run_post(int W, int H)
{
    const size_t N = (size_t)W * H;
    uint8_t* L = malloc(N);
    uint8_t* R = malloc(N);
    int16_t* d = malloc(N * 2);
    uint16_t* ids = malloc(N * 8);
    uint16_t* o16 = malloc(N * 2);
    uint16_t* o4 = malloc(N * 8);
    uint8_t* o8 = malloc(N * 4);
    float* ff = malloc(N * 4);
    float* lv = malloc(N * 4);
    for (size_t i = 0; i < N; ++i) {
        L[i] = (uint8_t)next();
        R[i] = (uint8_t)next();
        d[i] = (int16_t)((next() % 5 == 0) ? -1 : (int)(next() % 200));
        ids[4 * i] = (uint16_t)((next() % 7 == 0) ? 0xFFFA : 1 + next() % 300);
        ids[4 * i + 1] = ids[4 * i + 2] = ids[4 * i + 3] = 0;
        ff[i] = (float)(next() % 400) * 0.5f;
    }
    EXPECT(orc_mip_level0_ref(ids, W, H, lv) == 0);
    EXPECT(orc_mip_level0_plain(d, W, H, lv) == 0);
    if (W >= 3 && H >= 3) EXPECT(orc_mip_next(lv, W, H, -1.0f, ff) == 0);
    const int P = W < H ? (W < 5 ? W : 5) : (H < 5 ? H : 5);
    EXPECT(orc_disp_denoise(0, d, W, H, P, 0.5f, 1.0f, 1, o16) == 0);
    EXPECT(orc_disp_denoise(1, ids, W, H, P, 0.3f, 2.0f, 0, o16) == 0);
    if (W >= 48) {
        static const int patches[4] = {5, 9, 11, 15};
        for (int k = 0; k < 4; ++k) {
            const int Pb = patches[k];
            if (H < Pb + 3) continue;
            EXPECT(orc_bm1(L, R, W, H, Pb, k & 1, 0, 64, k - 1, k != 3, 1.0f, 10.0f, 2.0f, NULL, 0, 0, o4) == 0);
            EXPECT(orc_bm1(L, R, W, H, Pb, 0, 0, 64, 0, 1, 0.0f, 0.0f, 0.0f, ids, W / 2 ? W / 2 : 1, H / 2 ? H / 2 : 1,
                           o4) == 0);
        }
    } else {
        EXPECT(orc_bm1(L, R, W, H, 5, 0, 0, 16, 0, 1, 0.0f, 0.0f, 0.0f, NULL, 0, 0, o4) != 0);
    }
    float rp[22] = {1.1f * W, 1.1f * W, 0.5f * W, 0.5f * H, -0.2f, 0.05f, 0.001f, -0.001f, 0.01f,
                    1.0f * W, 1.0f * W, 0.5f * W, 0.5f * H, 0.999f, -0.04f, 0.0f, 0.04f, 0.999f, 0.0f, 0.0f, 0.0f, 1.0f};
    EXPECT(orc_rectify(L, W, H, rp, W + 3, H + 1, o8) == 0);
    rp[13 + 8] = 0.0f; /* Z = 0 for the optical axis: infinite positions are outside */
    EXPECT(orc_rectify(L, W, H, rp, W, H, o8) == 0);
    static const uint8_t colors[28] = {0};
    for (int style = 0; style < 5; ++style) {
        EXPECT(orc_color_disparity(ids, 0, W, H, style, 1, 300, 7, colors, style == 0 || style == 2 ? (void*)o8 : (void*)o16) == 0);
        EXPECT(orc_color_disparity(ff, 2, W, H, style, 20, 100, 40000, colors, style == 0 || style == 2 ? (void*)o8 : (void*)o16) == 0);
    }
    EXPECT(orc_color_disparity(ids, 0, W, H, 0, 5, 5, 0, colors, o8) != 0);
    free(L); free(R); free(d); free(ids); free(o16); free(o4); free(o8); free(ff); free(lv);
}

int main(void)
{
    run_post(48, 8);    /* the smallest BlockMatcher1 image */
    run_post(97, 61);
    run_post(5, 5);
    run_post(131, 19);
    run_matchers(40, 30, 16);
    run_matchers(23, 17, 32);   /* D larger than the matchable width */
    run_matchers(64, 12, 8);    /* fewer rows than the window */
    run_cnn();
    /* rejected arguments */
    uint8_t img[16] = {0};
    EXPECT(orc_census(img, 4, 4, 8, 7, NULL) != 0);
    EXPECT(orc_bm(7, img, img, 4, 4, 8, 3, 3, 1, 1, 1, NULL, NULL, NULL, NULL, 1) != 0);
    if (fails) fprintf(stderr, "%d checks failed\n", fails);
    else printf("oracle sanitize run ok\n");
    return fails ? 1 : 0;
}
