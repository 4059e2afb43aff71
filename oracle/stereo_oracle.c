/*
 * CPU oracle for the stereo disparity hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * It is the checker for the HIP engine (tests/) and the CPU baseline timed by
 * bench.py.  The engine never links or calls it.  Arrays are row-major with
 * pitch == width.
 *
 * Plain specification shared by every matcher here (the reference fixes none of
 * these rules for its own kernels because its results depend on block tiling):
 *   - a left pixel (x, y) is matchable iff bx <= x < W-bx and by <= y < H-by,
 *     where bx/by are the aggregation radius plus the census radius;
 *   - candidate disparity d is valid iff 0 <= d < D and x-d >= bx, the right
 *     window being centred on (x-d, y);
 *   - winner-take-all keeps the best cost, the lowest d on ties (the reference
 *     scans d upward with a strict comparison: BlockMatcherNCC.cu:298,
 *     BlockMatcherAW1.cu:294, BlockMatcherAW2.cu:167);
 *   - the right-to-left map takes, for right pixel xr, the best of the costs
 *     C(xr+d, y, d) of the same cost volume, lowest d on ties;
 *   - the checked map keeps dL where |dR(x-dL) - dL| <= tol, else -1.
 */
#include "stereo_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_version(void) { return 3; }

static inline int popc64(uint64_t v) { return __builtin_popcountll(v); }

/* ---------------------------------------------------------------------------------
 * Census.  Reference: CensusTransform.cu:103-119 compares the 8 neighbours of a 3x3
 * window with strict '>' against the centre and packs them raster-ordered, MSB
 * first, into a uint8; only pixels with IsInside(x, y, 1, 1) are written
 * (CensusTransform.cu:64-68).  The generalisation keeps strict '>', raster order
 * and MSB-first: raster position k of an n = cw*ch window maps to bit n-1-k (the
 * centre bit is always 0).  Pixels whose window leaves the image are 0.
 * --------------------------------------------------------------------------------- */
int orc_census(const uint8_t* img, int W, int H, int cw, int ch, uint64_t* out)
{
    if (!img || !out || cw < 1 || ch < 1 || (cw & 1) == 0 || (ch & 1) == 0 || cw * ch > 64) return -1;
    const int rx = cw / 2, ry = ch / 2, n = cw * ch;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            uint64_t v = 0;
            if (x >= rx && x < W - rx && y >= ry && y < H - ry) {
                const int c = img[(size_t)y * W + x];
                int k = 0;
                for (int dy = -ry; dy <= ry; ++dy)
                    for (int dx = -rx; dx <= rx; ++dx, ++k)
                        if (img[(size_t)(y + dy) * W + (x + dx)] > c) v |= 1ull << (n - 1 - k);
            }
            out[(size_t)y * W + x] = v;
        }
    return 0;
}

int orc_census3x3_ref(const uint8_t* img, int W, int H, uint8_t* out)
{
    static const int bits[3][3] = {{0x80, 0x40, 0x20}, {0x10, 0, 0x08}, {0x04, 0x02, 0x01}};
    if (!img || !out) return -1;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            int v = 0;
            if (x >= 1 && x < W - 1 && y >= 1 && y < H - 1) {
                const int c = img[(size_t)y * W + x];
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx)
                        if (bits[dy + 1][dx + 1] && img[(size_t)(y + dy) * W + (x + dx)] > c) v |= bits[dy + 1][dx + 1];
            }
            out[(size_t)y * W + x] = (uint8_t)v;
        }
    return 0;
}

/* ---------------------------------------------------------------------------------
 * Per-pixel matching costs.  AbsoluteDifference: BlockMatcherNCC.cu:147-151,
 * BlockMatcherAW2.cu:129-133.  CensusDifference (popcount of XOR):
 * BlockMatcherNCC.cu:153-158 (here over 64-bit descriptors).
 * --------------------------------------------------------------------------------- */
static inline int cost_at(int kind, const void* L, const void* R, size_t iL, size_t iR)
{
    if (kind == ORC_COST_SAD) {
        const int a = ((const uint8_t*)L)[iL], b = ((const uint8_t*)R)[iR];
        return a > b ? a - b : b - a;
    }
    return popc64(((const uint64_t*)L)[iL] ^ ((const uint64_t*)R)[iR]);
}

/* Horizontal running box sum of one cost row for disparity d over output columns
 * [xs, xe); h[x] receives sum_{i=-ax..ax} c(x+i, y, d).  Follows the slide-in /
 * slide-out update of AlgoArraySum (Kernel.Algo.ArraySum.h:149-189). */
static void hsum_row(int kind, const void* L, const void* R, int W, int y, int d, int ax, int xs,
                     int xe, int32_t* h)
{
    const size_t row = (size_t)y * W;
    int s = 0;
    for (int i = -ax; i <= ax; ++i) s += cost_at(kind, L, R, row + xs + i, row + xs + i - d);
    h[xs] = s;
    for (int x = xs + 1; x < xe; ++x) {
        s += cost_at(kind, L, R, row + x + ax, row + x + ax - d) -
             cost_at(kind, L, R, row + x - ax - 1, row + x - ax - 1 - d);
        h[x] = s;
    }
}

static void crosscheck_rows(const int16_t* dL, const int16_t* dR, int W, int y0, int y1, int tol,
                            int16_t* out)
{
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            const int d = dL[i];
            int16_t v = -1;
            if (d >= 0 && x - d >= 0) {
                const int r = dR[i - d];
                if (r >= 0 && abs(r - d) <= tol) v = (int16_t)d;
            }
            out[i] = v;
        }
}

int orc_crosscheck_plain(const int16_t* dL, const int16_t* dR, int W, int H, int tol, int16_t* out)
{
    if (!dL || !dR || !out) return -1;
    crosscheck_rows(dL, dR, W, 0, H, tol, out);
    return 0;
}

static int bm_args_ok(int kind, const void* L, const void* R, int W, int H, int D, int ww, int wh,
                      int cw, int ch)
{
    if (!L || !R || W <= 0 || H <= 0 || D <= 0 || D > 32767) return 0;
    if (kind != ORC_COST_SAD && kind != ORC_COST_CENSUS) return 0;
    if (ww < 1 || wh < 1 || (ww & 1) == 0 || (wh & 1) == 0) return 0;
    if (cw < 1 || ch < 1 || (cw & 1) == 0 || (ch & 1) == 0) return 0;
    return 1;
}

/* bins (optional, per left pixel): only candidates whose 8-disparity bin d >> 3 is set
 * take part, in both winner-take-alls -- the band search around an initial map and
 * of the learned proposal, whose bins are per 64 x 56 tile and expanded to pixels by
 * the caller.  The right view's candidate d of right pixel xr is the left pixel
 * xr + d's, so it obeys that left pixel's mask. */
static int bm_impl(int kind, const void* L, const void* R, int W, int H, int D, int ww, int wh, int cw,
                   int ch, int tol, const uint32_t* bins, int16_t* dL,
                   int16_t* dR, int16_t* out, int32_t* cost_out, int nthreads)
{
    if (!bm_args_ok(kind, L, R, W, H, D, ww, wh, cw, ch) || !dL || !dR) return -1;
    const int ax = ww / 2, ay = wh / 2;
    const int bx = ax + cw / 2, by = ay + ch / 2;
    const size_t N = (size_t)W * H;
    int32_t* bestL = (int32_t*)malloc(N * sizeof(int32_t));
    int32_t* bestR = (int32_t*)malloc(N * sizeof(int32_t));
    if (!bestL || !bestR) { free(bestL); free(bestR); return -2; }
    for (size_t i = 0; i < N; ++i) { bestL[i] = INT_MAX; bestR[i] = INT_MAX; dL[i] = -1; dR[i] = -1; }

    const int ylo = by, yhi = H - by;
    if (nthreads < 1) nthreads = 1;
    int rc = 0;
    if (yhi > ylo && W - bx > bx) {
        const int rows = yhi - ylo;
        int nb = nthreads * 4;
        if (nb > rows) nb = rows;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
        for (int b = 0; b < nb; ++b) {
            const int ya = ylo + (int)((long)rows * b / nb);
            const int yb = ylo + (int)((long)rows * (b + 1) / nb);
            const int nr = 2 * ay + 1;
            int32_t* ring = (int32_t*)malloc((size_t)nr * W * sizeof(int32_t));
            int32_t* V = (int32_t*)malloc((size_t)W * sizeof(int32_t));
            if (!ring || !V) { rc = -2; free(ring); free(V); continue; }
            for (int d = 0; d < D; ++d) {
                const int xs = bx + d, xe = W - bx;
                if (xs >= xe) break;
                for (int j = 0; j < nr; ++j) hsum_row(kind, L, R, W, ya - ay + j, d, ax, xs, xe, ring + (size_t)j * W);
                for (int x = xs; x < xe; ++x) {
                    int s = 0;
                    for (int j = 0; j < nr; ++j) s += ring[(size_t)j * W + x];
                    V[x] = s;
                }
                for (int y = ya; y < yb; ++y) {
                    if (y > ya) {
                        int32_t* slot = ring + (size_t)((y - ya - 1) % nr) * W; /* holds row y-ay-1 */
                        /* subtract the row leaving the window, add the row entering it */
                        for (int x = xs; x < xe; ++x) V[x] -= slot[x];
                        hsum_row(kind, L, R, W, y + ay, d, ax, xs, xe, slot);
                        for (int x = xs; x < xe; ++x) V[x] += slot[x];
                    }
                    const size_t row = (size_t)y * W;
                    for (int x = xs; x < xe; ++x) {
                        const int A = V[x];
                        if (bins && !((bins[row + x] >> (d >> 3)) & 1u)) continue;
                        if (A < bestL[row + x]) { bestL[row + x] = A; dL[row + x] = (int16_t)d; }
                        if (A < bestR[row + x - d]) { bestR[row + x - d] = A; dR[row + x - d] = (int16_t)d; }
                    }
                }
            }
            free(ring);
            free(V);
        }
    }
    if (out) crosscheck_rows(dL, dR, W, 0, H, tol, out);
    if (cost_out)
        for (size_t i = 0; i < N; ++i) cost_out[i] = dL[i] >= 0 ? bestL[i] : -1;
    free(bestL);
    free(bestR);
    return rc;
}

int orc_bm(int kind, const void* L, const void* R, int W, int H, int D, int ww, int wh, int cw,
           int ch, int tol, int16_t* dL, int16_t* dR, int16_t* out, int32_t* cost_out, int nthreads)
{
    return bm_impl(kind, L, R, W, H, D, ww, wh, cw, ch, tol, NULL, dL, dR, out, cost_out, nthreads);
}

int orc_bm_bins(int kind, const void* L, const void* R, int W, int H, int D, int ww, int wh, int cw,
                int ch, int tol, const uint32_t* bins, int16_t* dL, int16_t* dR, int16_t* out, int nthreads)
{
    if (!bins || D > 256) return -1;
    return bm_impl(kind, L, R, W, H, D, ww, wh, cw, ch, tol, bins, dL, dR, out, NULL, nthreads);
}

int orc_bm_brute(int kind, const void* L, const void* R, int W, int H, int D, int ww, int wh,
                 int cw, int ch, int tol, int16_t* dL, int16_t* dR, int16_t* out, int32_t* volume)
{
    if (!bm_args_ok(kind, L, R, W, H, D, ww, wh, cw, ch) || !dL || !dR) return -1;
    const int ax = ww / 2, ay = wh / 2;
    const int bx = ax + cw / 2, by = ay + ch / 2;
    const size_t N = (size_t)W * H;
    if (volume)
        for (size_t i = 0; i < N * (size_t)D; ++i) volume[i] = -1;
    for (size_t i = 0; i < N; ++i) { dL[i] = -1; dR[i] = -1; }
    int32_t* bestR = (int32_t*)malloc(N * sizeof(int32_t));
    if (!bestR) return -2;
    for (size_t i = 0; i < N; ++i) bestR[i] = INT_MAX;
    for (int y = by; y < H - by; ++y)
        for (int x = bx; x < W - bx; ++x) {
            int best = INT_MAX;
            for (int d = 0; d < D && x - d >= bx; ++d) {
                int A = 0;
                for (int j = -ay; j <= ay; ++j)
                    for (int i = -ax; i <= ax; ++i) {
                        const size_t p = (size_t)(y + j) * W + (x + i);
                        A += cost_at(kind, L, R, p, p - d);
                    }
                if (volume) volume[((size_t)y * W + x) * D + d] = A;
                if (A < best) { best = A; dL[(size_t)y * W + x] = (int16_t)d; }
                const size_t r = (size_t)y * W + (x - d);
                if (A < bestR[r] || (A == bestR[r] && d < dR[r])) { bestR[r] = A; dR[r] = (int16_t)d; }
            }
        }
    free(bestR);
    if (out) crosscheck_rows(dL, dR, W, 0, H, tol, out);
    return 0;
}

/* ---------------------------------------------------------------------------------
 * NCC.  Reference: BlockMatcherNCC.cu EvalMsd :338-373 (base mean / std-dev),
 * EvalSim :180-336 (ncc = (E[IJ] - muI muJ) / (sdI sdJ), 0 when the norm is small,
 * clamped at 0 before quantisation).  Restated exactly on integer box sums:
 *   cov = N*SIJ - SI*SJ, varI = N*SII - SI^2, varJ = N*SJJ - SJ^2,
 *   score = cov / sqrt(varI*varJ), and score = 0 when varI < N^2 or varJ < N^2
 *   (standard deviation below one grey level on either side).
 * Winner-take-all maximises q = floor(max(score, 0) * 2^23), lowest d on ties.
 * --------------------------------------------------------------------------------- */
static void box_stats(const uint8_t* img, int W, int H, int ax, int ay, int64_t* S, int64_t* SS)
{
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            int64_t s = 0, ss = 0;
            if (x >= ax && x < W - ax && y >= ay && y < H - ay)
                for (int j = -ay; j <= ay; ++j)
                    for (int i = -ax; i <= ax; ++i) {
                        const int v = img[(size_t)(y + j) * W + x + i];
                        s += v;
                        ss += v * v;
                    }
            S[(size_t)y * W + x] = s;
            SS[(size_t)y * W + x] = ss;
        }
}

/* NCC over the candidates of every left pixel p, or only those whose 8-disparity bin
 * (d >> 3) is set in bins[p] (the band search, MatcherUseDisparity BlockMatcherNCC.cu:894-1255,
 * with the engine's per-tile bins). */
static int ncc_impl(const uint8_t* L, const uint8_t* R, int W, int H, int D, int ww, int wh, int tol,
                    const uint32_t* bins, int16_t* dL, int16_t* dR, int16_t* out, double* score_out, double* volume,
                    int nthreads)
{
    if (!L || !R || !dL || !dR || W <= 0 || H <= 0 || D <= 0 || ww < 1 || wh < 1 || !(ww & 1) || !(wh & 1)) return -1;
    if (bins && D > 256) return -1;
    const int ax = ww / 2, ay = wh / 2, bx = ax, by = ay;
    const int64_t Nw = (int64_t)ww * wh;
    const size_t N = (size_t)W * H;
    int64_t* SI = (int64_t*)malloc(N * sizeof(int64_t));
    int64_t* SII = (int64_t*)malloc(N * sizeof(int64_t));
    int64_t* SJ = (int64_t*)malloc(N * sizeof(int64_t));
    int64_t* SJJ = (int64_t*)malloc(N * sizeof(int64_t));
    int64_t* bestL = (int64_t*)malloc(N * sizeof(int64_t));
    int64_t* bestR = (int64_t*)malloc(N * sizeof(int64_t));
    double* sc = (double*)malloc(N * sizeof(double));
    if (!SI || !SII || !SJ || !SJJ || !bestL || !bestR || !sc) {
        free(SI); free(SII); free(SJ); free(SJJ); free(bestL); free(bestR); free(sc);
        return -2;
    }
    box_stats(L, W, H, ax, ay, SI, SII);
    box_stats(R, W, H, ax, ay, SJ, SJJ);
    for (size_t i = 0; i < N; ++i) { bestL[i] = -1; bestR[i] = -1; dL[i] = -1; dR[i] = -1; sc[i] = NAN; }
    if (volume)
        for (size_t i = 0; i < N * (size_t)D; ++i) volume[i] = NAN;
    if (nthreads < 1) nthreads = 1;
    const int ylo = by, yhi = H - by;
    if (yhi > ylo) {
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
        for (int y = ylo; y < yhi; ++y) {
            int64_t* V = (int64_t*)malloc((size_t)W * sizeof(int64_t));
            for (int d = 0; d < D; ++d) {
                const int xs = bx + d, xe = W - bx;
                if (xs >= xe) break;
                /* column sums of I*J over the window rows, then a horizontal running sum */
                for (int x = xs - ax; x < xe + ax; ++x) {
                    int64_t s = 0;
                    for (int j = -ay; j <= ay; ++j) {
                        const size_t p = (size_t)(y + j) * W + x;
                        s += (int64_t)L[p] * R[p - d];
                    }
                    V[x] = s;
                }
                int64_t sij = 0;
                for (int i = -ax; i <= ax; ++i) sij += V[xs + i];
                for (int x = xs; x < xe; ++x) {
                    if (x > xs) sij += V[x + ax] - V[x - ax - 1];
                    const size_t pl = (size_t)y * W + x, pr = pl - d;
                    const int64_t vI = Nw * SII[pl] - SI[pl] * SI[pl];
                    const int64_t vJ = Nw * SJJ[pr] - SJ[pr] * SJ[pr];
                    double s = 0.0;
                    if (vI >= Nw * Nw && vJ >= Nw * Nw)
                        s = (double)(Nw * sij - SI[pl] * SJ[pr]) / sqrt((double)vI * (double)vJ);
                    if (volume) volume[pl * D + d] = s;
                    if (bins && !((bins[pl] >> (d >> 3)) & 1u)) continue;
                    const int64_t q = (int64_t)floor((s > 0.0 ? s : 0.0) * 8388608.0);
                    if (q > bestL[pl]) { bestL[pl] = q; dL[pl] = (int16_t)d; sc[pl] = s; }
                    if (q > bestR[pr]) { bestR[pr] = q; dR[pr] = (int16_t)d; }
                }
            }
            free(V);
        }
    }
    if (out) crosscheck_rows(dL, dR, W, 0, H, tol, out);
    if (score_out) memcpy(score_out, sc, N * sizeof(double));
    free(SI); free(SII); free(SJ); free(SJJ); free(bestL); free(bestR); free(sc);
    return 0;
}

int orc_ncc(const uint8_t* L, const uint8_t* R, int W, int H, int D, int ww, int wh, int tol,
            int16_t* dL, int16_t* dR, int16_t* out, double* score_out, double* volume, int nthreads)
{
    return ncc_impl(L, R, W, H, D, ww, wh, tol, NULL, dL, dR, out, score_out, volume, nthreads);
}

int orc_ncc_bins(const uint8_t* L, const uint8_t* R, int W, int H, int D, int ww, int wh, int tol,
                 const uint32_t* bins, int16_t* dL, int16_t* dR, int16_t* out, double* score_out, int nthreads)
{
    if (!bins) return -1;
    return ncc_impl(L, R, W, H, D, ww, wh, tol, bins, dL, dR, out, score_out, NULL, nthreads);
}

/* ---------------------------------------------------------------------------------
 * Adaptive windows.  Sub-patch sums S over s x s windows; the window value is the
 * centre sub-patch plus the two best of the four edge neighbours at distance s
 * (top, left, bottom, right) -- AlgoAdaptWnd_W32H16::SNeighborhood<Two_OutOf_Four>
 * (Kernel.Algo.AdaptWnd_W32H16.h:240-285, OrderedInsertMax :122-127) for AW1 with
 * similarity 255-|a-b| (BlockMatcherAW1.cu:191-195) and the maximum kept with
 * strict '>' (:294); AlgoAdaptWnd3D Two_OutOf_Four (Kernel.Algo.AdaptWnd3D.h:244-300,
 * OrderedInsert :127-131) for AW2 with |a-b| (BlockMatcherAW2.cu:129-133) and the
 * minimum kept with strict '<' (:167).  Border: bx = by = s + s/2.
 * --------------------------------------------------------------------------------- */
static inline void ins_max(int* f, int* s, int v)
{
    const int lo = *f < v ? *f : v;
    *s = *s > lo ? *s : lo;
    *f = *f > v ? *f : v;
}

static inline void ins_min(int* f, int* s, int v)
{
    const int hi = *f > v ? *f : v;
    *s = *s < hi ? *s : hi;
    *f = *f < v ? *f : v;
}

/* bins (optional, D <= 256): only candidates whose 8-disparity bin d >> 3 is set in
 * bins[p] of the left pixel p take part, in both views -- the band search
 * (MatcherUseDisparity, BlockMatcherAW1.cu:1336, BlockMatcherAW2.cu:858) with the
 * engine's per-tile bins. */
static int aw_impl(int mode, const uint8_t* L, const uint8_t* R, int W, int H, int D, int s, int tol,
                   const uint32_t* bins, int16_t* dL, int16_t* dR, int16_t* out, int32_t* value_out,
                   int32_t* volume, int nthreads)
{
    if ((mode != 1 && mode != 2) || !L || !R || !dL || !dR || s < 1 || (s & 1) == 0 || D <= 0) return -1;
    if (bins && D > 256) return -1;
    const int h = s / 2, b = s + h;
    const size_t N = (size_t)W * H;
    int32_t* bestL = (int32_t*)malloc(N * sizeof(int32_t));
    int32_t* bestR = (int32_t*)malloc(N * sizeof(int32_t));
    if (!bestL || !bestR) { free(bestL); free(bestR); return -2; }
    const int init = (mode == 1) ? -32767 : 32767;
    for (size_t i = 0; i < N; ++i) { bestL[i] = init; bestR[i] = init; dL[i] = -1; dR[i] = -1; }
    if (volume)
        for (size_t i = 0; i < N * (size_t)D; ++i) volume[i] = -1;
    if (nthreads < 1) nthreads = 1;
    if (H - b > b && W - b > b) {
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
        for (int y = b; y < H - b; ++y) {
            /* S rows y-s, y, y+s over all columns that a valid window touches */
            int32_t* S = (int32_t*)malloc((size_t)3 * W * sizeof(int32_t));
            for (int d = 0; d < D; ++d) {
                const int xs = b + d, xe = W - b;
                if (xs >= xe) break;
                for (int r = 0; r < 3; ++r) {
                    const int yc = y + (r - 1) * s;
                    for (int x = xs - s; x < xe + s; ++x) {
                        int acc = 0;
                        for (int j = -h; j <= h; ++j)
                            for (int i = -h; i <= h; ++i) {
                                const size_t p = (size_t)(yc + j) * W + x + i;
                                const int a = L[p], c = R[p - d];
                                const int ad = a > c ? a - c : c - a;
                                acc += (mode == 1) ? 255 - ad : ad;
                            }
                        S[(size_t)r * W + x] = acc;
                    }
                }
                for (int x = xs; x < xe; ++x) {
                    int v;
                    if (mode == 1) {
                        int m1 = S[x], m2 = -32767;              /* top    (x,   y-s) */
                        ins_max(&m1, &m2, S[(size_t)W + x - s]);  /* left   (x-s, y)   */
                        ins_max(&m1, &m2, S[(size_t)2 * W + x]);  /* bottom (x,   y+s) */
                        ins_max(&m1, &m2, S[(size_t)W + x + s]);  /* right  (x+s, y)   */
                        v = m1 + m2 + S[(size_t)W + x];
                    } else {
                        int m1 = S[x], m2 = 32767;
                        ins_min(&m1, &m2, S[(size_t)W + x - s]);
                        ins_min(&m1, &m2, S[(size_t)2 * W + x]);
                        ins_min(&m1, &m2, S[(size_t)W + x + s]);
                        v = m1 + m2 + S[(size_t)W + x];
                    }
                    const size_t pl = (size_t)y * W + x, pr = pl - d;
                    if (volume) volume[pl * D + d] = v;
                    if (bins && !((bins[pl] >> (d >> 3)) & 1u)) continue;
                    const int betterL = (mode == 1) ? (v > bestL[pl]) : (v < bestL[pl]);
                    if (betterL) { bestL[pl] = v; dL[pl] = (int16_t)d; }
                    const int betterR = (mode == 1) ? (v > bestR[pr]) : (v < bestR[pr]);
                    if (betterR) { bestR[pr] = v; dR[pr] = (int16_t)d; }
                }
            }
            free(S);
        }
    }
    if (out) crosscheck_rows(dL, dR, W, 0, H, tol, out);
    if (value_out)
        for (size_t i = 0; i < N; ++i) value_out[i] = dL[i] >= 0 ? bestL[i] : -1;
    free(bestL);
    free(bestR);
    return 0;
}

int orc_aw(int mode, const uint8_t* L, const uint8_t* R, int W, int H, int D, int s, int tol,
           int16_t* dL, int16_t* dR, int16_t* out, int32_t* value_out, int32_t* volume, int nthreads)
{
    return aw_impl(mode, L, R, W, H, D, s, tol, NULL, dL, dR, out, value_out, volume, nthreads);
}

int orc_aw_bins(int mode, const uint8_t* L, const uint8_t* R, int W, int H, int D, int s, int tol,
                const uint32_t* bins, int16_t* dL, int16_t* dR, int16_t* out, int32_t* value_out, int nthreads)
{
    if (!bins) return -1;
    return aw_impl(mode, L, R, W, H, D, s, tol, bins, dL, dR, out, value_out, NULL, nthreads);
}

/* ---------------------------------------------------------------------------------
 * Reference cross-check: CrossCheckDisparityMaps.cu:62-109 (per pixel) and the
 * host range reduction :135-192.  Codes outside [First, Last] pass through; the
 * right position is MapPixelPos(x, id-First, LtoR) = LtoR ? x-d-off : x+d+off
 * (inferred from the test-patch addressing of BlockMatcherNCC.cu:792-823);
 * outside the image or |id_R - id_L| > thresh gives Inconsistent; kept ids update
 * the global min/max that start at Last/First.
 * --------------------------------------------------------------------------------- */
int orc_crosscheck_ref(const uint16_t* lr, const uint16_t* rl, int W, int H, int thresh, int ltor,
                       int offset, uint16_t* out, uint16_t* dmin, uint16_t* dmax)
{
    if (!lr || !rl || !out) return -1;
    unsigned mn = ORC_DISP_LAST, mx = ORC_DISP_FIRST;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = ((size_t)y * W + x) * 4;
            const unsigned id = lr[i];
            memcpy(out + i, lr + i, 4 * sizeof(uint16_t));
            if (id < ORC_DISP_FIRST || id > ORC_DISP_LAST) continue;
            const int d = (int)id - ORC_DISP_FIRST;
            const int xr = ltor ? x - d - offset : x + d + offset;
            if (xr < 0 || xr >= W) { out[i] = ORC_DISP_INCONSISTENT; continue; }
            const int idr = rl[((size_t)y * W + xr) * 4];
            if (abs(idr - (int)id) <= thresh) {
                if (id < mn) mn = id;
                if (id > mx) mx = id;
            } else {
                out[i] = ORC_DISP_INCONSISTENT;
            }
        }
    if (dmin) *dmin = (uint16_t)mn;
    if (dmax) *dmax = (uint16_t)mx;
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Disparity mip maps (CreateDispMipMaps.cu).                                             */

int orc_mip_level0_ref(const uint16_t* rgba, int W, int H, float* out)
{
    /* CopyDispChannel (CreateDispMipMaps.cu:66-94): ids in [First, Last] as float, the
     * rest Unknown */
    if (!rgba || !out || W < 1 || H < 1) return -1;
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        const uint16_t id = rgba[4 * i];
        out[i] = (id >= ORC_DISP_FIRST && id <= ORC_DISP_LAST) ? (float)id : (float)ORC_DISP_UNKNOWN;
    }
    return 0;
}

int orc_mip_level0_plain(const int16_t* d, int W, int H, float* out)
{
    if (!d || !out || W < 1 || H < 1) return -1;
    for (size_t i = 0; i < (size_t)W * H; ++i) out[i] = d[i] >= 0 ? (float)d[i] : -1.0f;
    return 0;
}

/* one [1 2 1] tap of three values in the reference's order: the first value, twice the
 * second, the third, each only when it is not `unknown`; sum / weight, or unknown */
static float mip_tap(float v0, float v1, float v2, float unknown)
{
    float cnt = 0.0f, sum = 0.0f;
    if (v0 != unknown) { cnt += 1.0f; sum += v0; }
    if (v1 != unknown) { cnt += 2.0f; sum += v1 * 2.0f; }
    if (v2 != unknown) { cnt += 1.0f; sum += v2; }
    return cnt > 0.0f ? sum / cnt : unknown;
}

int orc_mip_next(const float* src, int W, int H, float unknown, float* out)
{
    /* SmoothHorizontalLum (:97-150) to a ((W-1)/2) x H buffer, then SmoothVerticalLum
     * (:153-218) with the factor 0.5 on the normalised sum; NextLevelWidth/Height are
     * (n - 1) / 2 (DeviceSurfMipMap.h:78-86), so every 3-tap window lies inside */
    if (!src || !out || W < 3 || H < 3) return -1;
    const int Wn = (W - 1) / 2, Hn = (H - 1) / 2;
    float* tmp = (float*)malloc(sizeof(float) * (size_t)Wn * H);
    if (!tmp) return -2;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < Wn; ++x) {
            const float* r = src + (size_t)y * W + 2 * x;
            tmp[(size_t)y * Wn + x] = mip_tap(r[0], r[1], r[2], unknown);
        }
    for (int y = 0; y < Hn; ++y)
        for (int x = 0; x < Wn; ++x) {
            const float* c = tmp + (size_t)(2 * y) * Wn + x;
            const float v = mip_tap(c[0], c[Wn], c[2 * Wn], unknown);
            out[(size_t)y * Wn + x] = v != unknown ? 0.5f * v : unknown;
        }
    free(tmp);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Disparity de-noise (Statistics.DispDenoise.cu).                                        */

int orc_disp_denoise(int ref_format, const void* in, int W, int H, int P, float support_ratio, float mean_cutoff,
                     int fill, void* out)
{
    if (!in || !out || P < 1 || W < P || H < P) return -1;
    const uint16_t* rin = (const uint16_t*)in;
    const int16_t* pin = (const int16_t*)in;
    /* value of pixel (x, y): the id (reference) or d (plain) when valid, else -1 */
    #define ORC_DN_VAL(x, y) (ref_format ? ((rin[4 * ((size_t)(y) * W + (x))] >= ORC_DISP_FIRST && \
                                              rin[4 * ((size_t)(y) * W + (x))] <= ORC_DISP_LAST) ? \
                                              (int)rin[4 * ((size_t)(y) * W + (x))] : -1) \
                                           : (pin[(size_t)(y) * W + (x)] >= 0 ? (int)pin[(size_t)(y) * W + (x)] : -1))
    const int unknown = ref_format ? ORC_DISP_UNKNOWN : -1;
    const float need = (float)(P * P) * support_ratio;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            int cnt = 0, sum = 0;
            int64_t sum2 = 0;
            for (int j = 0; j < P; ++j) {
                int yy = y - P / 2 + j;
                if (yy < 0) yy += P;
                if (yy >= H) yy -= P;
                for (int i = 0; i < P; ++i) {
                    int xx = x - P / 2 + i;
                    if (xx < 0) xx += P;
                    if (xx >= W) xx -= P;
                    const int v = ORC_DN_VAL(xx, yy);
                    if (v >= 0) {
                        ++cnt;
                        sum += v;
                        sum2 += (int64_t)v * v;
                    }
                }
            }
            const int raw = ref_format ? (int)rin[4 * ((size_t)y * W + x)] : (int)pin[(size_t)y * W + x];
            const int v = ORC_DN_VAL(x, y);
            int res;
            if (!((float)cnt >= need)) {
                res = unknown;                       /* too little support (:303-304) */
            } else if (v >= 0) {
                /* the mean of the other valid values; keep it when the pixel is near it (:308-316) */
                const float fd = (float)v;
                const float mean = ((float)sum - fd) / (float)(cnt - 1);
                res = fabsf(fd - mean) < mean_cutoff ? (int)floorf(mean + 0.5f) : unknown;
            } else if (fill) {
                /* fill from the window when it is flat (:318-325) */
                const float mean = (float)sum / (float)cnt;
                const float mean2 = (float)sum2 / (float)cnt;
                const float sq = mean * mean;
                const float sd = sqrtf(fabsf(mean2 - sq));
                res = sd < mean_cutoff ? (int)floorf(mean + 0.5f) : raw;
            } else {
                res = raw;
            }
            if (ref_format) ((uint16_t*)out)[(size_t)y * W + x] = (uint16_t)res;
            else ((int16_t*)out)[(size_t)y * W + x] = (int16_t)res;
        }
    }
    #undef ORC_DN_VAL
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* BlockMatcher1, emulated block by block (CluTec.Cuda.MultiView/BlockMatcher1.cu).      */
/*                                                                                        */
/* The reference launches one 32-thread block per PCX x 4 result origins                 */
/* (AlgoArraySum<32, 2, P, P>, Kernel.Algo.ArraySum.h:38-60: PCX = ((32 - P + 1) / 8) * 8), */
/* blocks counted by EvalThreadConfig_ReadDepBlockSize (KernelDriverBase.cpp:166-197)     */
/* from a 33 x (P + 3) read area, full blocks only.  Which disparities a result sees, and where its test    */
/* patch lies, is decided per block; this function follows those decisions.              */

/* Similarity of two pixels (EvalSim, BlockMatcher1.cu:184-221). */
static inline int bm1_sim(int sim_type, int a, int b)
{
    if (sim_type == 0) return (255 - abs(a - b)) >> 1; /* AbsoluteDifference */
    float f1 = 2.0f * ((float)a / 255.0f - 0.5f);      /* SignedProduct */
    const float f2 = 2.0f * ((float)b / 255.0f - 0.5f);
    f1 *= f2;
    f1 = f1 * 127.0f;
    return (int)f1;
}

/* SimValueToFloat (:155-159) and SimFloatToInteger (:172-176) */
static inline float bm1_sim_float(int v, int P) { return ((float)v / (float)(P * P)) * (100.0f / 127.0f); }
static inline uint16_t bm1_sim_u16(float f)
{
    float t = (f / 100.0f + 1.0f) / 2.0f;
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return (uint16_t)(t * 65535.0f);
}

typedef struct {
    int16_t prev2, prev, max, next, next2, last;
} bm1_store;

/* One disparity of one result (the update lambda of EvalSim, :262-300). */
static inline void bm1_update(bm1_store* s, uint16_t* disp, int v, int d)
{
    const int above = v > s->max;
    const int behind1 = !above && (int)*disp + 1 == d + ORC_DISP_FIRST;
    const int behind2 = !above && (int)*disp + 2 == d + ORC_DISP_FIRST;
    if (above) {
        s->prev2 = s->prev; /* the value before the previous maximum, not the one 2 before d */
        s->prev = s->last;
        s->max = (int16_t)v;
        s->next = (int16_t)v;
        s->next2 = (int16_t)v;
        *disp = (uint16_t)(d + ORC_DISP_FIRST);
    }
    if (behind1) s->next = (int16_t)v;
    if (behind2) s->next2 = (int16_t)v;
    s->last = (int16_t)v;
}

int orc_bm1(const uint8_t* L, const uint8_t* R, int W, int H, int P, int sim_type, int dmin, int dmax, int offset,
            int ltor, float min_delta, float sim_thresh, float grad_thresh, const uint16_t* init, int iW, int iH,
            uint16_t* out)
{
    if (!L || !R || !out || (P != 5 && P != 9 && P != 11 && P != 15) || W < 48 || H < P + 3 || dmin < 0 ||
        dmax <= dmin || (init && (iW < 1 || iH < 1 || W / iW < 1)))
        return -1;
    const int CG = (32 - P + 1) / 8, PCX = CG * 8, PCY = 4, TBX = 33, TBY = P + 3, SDC = 16;
    const unsigned effW = (unsigned)W - (unsigned)(TBX - PCX), effH = (unsigned)H - (unsigned)(TBY - PCY);
    /* only full blocks: bOnlyFullBlocks defaults to true (KernelDriverBase.h:117-121), so
       the remainder strips at the right and bottom are never launched */
    const int nbx = (int)(effW / PCX), nby = (int)(effH / PCY);
    const int NR = PCX * PCY;
#pragma omp parallel for schedule(dynamic, 1)
    for (int by = 0; by < nby; ++by) {
        bm1_store st[24 * 4];
        uint16_t disp[24 * 4];
        for (int bx = 0; bx < nbx; ++bx) {
            const int X0 = bx * PCX, Y0 = by * PCY;
            if (!(X0 + P <= W && Y0 + P <= H)) continue; /* :626-630 */
            int any_eval = 0;
            for (int k = 0; k < NR; ++k) {
                st[k] = (bm1_store){0, 0, -32767, 0, 0, 0};
                disp[k] = ORC_DISP_UNKNOWN;
            }
            int dlo = ORC_DISP_LAST, dhi = ORC_DISP_FIRST;
            if (init) {
                /* MatcherUseDisparity (:912-963).  Thread t holds the results
                   ((t / 4) * CG + r, t % 4), r < CG.  A written result takes the initial id at
                   its centre over the scale (positions outside the initial map read as
                   Unknown): an id up to Last scaled (16-bit), a code above Last as it is.  The
                   thread's plain min / max of the scaled values (from Last / First) are then
                   reduced over the warp with MinDisp / MaxDisp, which drop invalid
                   operands: the result is the min / max over the threads whose own value is
                   valid. */
                const int scale = W / iW;
                int red_lo = -1, red_hi = -1;
                for (int t = 0; t < 32; ++t) {
                    int tlo = ORC_DISP_LAST, thi = ORC_DISP_FIRST;
                    for (int r = 0; r < CG; ++r) {
                        const int ix = (t / 4) * CG + r, iy = t % 4, k = iy * PCX + ix;
                        if (!(X0 + ix + P < W && Y0 + iy + P < H)) continue;
                        const int px = (X0 + ix + P / 2) / scale, py = (Y0 + iy + P / 2) / scale;
                        const uint16_t id = (px < iW && py < iH) ? init[4 * ((size_t)py * iW + px)] : ORC_DISP_UNKNOWN;
                        if (id <= ORC_DISP_LAST) {
                            const uint16_t v = (uint16_t)(scale * id);
                            if (v < tlo) tlo = v;
                            if (v > thi) thi = v;
                            disp[k] = v;
                        } else {
                            disp[k] = id;
                        }
                    }
                    const int vlo = tlo >= ORC_DISP_FIRST && tlo <= ORC_DISP_LAST;
                    const int vhi = thi >= ORC_DISP_FIRST && thi <= ORC_DISP_LAST;
                    if (vlo) red_lo = red_lo < 0 ? tlo : (tlo < red_lo ? tlo : red_lo);
                    if (vhi) red_hi = red_hi < 0 ? thi : (thi > red_hi ? thi : red_hi);
                }
                dlo = red_lo; /* -1: no thread holds a valid value */
                dhi = red_hi;
            }
            for (int k = 0; k < NR && !init; ++k) {
                const int ix = k % PCX, iy = k / PCX;
                /* EvalGrad (:325-361): the standard deviation of the left patch */
                int64_t s = 0, s2 = 0;
                for (int j = 0; j < P; ++j)
                    for (int i = 0; i < P; ++i) {
                        const int x = X0 + ix + i, y = Y0 + iy + j;
                        const int v = (x < W && y < H) ? L[(size_t)y * W + x] : 0;
                        s += v;
                        s2 += v * v;
                    }
                const float mean = (float)s / (float)(P * P);
                const float mean2 = (float)s2 / (float)(P * P);
                const float sq = mean * mean;
                const float sd = sqrtf(fabsf(mean2 - sq)) / 255.0f * 100.0f;
                if (sd < grad_thresh) disp[k] = ORC_DISP_CANNOT_EVALUATE;
                else any_eval = 1;
            }
            int d_start, d_end_excl, lastgroup_cap;
            if (init) {
                /* the block's range of valid scaled values (:955-963, the warp's shuffle
                   reduction), else every result keeps its initial code (:967-999) */
                if (dlo < 0 || dhi < 0 || dlo > dhi) {
                    for (int k = 0; k < NR; ++k) {
                        const int ix = k % PCX, iy = k / PCX;
                        if (X0 + ix + P < W && Y0 + iy + P < H) {
                            uint16_t* o = out + 4 * ((size_t)(Y0 + iy + P / 2) * W + X0 + ix + P / 2);
                            o[0] = disp[k];
                            o[1] = o[2] = o[3] = 0;
                        }
                    }
                    continue;
                }
                d_start = dlo < 4 ? 0 : dlo - 4;
                d_end_excl = dhi + 4 + 1; /* iDispIdx <= iDispEnd */
                lastgroup_cap = d_end_excl - d_start < SDC ? d_end_excl - d_start : SDC; /* :1088 */
                /* the loop variables are 16-bit: an end beyond 32767 evaluates nothing */
                any_eval = dhi + 4 <= 32767;
            } else {
                const int maxd = ltor ? X0 - offset + 1 : W - X0 - offset; /* MaxDisparity */
                d_start = dmin;
                d_end_excl = dmax < maxd ? dmax : maxd;
                lastgroup_cap = SDC;
            }
            if (any_eval) {
                for (int D0 = d_start; D0 < d_end_excl; D0 += SDC) {
                    const int tx = (ltor ? X0 - D0 - offset : X0 + D0 + offset) - ltor * (SDC - 1);
                    int n, base;
                    if (tx >= 0 && Y0 >= 0 && tx + TBX + SDC - 1 <= W && Y0 + TBY <= H) {
                        n = lastgroup_cap;
                        base = tx + ltor * (SDC - 1);
                    } else {
                        if (!(Y0 >= 0 && Y0 + TBY <= H)) continue;
                        int dx, sx;
                        if (tx >= 0) {
                            dx = W - tx - TBX;
                            sx = tx;
                        } else {
                            dx = tx + (SDC - 1);
                            sx = 0;
                        }
                        if (dx < 0) continue;
                        n = dx + 1;
                        base = sx + ltor * (n - 1);
                    }
                    for (int r = 0; r < n; ++r) {
                        const int d = D0 + r, t0 = ltor ? base - r : base + r;
                        for (int k = 0; k < NR; ++k) {
                            if (disp[k] > ORC_DISP_LAST) continue;
                            const int ix = k % PCX, iy = k / PCX;
                            int v = 0;
                            for (int j = 0; j < P; ++j) {
                                const int y = Y0 + iy + j;
                                for (int i = 0; i < P; ++i) {
                                    const int x = X0 + ix + i;
                                    const int a = x < W ? L[(size_t)y * W + x] : 0;
                                    v += bm1_sim(sim_type, a, R[(size_t)y * W + t0 + ix + i]);
                                }
                            }
                            bm1_update(&st[k], &disp[k], (int16_t)v, d);
                        }
                    }
                }
            }
            /* WriteDisparity (:447-575) */
            for (int k = 0; k < NR; ++k) {
                const int ix = k % PCX, iy = k / PCX;
                if (!(X0 + ix + P < W && Y0 + iy + P < H)) continue;
                uint16_t* o = out + 4 * ((size_t)(Y0 + iy + P / 2) * W + X0 + ix + P / 2);
                const int dv = disp[k];
                if (dv > ORC_DISP_LAST || dv == ORC_DISP_UNKNOWN) {
                    o[0] = (uint16_t)dv;
                    o[1] = o[2] = o[3] = 0;
                    continue;
                }
                const bm1_store* s = &st[k];
                const float fmax = bm1_sim_float(s->max, P);
                const float fp2 = bm1_sim_float(s->prev2, P), fp = bm1_sim_float(s->prev, P);
                const float fn = bm1_sim_float(s->next, P), fn2 = bm1_sim_float(s->next2, P);
                const float dprev = fmaxf(fabsf(fp - fmax), fabsf(fp2 - fmax));
                const float dnext = fmaxf(fabsf(fn - fmax), fabsf(fn2 - fmax));
                const int specific = dprev >= min_delta && dnext >= min_delta;
                const int above = fmax >= sim_thresh;
                uint16_t code;
                if (above) {
                    const int sat = dv <= ORC_DISP_FIRST + dmin + 1 || dv >= ORC_DISP_FIRST + dmax - 1;
                    code = sat ? ORC_DISP_SATURATED : (specific ? (uint16_t)dv : ORC_DISP_NOT_SPECIFIC);
                } else {
                    code = ORC_DISP_NOT_FOUND;
                }
                o[0] = code;
                o[1] = bm1_sim_u16(fmax);
                o[2] = bm1_sim_u16(dprev);
                o[3] = bm1_sim_u16(dnext);
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Image rectification (CluTec.Cuda.ImgProc/RectifyImage.cu) and disparity colouring     */
/* (CluTec.Cuda.MultiView/ColorDisparityImage.cu).                                        */

/* NormFloatTo<T>: clamp to [0, 1] (NaN to 0), scale by the type's maximum, round half up. */
static inline uint32_t orc_norm_to(float f, float maxv)
{
    f = fminf(fmaxf(f, 0.0f), 1.0f);
    return (uint32_t)(f * maxv + 0.5f);
}

/* The source position of target pixel (u, v): rectified pixel -> normalised ray ->
 * rotation into the source camera -> OpenCV distortion -> source pixel (MapPixelTrgToSrc
 * and DistortPX, RectifyImage.cu:111-121,150-158). */
static void rect_map(const float* p, int u, int v, float* sx, float* sy)
{
    const float x = ((float)u - p[11]) / p[9], y = ((float)v - p[12]) / p[10];
    const float* R = p + 13;
    const float X = R[0] * x + R[1] * y + R[2];
    const float Y = R[3] * x + R[4] * y + R[5];
    const float Z = R[6] * x + R[7] * y + R[8];
    const float xn = X / Z, yn = Y / Z;
    const float r2 = xn * xn + yn * yn;
    const float radial = 1.0f + r2 * (p[4] + r2 * (p[5] + r2 * p[8]));
    const float xy = xn * yn;
    const float xd = xn * radial + 2.0f * p[6] * xy + p[7] * (r2 + 2.0f * xn * xn);
    const float yd = yn * radial + p[6] * (r2 + 2.0f * yn * yn) + 2.0f * p[7] * xy;
    *sx = p[0] * xd + p[2];
    *sy = p[1] * yd + p[3];
}

/* params: src fx, fy, cx, cy, k1, k2, p1, p2, k3, rect fx, fy, cx, cy, R (9, rectified to
 * source, row-major): 22 floats.  CopyPixel (:64-107): bilinear when the 2x2 block at the
 * floor lies inside the source, else the nearest pixel when inside, else 0. */
int orc_rectify(const uint8_t* src, int sW, int sH, const float* params, int W, int H, uint8_t* out)
{
    if (!src || !params || !out || sW < 1 || sH < 1 || W < 1 || H < 1) return -1;
#pragma omp parallel for schedule(static)
    for (int v = 0; v < H; ++v)
        for (int u = 0; u < W; ++u) {
            float sx, sy;
            rect_map(params, u, v, &sx, &sy);
            const float fx0 = floorf(sx), fy0 = floorf(sy);
            const float nx = floorf(sx + 0.5f), ny = floorf(sy + 0.5f);
            uint8_t o = 0;
            /* positions beyond +-2^30 (or NaN) are outside */
            const int ok = fabsf(fx0) < 1073741824.0f && fabsf(fy0) < 1073741824.0f;
            const int ix = ok ? (int)fx0 : -2, iy = ok ? (int)fy0 : -2;
            const int okn = fabsf(nx) < 1073741824.0f && fabsf(ny) < 1073741824.0f;
            const int jx = okn ? (int)nx : -1, jy = okn ? (int)ny : -1;
            if (ix >= 0 && iy >= 0 && ix + 2 <= sW && iy + 2 <= sH) {
                const float ax = sx - fx0, ay = sy - fy0;
                const uint8_t* s0 = src + (size_t)iy * sW + ix;
                const uint8_t* s1 = s0 + sW;
                float f = (1.0f - ax) * (1.0f - ay) * ((float)s0[0] / 255.0f);
                f += ax * (1.0f - ay) * ((float)s0[1] / 255.0f);
                f += (1.0f - ax) * ay * ((float)s1[0] / 255.0f);
                f += ax * ay * ((float)s1[1] / 255.0f);
                o = (uint8_t)orc_norm_to(f, 255.0f);
            } else if (jx >= 0 && jy >= 0 && jx < sW && jy < sH) {
                o = src[(size_t)jy * sW + jx];
            }
            out[(size_t)v * W + u] = o;
        }
    return 0;
}

/* ColorDisparityImage::Kernel::Process (ColorDisparityImage.cu:45-270).  disp: fmt 0
 * (4 x uint16 ids), 1 (uint16 ids), 2 (float ids); style 0 Analysis (RGBA8), 1 Gray
 * (uint16), 2 Color (RGBA8), 3 DispAbs (uint16), 4 DispRel (uint16).  colors: RGBA8 for
 * Unknown, CannotEvaluate, Saturated, NotSpecific, NotFound, NotUnique, Inconsistent. */
int orc_color_disparity(const void* disp, int fmt, int W, int H, int style, int dmin, int dmax, int inf_offset,
                        const uint8_t* colors, void* out)
{
    if (!disp || !out || !colors || W < 1 || H < 1 || fmt < 0 || fmt > 2 || style < 0 || style > 4 || dmax <= dmin)
        return -1;
    const float range = (float)(dmax - dmin);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            float v;    /* the id as the pixel type holds it */
            int valid;
            float fd;   /* (id - min) / range */
            if (fmt == 2) {
                v = ((const float*)disp)[i];
                valid = v >= (float)ORC_DISP_FIRST && v <= (float)ORC_DISP_LAST;
                fd = (v - (float)dmin) / range;
            } else {
                const int id = fmt == 0 ? ((const uint16_t*)disp)[4 * i] : ((const uint16_t*)disp)[i];
                v = (float)id;
                valid = id >= ORC_DISP_FIRST && id <= ORC_DISP_LAST;
                fd = (float)(id - dmin) / range;
            }
            uint8_t* rgba = (uint8_t*)out + 4 * i;
            uint16_t* g = (uint16_t*)out + i;
            switch (style) {
            case 0: {
                static const int codes[7] = {ORC_DISP_UNKNOWN, ORC_DISP_CANNOT_EVALUATE, ORC_DISP_SATURATED,
                                             ORC_DISP_NOT_SPECIFIC, ORC_DISP_NOT_FOUND, ORC_DISP_NOT_UNIQUE,
                                             ORC_DISP_INCONSISTENT};
                if (valid) {
                    const uint8_t c = (uint8_t)orc_norm_to(fd, 255.0f);
                    rgba[0] = rgba[1] = rgba[2] = c;
                    rgba[3] = 255;
                } else {
                    int k = 0;
                    while (k < 7 && v != (float)codes[k]) ++k;
                    if (k < 7) {
                        memcpy(rgba, colors + 4 * k, 4);
                    } else {
                        rgba[0] = 255;
                        rgba[1] = 0;
                        rgba[2] = 0;
                        rgba[3] = 255;
                    }
                }
                break;
            }
            case 1:
                *g = valid ? (uint16_t)orc_norm_to(fd, 65535.0f) : 0;
                break;
            case 2:
                if (valid) {
                    /* NormFloatToColor (Conversion.h:56-71) */
                    const float fB = 1.0f - fminf(fd / 0.7f, 1.0f);
                    const float fG = 1.0f - fminf(fabsf(fd - 0.5f) / 0.59f, 1.0f);
                    const float fR = 1.0f - fminf(fabsf(fd - 0.8f) / 0.7f, 1.0f);
                    const float fM = fmaxf(fmaxf(fR, fG), fB);
                    rgba[0] = (uint8_t)orc_norm_to(fR / fM, 255.0f);
                    rgba[1] = (uint8_t)orc_norm_to(fG / fM, 255.0f);
                    rgba[2] = (uint8_t)orc_norm_to(fB / fM, 255.0f);
                } else {
                    rgba[0] = rgba[1] = rgba[2] = 0;
                }
                rgba[3] = 255;
                break;
            case 3:
                *g = valid ? (uint16_t)v : 0;
                break;
            default:
                if (!valid) {
                    *g = 0;
                } else if (fmt == 2) {
                    /* float arithmetic, converted to 16 bits with truncation, saturating */
                    const float r = ((float)0x7FFF + (float)inf_offset) - (v - (float)ORC_DISP_FIRST);
                    *g = !(r > 0.0f) ? 0 : (r >= 65535.0f ? 65535 : (uint16_t)r);
                } else {
                    *g = (uint16_t)((0x7FFF + inf_offset) - ((int)v - ORC_DISP_FIRST));
                }
                break;
            }
        }
    return 0;
}
