/* CPU oracle (test infrastructure only): the reference's NCC and adaptive-window AW1
 * operators, emulated block by block.
 *
 *   CluTec.Cuda.MultiView/BlockMatcherNCC.cu   (Matcher :614-884, MatcherUseDisparity :894-1255)
 *   CluTec.Cuda.MultiView/BlockMatcherAW1.cu   (Matcher :657-913, MatcherUseDisparity :923-1263)
 *
 * Both launch one 32-thread block per tile of result origins, only full blocks
 * (EvalThreadConfig_ReadDepBlockSize, KernelDriverBase.cpp:166-197, bOnlyFullBlocks
 * defaulting to true at KernelDriverBase.h:117-121).  A block decides which disparities
 * its results see -- groups of 16 from the range start while the group start is below
 * min(Max, MaxDisparity(X0)), a full group when its 16 + TBX - 1 test columns lie in the
 * right image, else a narrower group placed as the reference places it -- and each result
 * tracks the maximum of its similarity over those disparities in order, with the values
 * around it.  The similarities:
 *   NCC: (SIJ / N - mean_I mean_J) / (sd_I sd_J) from exact integer window sums with the
 *        reference's float formula (EvalSim :180-336, EvalMsd :338-373), clamped at 0 and
 *        truncated to a byte, 0 when sd_I sd_J < 1;
 *   AW1: the centre sub-patch plus the two largest of its four edge neighbours
 *        (Kernel.Algo.AdaptWnd_W32H16.h:240-285) of 255 - |a - b| or the signed product
 *        (BlockMatcherAW1.cu:191-222).
 * Floating point is evaluated as written: one IEEE single operation per operator, no
 * contraction (the Makefile builds with -ffp-contract=off), correctly rounded division
 * and square root; CUDA's max / min of floats are fmaxf / fminf.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "stereo_oracle.h"

#define SDC 16 /* SubDispCount */

/* The group of disparities at D0 of a block at (X0, Y0) with tbx x tby read area
 * (Matcher :789-876 of both operators).  n = 0: nothing evaluated.  Disparity D0 + r
 * (r < n) tests the right columns starting at base - r (left to right) or base + r. */
static void ref_group(int X0, int Y0, int D0, int W, int H, int tbx, int tby, int offset, int ltor, int cap, int* n,
                      int* base)
{
    const int tx = (ltor ? X0 - D0 - offset : X0 + D0 + offset) - ltor * (SDC - 1);
    *n = 0;
    *base = 0;
    if (tx >= 0 && Y0 >= 0 && tx + tbx + SDC - 1 <= W && Y0 + tby <= H) {
        *n = cap;
        *base = tx + ltor * (SDC - 1);
        return;
    }
    if (!(Y0 >= 0 && Y0 + tby <= H)) return;
    int dx, sx;
    if (tx >= 0) {
        dx = W - tx - tbx;
        sx = tx;
    } else {
        dx = tx + (SDC - 1);
        sx = 0;
    }
    if (dx < 0) return;
    *n = dx + 1;
    *base = sx + ltor * (*n - 1);
}

/* MatcherUseDisparity's block range (NCC :960-1001, AW1 :970-1011).  res_of_thread(t, r)
 * gives the result index of the reference thread t's r-th result (or -1), written[k] says
 * whether result k is written (and so reads the initial map).  Fills disp[] with the
 * initial values and returns 1 with [*lo, *hi] when the reduction is valid. */
typedef int (*res_map_fn)(int t, int r, const int* geo);

static int ref_band(const uint16_t* init, int iW, int iH, int W, int X0, int Y0, int half, int nthr, int per_thr,
                    res_map_fn map, const int* geo, const uint8_t* written, uint16_t* disp, int* lo, int* hi)
{
    const int scale = W / iW;
    const int pcx = geo[0];
    int red_lo = -1, red_hi = -1;
    for (int t = 0; t < nthr; ++t) {
        int tlo = ORC_DISP_LAST, thi = ORC_DISP_FIRST;
        for (int r = 0; r < per_thr; ++r) {
            const int k = map(t, r, geo);
            if (k < 0 || !written[k]) continue;
            const int ix = k % pcx, iy = k / pcx;
            const int px = (X0 + ix + half) / scale, py = (Y0 + iy + half) / scale;
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
        /* MinDisp / MaxDisp (:418-457) keep only valid operands over the warp */
        if (tlo >= ORC_DISP_FIRST && tlo <= ORC_DISP_LAST) red_lo = red_lo < 0 ? tlo : (tlo < red_lo ? tlo : red_lo);
        if (thi >= ORC_DISP_FIRST && thi <= ORC_DISP_LAST) red_hi = red_hi < 0 ? thi : (thi > red_hi ? thi : red_hi);
    }
    *lo = red_lo;
    *hi = red_hi;
    return !(red_lo < 0 || red_hi < 0 || red_lo > red_hi);
}

typedef struct {
    int prev2, prev, max, next, next2, last;
} ref_store;

/* The peak update of one result at disparity d (NCC EvalSim :280-333, AW1 EvalSimAW :262-315):
 * a strictly larger value becomes the maximum, with the value before it (and the one before
 * the previous maximum); the next two values after the maximum are recorded. */
static inline void ref_update(ref_store* s, uint16_t* disp, int v, int d)
{
    const int above = v > s->max;
    const int behind1 = !above && (int)*disp + 1 == d + ORC_DISP_FIRST;
    const int behind2 = !above && (int)*disp + 2 == d + ORC_DISP_FIRST;
    if (above) {
        s->prev2 = s->prev;
        s->prev = s->last;
        s->max = v;
        s->next = v;
        s->next2 = v;
        *disp = (uint16_t)(d + ORC_DISP_FIRST);
    }
    if (behind1) s->next = v;
    if (behind2) s->next2 = v;
    s->last = v;
}

static inline void put4(uint16_t* o, int a, int b, int c, int d)
{
    o[0] = (uint16_t)a;
    o[1] = (uint16_t)b;
    o[2] = (uint16_t)c;
    o[3] = (uint16_t)d;
}

/* ---------------------------------------------------------------------------------- */
/* NCC.  AlgoArraySum_W32H16<32, P, P> (Kernel.Algo.ArraySum_W32H16.h:49-100): 16 result  */
/* rows, two column groups of CG = (33 - P) / 2 results, PCX = 2 CG; the read area is     */
/* 33 x (15 + P) (TestBlockSizeX = 32 + 1).  Reference thread t holds the results         */
/* ((t / 16) CG + r, t % 16), r < CG.                                                     */

static int ncc_res_map(int t, int r, const int* geo)
{
    const int pcx = geo[0], cg = pcx / 2;
    if (r >= cg) return -1;
    return (t % 16) * pcx + (t / 16) * cg + r;
}

int orc_ncc_ref(const uint8_t* L, const uint8_t* R, int W, int H, int P, int dmin, int dmax, int offset, int ltor,
                int min_delta, float ncc_thresh, float grad_thresh, const uint16_t* init, int iW, int iH,
                uint16_t* out)
{
    if (!L || !R || !out || (P != 5 && P != 9 && P != 11) || W < 48 || H < 15 + P || dmin < 0 || dmax <= dmin ||
        dmax > 4096 || offset < -4096 || offset > 4096 || (init && (iW < 1 || iH < 1 || iW > W || iH > H)))
        return -1;
    const int CG = (32 - P + 1) / 2, PCX = 2 * CG, PCY = 16, TBX = 33, TBY = PCY + P - 1;
    const int nbx = (W - (TBX - PCX)) / PCX, nby = (H - (TBY - PCY)) / PCY;
    const int NR = PCX * PCY;
    const float fN = (float)(P * P);
    const int geo[1] = {PCX};
#pragma omp parallel for schedule(dynamic, 1)
    for (int by = 0; by < nby; ++by) {
        uint8_t lc[30][33];
        uint16_t disp[28 * 16];
        uint8_t written[28 * 16];
        ref_store st[28 * 16];
        float bmean[28 * 16], bsd[28 * 16];
        int c1[16][32], c2[16][32], c3[16][32];
        for (int bx = 0; bx < nbx; ++bx) {
            const int X0 = bx * PCX, Y0 = by * PCY;
            if (!(X0 + P <= W && Y0 + P <= H)) continue; /* :657-661 */
            for (int y = 0; y < TBY; ++y)
                for (int x = 0; x < TBX; ++x)
                    lc[y][x] = (X0 + x < W && Y0 + y < H) ? L[(size_t)(Y0 + y) * W + X0 + x] : 0;
            for (int k = 0; k < NR; ++k) {
                const int ix = k % PCX, iy = k / PCX;
                written[k] = X0 + ix + P < W && Y0 + iy + P < H; /* IsInside(x + P, y + P) */
                disp[k] = ORC_DISP_UNKNOWN;
                st[k] = (ref_store){0, 0, 0, 0, 0, 0};
            }
            int d_start, d_stop, cap, any = 1;
            if (!init) {
                /* EvalGrad (:388-410): mean |L(x) - L(x + 1)| over the window */
                any = 0;
                for (int k = 0; k < NR; ++k) {
                    const int ix = k % PCX, iy = k / PCX;
                    int g = 0;
                    for (int j = 0; j < P; ++j)
                        for (int i = 0; i < P; ++i) g += abs((int)lc[iy + j][ix + i] - (int)lc[iy + j][ix + i + 1]);
                    const float fv = (float)g / fN;
                    if (fv < grad_thresh) disp[k] = ORC_DISP_CANNOT_EVALUATE;
                    else any = 1;
                }
                if (!any) { /* __syncthreads_or early-out (:754-775) */
                    for (int k = 0; k < NR; ++k)
                        if (written[k]) {
                            const int ix = k % PCX, iy = k / PCX;
                            put4(out + 4 * ((size_t)(Y0 + iy + P / 2) * W + X0 + ix + P / 2), ORC_DISP_CANNOT_EVALUATE,
                                 0, 0, 0);
                        }
                    continue;
                }
                const int maxd = ltor ? X0 - offset + 1 : W - X0 - offset; /* MaxDisparity */
                d_start = dmin;
                d_stop = dmax < maxd ? dmax : maxd;
                cap = SDC;
            } else {
                int lo, hi;
                if (!ref_band(init, iW, iH, W, X0, Y0, P / 2, 32, CG, ncc_res_map, geo, written, disp, &lo, &hi)) {
                    for (int k = 0; k < NR; ++k)
                        if (written[k]) {
                            const int ix = k % PCX, iy = k / PCX;
                            put4(out + 4 * ((size_t)(Y0 + iy + P / 2) * W + X0 + ix + P / 2), disp[k], 0, 0, 0);
                        }
                    continue;
                }
                d_start = lo < 4 ? 0 : lo - 4;
                const int d_end = hi + 4; /* inclusive (:1047-1048) */
                any = d_end <= 32767;     /* the 16-bit loop variables wrap beyond */
                d_stop = d_end + 1;
                cap = d_end - d_start + 1 < SDC ? d_end - d_start + 1 : SDC; /* :1062 */
            }
            /* EvalMsd (:338-373): the base window's mean and standard deviation */
            for (int k = 0; k < NR; ++k) {
                const int ix = k % PCX, iy = k / PCX;
                int s = 0, s2 = 0;
                for (int j = 0; j < P; ++j)
                    for (int i = 0; i < P; ++i) {
                        const int v = lc[iy + j][ix + i];
                        s += v;
                        s2 += v * v;
                    }
                const float m = (float)s / fN, m2 = (float)s2 / fN;
                const float sq = m * m;
                bmean[k] = m;
                bsd[k] = sqrtf(m2 - sq);
            }
            for (int D0 = d_start; any && D0 < d_stop; D0 += SDC) {
                int n, base;
                ref_group(X0, Y0, D0, W, H, TBX, TBY, offset, ltor, cap, &n, &base);
                for (int r = 0; r < n; ++r) {
                    const int d = D0 + r, t0 = ltor ? base - r : base + r;
                    /* window sums of R, R^2 and L R, separably: P-row column sums, then
                       P-column sums (exact integers, so the order does not matter) */
                    for (int y = 0; y < PCY; ++y)
                        for (int x = 0; x < PCX + P - 1; ++x) {
                            int a1 = 0, a2 = 0, a3 = 0;
                            for (int j = 0; j < P; ++j) {
                                const int v = R[(size_t)(Y0 + y + j) * W + t0 + x];
                                a1 += v;
                                a2 += v * v;
                                a3 += (int)lc[y + j][x] * v;
                            }
                            c1[y][x] = a1;
                            c2[y][x] = a2;
                            c3[y][x] = a3;
                        }
                    for (int k = 0; k < NR; ++k) {
                        if (!written[k] || disp[k] > ORC_DISP_LAST) continue;
                        const int ix = k % PCX, iy = k / PCX;
                        int s1 = 0, s2 = 0, s3 = 0;
                        for (int i = 0; i < P; ++i) {
                            s1 += c1[iy][ix + i];
                            s2 += c2[iy][ix + i];
                            s3 += c3[iy][ix + i];
                        }
                        const float cross = (float)s3 / fN;
                        const float tm = (float)s1 / fN, tm2 = (float)s2 / fN;
                        const float tsq = tm * tm;
                        const float tsd = sqrtf(tm2 - tsq);
                        const float norm = bsd[k] * tsd;
                        float ncc;
                        if (norm < 1.0f) {
                            ncc = 0.0f;
                        } else {
                            const float mm = bmean[k] * tm;
                            ncc = (cross - mm) / norm;
                        }
                        ncc = fmaxf(0.0f, ncc);
                        const int u = (int)(uint8_t)fminf(255.0f, ncc * 255.0f);
                        ref_update(&st[k], &disp[k], u, d);
                    }
                }
            }
            /* WriteDisparity (:491-592) */
            for (int k = 0; k < NR; ++k) {
                if (!written[k]) continue;
                const int ix = k % PCX, iy = k / PCX;
                uint16_t* o = out + 4 * ((size_t)(Y0 + iy + P / 2) * W + X0 + ix + P / 2);
                const int dv = disp[k];
                if (dv > ORC_DISP_LAST || dv == ORC_DISP_UNKNOWN) {
                    put4(o, dv, 0, 0, 0);
                    continue;
                }
                const ref_store* s = &st[k];
                const float fmax = (float)s->max / 255.0f;
                const int above = fmax >= ncc_thresh;
                const int d1a = abs(s->prev - s->max), d1b = abs(s->prev2 - s->max);
                const int d2a = abs(s->next - s->max), d2b = abs(s->next2 - s->max);
                const int diff1 = d1a > d1b ? d1a : d1b, diff2 = d2a > d2b ? d2a : d2b;
                int code;
                if (above) {
                    const int sat = dv <= ORC_DISP_FIRST + dmin + 1 || dv >= ORC_DISP_FIRST + dmax - 1;
                    code = sat ? ORC_DISP_SATURATED
                               : ((diff1 >= min_delta && diff2 >= min_delta) ? dv : ORC_DISP_NOT_SPECIFIC);
                } else {
                    code = ORC_DISP_NOT_FOUND;
                }
                put4(o, code, s->max, diff1, diff2);
            }
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------------------- */
/* AW1.  AlgoAdaptWnd_W32H16<32, s, s> (Kernel.Algo.AdaptWnd_W32H16.h:46-140): sub-patch   */
/* sums over the 2 CG = 2 ((33 - s) / 2) left columns and 16 top rows of a 32 x (PCY + 3s */
/* - 1) read area; PCX = 2 CG - 2 s by PCY = 16 - 2 s results (18 x 6 for s = 5, 24 x 10 */
/* for s = 3); reference thread t holds the results r 32 + t (row-major) for r below       */
/* EvalLoopCount = ceil(PCX PCY / 32).                                                    */

static int aw_res_map(int t, int r, const int* geo)
{
    const int k = r * 32 + t;
    return k < geo[1] ? k : -1;
}

/* EvalSim (:191-222) */
static inline int aw1_sim(int sim_type, int a, int b)
{
    if (sim_type == 0) return 255 - abs(a - b);
    float f1 = 2.0f * ((float)a / 255.0f - 0.5f);
    const float f2 = 2.0f * ((float)b / 255.0f - 0.5f);
    f1 *= f2;
    f1 = f1 * 255.0f;
    return (int)(int16_t)(int)f1; /* toward zero, into short */
}

/* SimValueToFloat (:155-159), SimFloatToInteger (:172-176) */
static inline float aw1_sim_float(int v, int ne) { return ((float)v / (float)ne) * (100.0f / 255.0f); }
static inline uint16_t aw1_sim_u16(float f)
{
    float t = (f / 100.0f + 1.0f) / 2.0f;
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return (uint16_t)(t * 65535.0f);
}

int orc_aw1_ref(const uint8_t* L, const uint8_t* R, int W, int H, int s, int sim_type, int dmin, int dmax, int offset,
                int ltor, float min_delta, float sim_thresh, float grad_thresh, const uint16_t* init, int iW, int iH,
                uint16_t* out)
{
    if (!L || !R || !out || (s != 3 && s != 5) || (sim_type != 0 && sim_type != 1) || W < 48 || dmin < 0 ||
        dmax <= dmin || dmax > 4096 || offset < -4096 || offset > 4096 ||
        (init && (iW < 1 || iH < 1 || iW > W || iH > H)))
        return -1;
    const int CG = (32 - s + 1) / 2, SPCX = 2 * CG, P3 = 3 * s;
    const int PCX = SPCX - 2 * s, PCY = 16 - 2 * s, TBX = 32, TBY = PCY + P3 - 1, NE = 3 * s * s;
    if (H < TBY) return -1;
    const int NR = PCX * PCY, ELC = (NR + 31) / 32;
    const int nbx = (W - (TBX - PCX)) / PCX, nby = (H - (TBY - PCY)) / PCY;
    const int geo[2] = {PCX, NR};
#pragma omp parallel for schedule(dynamic, 1)
    for (int by = 0; by < nby; ++by) {
        uint8_t lc[20][32];
        int col[16][32], sp[16][32];
        uint16_t disp[240];
        uint8_t written[240];
        ref_store st[240];
        for (int bx = 0; bx < nbx; ++bx) {
            const int X0 = bx * PCX, Y0 = by * PCY;
            if (!(X0 + P3 <= W && Y0 + P3 <= H)) continue; /* :680-684 */
            for (int y = 0; y < TBY; ++y)
                for (int x = 0; x < TBX; ++x)
                    lc[y][x] = (X0 + x < W && Y0 + y < H) ? L[(size_t)(Y0 + y) * W + X0 + x] : 0;
            for (int k = 0; k < NR; ++k) {
                const int ix = k % PCX, iy = k / PCX;
                written[k] = X0 + ix + P3 < W && Y0 + iy + P3 < H;
                disp[k] = ORC_DISP_UNKNOWN;
                st[k] = (ref_store){0, 0, -32767, 0, 0, 0};
            }
            int d_start, d_stop, cap, any = 1;
            if (!init) {
                /* EvalGrad (:331-417): the P3 x P3 window's standard deviation, truncated to
                   an int, in % of 255 */
                any = 0;
                for (int k = 0; k < NR; ++k) {
                    const int ix = k % PCX, iy = k / PCX;
                    int sum = 0, sum2 = 0;
                    for (int j = 0; j < P3; ++j)
                        for (int i = 0; i < P3; ++i) {
                            const int v = lc[iy + j][ix + i];
                            sum += v;
                            sum2 += v * v;
                        }
                    const float m = (float)sum / (float)(P3 * P3), m2 = (float)sum2 / (float)(P3 * P3);
                    const float sq = m * m;
                    const int isd = (int)sqrtf(fabsf(m2 - sq));
                    const float sd = (float)isd / 255.0f * 100.0f;
                    if (sd < grad_thresh) disp[k] = ORC_DISP_CANNOT_EVALUATE;
                    else any = 1;
                }
                if (!any) {
                    for (int k = 0; k < NR; ++k)
                        if (written[k]) {
                            const int ix = k % PCX, iy = k / PCX;
                            put4(out + 4 * ((size_t)(Y0 + iy + P3 / 2) * W + X0 + ix + P3 / 2),
                                 ORC_DISP_CANNOT_EVALUATE, 0, 0, 0);
                        }
                    continue;
                }
                const int maxd = ltor ? X0 - offset + 1 : W - X0 - offset;
                d_start = dmin;
                d_stop = dmax < maxd ? dmax : maxd;
                cap = SDC;
            } else {
                int lo, hi;
                if (!ref_band(init, iW, iH, W, X0, Y0, P3 / 2, 32, ELC, aw_res_map, geo, written, disp, &lo, &hi)) {
                    for (int k = 0; k < NR; ++k)
                        if (written[k]) {
                            const int ix = k % PCX, iy = k / PCX;
                            put4(out + 4 * ((size_t)(Y0 + iy + P3 / 2) * W + X0 + ix + P3 / 2), disp[k], 0, 0, 0);
                        }
                    continue;
                }
                d_start = lo < 4 ? 0 : lo - 4;
                const int d_end = hi + 4;
                any = d_end <= 32767;
                d_stop = d_end + 1;
                cap = d_end - d_start + 1 < SDC ? d_end - d_start + 1 : SDC;
            }
            for (int D0 = d_start; any && D0 < d_stop; D0 += SDC) {
                int n, base;
                ref_group(X0, Y0, D0, W, H, TBX, TBY, offset, ltor, cap, &n, &base);
                for (int r = 0; r < n; ++r) {
                    const int d = D0 + r, t0 = ltor ? base - r : base + r;
                    /* sub-patch sums with their top-left at (x, y), x < 2 CG, y < 16
                       (SumCacheStore, Kernel.Algo.ArraySum_W32H16.h:277-329) */
                    for (int y = 0; y < 16; ++y)
                        for (int x = 0; x < 32; ++x) {
                            int c = 0;
                            for (int j = 0; j < s; ++j)
                                c += aw1_sim(sim_type, lc[y + j][x], R[(size_t)(Y0 + y + j) * W + t0 + x]);
                            col[y][x] = c;
                        }
                    for (int y = 0; y < 16; ++y)
                        for (int x = 0; x < SPCX; ++x) {
                            int c = 0;
                            for (int i = 0; i < s; ++i) c += col[y][x + i];
                            sp[y][x] = c;
                        }
                    for (int k = 0; k < NR; ++k) {
                        if (disp[k] > ORC_DISP_LAST) continue;
                        const int ix = k % PCX, iy = k / PCX;
                        /* Two_OutOf_Four (Kernel.Algo.AdaptWnd_W32H16.h:240-285) */
                        int m1 = sp[iy][ix + s], m2 = -32767, v;
                        v = sp[iy + s][ix];
                        m2 = m2 > (m1 < v ? m1 : v) ? m2 : (m1 < v ? m1 : v);
                        m1 = m1 > v ? m1 : v;
                        v = sp[iy + 2 * s][ix + s];
                        m2 = m2 > (m1 < v ? m1 : v) ? m2 : (m1 < v ? m1 : v);
                        m1 = m1 > v ? m1 : v;
                        v = sp[iy + s][ix + 2 * s];
                        m2 = m2 > (m1 < v ? m1 : v) ? m2 : (m1 < v ? m1 : v);
                        m1 = m1 > v ? m1 : v;
                        const int aw = (int)(int16_t)(m1 + m2 + sp[iy + s][ix + s]);
                        ref_update(&st[k], &disp[k], aw, d);
                    }
                }
            }
            /* WriteDisparity (:509-635) */
            for (int k = 0; k < NR; ++k) {
                if (!written[k]) continue;
                const int ix = k % PCX, iy = k / PCX;
                uint16_t* o = out + 4 * ((size_t)(Y0 + iy + P3 / 2) * W + X0 + ix + P3 / 2);
                const int dv = disp[k];
                if (dv > ORC_DISP_LAST || dv == ORC_DISP_UNKNOWN) {
                    put4(o, dv, 0, 0, 0);
                    continue;
                }
                const ref_store* q = &st[k];
                const float fmax = aw1_sim_float(q->max, NE);
                const float fp2 = aw1_sim_float(q->prev2, NE), fp = aw1_sim_float(q->prev, NE);
                const float fn = aw1_sim_float(q->next, NE), fn2 = aw1_sim_float(q->next2, NE);
                const float dprev = fmaxf(fabsf(fp - fmax), fabsf(fp2 - fmax));
                const float dnext = fmaxf(fabsf(fn - fmax), fabsf(fn2 - fmax));
                const int specific = dprev >= min_delta && dnext >= min_delta;
                int code;
                if (fmax >= sim_thresh) {
                    const int sat = dv <= ORC_DISP_FIRST + dmin + 1 || dv >= ORC_DISP_FIRST + dmax - 1;
                    code = sat ? ORC_DISP_SATURATED : (specific ? dv : ORC_DISP_NOT_SPECIFIC);
                } else {
                    code = ORC_DISP_NOT_FOUND;
                }
                put4(o, code, aw1_sim_u16(fmax), aw1_sim_u16(dprev), aw1_sim_u16(dnext));
            }
        }
    }
    return 0;
}
