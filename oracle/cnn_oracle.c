/* CPU oracle for the learned disparity-proposal network -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, the smoke check and the benchmark's CPU leg use this file, as the
 * checker.  The engine's network runs on the GPU (module.clutec.cuda_amd/csrc/cnn.hip).
 *
 * The network has no counterpart in the reference (it replaces the disparity input
 * of MatcherUseDisparity, BlockMatcherNCC.cu:894-1255, with a learned proposal).
 * This file restates the textbook operators it is built from, in FP64:
 *   3x3 convolution, zero padding, dilation r, NHWC activations, weights
 *   [Cout][3][3][Cin]:
 *     y[n,y,x,o] = b[o] + sum_{ky,kx,i} w[o,ky,kx,i] * x[n, y+(ky-1)r, x+(kx-1)r, i]
 *   and its gradients:
 *     dx[n,y,x,i] = sum_{ky,kx,o} w[o,ky,kx,i] * dy[n, y-(ky-1)r, x-(kx-1)r, o]
 *     dw[o,ky,kx,i] = sum_{n,y,x} dy[n,y,x,o] * x[n, y+(ky-1)r, x+(kx-1)r, i]
 *     db[o] = sum_{n,y,x} dy[n,y,x,o]
 * Pinned by central finite differences of a scalar loss and by an independent numpy
 * formulation in the tests.
 */
#include "cnn_oracle.h"

#include <stdlib.h>
#include <string.h>

int orc_cnn_version(void) { return 1; }

int orc_conv3x3_fwd(const float* x, int N, int H, int W, int Cin, const float* w, const float* b, int Cout, int dil,
                    float* y)
{
    if (!x || !w || !y || N <= 0 || H <= 0 || W <= 0 || Cin <= 0 || Cout <= 0 || dil <= 0) return -1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int n = 0; n < N; ++n)
        for (int yy = 0; yy < H; ++yy) {
            double* acc = (double*)malloc(sizeof(double) * Cout);
            for (int xx = 0; xx < W; ++xx) {
                for (int o = 0; o < Cout; ++o) acc[o] = b ? b[o] : 0.0;
                for (int ky = 0; ky < 3; ++ky) {
                    const int sy = yy + (ky - 1) * dil;
                    if (sy < 0 || sy >= H) continue;
                    for (int kx = 0; kx < 3; ++kx) {
                        const int sx = xx + (kx - 1) * dil;
                        if (sx < 0 || sx >= W) continue;
                        const float* xi = x + (((size_t)n * H + sy) * W + sx) * Cin;
                        for (int o = 0; o < Cout; ++o) {
                            const float* wo = w + ((size_t)o * 9 + ky * 3 + kx) * Cin;
                            double s = 0.0;
                            for (int i = 0; i < Cin; ++i) s += (double)wo[i] * (double)xi[i];
                            acc[o] += s;
                        }
                    }
                }
                float* yo = y + (((size_t)n * H + yy) * W + xx) * Cout;
                for (int o = 0; o < Cout; ++o) yo[o] = (float)acc[o];
            }
            free(acc);
        }
    return 0;
}

int orc_conv3x3_bwd(const float* x, int N, int H, int W, int Cin, const float* w, const float* dy, int Cout, int dil,
                    float* dx, double* dw, double* db)
{
    if (!x || !w || !dy || N <= 0 || H <= 0 || W <= 0 || Cin <= 0 || Cout <= 0 || dil <= 0) return -1;
    if (dx) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int n = 0; n < N; ++n)
            for (int yy = 0; yy < H; ++yy) {
                double* acc = (double*)malloc(sizeof(double) * Cin);
                for (int xx = 0; xx < W; ++xx) {
                    for (int i = 0; i < Cin; ++i) acc[i] = 0.0;
                    for (int ky = 0; ky < 3; ++ky) {
                        const int sy = yy - (ky - 1) * dil;
                        if (sy < 0 || sy >= H) continue;
                        for (int kx = 0; kx < 3; ++kx) {
                            const int sx = xx - (kx - 1) * dil;
                            if (sx < 0 || sx >= W) continue;
                            const float* g = dy + (((size_t)n * H + sy) * W + sx) * Cout;
                            for (int o = 0; o < Cout; ++o) {
                                const float* wo = w + ((size_t)o * 9 + ky * 3 + kx) * Cin;
                                const double go = g[o];
                                for (int i = 0; i < Cin; ++i) acc[i] += go * (double)wo[i];
                            }
                        }
                    }
                    float* o = dx + (((size_t)n * H + yy) * W + xx) * Cin;
                    for (int i = 0; i < Cin; ++i) o[i] = (float)acc[i];
                }
                free(acc);
            }
    }
    if (dw) {
        const int K = 9 * Cin;
#pragma omp parallel for schedule(static)
        for (int o = 0; o < Cout; ++o) {
            for (int k = 0; k < K; ++k) dw[(size_t)o * K + k] = 0.0;
            double bs = 0.0;
            for (int n = 0; n < N; ++n)
                for (int yy = 0; yy < H; ++yy)
                    for (int xx = 0; xx < W; ++xx) {
                        const double g = dy[(((size_t)n * H + yy) * W + xx) * Cout + o];
                        bs += g;
                        if (g == 0.0) continue;
                        for (int ky = 0; ky < 3; ++ky) {
                            const int sy = yy + (ky - 1) * dil;
                            if (sy < 0 || sy >= H) continue;
                            for (int kx = 0; kx < 3; ++kx) {
                                const int sx = xx + (kx - 1) * dil;
                                if (sx < 0 || sx >= W) continue;
                                const float* xi = x + (((size_t)n * H + sy) * W + sx) * Cin;
                                double* d = dw + (size_t)o * K + (ky * 3 + kx) * Cin;
                                for (int i = 0; i < Cin; ++i) d[i] += g * (double)xi[i];
                            }
                        }
                    }
            if (db) db[o] = bs;
        }
    }
    return 0;
}
