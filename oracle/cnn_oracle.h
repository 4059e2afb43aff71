/* CPU oracle for the learned disparity-proposal network -- TEST INFRASTRUCTURE ONLY. */
#ifndef CLUTEC_CNN_ORACLE_H
#define CLUTEC_CNN_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif
int orc_cnn_version(void);
/* 3x3 convolution, zero padding, dilation dil; x [N][H][W][Cin], w [Cout][3][3][Cin],
 * b [Cout] (may be NULL), y [N][H][W][Cout]; FP64 accumulation. */
int orc_conv3x3_fwd(const float* x, int N, int H, int W, int Cin, const float* w, const float* b, int Cout, int dil,
                    float* y);
/* Gradients of the same convolution for upstream gradient dy [N][H][W][Cout]:
 * dx [N][H][W][Cin] (may be NULL), dw [Cout][3][3][Cin] and db [Cout] in FP64 (may be NULL). */
int orc_conv3x3_bwd(const float* x, int N, int H, int W, int Cin, const float* w, const float* dy, int Cout, int dil,
                    float* dx, double* dw, double* db);
#ifdef __cplusplus
}
#endif
#endif
