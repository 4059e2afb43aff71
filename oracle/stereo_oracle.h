/*
 * CPU oracle for the stereo disparity hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP engine is
 * compared against and the CPU baseline that bench.py times beside it.  Nothing
 * in the shipped engine links, loads or calls it.
 *
 * Every function restates the arithmetic of the reference operator it names
 * (file:line in the .c file) on plain row-major arrays (pitch == width).
 */
#ifndef CLUTEC_STEREO_ORACLE_H
#define CLUTEC_STEREO_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_COST_SAD = 0, ORC_COST_CENSUS = 1 };

/* Disparity identifiers, values of EDisparityId (CluTec.Cuda.MultiView/DisparityId.h:33-45). */
enum {
    ORC_DISP_UNKNOWN = 0,
    ORC_DISP_FIRST = 1,
    ORC_DISP_LAST = 0xFFF0 - 1,
    ORC_DISP_CANNOT_EVALUATE = 0xFFFF,
    ORC_DISP_SATURATED = 0xFFFF - 1,
    ORC_DISP_NOT_FOUND = 0xFFFF - 2,
    ORC_DISP_NOT_SPECIFIC = 0xFFFF - 3,
    ORC_DISP_NOT_UNIQUE = 0xFFFF - 4,
    ORC_DISP_INCONSISTENT = 0xFFFF - 5
};

int orc_version(void);

/* Census with a cw x ch window (odd sizes, cw*ch <= 64) into uint64. */
int orc_census(const uint8_t* img, int W, int H, int cw, int ch, uint64_t* out);
/* Reference 3x3 census into uint8. */
int orc_census3x3_ref(const uint8_t* img, int W, int H, uint8_t* out);

/* SAD (L,R uint8) or census-Hamming (L,R uint64) block matcher, plain spec,
 * separable running sums.  Outputs int16 maps (-1 invalid).  cost_out (optional)
 * receives the winning aggregated cost of the left map (-1 where invalid). */
int orc_bm(int cost_kind, const void* L, const void* R, int W, int H, int D, int ww, int wh,
           int cw, int ch, int tol, int16_t* dL, int16_t* dR, int16_t* out, int32_t* cost_out,
           int nthreads);

/* Same specification evaluated by direct 2-D window sums (no running sums).
 * volume (optional, W*H*D int32, index (y*W+x)*D+d) receives every aggregated cost,
 * -1 for invalid (pixel, d) combinations. */
/* orc_bm restricted to candidates whose 8-disparity bin d >> 3 is set in bins[p]
 * of the left pixel p (D <= 256). */
int orc_bm_bins(int kind, const void* L, const void* R, int W, int H, int D, int ww, int wh, int cw,
                int ch, int tol, const uint32_t* bins, int16_t* dL, int16_t* dR, int16_t* out, int nthreads);
int orc_bm_brute(int cost_kind, const void* L, const void* R, int W, int H, int D, int ww, int wh,
                 int cw, int ch, int tol, int16_t* dL, int16_t* dR, int16_t* out, int32_t* volume);

/* Normalised cross-correlation, FP64.  score_out (optional) = winning score of the
 * left map (NaN where invalid); volume (optional, (y*W+x)*D+d) = every score. */
int orc_ncc(const uint8_t* L, const uint8_t* R, int W, int H, int D, int ww, int wh, int tol,
            int16_t* dL, int16_t* dR, int16_t* out, double* score_out, double* volume, int nthreads);
/* orc_ncc restricted to candidates whose 8-disparity bin d >> 3 is set in bins[p] of the
 * left pixel p (D <= 256): the band search. */
int orc_ncc_bins(const uint8_t* L, const uint8_t* R, int W, int H, int D, int ww, int wh, int tol,
                 const uint32_t* bins, int16_t* dL, int16_t* dR, int16_t* out, double* score_out, int nthreads);

/* Adaptive-window matchers.  mode 1 = AW1 similarity (255-|a-b|, two best of four
 * neighbouring sub-patches, maximise); mode 2 = AW2 SAD (two smallest, minimise).
 * s = sub-patch size (5 for the 15x15 window).  value_out (optional) = winning
 * aggregated value of the left map (-1 invalid); volume optional as above. */
int orc_aw(int mode, const uint8_t* L, const uint8_t* R, int W, int H, int D, int s, int tol,
           int16_t* dL, int16_t* dR, int16_t* out, int32_t* value_out, int32_t* volume, int nthreads);
/* orc_aw restricted to candidates whose 8-disparity bin d >> 3 is set in bins[p] of the
 * left pixel p (D <= 256): the band search. */
int orc_aw_bins(int mode, const uint8_t* L, const uint8_t* R, int W, int H, int D, int s, int tol,
                const uint32_t* bins, int16_t* dL, int16_t* dR, int16_t* out, int32_t* value_out, int nthreads);

/* Plain left-right check of two int16 maps. */
int orc_crosscheck_plain(const int16_t* dL, const int16_t* dR, int W, int H, int tol, int16_t* out);

/* Reference cross-check of two 4 x uint16 disparity-id maps. */
int orc_crosscheck_ref(const uint16_t* lr, const uint16_t* rl, int W, int H, int thresh, int ltor,
                       int offset, uint16_t* out, uint16_t* dmin, uint16_t* dmax);

/* CreateDispMipMaps (CreateDispMipMaps.cu).  Level 0 as float: the reference map's
 * channel 0 id when First <= id <= Last, else Unknown (0) (CopyDispChannel :66-94); the
 * plain map's d >= 0, else -1. */
int orc_mip_level0_ref(const uint16_t* rgba, int W, int H, float* out);
int orc_mip_level0_plain(const int16_t* d, int W, int H, float* out);
/* The next level, ((W-1)/2) x ((H-1)/2): [1 2 1] horizontally then vertically, values
 * equal to `unknown` skipped and the weights of the rest normalised, the vertical pass
 * halving (SmoothHorizontalLum :97-150, SmoothVerticalLum :153-218). */
int orc_mip_next(const float* src, int W, int H, float unknown, float* out);

/* Statistics::DispDenoise::FilterNoBorder (Statistics.DispDenoise.cu:110-330) on the whole
 * image: a P x P window (offsets -P/2 .. P-1-P/2) whose positions outside the image are
 * moved by P into it (:160-165); count, sum and sum of squares of the valid values; see
 * the .c file for the decision.  ref_format 1: 4 x uint16 EDisparityId pixels in, uint16
 * ids out (valid: First..Last, Unknown 0); 0: int16 d in and out (valid: d >= 0,
 * unknown -1).  W, H >= P. */
int orc_disp_denoise(int ref_format, const void* in, int W, int H, int P, float support_ratio, float mean_cutoff,
                     int fill, void* out);

/* BlockMatcher1 (BlockMatcher1.cu:590-847, and MatcherUseDisparity :857-1203 when init is
 * given), emulating the reference's block tiling: 32-thread blocks of PCX x 4 result
 * origins, P in {5, 9, 11, 15}, sim_type 0 (AbsoluteDifference) or 1 (SignedProduct),
 * disparities dmin <= d < dmax with offset, ltor 1 for left to right.  out is W*H*4
 * uint16 (EDisparityId, similarity max, prev delta, next delta); pixels the reference
 * does not write are left as they are.  init: an iW x iH map of 4 x uint16 ids with
 * W / iW >= 1, or NULL.  Needs W >= 48 and H >= P + 3. */
int orc_bm1(const uint8_t* L, const uint8_t* R, int W, int H, int P, int sim_type, int dmin, int dmax, int offset,
            int ltor, float min_delta, float sim_thresh, float grad_thresh, const uint16_t* init, int iW, int iH,
            uint16_t* out);

/* The reference's BlockMatcherNCC (BlockMatcherNCC.cu:614-884, and MatcherUseDisparity
 * :894-1255 when init is given), emulating its block tiling: 32-thread blocks of
 * (33 - P) / 2 * 2 x 16 result origins from a 33 x (15 + P) read area, P in {5, 9, 11}.
 * The similarity is the reference's float NCC truncated to a byte; the WTA takes a
 * strictly larger byte.  out is W*H*4 uint16: EDisparityId code, the byte maximum and the
 * two neighbour deltas (iMinDeltaThresh applies to them, fNccThresh to max / 255);
 * unwritten pixels are left as they are.  Needs W >= 48 and H >= 15 + P. */
int orc_ncc_ref(const uint8_t* L, const uint8_t* R, int W, int H, int P, int dmin, int dmax, int offset, int ltor,
                int min_delta, float ncc_thresh, float grad_thresh, const uint16_t* init, int iW, int iH,
                uint16_t* out);

/* The reference's BlockMatcherAW1 (BlockMatcherAW1.cu:657-913, MatcherUseDisparity
 * :923-1263), emulating its block tiling: sub-patch size s in {5, 3} (EConfig Patch_15x15,
 * Patch_9x9), blocks of 18 x 6 or 24 x 10 result origins from a 32 x 20 or 32 x 18 read
 * area, sim_type 0 (255 - |a - b|) or 1 (signed product).  Output as orc_bm1, with the
 * similarity channels in % of the 3 s^2-pixel window's maximum.  Needs W >= 48 and
 * H >= 20 (s = 5) or 18 (s = 3). */
int orc_aw1_ref(const uint8_t* L, const uint8_t* R, int W, int H, int s, int sim_type, int dmin, int dmax, int offset,
                int ltor, float min_delta, float sim_thresh, float grad_thresh, const uint16_t* init, int iW, int iH,
                uint16_t* out);

/* RectifyImage (RectifyImage.cu:64-158): target pixel -> source position through the
 * rectified camera (fx, fy, cx, cy), a rotation into the source camera and the source
 * camera with OpenCV distortion; bilinear / nearest / 0 sampling.  params: src fx, fy, cx,
 * cy, k1, k2, p1, p2, k3, rect fx, fy, cx, cy, R[9] (row-major, rectified to source). */
int orc_rectify(const uint8_t* src, int sW, int sH, const float* params, int W, int H, uint8_t* out);
/* ColorDisparityImage (ColorDisparityImage.cu:45-270).  fmt 0: 4 x uint16 ids, 1: uint16
 * ids, 2: float ids; style 0 Analysis / 2 Color: RGBA8 out, 1 Gray / 3 DispAbs / 4 DispRel:
 * uint16 out.  colors: 7 RGBA8 entries (Unknown, CannotEvaluate, Saturated, NotSpecific,
 * NotFound, NotUnique, Inconsistent). */
int orc_color_disparity(const void* disp, int fmt, int W, int H, int style, int dmin, int dmax, int inf_offset,
                        const uint8_t* colors, void* out);

#ifdef __cplusplus
}
#endif
#endif
