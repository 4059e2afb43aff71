// Launch interfaces of the HIP kernels (host side of each .hip translation unit).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace stereo {

// ---- census ----------------------------------------------------------------
bool census_supported(int cw, int ch, int ref3x3);
hipError_t launch_census(const View& src, const View& dst, int cw, int ch, int ref3x3, hipStream_t s);
// both images of a pair in one launch (src1/dst1 may be null for a single image)
hipError_t launch_census_pair(const View& src0, const View& dst0, const View* src1, const View* dst1, int cw, int ch,
                              hipStream_t s);

// ---- fused SAD / census matcher -------------------------------------------------
enum { BM_COST_SAD = 0, BM_COST_CENSUS = 1 };
constexpr int BM_NPART = 320;     // right-view partial keys per tile row
constexpr int BM_PAD_LEFT = 320;  // readable columns left of x = 0 in matcher inputs
constexpr int BM_PAD_RIGHT = 128; // readable columns right of x = W-1
constexpr int BM_MAX_SEG = 8;     // disparity segments per tile (one workgroup each)
constexpr int BM_MAX_WIDTH = 16384; // one image row of right-view keys fits in LDS

struct BmParams {
    View L, R;         // uint8 grey (SAD) or uint64 census (CENSUS)
    uint32_t* lkey;    // nseg*W*H best left keys per disparity segment (pitch W)
    uint32_t* rpart;   // nseg*ntx*H*BM_NPART right-view partial keys
    int W, H, D;
    int bx, by;        // matchable border
    int ntx, nseg, seg_len;
};

struct BmFinishParams {
    const uint32_t* lkey;   // nseg*W*H
    const uint32_t* rpart;  // nseg*ntx*H*BM_NPART
    View disp;     // int16 checked map (optional: ptr == nullptr)
    View disp_l;   // int16 left map (optional)
    View disp_r;   // int16 right map (optional)
    View dispid;   // 4 x uint16 reference pixels (optional)
    int W, H, D, ntx;
    int nseg, seg_len;
    int bx, by, tol;
    int value_base;  // != 0: the 4 x uint16 output carries value_base - cost (AW1 similarity)
};

bool bm_supported(int kind, int ww, int wh, int D);
void bm_grid(int W, int H, int D, int* ntx, int* nty, int* nseg, int* seg_len);
hipError_t launch_bm(int kind, int ww, const BmParams& p, hipStream_t s);
hipError_t launch_bm_finish(const BmFinishParams& f, hipStream_t s);

// ---- reference-format cross check ------------------------------------------------
// range[0..1] receive min / max of the kept ids (device memory, reset by the launch).
hipError_t launch_crosscheck_ref(const View& out, const View& lr, const View& rl, int thresh, int ltor, int offset,
                                 uint32_t* range, hipStream_t s);

// ---- NCC matcher ------------------------------------------------------------------
struct NccParams {
    View L, R;             // uint8 grey, padded views like the SAD matcher's inputs
    const int2* lstats;    // per-pixel (S, float bits of 2^23/sqrt(var) or 0)
    const int2* rstats;    // per-pixel (S, float bits of 1/sqrt(var) or 0)
    int stats_pitch;       // elements per stats row
    uint32_t* lkey;        // nseg*W*H
    uint32_t* rpart;       // nseg*ntx*H*BM_NPART
    int W, H, D;
    int bx, by;
    int ntx, nseg, seg_len;
};

struct NccScoreParams {
    View L, R;             // uint8 grey
    const uint32_t* lkey;  // nseg*W*H winner keys of the matcher
    View score;            // float winning score (NaN where no match)
    const int2* lstats;
    const int2* rstats;
    int stats_pitch;
    int W, H, ax, ay, bx, by, nseg;
};

bool ncc_supported(int ww, int wh, int D);
void ncc_grid(int W, int H, int D, int* ntx, int* nty, int* nseg, int* seg_len);
hipError_t launch_ncc_stats(const View& img, int2* out, int pitch_px, int ww, int wh, float scale, hipStream_t s);
hipError_t launch_ncc(int ww, const NccParams& p, hipStream_t s);
hipError_t launch_ncc_score(const NccScoreParams& f, hipStream_t s);

// ---- adaptive-window matchers -------------------------------------------------------
struct AwParams {
    View L, R;             // uint8 grey, padded views like the SAD matcher's inputs
    uint32_t* lkey;        // nseg*W*H
    uint32_t* rpart;       // nseg*ntx*H*BM_NPART
    int W, H, D;
    int bx, by;
    int ntx, nseg, seg_len;
};

bool aw_supported(int sub, int D);
void aw_grid(int W, int H, int D, int* ntx, int* nty, int* nseg, int* seg_len);
hipError_t launch_aw(int sub, const AwParams& p, hipStream_t s);

}  // namespace stereo
