// Launch interfaces of the HIP kernels (host side of each .hip translation unit).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace stereo {

// ---- census ----------------------------------------------------------------
bool census_supported(int cw, int ch, int ref3x3);
hipError_t launch_census(const View& src, const View& dst, int cw, int ch, int ref3x3, hipStream_t s);
// both images of a pair in one launch (src1/dst1 may be null for a single image);
// q0/q1, when given, receive the 4 x 4 box means (sum + 8) >> 4 of the images
// One image of a batch: source, padded destination, and its first workgroup in the
// batched launch (census tiles of 64 x 32 or rows, counted over the images before it).
struct ImageJob {
    View src, dst;
    int base;
};
int census_tiles(int w, int h);
// Census of every image of a batch in one launch; jobs in device memory, in order of base.
hipError_t launch_census_batch(const ImageJob* jobs, int njobs, int total_tiles, int cw, int ch, hipStream_t s);
// Copies every grey image of a batch into its padded input in one launch (one workgroup per row).
hipError_t launch_pad_copy_batch(const ImageJob* jobs, int njobs, int total_rows, hipStream_t s);
// Copies two grey images of one size into padded matcher inputs (one launch).
hipError_t launch_pad_copy_pair(const View& src0, const View& dst0, const View& src1, const View& dst1, hipStream_t s);
// The matchers' census of one or two images (and optionally their 4 x 4 box means), in the
// matchers' own bit order: only Hamming distances between such descriptors are meaningful.
hipError_t launch_census_pair(const View& src0, const View& dst0, const View* src1, const View* dst1, int cw, int ch,
                              hipStream_t s, const View* q0 = nullptr, const View* q1 = nullptr);

// ---- fused SAD / census matcher -------------------------------------------------
enum { BM_COST_SAD = 0, BM_COST_CENSUS = 1 };
constexpr int BM_NPART = 320;     // right-view partial keys per tile row
constexpr int BM_PAD_LEFT = 320;  // readable columns left of x = 0 in matcher inputs
constexpr int BM_PAD_RIGHT = 128; // readable columns right of x = W-1
constexpr int BM_MAX_SEG = 8;     // disparity segments per tile (one workgroup each)
constexpr int BM_MAX_WIDTH = 16384; // one image row of right-view keys fits in LDS

struct BmParams {
    View L, R;         // uint8 grey (SAD) or uint64 census (CENSUS)
    uint32_t* lkey;    // W*H best left keys (pitch W), all zero at the launch
    uint32_t* rkey;    // W*H best right keys (pitch W), all zero at the launch
    int W, H, D;
    int bx, by;        // matchable border
    int ntx, nseg, seg_len;
};

// The learned proposal's network input, written by the merge of the coarse census
// matcher row by row right after the row's checked map (cnn_input_row); x == nullptr:
// not written.
constexpr int CNN_MAX_W4 = 4096;  // widest quarter-resolution row of the network input

struct CnnInputTarget {
    View L4, R4;    // the quarter-resolution pair
    uint16_t* x;    // bf16 network input of frame n
    float* dfill;   // the hole-filled coarse map (float)
    int n;
};

struct BmFinishParams {
    // the NCC and adaptive-window matchers' keys: per disparity segment
    const uint32_t* lkey;   // nseg*W*H
    const uint32_t* rpart;  // nseg*ntx*H*BM_NPART
    // the SAD / census matcher's key planes (lplane != nullptr): read, then cleared
    uint32_t* lplane = nullptr;  // W*H
    uint32_t* rplane = nullptr;  // W*H
    View disp;     // int16 checked map (optional: ptr == nullptr)
    View disp_l;   // int16 left map (optional)
    View disp_r;   // int16 right map (optional)
    View dispid;   // 4 x uint16 reference pixels (optional)
    int W, H, D, ntx;
    int nseg, seg_len;
    int bx, by, tol;
    int value_base;  // != 0: the 4 x uint16 output carries value_base - cost (AW1 similarity)
    CnnInputTarget cnn{};  // with the key planes: the network input from the checked map `disp`
};

bool bm_supported(int kind, int ww, int wh, int D);
void bm_grid(int W, int H, int D, int* ntx, int* nty, int* nseg, int* seg_len);
hipError_t launch_bm(int kind, int ww, const BmParams& p, hipStream_t s);
hipError_t launch_bm_finish(const BmFinishParams& f, hipStream_t s);
// Clears n words of key planes (on a stream; safe under graph capture).
hipError_t launch_bm_clear(uint32_t* p, size_t n, hipStream_t s);

// Device address of the no-allocation entry point's scratch array on the current device,
// looked up once per device (csrc/noalloc.hip); called by stereo_create.
hipError_t noalloc_setup(int device);

// ---- work-list matcher: many frames, per-tile disparity ranges ---------------------
// A frame is one stereo pair; its tiles are 64 x BM_TILE_H pixels; every tile has a
// disparity range [lo, hi) cut into items (one workgroup's work each).  Every item folds
// its left winners and right-view partial maxima into its frame's two key planes (W x H
// each, at plane_off in lplane and rplane) with atomicMax; the merge launch reads the
// planes and clears them, so they are zero at every matcher launch.  The matcher launch
// is persistent: a fixed grid walks the item list, whose length may be produced on the
// device.
constexpr int BM_TILE_W = 64;
constexpr int BM_TILE_H = 56;

struct BmFrame {
    View L, R;          // padded inputs (uint8 grey or uint64 census)
    View disp, disp_l, disp_r, dispid;  // outputs (ptr == nullptr: not written)
    int W, H, D;
    int bx, by, tol;
    int value_base;
    int ntx, nty;
    int tile_base;      // index of tile (0, 0) in the tile table
    int row_base;       // first finish row of this frame
    size_t plane_off;   // element offset of the frame's key planes (row stride W)
};

struct BmTile {
    int lo, hi;         // disparity range searched for the tile's left pixels
    int first_item, nitems;
};

struct BmItem {
    int frame, tx, ty;
    int ds, de;
};

struct BmListParams {
    const BmFrame* frames;
    const BmTile* tiles;
    const BmItem* items;
    const int* nitems;  // device-side item count
    const int* sched;   // order in which the persistent grid takes the items (nullptr: 0, 1, ...)
    int* counter;       // items handed out after each workgroup's first (item grid + counter):
                        // 0 before the matcher launch; the finish launch resets it, so a
                        // matcher launch is always followed by its finish
    uint32_t* lplane;   // every frame's left-key plane, zero at the matcher launch
    uint32_t* rplane;   // every frame's right-key plane, zero at the matcher launch
    int nframes;
    int total_rows;     // sum of the frames' heights
};

hipError_t launch_bm_list(int kind, int ww, const BmListParams& p, int grid, hipStream_t s);
hipError_t launch_bm_list_finish(const BmListParams& p, int max_width, hipStream_t s);
// widest frame the work-list merge takes (one LDS row of right keys)
constexpr int BM_LIST_MAX_WIDTH = BM_MAX_WIDTH;

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
    // band search: searched 8-disparity bins per 64 x BM_TILE_H tile ([bins_nty][ntx]);
    // nullptr for the full range
    const uint32_t* bins;
    int bins_nty;
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
// Both images of a pair (one size) in one launch.
hipError_t launch_ncc_stats_pair(const View& img0, int2* out0, float scale0, const View& img1, int2* out1, float scale1,
                                 int pitch_px, int ww, int wh, hipStream_t s);
hipError_t launch_ncc(int ww, const NccParams& p, hipStream_t s);
hipError_t launch_ncc_score(const NccScoreParams& f, hipStream_t s);

// ---- adaptive-window matchers -------------------------------------------------------
struct AwParams {
    View L, R;             // uint8 grey, padded views like the SAD matcher's inputs
    uint32_t* lkey;        // W*H left-key plane, zero at the launch (every segment folds in)
    uint32_t* rkey;        // W*H right-key plane, zero at the launch
    int W, H, D;
    int bx, by;
    int ntx, nseg, seg_len;
    const uint32_t* bins;  // band search: per 64 x BM_TILE_H tile, the 8-disparity bins to search (or null)
    int bins_nty;          // tile rows of bins
};

// ---- learned proposal network -------------------------------------------------------
enum { CNN_EPI_RELU_BF16 = 0, CNN_EPI_F32 = 1, CNN_EPI_MASK_BF16 = 2 };
// 3x3 convolution tiles: CNN_TILE_ROWS rows x 64 pixels per workgroup; the staged
// input tile (CNN_TILE_ROWS + 2 dil) x (64 + 2 dil) x (2 cin + 8) bytes stays below
// 64 KiB of LDS for dil <= CNN_MAX_DIL
constexpr int CNN_TILE_ROWS = 4;
constexpr int CNN_MAX_DIL = 4;

struct CnnConvArgs {
    const uint16_t* x;     // [N][H][W][cin] bf16
    const uint16_t* w;     // [cout][9][cin] bf16
    const float* bias;     // [cout] or nullptr
    const uint16_t* mask;  // [N][H][W][cout] bf16 (CNN_EPI_MASK_BF16)
    void* y;               // [N][H][W][cout] bf16 or float
    int N, H, W, dil, sign;  // sign -1: tap offsets negated (data gradient)
};

struct CnnDwArgs {
    const uint16_t* x;     // [N][H][W][cin] layer input
    const uint16_t* dy;    // [N][H][W][cout] gradient of the layer output
    float* part;           // cnn_dw_parts() x (cout*9*cin + cout) partial sums
    int N, H, W, dil;
};

struct CnnBandArgs {
    const float* out;      // network output [H4][W4][16]
    const float* dfill;    // filled coarse disparity [H4][W4]
    BmTile* tiles;         // [nty][ntx]
    BmItem* items;
    int* nitems;
    int* list_counter;     // item counter of the list matcher, set to 0 with the item count
    int* tile_evals;       // [nty][ntx] disparities searched per tile | fallback << 16
    uint32_t* tile_mask;   // [nty][ntx] searched 8-disparity bins
    int W4, H4, ntx, nty, D;
    int band, chunk;
    int centre_residual;   // 1: centre = dfill + network residual, 0: dfill
    float logit_thresh;
};

hipError_t launch_cnn_conv(int cin, int cout, int epi, const CnnConvArgs& a, hipStream_t s);
// Inference of the whole 16-channel network (dilations 1, 2, 4, 2, 1) in one launch:
// x [N][H][W][16] bf16 in, out [N][H][W][16] FP32 (bias added, no ReLU on the last layer).
struct CnnFusedArgs {
    const uint16_t* x;
    const uint16_t* w[5];  // bf16 [16][9][16] per layer
    const float* bias[5];  // [16] per layer
    float* out;
    int N, H, W;
};
hipError_t launch_cnn_fused_forward(const CnnFusedArgs& a, hipStream_t s);
int cnn_dw_parts(const CnnDwArgs& a);
hipError_t launch_cnn_dw(int cin, int cout, const CnnDwArgs& a, float* out, hipStream_t s);
// clabel: per-pixel confidence target (0 or 1, NaN: from |e| <= tau where label is set)
hipError_t launch_cnn_loss(const float* out, const float* label, const float* clabel, const float* dfill, uint16_t* g,
                           float* part, int P, float scale, float lambda, float tau, hipStream_t s);
hipError_t launch_cnn_labels(const View& disp, int n, int W4, int H4, float* labels, hipStream_t s);
hipError_t launch_cnn_conf_labels(const View& conf, int n, float* out, hipStream_t s);
hipError_t launch_cnn_adam(float* param, const float* grad, float* m, float* v, int n, int cin, int cout, float lr,
                           float b1, float b2, float eps, float c1, float c2, uint16_t* wf, uint16_t* wb, hipStream_t s);
hipError_t launch_cnn_input(const View& L4, const View& R4, const View& dc, int n, uint16_t* x, float* dfill,
                            hipStream_t s);
hipError_t launch_band_items(const CnnBandArgs& a, hipStream_t s);
// Band search around a caller's initial disparity map (the role of MatcherUseDisparity):
// per 64 x BM_TILE_H tile, the 8-disparity bins that cover [min - 4, max + 4] of
// scale x d over the tile's pixels, d read at (x / scale, y / scale); all bins when the
// tile has no valid value.  Then the items as launch_band_items.
struct InitBandArgs {
    View init;             // LUM_I16 (d, -1 invalid) or RGBA_U16 (EDisparityId in channel 0)
    int ref_format;        // 1: RGBA_U16
    int scale;             // W / init width
    int W, H;
};
hipError_t launch_init_band(const InitBandArgs& ib, const CnnBandArgs& a, hipStream_t s);
// only the per-tile bin masks of the band search (a.tile_mask, a.tile_evals)
hipError_t launch_init_band_mask(const InitBandArgs& ib, const CnnBandArgs& a, hipStream_t s);

bool aw_supported(int sub, int D);
void aw_grid(int W, int H, int D, int* ntx, int* nty, int* nseg, int* seg_len);
hipError_t launch_aw(int sub, const AwParams& p, hipStream_t s);

}  // namespace stereo
