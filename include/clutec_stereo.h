/*
 * clutec_stereo.h -- C ABI of the MI355X (gfx950) stereo disparity engine.
 *
 * Plain C, plain pointers and sizes.  Every entry point names the reference
 * interface it replaces (paths relative to the module.CluTec.Cuda tree).  The C++
 * driver layer in include/clutec/ wraps these calls in the reference's class
 * shapes (CKernelDriverBase, CDriver::Configure / Process, ProcessKernel).
 *
 * Threading: a context belongs to one device; drivers keep their parameters in
 * host memory and pass them as kernel arguments, so two drivers may run
 * concurrently on different streams.  One driver must not be used from two
 * streams at once (its workspace is shared between its calls).
 * Process calls are asynchronous on the given stream; none of them allocates.
 */
#ifndef CLUTEC_STEREO_H
#define CLUTEC_STEREO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STEREO_ABI_VERSION 1

/* Status codes.  The reference throws CIException; each code names the throw
 * site whose condition it reports. */
typedef enum stereo_status {
    STEREO_OK = 0,
    STEREO_ERROR_INVALID_ARGUMENT = 1,     /* null pointers, sizes out of range */
    STEREO_ERROR_INVALID_CONFIG = 2,       /* "Invalid algorithm configuration." BlockMatcherNCC.cu:1314-1316 */
    STEREO_ERROR_UNSUPPORTED_PIXEL_TYPE = 3, /* "Pixel types of given images not supported" BlockMatcherNCC.cu:1418, CensusTransform.cu:152-155 */
    STEREO_ERROR_FORMAT_MISMATCH = 4,      /* CrossCheckDisparityMaps.cu:139-147, CensusTransform.cu:140-143 */
    STEREO_ERROR_INVALID_DEVICE = 5,       /* CDevice::Set, Device.cpp:66-74 */
    STEREO_ERROR_INVALID_FORMAT = 6,       /* "Invalid image format" KernelDriverBase.cpp:52-55 */
    STEREO_ERROR_HIP = 7,                  /* runtime error, EnsureOK Api.cpp:86-94 */
    STEREO_ERROR_OUT_OF_MEMORY = 8,
    STEREO_ERROR_NOT_READY = 9             /* result requested before the work finished */
} stereo_status;

/* Pixel types (the subset of Clu::EPixelType x EDataType this path uses). */
typedef enum stereo_pixel_type {
    STEREO_PIXEL_LUM_U8 = 1,    /* TPixel_Lum_UInt8: grey images, reference census */
    STEREO_PIXEL_LUM_U16 = 2,
    STEREO_PIXEL_LUM_I16 = 3,   /* plain disparity map, -1 = invalid */
    STEREO_PIXEL_LUM_U32 = 4,
    STEREO_PIXEL_LUM_U64 = 5,   /* wide census descriptors */
    STEREO_PIXEL_LUM_F32 = 6,   /* NCC score maps */
    STEREO_PIXEL_RGBA_U16 = 7,  /* _SDisparityConfig::TPixelEx: 4 x uint16, EDisparityId in channel 0 */
    STEREO_PIXEL_RGBA_U8 = 8    /* TPixel_RGBA_UInt8: colour images (r, g, b, a bytes) */
} stereo_pixel_type;

/* Image format (Clu::SImageFormat: width, height, pixel type). */
typedef struct stereo_image_format {
    int32_t width;
    int32_t height;
    int32_t pixel_type;
} stereo_image_format;

/* Non-owning device image view (the role of _CDeviceImage, DeviceImage.h:208-229):
 * base pointer, row pitch in bytes, format.  Any device allocation can be
 * wrapped, e.g. a torch tensor's data_ptr() and stride. */
typedef struct stereo_image {
    void* data;
    size_t pitch;
    stereo_image_format format;
} stereo_image;

typedef struct stereo_context stereo_context;
typedef struct stereo_driver stereo_driver;
typedef void* stereo_stream; /* hipStream_t; NULL = default stream */

/* ---- library and device ------------------------------------------------ */

int stereo_abi_version(void);

/* Replaces CDevice::Set / MakeCurrent (Device.cpp:61-107). */
stereo_status stereo_create(stereo_context** ctx, int device);
void stereo_destroy(stereo_context* ctx);
/* Message of the last failing call on this context (or of the failing
 * stereo_create when ctx is NULL). */
const char* stereo_last_error(const stereo_context* ctx);
const char* stereo_status_string(stereo_status s);
/* Enables hipEvent timing of process calls (KernelDriverBase.h:139-148). */
stereo_status stereo_set_timing(stereo_context* ctx, int enable);
stereo_status stereo_synchronize(stereo_context* ctx, stereo_stream stream);

/* ---- images (CDeviceImage, DeviceImage.h / DeviceImage.cpp:82-99) --------- */

size_t stereo_pixel_size(int32_t pixel_type);
stereo_status stereo_image_alloc(stereo_context* ctx, stereo_image* img, int32_t width, int32_t height,
                                 int32_t pixel_type);
stereo_status stereo_image_free(stereo_context* ctx, stereo_image* img);
/* CDeviceImage::CopyFrom: host rows (host_pitch bytes apart) -> device.  Synchronous;
 * the rows pass through a pinned staging buffer of the context (8 MiB, allocated on the
 * first transfer), so transfers on one context are serialised. */
stereo_status stereo_image_upload(stereo_context* ctx, stereo_image* img, const void* host, size_t host_pitch,
                                  stereo_stream stream);
/* CDeviceImage::CopyInto: device -> host rows, as stereo_image_upload. */
stereo_status stereo_image_download(stereo_context* ctx, const stereo_image* img, void* host, size_t host_pitch,
                                    stereo_stream stream);

/* Clu::Cuda::Malloc / Free / MemCpy (Api.h:107-229): raw device buffers, e.g. for
 * the network's kernel-level entry points.  kind: 0 host->device, 1 device->host,
 * 2 device->device.  The copy is synchronous. */
stereo_status stereo_device_alloc(stereo_context* ctx, void** ptr, size_t bytes);
stereo_status stereo_device_free(stereo_context* ctx, void* ptr);
stereo_status stereo_memcpy(stereo_context* ctx, void* dst, const void* src, size_t bytes, int32_t kind);

/* ---- drivers (CKernelDriverBase, KernelDriverBase.h:47-177) ---------------- */

stereo_status stereo_driver_destroy(stereo_driver* drv);
const char* stereo_driver_name(const stereo_driver* drv);
/* LastProcessTime (KernelDriverBase.h:168-171): ms of the last process call;
 * waits for it.  Requires stereo_set_timing(ctx, 1) before the call. */
stereo_status stereo_driver_last_time_ms(stereo_driver* drv, double* ms);

/* ---- CensusTransform (CensusTransform.h:46-66) ---------------------------- */

typedef struct stereo_census_params {
    int32_t window_w;     /* odd, window_w * window_h <= 64 */
    int32_t window_h;
    int32_t reference_3x3; /* 1: the reference 3x3 transform into LUM_U8 (CensusTransform.cu:59-120) */
} stereo_census_params;

/* CensusTransform::CDriver::Configure (CensusTransform.cu:126-135). */
stereo_status stereo_census_configure(stereo_context* ctx, stereo_driver** drv, const stereo_image_format* fmt,
                                      const stereo_census_params* params);
/* CensusTransform::CDriver::Process (CensusTransform.cu:138-157): trg LUM_U64
 * (LUM_U8 for the 3x3 reference transform), src LUM_U8, same size. */
stereo_status stereo_census_process(stereo_driver* drv, stereo_image* trg, const stereo_image* src,
                                    stereo_stream stream);

/* ---- block matchers ------------------------------------------------------ */

typedef enum stereo_algorithm {
    STEREO_ALGO_SAD = 0,    /* BlockMatcherSAD1: integer SAD */
    STEREO_ALGO_CENSUS = 1, /* census + Hamming cost (CensusDifference, BlockMatcherNCC.cu:153-158) */
    STEREO_ALGO_NCC = 2,    /* BlockMatcherNCC */
    STEREO_ALGO_AW1 = 3,    /* BlockMatcherAW1: adaptive window, similarity 255-|a-b| */
    STEREO_ALGO_AW2 = 4     /* BlockMatcherAW2: adaptive window, SAD */
} stereo_algorithm;

typedef struct stereo_matcher_params {
    int32_t algorithm;       /* stereo_algorithm */
    int32_t window_w;        /* aggregation window (odd); AW1/AW2: sub-patch size (3 or 5) */
    int32_t window_h;
    int32_t census_w;        /* STEREO_ALGO_CENSUS: census window (odd, w*h <= 64); else ignored */
    int32_t census_h;
    int32_t num_disparities; /* D: candidates 0..D-1, multiple of 8, <= 256 */
    int32_t lr_tolerance;    /* |dR(x-dL) - dL| <= tol keeps the pixel */
    int32_t use_disp_input;  /* 1: also configure stereo_matcher_process_init (SParameter::iUseDispInput,
                                BlockMatcherNCC.cu:1330); every algorithm */
    int32_t reserved[4];     /* must be 0 */
} stereo_matcher_params;

/* Checks a configuration without touching a device.  Returns STEREO_OK or
 * STEREO_ERROR_INVALID_CONFIG / INVALID_FORMAT with a message in msg. */
stereo_status stereo_matcher_validate(const stereo_image_format* fmt, const stereo_matcher_params* params,
                                      char* msg, size_t msg_len);

/* Configure of BlockMatcherSAD1 (BlockMatcherSAD1.cu:492-515), BlockMatcherNCC
 * (BlockMatcherNCC.cu:1295-1318), BlockMatcherAW1 (BlockMatcherAW1.cu:1303-1397),
 * BlockMatcherAW2 (BlockMatcherAW2.cu), selected by params->algorithm.  Allocates
 * the driver's workspace. */
stereo_status stereo_matcher_configure(stereo_context* ctx, stereo_driver** drv, const stereo_image_format* fmt,
                                       const stereo_matcher_params* params);

/* Process of the matchers (BlockMatcherNCC.h:137-143, BlockMatcherAW1.h:151-157,
 * BlockMatcherSAD1.h:121-125) fused with the left-right pass and the cross check
 * (CrossCheckDisparityMaps::CDriver::Process, CrossCheckDisparityMaps.cu:135-192).
 *   left, right : LUM_U8 rectified grey images of the configured size.
 *   disp        : checked left map, LUM_I16 (-1 invalid) or RGBA_U16 (EDisparityId
 *                 codes: First+d, Inconsistent, NotFound, Unknown at the border).
 *   disp_lr     : optional (NULL) LUM_I16 left-to-right map before the check.
 *   disp_rl     : optional (NULL) LUM_I16 right-to-left map.
 * Any of the three outputs may be NULL. */
stereo_status stereo_matcher_process(stereo_driver* drv, stereo_image* disp, stereo_image* disp_lr,
                                     stereo_image* disp_rl, const stereo_image* left, const stereo_image* right,
                                     stereo_stream stream);

/* MatcherUseDisparity (BlockMatcherNCC.cu:894-1255), the Process overload with an initial
 * disparity map (BlockMatcherNCC.h:137-143): disp_init is LUM_I16 (d, -1 invalid) or
 * RGBA_U16 (EDisparityId in channel 0) of any size up to the images'; scale = width /
 * its width.  Each 64x56 tile searches the 8-disparity bins covering
 * [min - 4, max + 4] of scale x d over its pixels, d read at (x / scale, y / scale);
 * a tile with no valid value searches the full range.  Outputs as
 * stereo_matcher_process.  Requires use_disp_input at configure. */
stereo_status stereo_matcher_process_init(stereo_driver* drv, stereo_image* disp, stereo_image* disp_lr,
                                          stereo_image* disp_rl, const stereo_image* left, const stereo_image* right,
                                          const stereo_image* disp_init, stereo_stream stream);

/* Winning NCC score of every left pixel for the last process call on the same
 * images (FP32, NaN where the left map has no match).  Scores carry the exact
 * integer covariance through one float conversion and two multiplies: relative
 * error below 1e-6 against FP64.  The score the reference stores quantised to
 * uint8 in channel 1 of its disparity pixel (BlockMatcherNCC.cu:321-323, :563). */
stereo_status stereo_matcher_scores(stereo_driver* drv, stereo_image* score, const stereo_image* left,
                                    const stereo_image* right, stereo_stream stream);

/* Per-stage device time (ms) of the last timed matcher process call: stage 0 =
 * input preparation (census of both images, or the padded copy for SAD), stage 1
 * = the fused matching kernel, stage 2 = right-view merge and check.  Extends
 * LastProcessTime (KernelDriverBase.h:168-171) to the stages of the fused path.
 * Waits for the call; requires stereo_set_timing(ctx, 1). */
stereo_status stereo_matcher_stage_times(stereo_driver* drv, double* ms, int32_t n);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * stereo_matcher_process for the SAD and census matchers with no configured driver and no
 * allocation.  params as for stereo_matcher_configure (algorithm SAD or CENSUS,
 * use_disp_input 0; the frame size is the left image's).  The padded inputs and the two key
 * planes live in a library-owned device array of STEREO_NOALLOC_SCRATCH_BYTES per GPU, set
 * up by stereo_create and held until the process ends; frames whose workspace exceeds it
 * (2560x1440 census, for instance; 1920x1080 census and 3200x1800 SAD fit, at any D) are
 * STEREO_ERROR_INVALID_FORMAT.  Outputs equal
 * stereo_matcher_process's.  Every call on one GPU shares the array: calls that overlap on
 * one GPU (two streams, two threads) race; calls on one stream are ordered. */
#define STEREO_NOALLOC_SCRATCH_BYTES (64u << 20)
stereo_status stereo_matcher_process_noalloc(stereo_context* ctx, const stereo_matcher_params* params,
                                             stereo_image* disp, stereo_image* disp_lr, stereo_image* disp_rl,
                                             const stereo_image* left, const stereo_image* right,
                                             stereo_stream stream);

/* ---- CrossCheckDisparityMaps (CrossCheckDisparityMaps.h:60-111) ------------ */

typedef struct stereo_crosscheck_params {
    int32_t disp_delta_thresh; /* iDispDeltaThresh */
    int32_t left_to_right;     /* iIsLeftToRight */
    int32_t disp_offset;       /* _SDisparityConfig offset added by MapPixelPos */
    int32_t reserved;
} stereo_crosscheck_params;

/* CrossCheckDisparityMaps::CDriver::Configure (CrossCheckDisparityMaps.cu:115-132). */
stereo_status stereo_crosscheck_configure(stereo_context* ctx, stereo_driver** drv, const stereo_image_format* fmt,
                                          const stereo_crosscheck_params* params);
/* CrossCheckDisparityMaps::CDriver::Process (CrossCheckDisparityMaps.cu:135-192):
 * RGBA_U16 maps in and out.  Asynchronous; the range is available through
 * stereo_crosscheck_range after the stream has reached this call. */
stereo_status stereo_crosscheck_process(stereo_driver* drv, stereo_image* result, const stereo_image* disp_lr,
                                        const stereo_image* disp_rl, stereo_stream stream);
/* DispMin / DispMax (CrossCheckDisparityMaps.h:96-104); waits for the last process. */
stereo_status stereo_crosscheck_range(stereo_driver* drv, uint16_t* disp_min, uint16_t* disp_max);

/* ---- batched matcher: many pairs of uneven size in one persistent launch ---
 *
 * The reference processes one pair per Process call (BlockMatcherSAD1.h:121-125,
 * one CDriver per image format, KernelDriverBase.h:47-177).  A batch driver is
 * configured once for a list of pair formats and disparity counts; each process
 * call runs the census / padding of every pair, one persistent matching launch
 * over the work items of all pairs, and one merge-and-check launch, and on the
 * second call with the same images replays them as a captured HIP graph.  Every
 * pair's maps equal what stereo_matcher_process gives for that pair alone. */

typedef struct stereo_batch_tuning {
    int32_t chunk;       /* disparities per work item (multiple of 8); 0 = autotuned or default */
    int32_t grid;        /* workgroups of the persistent launch; 0 = 2 per CU */
    int32_t order;       /* 0: items longest first, 1: raster order */
    int32_t use_graph;   /* 1: capture the launch sequence and replay it */
} stereo_batch_tuning;

/* params: algorithm SAD or CENSUS, windows and tolerance as for the matcher;
 * params->num_disparities is ignored, num_disparities[i] gives pair i's D.  Widths up
 * to 16384 (STEREO_ERROR_INVALID_FORMAT above). */
stereo_status stereo_batch_configure(stereo_context* ctx, stereo_driver** drv, int32_t npairs,
                                     const stereo_image_format* fmts, const int32_t* num_disparities,
                                     const stereo_matcher_params* params, const stereo_batch_tuning* tuning);
/* disp[i], left[i], right[i] for i < npairs (disp LUM_I16); disp_lr / disp_rl
 * may be NULL or arrays of npairs (entries with data == NULL are skipped). */
stereo_status stereo_batch_process(stereo_driver* drv, stereo_image* disp, stereo_image* disp_lr,
                                   stereo_image* disp_rl, const stereo_image* left, const stereo_image* right,
                                   stereo_stream stream);
/* Times the candidate tunings on the given images and keeps the fastest: one work-item
 * length (32, 64, 128, 256) for every pair with the grid (1, 2, 3 workgroups per CU) and the
 * item order, then the item length of each class of pairs (disparity count x tile count,
 * as in stereo_batch_tuning_table), largest classes first.  The tile is 64 x 56 for every
 * class.  Returns the overall tuning in *best (may be NULL).  Blocks. */
stereo_status stereo_batch_autotune(stereo_driver* drv, const stereo_image* left, const stereo_image* right,
                                    stereo_batch_tuning* best);
stereo_status stereo_batch_get_tuning(const stereo_driver* drv, stereo_batch_tuning* tuning);
/* The tuning as a JSON table for a later process: {"matcher", "window", "grid", "order",
 * "use_graph", "chunk", "classes": [{"disparities", "tiles", "tile", "chunk"}, ...]}, one class
 * per disparity count and tile count ("small" below 64 tiles of 64 x 56, "medium" below 256,
 * "large") among the driver's pairs, each with its work-item length.  buf may be NULL;
 * *needed (may be NULL) receives the text's length + 1. */
stereo_status stereo_batch_tuning_table(const stereo_driver* drv, char* buf, size_t n, size_t* needed);
/* Applies a table written by stereo_batch_tuning_table: grid, order, use_graph, the default
 * work-item length, and the length of every listed class the driver's pairs have (the others
 * take the default).  Malformed tables and lengths outside 32..256 are
 * STEREO_ERROR_INVALID_CONFIG, and the previous tuning stays. */
stereo_status stereo_batch_apply_tuning_table(stereo_driver* drv, const char* json);

/* ---- learned disparity proposal ----------------------------------------------
 *
 * No counterpart in the reference; it takes the place of the initial disparity map
 * of MatcherUseDisparity (BlockMatcherNCC.cu:894-1255, BlockMatcherAW1.cu:923-1263),
 * which searches a band around a given disparity.  A 5-layer 3x3 convolutional
 * network (BF16 MFMA, FP32 accumulation) sees the pair at 1/4 resolution -- the left
 * image and the right image warped by a census match at that resolution -- and
 * proposes a centre residual and a confidence per quarter pixel.  Parameters: 7554
 * FP32 values, layer by layer: weights [cout][3][3][cin] then bias [cout], for
 * (cin, cout, dilation) = (2,16,1) (16,16,2) (16,16,4) (16,16,2) (16,2,1). */

#define STEREO_CNN_PARAMS 7554

typedef struct stereo_cnn stereo_cnn;

/* A network with buffers for up to max_frames inputs of W4 x H4. */
stereo_status stereo_cnn_create(stereo_context* ctx, stereo_cnn** cnn, int32_t W4, int32_t H4, int32_t max_frames);
void stereo_cnn_destroy(stereo_cnn* cnn);
stereo_status stereo_cnn_set_params(stereo_cnn* cnn, const float* host, size_t n); /* resets Adam */
stereo_status stereo_cnn_get_params(stereo_cnn* cnn, float* host, size_t n);
/* Gradients of the loss of the last train step, same layout as the parameters. */
stereo_status stereo_cnn_get_grads(stereo_cnn* cnn, float* host, size_t n);
/* Training state for checkpoint / resume: a versioned blob of the FP32 parameters,
 * the Adam moments and the step count (layout: 8-byte magic "CLSTCNN1", uint32
 * version 1, uint32 step, then 3 x STEREO_CNN_PARAMS floats: parameters, first and
 * second moments, each in the public parameter layout).  Loading sets the bf16
 * weight copies used by inference. */
size_t stereo_cnn_state_size(void);
stereo_status stereo_cnn_save_state(stereo_cnn* cnn, void* host, size_t n);
stereo_status stereo_cnn_load_state(stereo_cnn* cnn, const void* host, size_t n);
/* Input slot `index`: quarter-resolution L4, R4 (LUM_U8, W4 x H4) and the coarse
 * disparity (LUM_I16, -1 invalid) that warps R4. */
stereo_status stereo_cnn_set_input(stereo_cnn* cnn, int32_t index, const stereo_image* L4, const stereo_image* R4,
                                   const stereo_image* coarse, stereo_stream stream);
/* Labels of slot `index` from a full-resolution checked map (LUM_I16, 4 W4 x 4 H4). */
stereo_status stereo_cnn_set_labels(stereo_cnn* cnn, int32_t index, const stereo_image* disp, stereo_stream stream);
/* Confidence targets of slot `index` (LUM_U8, W4 x H4): 1 where the quarter pixel's
 * band search agrees with the full range, 0 where it does not, other values leave
 * the pixel without a target.  Without them the target is |residual error| <= 2. */
stereo_status stereo_cnn_set_conf_labels(stereo_cnn* cnn, int32_t index, const stereo_image* conf,
                                         stereo_stream stream);
/* Forward over slots [0, n): out (device, may be NULL) receives n x H4 x W4 x 2 floats
 * (centre residual in quarter pixels, confidence logit). */
stereo_status stereo_cnn_forward(stereo_cnn* cnn, int32_t n, float* out, stereo_stream stream);
/* One Adam step (lr, beta1 0.9, beta2 0.999, eps 1e-8) on slots [0, n).  loss
 * (host, may be NULL) receives {mean Huber, mean BCE, labelled pixels, fraction of
 * labelled pixels within tau}; waits for the step when loss is given. */
stereo_status stereo_cnn_train_step(stereo_cnn* cnn, int32_t n, float lr, float* loss, stereo_stream stream);

/* Kernel-level entry points of the network's convolutions (device pointers, NHWC
 * bf16 activations with cin, cout in {16, 32}, 16-byte aligned; dilation 1 to 4, larger ones are
 * STEREO_ERROR_INVALID_CONFIG).  y = conv(x, w) + bias with w [cout][9][cin];
 * relu != 0 writes bf16 max(y, 0), else FP32 y. */
stereo_status stereo_conv3x3_bf16(stereo_context* ctx, const uint16_t* x, int32_t N, int32_t H, int32_t W,
                                  int32_t cin, const uint16_t* w, const float* bias, int32_t cout, int32_t dil,
                                  int32_t relu, void* y, stereo_stream stream);
/* Gradients for upstream dy [N][H][W][cout] (bf16): dx [N][H][W][cin] (FP32, may be
 * NULL) from wt = w transposed to [cin][9][cout]; dwdb = cout*9*cin weight
 * gradients then cout bias gradients (FP32 device buffer). */
stereo_status stereo_conv3x3_bwd_bf16(stereo_context* ctx, const uint16_t* x, int32_t N, int32_t H, int32_t W,
                                      int32_t cin, const uint16_t* wt, const uint16_t* dy, int32_t cout, int32_t dil,
                                      float* dx, float* dwdb, stereo_stream stream);

typedef struct stereo_learned_params {
    stereo_matcher_params matcher; /* full-resolution census matcher (algorithm CENSUS) */
    int32_t coarse_disparities;    /* D of the quarter-resolution matcher (default D/4) */
    int32_t band;                  /* the searched band reaches band/2 + 4 beyond each centre (8) */
    int32_t chunk;                 /* disparities per work item (64) */
    float logit_thresh;            /* tiles with a smaller confidence logit search the full range */
    int32_t centre_residual;       /* 0: centres from the quarter-resolution census map (default);
                                      1: plus the network's centre residual */
    int32_t graph_replay;          /* 1: capture the stages once per pair of input images and replay
                                      them as a HIP graph; 0 (default): launch them directly, which
                                      measured 3-4 % faster per frame on gfx950 */
    int32_t reserved[2];           /* must be 0 */
} stereo_learned_params;

/* The learned-proposal census matcher for LUM_U8 pairs of fmt (width and height
 * multiples of 4, width up to 16384); cnn_params = STEREO_CNN_PARAMS trained values. */
stereo_status stereo_learned_configure(stereo_context* ctx, stereo_driver** drv, const stereo_image_format* fmt,
                                       const stereo_learned_params* params, const float* cnn_params, size_t n);
/* Outputs as stereo_matcher_process. */
stereo_status stereo_learned_process(stereo_driver* drv, stereo_image* disp, stereo_image* disp_lr,
                                     stereo_image* disp_rl, const stereo_image* left, const stereo_image* right,
                                     stereo_stream stream);
/* Of the last process call (waits for it): disparity evaluations searched at full
 * resolution, tiles, tiles that fell back to the full range, and with timing on
 * the stage times in ms {census, coarse match, network, band match, merge/check}.
 * Untimed process calls replay a captured HIP graph of the whole sequence (one
 * per pair of input buffers); timed calls run the stages one by one. */
stereo_status stereo_learned_stats(stereo_driver* drv, int64_t* evals, int32_t* tiles, int32_t* fallback_tiles,
                                   double* stage_ms, int32_t n_stages);
/* Disparity range [lo, hi) searched by each 64 x 56 tile in the last process call,
 * row-major tiles, lohi[2 t], lohi[2 t + 1]; n >= 2 * tiles.  Waits for the call. */
stereo_status stereo_learned_tile_ranges(stereo_driver* drv, int32_t* lohi, int32_t n);
/* The searched disparities of each tile in the last process call as a mask of
 * 8-disparity bins (bit b: d in [8b, 8b + 8)); n >= tiles.  Waits for the call. */
stereo_status stereo_learned_tile_masks(stereo_driver* drv, uint32_t* masks, int32_t n);

/* ---- disparity mip maps (CreateDispMipMaps, CreateDispMipMaps.h:40-70) ------- */

/* CreateDispMipMaps::CDriver::Run (CreateDispMipMaps.cu:250-330): levels[0..nlevels) are
 * LUM_F32 images, levels[0] the size of disp, level i+1 ((w-1)/2) x ((h-1)/2) of level i.
 * Level 0 is the map as float: for RGBA_U16 the channel-0 id when First <= id <= Last,
 * else Unknown (0); for LUM_I16 d >= 0, else -1.  Each next level is a [1 2 1] filter
 * along x and then along y that skips the unknown marker and divides by the weights it
 * kept, the vertical pass halving the value (CreateDispMipMaps.cu:97-218).  Levels that
 * do not fit are STEREO_ERROR_INVALID_CONFIG, as the reference's "Mip-map level count
 * larger than the maximal possible level" (:262-265). */
stereo_status stereo_disp_mipmaps(stereo_context* ctx, const stereo_image* disp, stereo_image* levels, int32_t nlevels,
                                  stereo_stream stream);

/* ---- BlockMatcher1 (BlockMatcher1.h:40-180) ---------------------------------------- */

typedef struct stereo_bm1_params {
    int32_t patch;            /* EConfig: 5, 9, 11 or 15 */
    int32_t sim_type;         /* ESimType: 0 AbsoluteDifference, 1 SignedProduct */
    int32_t disp_min;         /* _SDisparityConfig: disparities disp_min <= d < disp_max */
    int32_t disp_max;         /*   (<= 4096) */
    int32_t disp_offset;      /*   and the pixel offset of MapPixelPos */
    float min_delta_thresh;   /* fMinDeltaThresh */
    float sim_thresh;         /* fSimThresh */
    float grad_thresh;        /* fGradThresh */
    int32_t use_disp_input;   /* iUseDispInput: MatcherUseDisparity around an initial map */
    int32_t left_to_right;    /* bLeftToRight */
    int32_t reserved[4];      /* must be 0 */
} stereo_bm1_params;

/* CDriver::Configure (BlockMatcher1.cu:1218-1268): fmt is the LUM_U8 pair format, at least
 * 48 wide and patch + 3 tall. */
stereo_status stereo_bm1_configure(stereo_context* ctx, stereo_driver** drv, const stereo_image_format* fmt,
                                   const stereo_bm1_params* params);
/* CDriver::Process (BlockMatcher1.cu:1360-1400) with the reference's block-tiled results:
 * 32-thread blocks of PCX x 4 result origins (PCX = 8 * ((33 - patch) / 8)), each with the
 * disparity groups, test patches and edge rules of the reference (Matcher :590-847,
 * MatcherUseDisparity :857-1203).  disp: RGBA_U16 (EDisparityId, similarity maximum and the
 * two neighbour deltas as 16-bit fractions); pixels the reference does not write are left
 * as they are.  disp_init: an RGBA_U16 id map of at most the image size (scale = width /
 * its width; positions outside it read as Unknown), required iff use_disp_input. */
stereo_status stereo_bm1_process(stereo_driver* drv, stereo_image* disp, const stereo_image* left,
                                 const stereo_image* right, const stereo_image* disp_init, stereo_stream stream);

/* ---- BlockMatcherNCC and BlockMatcherAW1 with the reference's results ------------------
 *
 * The reference operators themselves, as opposed to the fused plain-specification matcher
 * above (stereo_matcher_* with STEREO_ALGO_NCC / STEREO_ALGO_AW1): one direction, the
 * reference's block tiling (32-thread blocks of result origins, full blocks only), its
 * disparity groups and test-patch placement, its winner rule (a strictly larger value
 * wins, the lowest disparity on ties), its gates and codes, and its four output channels.
 * disp is RGBA_U16; pixels the reference does not write are left as they are.  disp_init
 * (use_disp_input): an RGBA_U16 id map of at most the image size, scale = width / its
 * width, positions outside it read as Unknown (MatcherUseDisparity). */

/* BlockMatcherNCC::_SParameter (BlockMatcherNCC.h:83-110). */
typedef struct stereo_ncc_params {
    int32_t patch;            /* EConfig: 11 (Patch_11x11), 9 or 5 */
    int32_t disp_min;         /* _SDisparityConfig: disparities disp_min <= d < disp_max (<= 4096) */
    int32_t disp_max;
    int32_t disp_offset;      /*   and the pixel offset of MapPixelPos */
    int32_t min_delta_thresh; /* iMinDeltaThresh: both byte drops around the maximum at least this */
    float ncc_thresh;         /* fNccThresh: byte maximum / 255 at least this, else NotFound */
    float grad_thresh;        /* fGradThresh: mean |L(x) - L(x + 1)| below it is CannotEvaluate */
    int32_t use_disp_input;   /* iUseDispInput: MatcherUseDisparity around an initial map */
    int32_t left_to_right;    /* bLeftToRight */
    int32_t reserved[4];      /* must be 0 */
} stereo_ncc_params;

/* CDriver::Configure (BlockMatcherNCC.cu:1270-1318): fmt is the LUM_U8 pair format, at least
 * 48 wide and 15 + patch tall. */
stereo_status stereo_ncc_configure(stereo_context* ctx, stereo_driver** drv, const stereo_image_format* fmt,
                                   const stereo_ncc_params* params);
/* CDriver::Process (BlockMatcherNCC.cu:1393-1420; Matcher :614-884, MatcherUseDisparity
 * :894-1255).  disp channels: EDisparityId (First + d, NotFound, NotSpecific, Saturated,
 * CannotEvaluate, or Unknown where no byte score exceeds 0), the byte NCC maximum
 * min(255, 255 max(0, ncc)) of the reference's float formula, and the largest byte drops to
 * the two disparities before and after the maximum. */
stereo_status stereo_ncc_process(stereo_driver* drv, stereo_image* disp, const stereo_image* left,
                                 const stereo_image* right, const stereo_image* disp_init, stereo_stream stream);

/* BlockMatcherAW1::_SParameter (BlockMatcherAW1.h:84-124). */
typedef struct stereo_aw1_params {
    int32_t patch;           /* EConfig: 15 (Patch_15x15, sub-patches of 5) or 9 (sub-patches of 3) */
    int32_t sim_type;        /* ESimType: 0 AbsoluteDifference, 1 SignedProduct */
    int32_t disp_min;        /* _SDisparityConfig, as stereo_ncc_params */
    int32_t disp_max;
    int32_t disp_offset;
    float min_delta_thresh;  /* fMinDeltaThresh, in % */
    float sim_thresh;        /* fSimThresh, in % */
    float grad_thresh;       /* fGradThresh: the window's standard deviation, in % of 255 */
    int32_t use_disp_input;  /* iUseDispInput */
    int32_t left_to_right;   /* bLeftToRight */
    int32_t reserved[4];     /* must be 0 */
} stereo_aw1_params;

/* CDriver::Configure (BlockMatcherAW1.cu:1270-1300): fmt is the LUM_U8 pair format, at least
 * 48 wide and 20 (patch 15) or 18 (patch 9) tall. */
stereo_status stereo_aw1_configure(stereo_context* ctx, stereo_driver** drv, const stereo_image_format* fmt,
                                   const stereo_aw1_params* params);
/* CDriver::Process (BlockMatcherAW1.cu:1380-1420; Matcher :657-913, MatcherUseDisparity
 * :923-1263).  disp channels: EDisparityId as for NCC, then the similarity maximum and the two
 * neighbour deltas as 16-bit fractions of [-100 %, 100 %] (SimFloatToInteger :172-176). */
stereo_status stereo_aw1_process(stereo_driver* drv, stereo_image* disp, const stereo_image* left,
                                 const stereo_image* right, const stereo_image* disp_init, stereo_stream stream);

/* ---- rectification (RectifyImage.h:40-90) ------------------------------------------- */

/* The camera model of the rectification: the rectified camera's pixel becomes a ray,
 * rotated into the source camera (R, row-major, rectified to source), projected, distorted
 * with the OpenCV model and mapped through the source camera.  It stands in for the
 * reference's _CPinhole / _CDistortion<OpenCV> / _CFrame3D parameters, whose library is
 * not part of the reference. */
typedef struct stereo_rectify_params {
    float src_fx, src_fy, src_cx, src_cy;      /* source camera, pixels */
    float k1, k2, p1, p2, k3;                  /* its OpenCV distortion */
    float rect_fx, rect_fy, rect_cx, rect_cy;  /* rectified camera, pixels */
    float R[9];
    int32_t reserved[4];                       /* must be 0 */
} stereo_rectify_params;

/* CDriver::Configure (RectifyImage.cu:166-203): fmt is the rectified (target) LUM_U8 image. */
stereo_status stereo_rectify_configure(stereo_context* ctx, stereo_driver** drv, const stereo_image_format* fmt,
                                       const stereo_rectify_params* params);
/* CDriver::Process (RectifyImage.cu:205-232, RectifyUndistort :140-159): per target pixel,
 * bilinear over the 2x2 source block at the floor of its source position when the block is
 * inside (values / 255, result rounded), else the nearest source pixel when inside, else 0.
 * The source is LUM_U8 of any size. */
stereo_status stereo_rectify_process(stereo_driver* drv, stereo_image* target, const stereo_image* source,
                                     stereo_stream stream);

/* ---- disparity colouring (ColorDisparityImage.h:40-125) ------------------------------ */

enum {
    STEREO_COLOR_ANALYSIS = 0, /* RGBA_U8: grey ids, palette for the codes */
    STEREO_COLOR_GRAY = 1,     /* LUM_U16 */
    STEREO_COLOR_COLOR = 2,    /* RGBA_U8: colour ramp */
    STEREO_COLOR_DISP_ABS = 3, /* LUM_U16: the id */
    STEREO_COLOR_DISP_REL = 4  /* LUM_U16: (0x7FFF + inf offset) - (id - First) */
};

typedef struct stereo_color_params {
    uint16_t disp_min, disp_max;  /* uDispMin, uDispMax: the ids mapped to 0 and 1 */
    uint16_t disp_inf_offset;     /* uDispInfOffset (DispRel) */
    uint16_t pad;
    /* RGBA for Unknown, CannotEvaluate, Saturated, NotSpecific, NotFound, NotUnique,
     * Inconsistent (SParameter defaults, ColorDisparityImage.h:76-84) */
    uint8_t colors[7][4];
    int32_t reserved[4];          /* must be 0 */
} stereo_color_params;

/* CDriver::Configure: fmt is the disparity image, RGBA_U16 (EDisparityId), LUM_U16 or LUM_F32. */
stereo_status stereo_color_configure(stereo_context* ctx, stereo_driver** drv, const stereo_image_format* fmt,
                                     const stereo_color_params* params);
/* CDriver::Process (ColorDisparityImage.cu:330-410) for one STEREO_COLOR_* style; out is
 * RGBA_U8 (Analysis, Color) or LUM_U16 (Gray, DispAbs, DispRel) of the same size. */
stereo_status stereo_color_process(stereo_driver* drv, stereo_image* out, const stereo_image* disp, int32_t style,
                                   stereo_stream stream);

/* ---- disparity de-noise (Statistics::DispDenoise, Statistics.DispDenoise.h:40-190) ---- */

typedef struct stereo_denoise_params {
    int32_t patch;         /* window P: 3, 5, 7, 9, 11 or 16 (EConfig::Patch_PxP) */
    float support_ratio;   /* fSupportRatio: at least P*P*ratio valid values (0.5) */
    float mean_cutoff;     /* fMeanCutOff (1.0) */
    int32_t fill;          /* iDoFill: fill invalid pixels in flat windows (1) */
    int32_t reserved[4];   /* must be 0 */
} stereo_denoise_params;

/* CDriver::Configure (Statistics.DispDenoise.cu:360-380): fmt is the input map, RGBA_U16
 * (EDisparityId in channel 0) or LUM_I16 (d, -1 invalid), at least P x P. */
stereo_status stereo_denoise_configure(stereo_context* ctx, stereo_driver** drv, const stereo_image_format* fmt,
                                       const stereo_denoise_params* params);
/* CDriver::Process (FilterNoBorder, Statistics.DispDenoise.cu:110-330) on the whole image:
 * per pixel, a P x P window (offsets -P/2 .. P-1-P/2, positions outside the image moved by
 * P into it) gives the count, sum and sum of squares of the valid values.  Fewer than
 * P*P*ratio valid: Unknown (0, or -1).  A valid pixel v: m = (sum - v) / (count - 1) in
 * FP32; floor(m + 0.5) when |v - m| < cut-off, else Unknown.  An invalid pixel (with fill):
 * m = sum / count, m2 = sumsq / count; floor(m + 0.5) when sqrt(|m2 - m*m|) < cut-off,
 * else unchanged.  out: LUM_U16 ids for RGBA_U16 input, LUM_I16 for LUM_I16. */
stereo_status stereo_denoise_process(stereo_driver* drv, stereo_image* out, const stereo_image* in,
                                     stereo_stream stream);

/* ---- device calibration ------------------------------------------------------ */

/* INT32 VALU issue rate of the device in tera lane-operations per second, measured
 * with independent chains of one instruction on every SIMD (op 0: v_add_u32, op 1:
 * v_bcnt_u32_b32); the roofline of the integer matchers uses the lower of this and
 * the specified peak.  No counterpart in the reference.  Synchronous. */
stereo_status stereo_valu_peak(stereo_context* ctx, int32_t op, double* tera_ops);
/* The number of float bit patterns (NaNs excepted) whose bf16 from the device's packed
 * conversion instruction, which the one-launch network uses, differs from integer
 * round-to-nearest-even, which its other kernels and CPU reference use; 0 on gfx950.
 * No counterpart in the reference.  Synchronous. */
stereo_status stereo_bf16_rounding_check(stereo_context* ctx, uint64_t* mismatches);

/* ---- synthetic workloads (host only, no device needed) ------------------- */

/* Seeded stereo pair: L, R (W*H bytes, pitch W) and optional ground truth
 * (int16, pitch W).  kind: 0 constant rectangles, 1 slanted planes, 2 mixed.
 * Returns 0 on success. */
int stereo_synth_pair(uint64_t seed, int W, int H, int kind, int nregions, int dmax, int dbg, uint8_t* L,
                      uint8_t* R, int16_t* gt);

#ifdef __cplusplus
}
#endif

#endif /* CLUTEC_STEREO_H */
