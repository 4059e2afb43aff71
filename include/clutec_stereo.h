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
    STEREO_PIXEL_RGBA_U16 = 7   /* _SDisparityConfig::TPixelEx: 4 x uint16, EDisparityId in channel 0 */
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
/* CDeviceImage::CopyFrom: host rows (host_pitch bytes apart) -> device. */
stereo_status stereo_image_upload(stereo_context* ctx, stereo_image* img, const void* host, size_t host_pitch,
                                  stereo_stream stream);
/* CDeviceImage::CopyInto: device -> host rows. */
stereo_status stereo_image_download(stereo_context* ctx, const stereo_image* img, void* host, size_t host_pitch,
                                    stereo_stream stream);

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
    int32_t reserved[5];     /* must be 0 */
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
