// C++ driver layer over the C ABI of the gfx950 stereo engine.
//
// Keeps the call shape of the reference's MultiView drivers: a device, device images,
// one CDriver per operator with SParameter::Set(...), Configure(device, format, pars),
// Process(...), and ProcessKernel(driver, args...) that times the call and returns the
// milliseconds (KernelDriverBase.h:47-200).  Errors come back as exceptions carrying
// the C ABI status and message (EnsureOK, Api.cpp:86-94).  Header only; link
// libclutec_stereo.so.
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../clutec_stereo.h"

namespace Clu
{
namespace Cuda
{

class CException : public std::runtime_error
{
public:
    CException(stereo_status st, const std::string& msg) : std::runtime_error(msg), m_status(st) {}
    stereo_status Status() const { return m_status; }

private:
    stereo_status m_status;
};

inline void EnsureOK(stereo_status st, const stereo_context* ctx, const char* what)
{
    if (st != STEREO_OK) {
        const char* m = stereo_last_error(ctx);
        throw CException(st, std::string(what) + ": " + (m && *m ? m : stereo_status_string(st)));
    }
}

// Pixel types used on the path (Clu::EPixelType x EDataType subset).
enum class EPixelType : int32_t
{
    Lum_UInt8 = STEREO_PIXEL_LUM_U8,
    Lum_UInt16 = STEREO_PIXEL_LUM_U16,
    Lum_Int16 = STEREO_PIXEL_LUM_I16,
    Lum_UInt32 = STEREO_PIXEL_LUM_U32,
    Lum_UInt64 = STEREO_PIXEL_LUM_U64,
    Lum_Single = STEREO_PIXEL_LUM_F32,
    RGBA_UInt16 = STEREO_PIXEL_RGBA_U16,
    RGBA_UInt8 = STEREO_PIXEL_RGBA_U8,
};

// Clu::SImageFormat subset: width, height, pixel type.
struct SImageFormat
{
    int32_t iWidth = 0, iHeight = 0;
    EPixelType ePixelType = EPixelType::Lum_UInt8;

    SImageFormat() = default;
    SImageFormat(int32_t w, int32_t h, EPixelType t) : iWidth(w), iHeight(h), ePixelType(t) {}
    bool IsValid() const { return iWidth > 0 && iHeight > 0; }
    stereo_image_format Raw() const { return stereo_image_format{iWidth, iHeight, (int32_t)ePixelType}; }
};

// EDisparityId (DisparityId.h:33-45).
enum class EDisparityId : uint16_t
{
    First = 1,
    Last = 0xFFF0 - 1,
    Unknown = 0,
    CannotEvaluate = 0xFFFF,
    Saturated = 0xFFFF - 1,
    NotFound = 0xFFFF - 2,
    NotSpecific = 0xFFFF - 3,
    NotUnique = 0xFFFF - 4,
    Inconsistent = 0xFFFF - 5,
};

// _SDisparityConfig subset (offset added by MapPixelPos).
struct _SDisparityConfig
{
    int32_t iMin = 0, iMax = 0, iOffset = 0;
    int32_t Min() const { return iMin; }
    int32_t Max() const { return iMax; }
    int32_t Range() const { return iMax - iMin; }
    int32_t MapPixelPos(int32_t x, int32_t d, int32_t ltor) const { return ltor ? x - d - iOffset : x + d + iOffset; }
};

class CDevice
{
public:
    CDevice() = default;
    explicit CDevice(int iIdx) { Set(iIdx); }
    ~CDevice()
    {
        if (m_ctx) stereo_destroy(m_ctx);
    }
    CDevice(const CDevice&) = delete;
    CDevice& operator=(const CDevice&) = delete;

    void Set(int iIdx)
    {
        if (m_ctx) stereo_destroy(m_ctx);
        m_ctx = nullptr;
        EnsureOK(stereo_create(&m_ctx, iIdx), nullptr, "CDevice::Set");
    }
    stereo_context* Context() const { return m_ctx; }
    void SetTiming(bool b) const { EnsureOK(stereo_set_timing(m_ctx, b ? 1 : 0), m_ctx, "CDevice::SetTiming"); }
    void Synchronize() const { EnsureOK(stereo_synchronize(m_ctx, nullptr), m_ctx, "CDevice::Synchronize"); }

private:
    stereo_context* m_ctx = nullptr;
};

// Owning pitched device image (CDeviceImage, DeviceImage.h).
class CDeviceImage
{
public:
    CDeviceImage() = default;
    CDeviceImage(const CDevice& dev, const SImageFormat& f) { Create(dev, f); }
    ~CDeviceImage() { Destroy(); }
    CDeviceImage(const CDeviceImage&) = delete;
    CDeviceImage& operator=(const CDeviceImage&) = delete;
    CDeviceImage(CDeviceImage&& o) noexcept : m_ctx(o.m_ctx), m_img(o.m_img) { o.m_img.data = nullptr; }
    CDeviceImage& operator=(CDeviceImage&& o) noexcept
    {
        if (this != &o) {
            Destroy();
            m_ctx = o.m_ctx;
            m_img = o.m_img;
            o.m_img.data = nullptr;
        }
        return *this;
    }

    void Create(const CDevice& dev, const SImageFormat& f)
    {
        Destroy();
        m_ctx = dev.Context();
        EnsureOK(stereo_image_alloc(m_ctx, &m_img, f.iWidth, f.iHeight, (int32_t)f.ePixelType), m_ctx,
                 "CDeviceImage::Create");
    }
    void Destroy()
    {
        if (m_img.data) stereo_image_free(m_ctx, &m_img);
        m_img.data = nullptr;
    }
    void CopyFrom(const void* host, size_t hostPitch)
    {
        EnsureOK(stereo_image_upload(m_ctx, &m_img, host, hostPitch, nullptr), m_ctx, "CDeviceImage::CopyFrom");
    }
    void CopyInto(void* host, size_t hostPitch) const
    {
        EnsureOK(stereo_image_download(m_ctx, &m_img, host, hostPitch, nullptr), m_ctx, "CDeviceImage::CopyInto");
    }
    SImageFormat Format() const
    {
        return SImageFormat(m_img.format.width, m_img.format.height, (EPixelType)m_img.format.pixel_type);
    }
    stereo_image* Raw() { return &m_img; }
    const stereo_image* Raw() const { return &m_img; }

private:
    stereo_context* m_ctx = nullptr;
    stereo_image m_img{};
};

// KernelDriverBase.h:47-177: name, last process time, timing around Process.
class CKernelDriverBase
{
public:
    explicit CKernelDriverBase(const char* name) : m_sKernelName(name) {}
    virtual ~CKernelDriverBase()
    {
        if (m_drv) stereo_driver_destroy(m_drv);
    }
    CKernelDriverBase(const CKernelDriverBase&) = delete;
    CKernelDriverBase& operator=(const CKernelDriverBase&) = delete;

    const std::string& KernelName() const { return m_sKernelName; }
    double LastProcessTime() const { return m_dLastProcessTime; }
    void LogLastProcessTime() const { std::fprintf(stderr, "Kernel '%s': %gms\n", m_sKernelName.c_str(), m_dLastProcessTime); }

    void Begin() { m_bTimed = true; }
    void End()
    {
        if (!m_drv || !m_bTimed) return;
        double ms = 0.0;
        EnsureOK(stereo_driver_last_time_ms(m_drv, &ms), m_ctx, "End");
        m_dLastProcessTime = ms;
    }

protected:
    void Reset(stereo_context* ctx, stereo_driver* drv)
    {
        if (m_drv) stereo_driver_destroy(m_drv);
        m_ctx = ctx;
        m_drv = drv;
    }

    stereo_context* m_ctx = nullptr;
    stereo_driver* m_drv = nullptr;
    std::string m_sKernelName;
    double m_dLastProcessTime = 0.0;
    bool m_bTimed = false;
};

// ProcessKernel (KernelDriverBase.h:180-200): times one Process call and returns ms.
// Timing needs CDevice::SetTiming(true); otherwise 0 is returned.
template <typename TDriver, typename... TPars>
double ProcessKernel(TDriver& xDriver, TPars&&... xPars)
{
    xDriver.Begin();
    xDriver.Process(std::forward<TPars>(xPars)...);
    try {
        xDriver.End();
    } catch (const CException& e) {
        if (e.Status() != STEREO_ERROR_NOT_READY) throw;
    }
    return xDriver.LastProcessTime();
}

namespace CensusTransform
{
struct SParameter
{
    int32_t iWindowW = 9, iWindowH = 7;
    bool bReference3x3 = false;
    void Set(int32_t w, int32_t h, bool ref3x3 = false)
    {
        iWindowW = w;
        iWindowH = h;
        bReference3x3 = ref3x3;
    }
};

class CDriver : public CKernelDriverBase
{
public:
    CDriver() : CKernelDriverBase("Census Transform") {}
    void Configure(const CDevice& dev, const SImageFormat& fmt, const SParameter& pars = SParameter())
    {
        const stereo_image_format f = fmt.Raw();
        const stereo_census_params p{pars.iWindowW, pars.iWindowH, pars.bReference3x3 ? 1 : 0};
        stereo_driver* d = nullptr;
        EnsureOK(stereo_census_configure(dev.Context(), &d, &f, &p), dev.Context(), "CensusTransform::Configure");
        Reset(dev.Context(), d);
    }
    void Process(CDeviceImage& trg, const CDeviceImage& src)
    {
        EnsureOK(stereo_census_process(m_drv, trg.Raw(), src.Raw(), nullptr), m_ctx, "CensusTransform::Process");
    }
};
}  // namespace CensusTransform

// Shared implementation of the fused plain-specification matchers: one call runs the
// left-to-right and right-to-left winner-take-alls and the cross check
// (CrossCheckDisparityMaps.cu:62-109), writing the checked map as LUM_I16 (-1 invalid) or
// as RGBA_UInt16 disparity-id pixels, and optionally the two one-sided LUM_I16 maps.
// Their disparity range is [0, D) with no pixel offset.
class CMatcherDriver : public CKernelDriverBase
{
public:
    using CKernelDriverBase::CKernelDriverBase;

    void Process(CDeviceImage& disp, const CDeviceImage& L, const CDeviceImage& R)
    {
        EnsureOK(stereo_matcher_process(m_drv, disp.Raw(), nullptr, nullptr, L.Raw(), R.Raw(), nullptr), m_ctx,
                 "Matcher::Process");
    }
    void ProcessMaps(CDeviceImage& disp, CDeviceImage& lr, CDeviceImage& rl, const CDeviceImage& L, const CDeviceImage& R)
    {
        EnsureOK(stereo_matcher_process(m_drv, disp.Raw(), lr.Raw(), rl.Raw(), L.Raw(), R.Raw(), nullptr), m_ctx,
                 "Matcher::Process");
    }
    // With an initial disparity map (LUM_I16 or RGBA_UInt16 ids, any size up to the
    // images'): the band search of MatcherUseDisparity (BlockMatcherNCC.cu:894-1255) over
    // 8-disparity bins per 64 x 56 tile.  The driver must be configured with iUseDispInput.
    void Process(CDeviceImage& disp, const CDeviceImage& L, const CDeviceImage& R, const CDeviceImage& dispInit)
    {
        EnsureOK(stereo_matcher_process_init(m_drv, disp.Raw(), nullptr, nullptr, L.Raw(), R.Raw(), dispInit.Raw(),
                                             nullptr),
                 m_ctx, "Matcher::Process");
    }

protected:
    void Configure(const CDevice& dev, const SImageFormat& fmt, const stereo_matcher_params& p)
    {
        const stereo_image_format f = fmt.Raw();
        stereo_driver* d = nullptr;
        EnsureOK(stereo_matcher_configure(dev.Context(), &d, &f, &p), dev.Context(), "Matcher::Configure");
        Reset(dev.Context(), d);
    }
    static stereo_matcher_params Params(int32_t algo, int32_t ww, int32_t wh, int32_t D, int32_t tol, int32_t cw = 1,
                                        int32_t ch = 1, int32_t use_disp_input = 0)
    {
        stereo_matcher_params p{};
        p.use_disp_input = use_disp_input;
        p.algorithm = algo;
        p.window_w = ww;
        p.window_h = wh;
        p.census_w = cw;
        p.census_h = ch;
        p.num_disparities = D;
        p.lr_tolerance = tol;
        return p;
    }
    // The fused matchers search [0, D) from the left pixel's own column: a range that does
    // not start at 0, or a pixel offset, is refused rather than searched differently.
    static int32_t PlainRange(const _SDisparityConfig& c, const char* who)
    {
        if (c.Min() != 0 || c.iOffset != 0)
            throw CException(STEREO_ERROR_INVALID_CONFIG,
                             std::string(who) + ": the fused matcher searches disparities [0, D) with no offset; "
                                                "Min() = " + std::to_string(c.Min()) + ", offset = " +
                                 std::to_string(c.iOffset) + " (the reference-format BlockMatcherNCC / "
                                                             "BlockMatcherAW1 drivers take both)");
        return c.Range();
    }
    static void Inactive(bool active, const char* who, const char* what)
    {
        if (active)
            throw CException(STEREO_ERROR_INVALID_CONFIG,
                             std::string(who) + ": " + what + " is not applied by the fused matcher; pass an inactive value");
    }
};

// BlockMatcherSAD1 (BlockMatcherSAD1.h:40-140).  The reference operator does not compile
// (it reads AlgoSum::PatchCountSplitX, which AlgoArraySum lacks) and is excluded from its
// build, so no reference results exist; the driver keeps its SParameter::Set signature
// and runs the fused SAD matcher.  Its only reference window, Patch_16x16, is even and is
// refused; the fused windows 9x9, 7x7 and 5x5 follow it in the enum.  Thresholds must be
// inactive: fSadThresh > 255 (a mean SAD always passes), iMinDeltaThresh < 0, fGradThresh <= 0.
namespace BlockMatcherSAD1
{
enum class EConfig
{
    Patch_16x16 = 0,
    Patch_9x9,
    Patch_7x7,
    Patch_5x5,
};
struct _SParameter
{
    int32_t iOffset = 0;
    int32_t iDispRange = 64;
    float fSadThresh = 1e30f;
    int32_t iMinDeltaThresh = -1;
    float fGradThresh = 0.0f;
};
struct SParameter : public _SParameter
{
    EConfig eConfig = EConfig::Patch_9x9;
    int32_t iLrTolerance = 1;   // the fused check's |dR - dL| tolerance
    int32_t iUseDispInput = 0;  // 1: Process may take an initial disparity map
    void Set(int32_t _iOffset, int32_t _iDispRange, int32_t _iMinDeltaThresh, float _fSadThresh, float _fGradThresh,
             EConfig _eConfig)
    {
        iOffset = _iOffset;
        iDispRange = _iDispRange;
        iMinDeltaThresh = _iMinDeltaThresh;
        fSadThresh = _fSadThresh;
        fGradThresh = _fGradThresh;
        eConfig = _eConfig;
    }
};
class CDriver : public CMatcherDriver
{
public:
    CDriver() : CMatcherDriver("Block Matcher SAD 1") {}
    void Configure(const CDevice& dev, const SImageFormat& fmt, const SParameter& p)
    {
        static const int32_t sizes[] = {16, 9, 7, 5};
        const int i = static_cast<int>(p.eConfig);
        if (i < 1 || i > 3)
            throw CException(STEREO_ERROR_INVALID_CONFIG,
                             "BlockMatcherSAD1: Patch_16x16 (the reference's only window, even) is not built; use "
                             "Patch_9x9, Patch_7x7 or Patch_5x5");
        _SDisparityConfig c;
        c.iMin = 0;
        c.iMax = p.iDispRange;
        c.iOffset = p.iOffset;
        const int32_t D = PlainRange(c, "BlockMatcherSAD1");
        Inactive(!(p.fSadThresh > 255.0f), "BlockMatcherSAD1", "fSadThresh");
        Inactive(p.iMinDeltaThresh >= 0, "BlockMatcherSAD1", "iMinDeltaThresh");
        Inactive(p.fGradThresh > 0.0f, "BlockMatcherSAD1", "fGradThresh");
        CMatcherDriver::Configure(dev, fmt, Params(STEREO_ALGO_SAD, sizes[i], sizes[i], D, p.iLrTolerance, 1, 1,
                                                   p.iUseDispInput));
    }
};
}  // namespace BlockMatcherSAD1

// The census matcher: CensusDifference (BlockMatcherNCC.cu:153-158) over the fused matcher.
// The reference has no census block matcher of its own (its call sites are commented
// out), so the parameters are this engine's.
namespace BlockMatcherCensus
{
struct SParameter
{
    _SDisparityConfig xDispConfig;
    int32_t iCensusW = 9, iCensusH = 7, iPatchSize = 9, iLrTolerance = 1;
    int32_t iUseDispInput = 0;  // 1: Process may take an initial disparity map
    void Set(const _SDisparityConfig& cfg, int32_t cw = 9, int32_t ch = 7, int32_t patch = 9, int32_t tol = 1,
             int32_t useDispInput = 0)
    {
        xDispConfig = cfg;
        iCensusW = cw;
        iCensusH = ch;
        iPatchSize = patch;
        iLrTolerance = tol;
        iUseDispInput = useDispInput;
    }
};
class CDriver : public CMatcherDriver
{
public:
    CDriver() : CMatcherDriver("Block Matcher Census") {}
    void Configure(const CDevice& dev, const SImageFormat& fmt, const SParameter& p)
    {
        CMatcherDriver::Configure(dev, fmt,
                                  Params(STEREO_ALGO_CENSUS, p.iPatchSize, p.iPatchSize,
                                         PlainRange(p.xDispConfig, "BlockMatcherCensus"), p.iLrTolerance, p.iCensusW,
                                         p.iCensusH, p.iUseDispInput));
    }
};
}  // namespace BlockMatcherCensus

// Reference-format drivers of BlockMatcherNCC / BlockMatcherAW1: one direction, the
// reference's block tiling, decisions and RGBA_UInt16 channels (stereo_ncc_*, stereo_aw1_*).
class CRefMatcherDriver : public CKernelDriverBase
{
public:
    using CKernelDriverBase::CKernelDriverBase;
    // The reference's Process (BlockMatcherNCC.h:137-143, BlockMatcherAW1.h:151-157): dispInit
    // is read when configured with iUseDispInput; the debug surface is not written.
    void Process(CDeviceImage& disp, const CDeviceImage& L, const CDeviceImage& R, const CDeviceImage& dispInit,
                 const CDeviceImage& /*debug*/)
    {
        Run(disp, L, R, m_bUseInit ? dispInit.Raw() : nullptr);
    }
    void Process(CDeviceImage& disp, const CDeviceImage& L, const CDeviceImage& R, const CDeviceImage& dispInit)
    {
        Run(disp, L, R, m_bUseInit ? dispInit.Raw() : nullptr);
    }
    void Process(CDeviceImage& disp, const CDeviceImage& L, const CDeviceImage& R) { Run(disp, L, R, nullptr); }

protected:
    virtual void Run(CDeviceImage& disp, const CDeviceImage& L, const CDeviceImage& R, const stereo_image* init) = 0;
    bool m_bUseInit = false;
};

namespace BlockMatcherNCC
{
enum class EConfig
{
    Patch_11x11 = 0,
    Patch_9x9,
    Patch_5x5,
};
inline int32_t PatchSize(EConfig e)
{
    static const int32_t sizes[] = {11, 9, 5};
    const int i = static_cast<int>(e);
    return i >= 0 && i < 3 ? sizes[i] : 0;
}
// BlockMatcherNCC.h:83-111
struct _SParameter
{
    _SDisparityConfig xDispConfig;
    int32_t iMinDeltaThresh = 0;
    float fNccThresh = 0.0f;
    float fGradThresh = 0.0f;
    int32_t iUseDispInput = 0;
};
struct SParameter : public _SParameter
{
    EConfig eConfig = EConfig::Patch_11x11;
    bool bLeftToRight = true;
    void Set(const _SDisparityConfig& _xDispConfig, int32_t _iMinDeltaThresh, float _fNccThresh, float _fGradThresh,
             int32_t _iUseDispInput, bool _bLeftToRight, EConfig _eConfig)
    {
        xDispConfig = _xDispConfig;
        iMinDeltaThresh = _iMinDeltaThresh;
        fNccThresh = _fNccThresh;
        fGradThresh = _fGradThresh;
        iUseDispInput = _iUseDispInput;
        bLeftToRight = _bLeftToRight;
        eConfig = _eConfig;
    }
};
class CDriver : public CRefMatcherDriver
{
public:
    CDriver() : CRefMatcherDriver("Block Matcher NCC") {}
    void Configure(const CDevice& dev, const SImageFormat& fmt, const SParameter& p)
    {
        const int32_t P = PatchSize(p.eConfig);
        if (!P) throw CException(STEREO_ERROR_INVALID_CONFIG, "Invalid algorithm configuration.");
        const stereo_image_format f = fmt.Raw();
        stereo_ncc_params q{};
        q.patch = P;
        q.disp_min = p.xDispConfig.iMin;
        q.disp_max = p.xDispConfig.iMax;
        q.disp_offset = p.xDispConfig.iOffset;
        q.min_delta_thresh = p.iMinDeltaThresh;
        q.ncc_thresh = p.fNccThresh;
        q.grad_thresh = p.fGradThresh;
        q.use_disp_input = p.iUseDispInput ? 1 : 0;
        q.left_to_right = p.bLeftToRight ? 1 : 0;
        stereo_driver* d = nullptr;
        EnsureOK(stereo_ncc_configure(dev.Context(), &d, &f, &q), dev.Context(), "BlockMatcherNCC::Configure");
        Reset(dev.Context(), d);
        m_bUseInit = q.use_disp_input != 0;
    }

protected:
    void Run(CDeviceImage& disp, const CDeviceImage& L, const CDeviceImage& R, const stereo_image* init) override
    {
        EnsureOK(stereo_ncc_process(m_drv, disp.Raw(), L.Raw(), R.Raw(), init, nullptr), m_ctx,
                 "BlockMatcherNCC::Process");
    }
};

// The fused plain-specification NCC matcher: exact integer covariance, both directions and
// the check in one call (stereo_matcher_* with STEREO_ALGO_NCC).
struct SFusedParameter
{
    _SDisparityConfig xDispConfig;
    EConfig eConfig = EConfig::Patch_11x11;
    int32_t iLrTolerance = 1;
    int32_t iUseDispInput = 0;
    void Set(const _SDisparityConfig& cfg, EConfig config, int32_t tol = 1, int32_t useDispInput = 0)
    {
        xDispConfig = cfg;
        eConfig = config;
        iLrTolerance = tol;
        iUseDispInput = useDispInput;
    }
};
class CFusedDriver : public CMatcherDriver
{
public:
    CFusedDriver() : CMatcherDriver("Block Matcher NCC (fused)") {}
    void Configure(const CDevice& dev, const SImageFormat& fmt, const SFusedParameter& p)
    {
        const int32_t w = PatchSize(p.eConfig);
        if (!w) throw CException(STEREO_ERROR_INVALID_CONFIG, "Invalid algorithm configuration.");
        CMatcherDriver::Configure(dev, fmt, Params(STEREO_ALGO_NCC, w, w, PlainRange(p.xDispConfig, "BlockMatcherNCC"),
                                                   p.iLrTolerance, 1, 1, p.iUseDispInput));
    }
    // winning scores of the last Process (LUM_Single)
    void Scores(CDeviceImage& score, const CDeviceImage& L, const CDeviceImage& R)
    {
        EnsureOK(stereo_matcher_scores(m_drv, score.Raw(), L.Raw(), R.Raw(), nullptr), m_ctx, "BlockMatcherNCC::Scores");
    }
};
}  // namespace BlockMatcherNCC

namespace BlockMatcherAW1
{
enum class EConfig
{
    Patch_15x15 = 0,
    Patch_9x9,
};
enum class ESimType
{
    AbsoluteDifference = 0,
    SignedProduct,
};
inline int32_t PatchSize(EConfig e)
{
    return e == EConfig::Patch_15x15 ? 15 : e == EConfig::Patch_9x9 ? 9 : 0;
}
// BlockMatcherAW1.h:84-124
struct _SParameter
{
    _SDisparityConfig xDispConfig;
    float fMinDeltaThresh = 0.0f;
    float fSimThresh = 0.0f;
    float fGradThresh = 0.0f;
    int32_t iUseDispInput = 0;
};
struct SParameter : public _SParameter
{
    EConfig eConfig = EConfig::Patch_15x15;
    ESimType eSimType = ESimType::AbsoluteDifference;
    bool bLeftToRight = true;
    void Set(const _SDisparityConfig& _xDispConfig, float _fMinDeltaThresh, float _fSimThresh, float _fGradThresh,
             int32_t _iUseDispInput, bool _bLeftToRight, EConfig _eConfig, ESimType _eSimType)
    {
        xDispConfig = _xDispConfig;
        fMinDeltaThresh = _fMinDeltaThresh;
        fSimThresh = _fSimThresh;
        fGradThresh = _fGradThresh;
        iUseDispInput = _iUseDispInput;
        bLeftToRight = _bLeftToRight;
        eConfig = _eConfig;
        eSimType = _eSimType;
    }
};
class CDriver : public CRefMatcherDriver
{
public:
    CDriver() : CRefMatcherDriver("Block Matcher Adaptive Window 1") {}
    void Configure(const CDevice& dev, const SImageFormat& fmt, const SParameter& p)
    {
        const int32_t P = PatchSize(p.eConfig);
        if (!P) throw CException(STEREO_ERROR_INVALID_CONFIG, "Invalid algorithm configuration.");
        const stereo_image_format f = fmt.Raw();
        stereo_aw1_params q{};
        q.patch = P;
        q.sim_type = static_cast<int32_t>(p.eSimType);
        q.disp_min = p.xDispConfig.iMin;
        q.disp_max = p.xDispConfig.iMax;
        q.disp_offset = p.xDispConfig.iOffset;
        q.min_delta_thresh = p.fMinDeltaThresh;
        q.sim_thresh = p.fSimThresh;
        q.grad_thresh = p.fGradThresh;
        q.use_disp_input = p.iUseDispInput ? 1 : 0;
        q.left_to_right = p.bLeftToRight ? 1 : 0;
        stereo_driver* d = nullptr;
        EnsureOK(stereo_aw1_configure(dev.Context(), &d, &f, &q), dev.Context(), "BlockMatcherAW1::Configure");
        Reset(dev.Context(), d);
        m_bUseInit = q.use_disp_input != 0;
    }

protected:
    void Run(CDeviceImage& disp, const CDeviceImage& L, const CDeviceImage& R, const stereo_image* init) override
    {
        EnsureOK(stereo_aw1_process(m_drv, disp.Raw(), L.Raw(), R.Raw(), init, nullptr), m_ctx,
                 "BlockMatcherAW1::Process");
    }
};

// The fused plain-specification AW1 matcher (stereo_matcher_* with STEREO_ALGO_AW1):
// AbsoluteDifference similarity, both directions and the check in one call.
struct SFusedParameter
{
    _SDisparityConfig xDispConfig;
    EConfig eConfig = EConfig::Patch_15x15;
    int32_t iLrTolerance = 1;
    int32_t iUseDispInput = 0;
    void Set(const _SDisparityConfig& cfg, EConfig config, int32_t tol = 1, int32_t useDispInput = 0)
    {
        xDispConfig = cfg;
        eConfig = config;
        iLrTolerance = tol;
        iUseDispInput = useDispInput;
    }
};
class CFusedDriver : public CMatcherDriver
{
public:
    CFusedDriver() : CMatcherDriver("Block Matcher Adaptive Window 1 (fused)") {}
    void Configure(const CDevice& dev, const SImageFormat& fmt, const SFusedParameter& p)
    {
        const int32_t s = PatchSize(p.eConfig) / 3;
        if (!s) throw CException(STEREO_ERROR_INVALID_CONFIG, "Invalid algorithm configuration.");
        CMatcherDriver::Configure(dev, fmt, Params(STEREO_ALGO_AW1, s, s, PlainRange(p.xDispConfig, "BlockMatcherAW1"),
                                                   p.iLrTolerance, 1, 1, p.iUseDispInput));
    }
};
}  // namespace BlockMatcherAW1

// BlockMatcherAW2 (BlockMatcherAW2.h:40-140).  Excluded from the reference's build, whose
// band-search branch is empty (BlockMatcherAW2.cu:858-875); the driver keeps its
// SParameter::Set signature and runs the fused AW2 matcher (SAD, the two smallest
// neighbours), its band search included.  iOffset must be 0 and the thresholds inactive:
// fSadThresh > 255, iMinDeltaThresh < 0, fGradThresh <= 0.
namespace BlockMatcherAW2
{
using EConfig = BlockMatcherAW1::EConfig;
struct _SParameter
{
    int32_t iOffset = 0;
    int32_t iDispRange = 64;
    int32_t iMinDeltaThresh = -1;
    float fSadThresh = 1e30f;
    float fGradThresh = 0.0f;
    int32_t iUseDispInput = 0;
};
struct SParameter : public _SParameter
{
    EConfig eConfig = EConfig::Patch_15x15;
    int32_t iLrTolerance = 1;
    void Set(int32_t _iOffset, int32_t _iDispRange, int32_t _iMinDeltaThresh, float _fSadThresh, float _fGradThresh,
             int32_t _iUseDispInput, EConfig _eConfig)
    {
        iOffset = _iOffset;
        iDispRange = _iDispRange;
        iMinDeltaThresh = _iMinDeltaThresh;
        fSadThresh = _fSadThresh;
        fGradThresh = _fGradThresh;
        iUseDispInput = _iUseDispInput;
        eConfig = _eConfig;
    }
};
class CDriver : public CMatcherDriver
{
public:
    CDriver() : CMatcherDriver("Block Matcher Adaptive Window 2") {}
    void Configure(const CDevice& dev, const SImageFormat& fmt, const SParameter& p)
    {
        const int32_t s = BlockMatcherAW1::PatchSize(p.eConfig) / 3;
        if (!s) throw CException(STEREO_ERROR_INVALID_CONFIG, "Invalid algorithm configuration.");
        _SDisparityConfig c;
        c.iMin = 0;
        c.iMax = p.iDispRange;
        c.iOffset = p.iOffset;
        const int32_t D = PlainRange(c, "BlockMatcherAW2");
        Inactive(!(p.fSadThresh > 255.0f), "BlockMatcherAW2", "fSadThresh");
        Inactive(p.iMinDeltaThresh >= 0, "BlockMatcherAW2", "iMinDeltaThresh");
        Inactive(p.fGradThresh > 0.0f, "BlockMatcherAW2", "fGradThresh");
        m_bUseInit = p.iUseDispInput != 0;
        CMatcherDriver::Configure(dev, fmt, Params(STEREO_ALGO_AW2, s, s, D, p.iLrTolerance, 1, 1, p.iUseDispInput));
    }
    // The reference's Process (BlockMatcherAW2.h:129-134): dispInit is read when configured
    // with iUseDispInput.
    void Process(CDeviceImage& disp, const CDeviceImage& L, const CDeviceImage& R, const CDeviceImage& dispInit)
    {
        if (m_bUseInit) CMatcherDriver::Process(disp, L, R, dispInit);
        else CMatcherDriver::Process(disp, L, R);
    }
    using CMatcherDriver::Process;

private:
    bool m_bUseInit = false;
};
}  // namespace BlockMatcherAW2

// BlockMatcher1 (BlockMatcher1.h:40-180) with the reference's block-tiled results.
namespace BlockMatcher1
{
enum class EConfig
{
    Patch_5x5 = 0,
    Patch_9x9,
    Patch_11x11,
    Patch_15x15,
};
enum class ESimType
{
    AbsoluteDifference = 0,
    SignedProduct,
};
inline int32_t PatchSize(EConfig e)
{
    static const int32_t sizes[] = {5, 9, 11, 15};
    const int i = static_cast<int>(e);
    return i >= 0 && i < 4 ? sizes[i] : 0;
}
struct SParameter
{
    _SDisparityConfig xDispConfig;
    float fMinDeltaThresh = 0.0f;
    float fSimThresh = 0.0f;
    float fGradThresh = 0.0f;
    int32_t iUseDispInput = 0;
    bool bLeftToRight = true;
    EConfig eConfig = EConfig::Patch_9x9;
    ESimType eSimType = ESimType::AbsoluteDifference;
    void Set(const _SDisparityConfig& cfg, float minDelta, float simThresh, float gradThresh, int32_t useDispInput,
             bool leftToRight, EConfig config, ESimType simType)
    {
        xDispConfig = cfg;
        fMinDeltaThresh = minDelta;
        fSimThresh = simThresh;
        fGradThresh = gradThresh;
        iUseDispInput = useDispInput;
        bLeftToRight = leftToRight;
        eConfig = config;
        eSimType = simType;
    }
};
class CDriver : public CKernelDriverBase
{
public:
    CDriver() : CKernelDriverBase("Block Matcher Adaptive Window 1") {}
    void Configure(const CDevice& dev, const SImageFormat& fmt, const SParameter& p)
    {
        const stereo_image_format f = fmt.Raw();
        const stereo_bm1_params bp{PatchSize(p.eConfig), static_cast<int32_t>(p.eSimType), p.xDispConfig.iMin,
                                   p.xDispConfig.iMax, p.xDispConfig.iOffset, p.fMinDeltaThresh, p.fSimThresh,
                                   p.fGradThresh, p.iUseDispInput, p.bLeftToRight ? 1 : 0, {0, 0, 0, 0}};
        stereo_driver* d = nullptr;
        EnsureOK(stereo_bm1_configure(dev.Context(), &d, &f, &bp), dev.Context(), "BlockMatcher1::Configure");
        Reset(dev.Context(), d);
    }
    // Process(disp, L, R, dispInit): dispInit is used when iUseDispInput is set
    void Process(CDeviceImage& disp, const CDeviceImage& L, const CDeviceImage& R, const CDeviceImage* dispInit = nullptr)
    {
        EnsureOK(stereo_bm1_process(m_drv, disp.Raw(), L.Raw(), R.Raw(), dispInit ? dispInit->Raw() : nullptr, nullptr),
                 m_ctx, "BlockMatcher1::Process");
    }
};
}  // namespace BlockMatcher1

namespace CrossCheckDisparityMaps
{
struct SParameter
{
    _SDisparityConfig xDispConfig;
    int32_t iDispDeltaThresh = 1;
    int32_t iIsLeftToRight = 1;
    void Set(const _SDisparityConfig& cfg, int32_t thresh, int32_t ltor)
    {
        xDispConfig = cfg;
        iDispDeltaThresh = thresh;
        iIsLeftToRight = ltor;
    }
};
class CDriver : public CKernelDriverBase
{
public:
    CDriver() : CKernelDriverBase("Cross Check Disparity Maps") {}
    void Configure(const CDevice& dev, const SImageFormat& fmt, const SParameter& p)
    {
        const stereo_image_format f = fmt.Raw();
        const stereo_crosscheck_params cp{p.iDispDeltaThresh, p.iIsLeftToRight, p.xDispConfig.iOffset, 0};
        stereo_driver* d = nullptr;
        EnsureOK(stereo_crosscheck_configure(dev.Context(), &d, &f, &cp), dev.Context(), "CrossCheckDisparityMaps::Configure");
        Reset(dev.Context(), d);
    }
    void Process(CDeviceImage& result, const CDeviceImage& lr, const CDeviceImage& rl)
    {
        EnsureOK(stereo_crosscheck_process(m_drv, result.Raw(), lr.Raw(), rl.Raw(), nullptr), m_ctx,
                 "CrossCheckDisparityMaps::Process");
    }
    uint16_t DispMin()
    {
        uint16_t a = 0, b = 0;
        EnsureOK(stereo_crosscheck_range(m_drv, &a, &b), m_ctx, "DispMin");
        return a;
    }
    uint16_t DispMax()
    {
        uint16_t a = 0, b = 0;
        EnsureOK(stereo_crosscheck_range(m_drv, &a, &b), m_ctx, "DispMax");
        return b;
    }
};
}  // namespace CrossCheckDisparityMaps

// RectifyImage (RectifyImage.h:40-90).  The reference's parameters are CluTec.Core camera
// objects (_CPinhole, _CDistortion<OpenCV>, _CFrame3D), which are not part of the
// reference; _SParameter holds the same model as numbers: the source camera with its
// OpenCV distortion, the rectified camera, and the rotation from the rectified to the
// source camera (row-major).
namespace RectifyImage
{
struct _SParameter
{
    float fSrcFx = 1.0f, fSrcFy = 1.0f, fSrcCx = 0.0f, fSrcCy = 0.0f;
    float fK1 = 0.0f, fK2 = 0.0f, fP1 = 0.0f, fP2 = 0.0f, fK3 = 0.0f;
    float fRectFx = 1.0f, fRectFy = 1.0f, fRectCx = 0.0f, fRectCy = 0.0f;
    float mRot[9] = {1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f};
};
class CDriver : public CKernelDriverBase
{
public:
    CDriver() : CKernelDriverBase("Rectify Image") {}
    // fmt: the rectified (target) image
    void Configure(const CDevice& dev, const SImageFormat& fmt, const _SParameter& p)
    {
        const stereo_image_format f = fmt.Raw();
        stereo_rectify_params rp{p.fSrcFx, p.fSrcFy, p.fSrcCx, p.fSrcCy, p.fK1, p.fK2, p.fP1, p.fP2, p.fK3,
                                 p.fRectFx, p.fRectFy, p.fRectCx, p.fRectCy, {}, {0, 0, 0, 0}};
        for (int i = 0; i < 9; ++i) rp.R[i] = p.mRot[i];
        stereo_driver* d = nullptr;
        EnsureOK(stereo_rectify_configure(dev.Context(), &d, &f, &rp), dev.Context(), "RectifyImage::Configure");
        Reset(dev.Context(), d);
    }
    void Process(CDeviceImage& target, const CDeviceImage& source)
    {
        EnsureOK(stereo_rectify_process(m_drv, target.Raw(), source.Raw(), nullptr), m_ctx, "RectifyImage::Process");
    }
};
}  // namespace RectifyImage

// ColorDisparityImage (ColorDisparityImage.h:40-125).
namespace ColorDisparityImage
{
enum class EStyle
{
    Analysis = STEREO_COLOR_ANALYSIS,
    Gray = STEREO_COLOR_GRAY,
    Color = STEREO_COLOR_COLOR,
    DispAbs = STEREO_COLOR_DISP_ABS,
    DispRel = STEREO_COLOR_DISP_REL,
};
struct SParameter
{
    uint16_t uDispMin = (uint16_t)EDisparityId::First, uDispMax = (uint16_t)EDisparityId::Last;
    uint16_t uDispInfOffset = 0;
    uint8_t pixUnknown[4] = {10, 100, 180, 255};
    uint8_t pixCannotEvaluate[4] = {0, 150, 50, 255};
    uint8_t pixSaturated[4] = {0, 200, 200, 255};
    uint8_t pixNotSpecific[4] = {200, 0, 200, 255};
    uint8_t pixNotFound[4] = {200, 50, 50, 255};
    uint8_t pixNotUnique[4] = {240, 50, 150, 255};
    uint8_t pixInconsistent[4] = {100, 10, 10, 255};
    void Set(uint16_t dispMin, uint16_t dispMax, uint16_t dispInfOffset)
    {
        uDispMin = dispMin;
        uDispMax = dispMax;
        uDispInfOffset = dispInfOffset;
    }
};
class CDriver : public CKernelDriverBase
{
public:
    CDriver() : CKernelDriverBase("Color Disparity Image") {}
    // fmt: the disparity image (RGBA_UInt16, Lum_UInt16 or Lum_Single ids)
    void Configure(const CDevice& dev, const SImageFormat& fmt, const SParameter& p)
    {
        const stereo_image_format f = fmt.Raw();
        stereo_color_params cp{};
        cp.disp_min = p.uDispMin;
        cp.disp_max = p.uDispMax;
        cp.disp_inf_offset = p.uDispInfOffset;
        const uint8_t* pal[7] = {p.pixUnknown, p.pixCannotEvaluate, p.pixSaturated, p.pixNotSpecific,
                                 p.pixNotFound, p.pixNotUnique, p.pixInconsistent};
        for (int k = 0; k < 7; ++k)
            for (int c = 0; c < 4; ++c) cp.colors[k][c] = pal[k][c];
        stereo_driver* d = nullptr;
        EnsureOK(stereo_color_configure(dev.Context(), &d, &f, &cp), dev.Context(), "ColorDisparityImage::Configure");
        Reset(dev.Context(), d);
    }
    // color: RGBA_UInt8 for Analysis and Color, Lum_UInt16 for the other styles
    void Process(CDeviceImage& color, const CDeviceImage& disp, EStyle style)
    {
        EnsureOK(stereo_color_process(m_drv, color.Raw(), disp.Raw(), (int32_t)style, nullptr), m_ctx,
                 "ColorDisparityImage::Process");
    }
};
}  // namespace ColorDisparityImage

// Statistics::DispDenoise (Statistics.DispDenoise.h:45-190): the input is an RGBA_U16
// EDisparityId map (uint16 ids out, LUM_U16) or a LUM_I16 map (LUM_I16 out).
namespace Statistics
{
namespace DispDenoise
{
enum class EConfig
{
    Patch_16x16 = 0,
    Patch_11x11,
    Patch_9x9,
    Patch_7x7,
    Patch_5x5,
    Patch_3x3,
};

inline int32_t PatchSize(EConfig e)
{
    static const int32_t sizes[] = {16, 11, 9, 7, 5, 3};
    const int i = static_cast<int>(e);
    return i >= 0 && i < 6 ? sizes[i] : 0;
}

struct SParameter
{
    float fSupportRatio = 0.5f;
    float fMeanCutOff = 1.0f;
    int32_t iDoFill = 1;
    EConfig eConfig = EConfig::Patch_5x5;
    void Set(float ratio, float cutoff, int32_t fill, EConfig e)
    {
        fSupportRatio = ratio;
        fMeanCutOff = cutoff;
        iDoFill = fill;
        eConfig = e;
    }
};

class CDriver : public CKernelDriverBase
{
public:
    CDriver() : CKernelDriverBase("Statistics disparity de-noise Filter") {}
    void Configure(const CDevice& dev, const SImageFormat& fmt, const SParameter& p)
    {
        const stereo_image_format f = fmt.Raw();
        const stereo_denoise_params dp{PatchSize(p.eConfig), p.fSupportRatio, p.fMeanCutOff, p.iDoFill, {0, 0, 0, 0}};
        stereo_driver* d = nullptr;
        EnsureOK(stereo_denoise_configure(dev.Context(), &d, &f, &dp), dev.Context(), "DispDenoise::Configure");
        Reset(dev.Context(), d);
    }
    void Process(CDeviceImage& target, const CDeviceImage& source)
    {
        EnsureOK(stereo_denoise_process(m_drv, target.Raw(), source.Raw(), nullptr), m_ctx, "DispDenoise::Process");
    }
};
}  // namespace DispDenoise
}  // namespace Statistics

}  // namespace Cuda
}  // namespace Clu
