// stereo_run: the reference call sequence through the C++ driver layer.
//
//   1. generate a synthetic pair, 2. upload into device images, 3. Configure the
//   drivers, 4. ProcessKernel: matcher (census or SAD; L->R, R->L and check fused)
//   and, from the two one-sided maps in disparity-id form, the reference-style
//   cross-check, then Statistics::DispDenoise and ColorDisparityImage over the checked id
//   map and RectifyImage of the left image, 5. download, 6. write the maps and report
//   milliseconds.
//
// usage: stereo_run <algo:sad|census|ncc|aw1|aw2|bm1|aw1band|nccband|nccref|aw1ref> <W> <H> <D> <seed> <out_prefix> [runs]
// aw1band / nccband: the full-range map, taken at half resolution (d / 2), as the initial
// map of Process(disp, L, R, dispInit) with iUseDispInput; writes <out_prefix>_init.i16 and
// _band.i16 and stops there.
// nccref / aw1ref: the reference-format BlockMatcherNCC / BlockMatcherAW1 drivers configured
// with the reference's SParameter::Set over the disparity range [3, D) with pixel offset 2,
// Process(disp, L, R, dispInit, debug); writes <out_prefix>_ref.u16x4 and stops there.
// Otherwise writes <out_prefix>_checked.i16, _lr.i16, _rl.i16 (row-major int16, -1 invalid),
// _checkid.u16x4 (the cross-check's id map), _denoised.u16 (its de-noised ids),
// _color.rgba (its Analysis colouring) and _rectified.u8 (the rectified left image)
// and prints one line "ms <median> valid <fraction> range <min> <max>".
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "clutec/stereo.hpp"

using namespace Clu::Cuda;

static bool write_file(const std::string& path, const void* p, size_t n)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(p, 1, n, f) == n;
    std::fclose(f);
    return ok;
}

int main(int argc, char** argv)
{
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s <sad|census|ncc|aw1|aw2|bm1|aw1band|nccband|nccref|aw1ref> W H D seed out_prefix [runs]\n",
                     argv[0]);
        return 2;
    }
    const std::string algo = argv[1];
    const int W = std::atoi(argv[2]), H = std::atoi(argv[3]), D = std::atoi(argv[4]);
    const unsigned long long seed = std::strtoull(argv[5], nullptr, 10);
    const std::string out = argv[6];
    const int runs = argc > 7 ? std::max(1, std::atoi(argv[7])) : 5;

    try {
        std::vector<uint8_t> L((size_t)W * H), R((size_t)W * H);
        if (stereo_synth_pair(seed, W, H, 0, 8, D - 1, std::min(8, D - 1), L.data(), R.data(), nullptr) != 0)
            throw CException(STEREO_ERROR_INVALID_ARGUMENT, "stereo_synth_pair failed");

        CDevice dev(0);
        dev.SetTiming(true);
        const SImageFormat fmtIn(W, H, EPixelType::Lum_UInt8);
        const SImageFormat fmtDisp(W, H, EPixelType::Lum_Int16);
        const SImageFormat fmtId(W, H, EPixelType::RGBA_UInt16);
        CDeviceImage imgL(dev, fmtIn), imgR(dev, fmtIn);
        imgL.CopyFrom(L.data(), W);
        imgR.CopyFrom(R.data(), W);
        CDeviceImage disp(dev, fmtDisp), lr(dev, fmtDisp), rl(dev, fmtDisp);

        _SDisparityConfig dcfg;
        dcfg.iMin = 0;
        dcfg.iMax = D;
        std::vector<double> ms;

        auto run = [&](auto& drv) {
            for (int i = 0; i < runs; ++i) {
                drv.Begin();
                drv.ProcessMaps(disp, lr, rl, imgL, imgR);
                drv.End();
                ms.push_back(drv.LastProcessTime());
            }
            drv.LogLastProcessTime();
        };
        if (algo == "census") {
            BlockMatcherCensus::SParameter p;
            p.Set(dcfg, 9, 7, 9, 1);
            BlockMatcherCensus::CDriver drv;
            drv.Configure(dev, fmtIn, p);
            run(drv);
        } else if (algo == "sad") {
            // the reference's Set(iOffset, iDispRange, iMinDeltaThresh, fSadThresh, fGradThresh,
            // eConfig), thresholds inactive
            BlockMatcherSAD1::SParameter p;
            p.Set(0, D, -1, 1e30f, 0.0f, BlockMatcherSAD1::EConfig::Patch_9x9);
            BlockMatcherSAD1::CDriver drv;
            drv.Configure(dev, fmtIn, p);
            run(drv);
        } else if (algo == "ncc") {
            BlockMatcherNCC::SFusedParameter p;
            p.Set(dcfg, BlockMatcherNCC::EConfig::Patch_11x11, 1);
            BlockMatcherNCC::CFusedDriver drv;
            drv.Configure(dev, fmtIn, p);
            run(drv);
        } else if (algo == "aw1" || algo == "aw2") {
            if (algo == "aw1") {
                BlockMatcherAW1::SFusedParameter p;
                p.Set(dcfg, BlockMatcherAW1::EConfig::Patch_15x15, 1);
                BlockMatcherAW1::CFusedDriver drv;
                drv.Configure(dev, fmtIn, p);
                run(drv);
            } else {
                BlockMatcherAW2::SParameter p;
                p.Set(0, D, -1, 1e30f, 0.0f, 0, BlockMatcherAW2::EConfig::Patch_15x15);
                BlockMatcherAW2::CDriver drv;
                drv.Configure(dev, fmtIn, p);
                run(drv);
            }
        } else if (algo == "aw1band" || algo == "nccband") {
            // MatcherUseDisparity through the driver layer: a full-range pass, then the band
            // search around its map at half resolution
            const bool aw = algo == "aw1band";
            BlockMatcherAW1::SFusedParameter pa;
            BlockMatcherNCC::SFusedParameter pn;
            BlockMatcherAW1::CFusedDriver fa, ba;
            BlockMatcherNCC::CFusedDriver fn, bn;
            if (aw) {
                pa.Set(dcfg, BlockMatcherAW1::EConfig::Patch_9x9, 1, 0);
                fa.Configure(dev, fmtIn, pa);
                pa.Set(dcfg, BlockMatcherAW1::EConfig::Patch_9x9, 1, 1);
                ba.Configure(dev, fmtIn, pa);
                fa.Process(disp, imgL, imgR);
            } else {
                pn.Set(dcfg, BlockMatcherNCC::EConfig::Patch_9x9, 1, 0);
                fn.Configure(dev, fmtIn, pn);
                pn.Set(dcfg, BlockMatcherNCC::EConfig::Patch_9x9, 1, 1);
                bn.Configure(dev, fmtIn, pn);
                fn.Process(disp, imgL, imgR);
            }
            std::vector<int16_t> full((size_t)W * H);
            disp.CopyInto(full.data(), (size_t)W * 2);
            const int iw = W / 2, ih = H / 2;
            std::vector<int16_t> init((size_t)iw * ih);
            for (int y = 0; y < ih; ++y)
                for (int x = 0; x < iw; ++x) {
                    const int16_t d = full[(size_t)(2 * y) * W + 2 * x];
                    init[(size_t)y * iw + x] = d >= 0 ? (int16_t)(d / 2) : (int16_t)-1;
                }
            CDeviceImage dinit(dev, SImageFormat(iw, ih, EPixelType::Lum_Int16));
            dinit.CopyFrom(init.data(), (size_t)iw * 2);
            for (int i = 0; i < runs; ++i) {
                if (aw)
                    ms.push_back(ProcessKernel(ba, disp, imgL, imgR, dinit));
                else
                    ms.push_back(ProcessKernel(bn, disp, imgL, imgR, dinit));
            }
            std::vector<int16_t> band((size_t)W * H);
            disp.CopyInto(band.data(), (size_t)W * 2);
            if (!write_file(out + "_init.i16", init.data(), init.size() * 2) ||
                !write_file(out + "_band.i16", band.data(), band.size() * 2))
                return 1;
            std::sort(ms.begin(), ms.end());
            std::printf("ms %.4f\n", ms[ms.size() / 2]);
            return 0;
        } else if (algo == "nccref" || algo == "aw1ref") {
            // the reference operators through their own SParameter::Set, a range that does not
            // start at 0 and a pixel offset
            _SDisparityConfig rcfg;
            rcfg.iMin = 3;
            rcfg.iMax = D;
            rcfg.iOffset = 2;
            CDeviceImage ids(dev, fmtId), unused(dev, fmtId);
            std::vector<uint16_t> hids((size_t)W * H * 4, 0x5A5A);
            ids.CopyFrom(hids.data(), (size_t)W * 8);
            if (algo == "nccref") {
                BlockMatcherNCC::SParameter p;
                p.Set(rcfg, 2, 0.5f, 2.0f, 0, true, BlockMatcherNCC::EConfig::Patch_9x9);
                BlockMatcherNCC::CDriver drv;
                drv.Configure(dev, fmtIn, p);
                for (int i = 0; i < runs; ++i) ms.push_back(ProcessKernel(drv, ids, imgL, imgR, unused, unused));
            } else {
                BlockMatcherAW1::SParameter p;
                p.Set(rcfg, 0.5f, 60.0f, 2.0f, 0, true, BlockMatcherAW1::EConfig::Patch_15x15,
                      BlockMatcherAW1::ESimType::AbsoluteDifference);
                BlockMatcherAW1::CDriver drv;
                drv.Configure(dev, fmtIn, p);
                for (int i = 0; i < runs; ++i) ms.push_back(ProcessKernel(drv, ids, imgL, imgR, unused, unused));
            }
            ids.CopyInto(hids.data(), (size_t)W * 8);
            if (!write_file(out + "_ref.u16x4", hids.data(), hids.size() * 2)) return 1;
            std::sort(ms.begin(), ms.end());
            std::printf("ms %.4f\n", ms[ms.size() / 2]);
            return 0;
        } else if (algo == "bm1") {
            // BlockMatcher1 writes the reference's id map with similarity channels; the
            // program writes it to <out_prefix>_bm1.u16x4 and stops there
            BlockMatcher1::SParameter p;
            p.Set(dcfg, 1.0f, 60.0f, 2.0f, 0, true, BlockMatcher1::EConfig::Patch_9x9,
                  BlockMatcher1::ESimType::AbsoluteDifference);
            BlockMatcher1::CDriver drv;
            drv.Configure(dev, fmtIn, p);
            CDeviceImage ids(dev, fmtId);
            std::vector<uint16_t> hids((size_t)W * H * 4, 0x5A5A);
            ids.CopyFrom(hids.data(), (size_t)W * 8);
            for (int i = 0; i < runs; ++i) ms.push_back(ProcessKernel(drv, ids, imgL, imgR));
            ids.CopyInto(hids.data(), (size_t)W * 8);
            if (!write_file(out + "_bm1.u16x4", hids.data(), hids.size() * 2)) return 1;
            std::sort(ms.begin(), ms.end());
            std::printf("ms %.4f\n", ms[ms.size() / 2]);
            return 0;
        } else {
            std::fprintf(stderr, "unknown algorithm '%s'\n", algo.c_str());
            return 2;
        }

        // the reference's separate cross-check over disparity-id maps: the matcher's
        // one-sided maps converted to EDisparityId (First + d), checked L->R
        std::vector<int16_t> hlr((size_t)W * H), hrl((size_t)W * H), hchk((size_t)W * H);
        lr.CopyInto(hlr.data(), (size_t)W * 2);
        rl.CopyInto(hrl.data(), (size_t)W * 2);
        disp.CopyInto(hchk.data(), (size_t)W * 2);
        std::vector<uint16_t> idl((size_t)W * H * 4, 0), idr((size_t)W * H * 4, 0);
        for (size_t i = 0; i < (size_t)W * H; ++i) {
            idl[4 * i] = hlr[i] >= 0 ? (uint16_t)((int)EDisparityId::First + hlr[i]) : (uint16_t)EDisparityId::NotFound;
            idr[4 * i] = hrl[i] >= 0 ? (uint16_t)((int)EDisparityId::First + hrl[i]) : (uint16_t)EDisparityId::NotFound;
        }
        CDeviceImage dl(dev, fmtId), dr(dev, fmtId), dchk(dev, fmtId);
        dl.CopyFrom(idl.data(), (size_t)W * 8);
        dr.CopyFrom(idr.data(), (size_t)W * 8);
        CrossCheckDisparityMaps::SParameter cp;
        cp.Set(dcfg, 1, 1);
        CrossCheckDisparityMaps::CDriver cc;
        cc.Configure(dev, fmtId, cp);
        const double cms = ProcessKernel(cc, dchk, dl, dr);
        const int dmin = cc.DispMin(), dmax = cc.DispMax();

        // Statistics::DispDenoise over the cross-checked id map (5x5, defaults)
        Statistics::DispDenoise::SParameter np;
        Statistics::DispDenoise::CDriver dn;
        dn.Configure(dev, fmtId, np);
        CDeviceImage dden(dev, SImageFormat(W, H, EPixelType::Lum_UInt16));
        ProcessKernel(dn, dden, dchk);
        std::vector<uint16_t> hid((size_t)W * H * 4), hden((size_t)W * H);
        dchk.CopyInto(hid.data(), (size_t)W * 8);
        dden.CopyInto(hden.data(), (size_t)W * 2);

        // ColorDisparityImage (Analysis style) of the checked id map, and RectifyImage of the
        // left image by a small rotation about the optical axis
        ColorDisparityImage::SParameter colp;
        colp.Set(1, (uint16_t)(D + 1), 0);
        ColorDisparityImage::CDriver col;
        col.Configure(dev, fmtId, colp);
        CDeviceImage dcol(dev, SImageFormat(W, H, EPixelType::RGBA_UInt8));
        ProcessKernel(col, dcol, dchk, ColorDisparityImage::EStyle::Analysis);
        std::vector<uint8_t> hcol((size_t)W * H * 4), hrect((size_t)W * H);
        dcol.CopyInto(hcol.data(), (size_t)W * 4);
        RectifyImage::_SParameter rp;
        rp.fSrcFx = rp.fSrcFy = rp.fRectFx = rp.fRectFy = 0.8f * (float)W;
        rp.fSrcCx = rp.fRectCx = 0.5f * (float)(W - 1);
        rp.fSrcCy = rp.fRectCy = 0.5f * (float)(H - 1);
        rp.fK1 = -0.05f;
        const float c = 0.9998f, sn = 0.0199987f;  // about 1.15 degrees
        const float rot[9] = {c, -sn, 0.0f, sn, c, 0.0f, 0.0f, 0.0f, 1.0f};
        for (int i = 0; i < 9; ++i) rp.mRot[i] = rot[i];
        RectifyImage::CDriver rect;
        rect.Configure(dev, fmtIn, rp);
        CDeviceImage drect(dev, fmtIn);
        ProcessKernel(rect, drect, imgL);
        drect.CopyInto(hrect.data(), (size_t)W);

        size_t valid = 0;
        for (int16_t v : hchk) valid += v >= 0;
        std::sort(ms.begin(), ms.end());
        if (!write_file(out + "_checked.i16", hchk.data(), hchk.size() * 2) ||
            !write_file(out + "_lr.i16", hlr.data(), hlr.size() * 2) || !write_file(out + "_rl.i16", hrl.data(), hrl.size() * 2) ||
            !write_file(out + "_checkid.u16x4", hid.data(), hid.size() * 2) ||
            !write_file(out + "_denoised.u16", hden.data(), hden.size() * 2) ||
            !write_file(out + "_color.rgba", hcol.data(), hcol.size()) ||
            !write_file(out + "_rectified.u8", hrect.data(), hrect.size())) {
            std::fprintf(stderr, "cannot write %s_*.i16\n", out.c_str());
            return 1;
        }
        std::printf("ms %.4f valid %.6f range %d %d crosscheck_ms %.4f\n", ms[ms.size() / 2], (double)valid / hchk.size(),
                    dmin, dmax, cms);
    } catch (const CException& e) {
        std::fprintf(stderr, "error %d: %s\n", (int)e.Status(), e.what());
        return 1;
    }
    return 0;
}
