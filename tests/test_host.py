"""Host-side checks that need no GPU: the C ABI library loads and exports every
entry point its header declares, configuration validation mirrors the
reference's throw sites, and the synthetic workload generator is pinned."""
import hashlib
import json
import os
import re

import numpy as np
import pytest

import clutec_stereo as cs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "clutec_stereo.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(stereo_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_symbols_exported():
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(cs.lib, n), f"{n} declared in clutec_stereo.h but not exported"
    # the binding covers the whole header
    assert sorted(cs.EXPORTED) == names


def test_abi_version():
    assert cs.lib.stereo_abi_version() == 1


def test_status_strings():
    assert cs.lib.stereo_status_string(cs.ERROR_INVALID_CONFIG).decode() == "invalid algorithm configuration"
    assert cs.lib.stereo_pixel_size(cs.PIXEL_RGBA_U16) == 8
    assert cs.lib.stereo_pixel_size(99) == 0


@pytest.mark.parametrize("algo,win,cen,D,expect", [
    (cs.ALGO_SAD, (9, 9), (1, 1), 64, cs.OK),
    (cs.ALGO_SAD, (7, 7), (1, 1), 256, cs.OK),
    (cs.ALGO_CENSUS, (9, 9), (9, 7), 256, cs.OK),
    (cs.ALGO_CENSUS, (9, 9), (9, 9), 256, cs.ERROR_INVALID_CONFIG),   # 81 bits do not fit
    (cs.ALGO_SAD, (11, 11), (1, 1), 64, cs.ERROR_INVALID_CONFIG),
    (cs.ALGO_SAD, (9, 9), (1, 1), 300, cs.ERROR_INVALID_CONFIG),
    (cs.ALGO_SAD, (9, 9), (1, 1), 12, cs.ERROR_INVALID_CONFIG),
    (9, (9, 9), (1, 1), 64, cs.ERROR_INVALID_CONFIG),
])
def test_matcher_validate(algo, win, cen, D, expect):
    p = cs.matcher_params(algo, win, cen, D, 1)
    st, msg = cs.validate_matcher(640, 480, p)
    assert st == expect, msg
    if algo == 9:
        assert msg == "Invalid algorithm configuration."


def test_matcher_validate_band_search():
    # the band search around an initial map is configured for every matcher, up to the
    # matchers' own width limit
    for algo, win in [(cs.ALGO_SAD, (9, 9)), (cs.ALGO_CENSUS, (9, 9)), (cs.ALGO_NCC, (11, 11)),
                      (cs.ALGO_AW1, (5, 5)), (cs.ALGO_AW2, (3, 3))]:
        p = cs.matcher_params(algo, win, (9, 7), 128, 1, use_disp_input=True)
        st, msg = cs.validate_matcher(640, 480, p)
        assert st == cs.OK, msg
        assert cs.validate_matcher(16384 + 64, 480, p)[0] == cs.ERROR_INVALID_FORMAT
    p = cs.matcher_params(cs.ALGO_AW1, (5, 5), (9, 7), 128, 1)
    p.use_disp_input = 2
    assert cs.validate_matcher(640, 480, p)[0] == cs.ERROR_INVALID_CONFIG


def test_matcher_validate_formats():
    p = cs.matcher_params(cs.ALGO_SAD, (9, 9), (1, 1), 64, 1)
    assert cs.validate_matcher(8, 8, p)[0] == cs.ERROR_INVALID_FORMAT
    assert cs.validate_matcher(0, 100, p)[0] == cs.ERROR_INVALID_FORMAT
    p.reserved[2] = 5
    assert cs.validate_matcher(640, 480, p)[0] == cs.ERROR_INVALID_CONFIG
    p.reserved[2] = 0
    p.lr_tolerance = -1
    assert cs.validate_matcher(640, 480, p)[0] == cs.ERROR_INVALID_CONFIG


def test_create_without_device_reports_invalid_device():
    import ctypes as C
    h = C.c_void_p()
    st = cs.lib.stereo_create(C.byref(h), 0)
    if st == cs.OK:
        cs.lib.stereo_destroy(h)
        pytest.skip("a gfx950 device is present")
    assert st == cs.ERROR_INVALID_DEVICE
    assert cs.lib.stereo_last_error(None)
    assert cs.lib.stereo_create(C.byref(h), -1) == cs.ERROR_INVALID_DEVICE


def test_synth_deterministic_and_seeded():
    a = cs.synth_pair(3, 120, 80, 0, 5, 40, 8, with_gt=True)
    b = cs.synth_pair(3, 120, 80, 0, 5, 40, 8, with_gt=True)
    c = cs.synth_pair(4, 120, 80, 0, 5, 40, 8, with_gt=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[0], c[0])
    L, R, gt = a
    assert L.min() == 0 and L.max() == 255          # min/max rescale
    assert gt.min() >= 0 and gt.max() <= 40
    # forward warp: where the right pixel is not disoccluded it is L(x+d) plus small noise
    y, x = 40, 100
    d = int(gt[y, x])
    assert abs(int(R[y, x - d]) - int(L[y, x])) <= 12


def test_synth_golden_hashes():
    """Byte-exact pin of the generator (fixture written by tests/golden/make_synth_golden.py)."""
    path = os.path.join(ROOT, "tests", "golden", "synth_sha256.json")
    gold = json.load(open(path))
    for entry in gold:
        L, R, gt = cs.synth_pair(entry["seed"], entry["W"], entry["H"], entry["kind"], entry["nregions"],
                                 entry["dmax"], entry["dbg"], with_gt=True)
        h = hashlib.sha256(L.tobytes() + R.tobytes() + gt.tobytes()).hexdigest()
        assert h == entry["sha256"], entry


def test_synth_rejects_bad_args():
    import ctypes as C
    buf = np.zeros(100, np.uint8)
    assert cs.lib.stereo_synth_pair(0, 5, 5, 0, 1, 10, 1, buf.ctypes.data, buf.ctypes.data, None) != 0
    assert cs.lib.stereo_synth_pair(0, 50, 50, 7, 1, 10, 1, buf.ctypes.data, buf.ctypes.data, None) != 0


def test_cpp_driver_layer_builds_and_reports_errors(tmp_path):
    """The C++ driver header compiles on its own, and the host program built on it
    prints its usage and turns a missing device into a CException exit."""
    import subprocess
    src = tmp_path / "t.cpp"
    # the reference's own SParameter::Set calls (BlockMatcherNCC.h:95, BlockMatcherAW1.h:106,
    # BlockMatcherSAD1.h:86, BlockMatcherAW2.h:95) and Process shapes compile unchanged
    src.write_text(
        '#include "clutec/stereo.hpp"\n'
        'using namespace Clu::Cuda;\n'
        'int main(){ _SDisparityConfig c; c.iMin = 3; c.iMax = 64; c.iOffset = 2;\n'
        ' BlockMatcherNCC::SParameter p; p.Set(c, 2, 0.5f, 2.0f, 1, true, BlockMatcherNCC::EConfig::Patch_9x9);\n'
        ' BlockMatcherAW1::SParameter a; a.Set(c, 0.5f, 60.0f, 2.0f, 0, false, BlockMatcherAW1::EConfig::Patch_15x15,\n'
        '   BlockMatcherAW1::ESimType::SignedProduct);\n'
        ' BlockMatcherSAD1::SParameter s; s.Set(0, 64, -1, 1e30f, 0.0f, BlockMatcherSAD1::EConfig::Patch_9x9);\n'
        ' BlockMatcherAW2::SParameter q; q.Set(0, 64, -1, 1e30f, 0.0f, 1, BlockMatcherAW2::EConfig::Patch_9x9);\n'
        ' if (p.iUseDispInput + q.iUseDispInput == 7) { CDevice dev(0); CDeviceImage img;\n'
        '   BlockMatcherNCC::CDriver n; n.Configure(dev, img.Format(), p); n.Process(img, img, img, img, img);\n'
        '   BlockMatcherAW1::CDriver w; w.Configure(dev, img.Format(), a); w.Process(img, img, img, img, img);\n'
        '   BlockMatcherSAD1::CDriver d1; d1.Configure(dev, img.Format(), s); d1.Process(img, img, img);\n'
        '   BlockMatcherAW2::CDriver d2; d2.Configure(dev, img.Format(), q); d2.Process(img, img, img, img);\n'
        '   ProcessKernel(n, img, img, img, img, img); }\n'
        ' return p.iUseDispInput + q.iUseDispInput - 2; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    exe = os.path.join(os.path.dirname(cs.__file__), "lib", "stereo_run")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    ctx_ok = True
    try:
        cs.Context(0).close()
    except cs.StereoError:
        ctx_ok = False
    if not ctx_ok:
        r = subprocess.run([exe, "census", "64", "32", "16", "1", str(tmp_path / "o")], capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 1 and "error 5" in r.stderr


def test_kernel_isa_has_no_fused_or_in_place_dpp(tmp_path):
    """Row-shift DPP moves must stay standalone v_mov_b32_dpp into a fresh register:
    folded into arithmetic (v_add/v_sub/v_subrev_u32_dpp) or done in place they gave
    wrong window sums on gfx950.  The wave-shift maxima of the right-view chain are
    written as one inline-asm block and are allowed."""
    import subprocess
    csrc = os.path.join(os.path.dirname(cs.__file__), "csrc")
    bad = []
    for f in ("bm.hip", "ncc.hip", "aw.hip"):
        asm = tmp_path / (f + ".s")
        r = subprocess.run(["/opt/rocm/bin/hipcc", "-O3", "-std=c++17", "--offload-arch=gfx950", "--cuda-device-only",
                            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I",
                            os.path.join(ROOT, "include"), "-I", csrc, "-x", "hip", "-S", os.path.join(csrc, f), "-o",
                            str(asm)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        for line in asm.read_text().splitlines():
            m = re.match(r"\s*(v_\w+_dpp)\s+(v\d+),\s*(v\d+)", line)
            if not m or "wave_shr" in line:
                continue
            op, dst, src = m.groups()
            if op != "v_mov_b32_dpp" or dst == src:
                bad.append(f"{f}: {line.strip()}")
    assert not bad, "\n".join(bad[:10])
