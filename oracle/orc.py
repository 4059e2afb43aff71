"""ctypes wrapper of the CPU oracle (oracle/build/liborc_stereo.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / CPU baseline.  The engine never
uses it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liborc_stereo.so")

COST_SAD = 0
COST_CENSUS = 1

DISP_UNKNOWN = 0
DISP_FIRST = 1
DISP_LAST = 0xFFF0 - 1
DISP_CANNOT_EVALUATE = 0xFFFF
DISP_SATURATED = 0xFFFF - 1
DISP_NOT_FOUND = 0xFFFF - 2
DISP_NOT_SPECIFIC = 0xFFFF - 3
DISP_NOT_UNIQUE = 0xFFFF - 4
DISP_INCONSISTENT = 0xFFFF - 5


def build():
    """Compile the oracle with its Makefile (gcc, no GPU)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    i = C.c_int
    lib.orc_version.restype = i
    lib.orc_census.argtypes = [P, i, i, i, i, P]
    lib.orc_census3x3_ref.argtypes = [P, i, i, P]
    lib.orc_bm.argtypes = [i, P, P, i, i, i, i, i, i, i, i, P, P, P, P, i]
    lib.orc_bm_brute.argtypes = [i, P, P, i, i, i, i, i, i, i, i, P, P, P, P]
    lib.orc_bm_bins.argtypes = [i, P, P, i, i, i, i, i, i, i, i, P, P, P, P, i]
    lib.orc_bm_bins.restype = i
    lib.orc_ncc.argtypes = [P, P, i, i, i, i, i, i, P, P, P, P, P, i]
    lib.orc_ncc_bins.argtypes = [P, P, i, i, i, i, i, i, P, P, P, P, P, i]
    lib.orc_aw.argtypes = [i, P, P, i, i, i, i, i, P, P, P, P, P, i]
    lib.orc_aw_bins.argtypes = [i, P, P, i, i, i, i, i, P, P, P, P, P, i]
    lib.orc_crosscheck_plain.argtypes = [P, P, i, i, i, P]
    lib.orc_crosscheck_ref.argtypes = [P, P, i, i, i, i, i, P, P, P]
    lib.orc_bm1.argtypes = [P, P, i, i, i, i, i, i, i, i, C.c_float, C.c_float, C.c_float, P, i, i, P]
    lib.orc_ncc_ref.argtypes = [P, P, i, i, i, i, i, i, i, i, C.c_float, C.c_float, P, i, i, P]
    lib.orc_aw1_ref.argtypes = [P, P, i, i, i, i, i, i, i, i, C.c_float, C.c_float, C.c_float, P, i, i, P]
    lib.orc_rectify.argtypes = [P, i, i, P, i, i, P]
    lib.orc_color_disparity.argtypes = [P, i, i, i, i, i, i, i, P, P]
    lib.orc_mip_level0_ref.argtypes = [P, i, i, P]
    lib.orc_mip_level0_plain.argtypes = [P, i, i, P]
    lib.orc_mip_next.argtypes = [P, i, i, C.c_float, P]
    lib.orc_disp_denoise.argtypes = [i, P, i, i, i, C.c_float, C.c_float, i, P]
    for name in ["orc_census", "orc_census3x3_ref", "orc_bm", "orc_bm_brute", "orc_ncc", "orc_aw",
                 "orc_crosscheck_plain", "orc_crosscheck_ref", "orc_mip_level0_ref", "orc_mip_level0_plain",
                 "orc_mip_next", "orc_disp_denoise", "orc_bm1", "orc_ncc_ref", "orc_aw1_ref",
                 "orc_rectify", "orc_color_disparity", "orc_ncc_bins", "orc_aw_bins"]:
        getattr(lib, name).restype = i
    return lib


lib = _load()


def _p(a):
    return None if a is None else a.ctypes.data


def _chk(rc, name):
    if rc != 0:
        raise RuntimeError(f"{name} failed ({rc})")


def census(img, cw=9, ch=7):
    img = np.ascontiguousarray(img, np.uint8)
    H, W = img.shape
    out = np.empty((H, W), np.uint64)
    _chk(lib.orc_census(_p(img), W, H, cw, ch, _p(out)), "orc_census")
    return out


def census3x3_ref(img):
    img = np.ascontiguousarray(img, np.uint8)
    H, W = img.shape
    out = np.empty((H, W), np.uint8)
    _chk(lib.orc_census3x3_ref(_p(img), W, H, _p(out)), "orc_census3x3_ref")
    return out


def bm(kind, L, R, D, window=(9, 9), census_win=(1, 1), tol=1, nthreads=1, with_cost=False):
    """Plain-spec matcher.  SAD: L, R uint8.  Census: L, R uint64 descriptors and
    census_win the window that produced them (sets the border)."""
    dt = np.uint8 if kind == COST_SAD else np.uint64
    L = np.ascontiguousarray(L, dt)
    R = np.ascontiguousarray(R, dt)
    H, W = L.shape
    dL = np.empty((H, W), np.int16)
    dR = np.empty((H, W), np.int16)
    out = np.empty((H, W), np.int16)
    cost = np.empty((H, W), np.int32) if with_cost else None
    _chk(lib.orc_bm(kind, _p(L), _p(R), W, H, D, window[0], window[1], census_win[0], census_win[1], tol,
                    _p(dL), _p(dR), _p(out), _p(cost), nthreads), "orc_bm")
    return (out, dL, dR, cost) if with_cost else (out, dL, dR)


def bm_bins(kind, L, R, D, bins, window=(9, 9), census_win=(1, 1), tol=1, nthreads=1):
    """orc_bm restricted per left pixel to the 8-disparity bins set in bins (uint32)."""
    L = np.ascontiguousarray(L)
    R = np.ascontiguousarray(R)
    H, W = L.shape
    bins = np.ascontiguousarray(bins, np.uint32)
    dL = np.empty((H, W), np.int16)
    dR = np.empty((H, W), np.int16)
    out = np.empty((H, W), np.int16)
    _chk(lib.orc_bm_bins(kind, _p(L), _p(R), W, H, D, window[0], window[1], census_win[0], census_win[1], tol,
                         _p(bins), _p(dL), _p(dR), _p(out), nthreads), "orc_bm_bins")
    return out, dL, dR


def bm_brute(kind, L, R, D, window=(9, 9), census_win=(1, 1), tol=1, with_volume=False):
    dt = np.uint8 if kind == COST_SAD else np.uint64
    L = np.ascontiguousarray(L, dt)
    R = np.ascontiguousarray(R, dt)
    H, W = L.shape
    dL = np.empty((H, W), np.int16)
    dR = np.empty((H, W), np.int16)
    out = np.empty((H, W), np.int16)
    vol = np.empty((H, W, D), np.int32) if with_volume else None
    _chk(lib.orc_bm_brute(kind, _p(L), _p(R), W, H, D, window[0], window[1], census_win[0], census_win[1], tol,
                          _p(dL), _p(dR), _p(out), _p(vol)), "orc_bm_brute")
    return (out, dL, dR, vol) if with_volume else (out, dL, dR)


def ncc(L, R, D, window=(11, 11), tol=1, nthreads=1, with_volume=False):
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    H, W = L.shape
    dL = np.empty((H, W), np.int16)
    dR = np.empty((H, W), np.int16)
    out = np.empty((H, W), np.int16)
    sc = np.empty((H, W), np.float64)
    vol = np.empty((H, W, D), np.float64) if with_volume else None
    _chk(lib.orc_ncc(_p(L), _p(R), W, H, D, window[0], window[1], tol, _p(dL), _p(dR), _p(out), _p(sc), _p(vol),
                     nthreads), "orc_ncc")
    return (out, dL, dR, sc, vol) if with_volume else (out, dL, dR, sc)


def ncc_bins(L, R, D, bins, window=(11, 11), tol=1, nthreads=1):
    """orc_ncc restricted to the 8-disparity bins set in bins[y, x] (uint32) of each left pixel."""
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    H, W = L.shape
    bins = np.ascontiguousarray(bins, np.uint32)
    assert bins.shape == (H, W)
    dL = np.empty((H, W), np.int16)
    dR = np.empty((H, W), np.int16)
    out = np.empty((H, W), np.int16)
    sc = np.empty((H, W), np.float64)
    _chk(lib.orc_ncc_bins(_p(L), _p(R), W, H, D, window[0], window[1], tol, _p(bins), _p(dL), _p(dR), _p(out), _p(sc),
                          nthreads), "orc_ncc_bins")
    return out, dL, dR, sc


def aw(mode, L, R, D, sub=5, tol=1, nthreads=1, with_volume=False):
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    H, W = L.shape
    dL = np.empty((H, W), np.int16)
    dR = np.empty((H, W), np.int16)
    out = np.empty((H, W), np.int16)
    val = np.empty((H, W), np.int32)
    vol = np.empty((H, W, D), np.int32) if with_volume else None
    _chk(lib.orc_aw(mode, _p(L), _p(R), W, H, D, sub, tol, _p(dL), _p(dR), _p(out), _p(val), _p(vol), nthreads),
         "orc_aw")
    return (out, dL, dR, val, vol) if with_volume else (out, dL, dR, val)


def aw_bins(mode, L, R, D, bins, sub=5, tol=1, nthreads=1):
    """orc_aw restricted to the 8-disparity bins set in bins[y, x] (uint32) of each left pixel."""
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    H, W = L.shape
    bins = np.ascontiguousarray(bins, np.uint32)
    assert bins.shape == (H, W)
    dL = np.empty((H, W), np.int16)
    dR = np.empty((H, W), np.int16)
    out = np.empty((H, W), np.int16)
    val = np.empty((H, W), np.int32)
    _chk(lib.orc_aw_bins(mode, _p(L), _p(R), W, H, D, sub, tol, _p(bins), _p(dL), _p(dR), _p(out), _p(val), nthreads),
         "orc_aw_bins")
    return out, dL, dR, val


def crosscheck_plain(dL, dR, tol=1):
    dL = np.ascontiguousarray(dL, np.int16)
    dR = np.ascontiguousarray(dR, np.int16)
    H, W = dL.shape
    out = np.empty((H, W), np.int16)
    _chk(lib.orc_crosscheck_plain(_p(dL), _p(dR), W, H, tol, _p(out)), "orc_crosscheck_plain")
    return out


def crosscheck_ref(lr, rl, thresh=1, ltor=1, offset=0):
    lr = np.ascontiguousarray(lr, np.uint16)
    rl = np.ascontiguousarray(rl, np.uint16)
    H, W = lr.shape[:2]
    out = np.empty_like(lr)
    mn = np.zeros(1, np.uint16)
    mx = np.zeros(1, np.uint16)
    _chk(lib.orc_crosscheck_ref(_p(lr), _p(rl), W, H, thresh, ltor, offset, _p(out), _p(mn), _p(mx)),
         "orc_crosscheck_ref")
    return out, int(mn[0]), int(mx[0])


# ---- learned proposal network (FP64 restatement) --------------------------------

lib.orc_conv3x3_fwd.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                C.c_int, C.c_void_p]
lib.orc_conv3x3_bwd.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
lib.orc_conv3x3_fwd.restype = C.c_int
lib.orc_conv3x3_bwd.restype = C.c_int


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest, ties to even) and back to float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out


def conv3x3(x, w, b=None, dil=1):
    """x [N,H,W,Cin] float32, w [Cout,3,3,Cin] -> y [N,H,W,Cout] float32 (FP64 accumulation)."""
    x = np.ascontiguousarray(x, np.float32)
    w = np.ascontiguousarray(w, np.float32)
    N, H, W, Ci = x.shape
    Co = w.shape[0]
    assert w.shape == (Co, 3, 3, Ci)
    bb = None if b is None else np.ascontiguousarray(b, np.float32)
    y = np.empty((N, H, W, Co), np.float32)
    _chk(lib.orc_conv3x3_fwd(_p(x), N, H, W, Ci, _p(w), _p(bb), Co, dil, _p(y)), "orc_conv3x3_fwd")
    return y


def conv3x3_bwd(x, w, dy, dil=1, want_dx=True):
    """Gradients (dx float32 or None, dw float64 [Cout,3,3,Cin], db float64 [Cout])."""
    x = np.ascontiguousarray(x, np.float32)
    w = np.ascontiguousarray(w, np.float32)
    dy = np.ascontiguousarray(dy, np.float32)
    N, H, W, Ci = x.shape
    Co = w.shape[0]
    dx = np.empty_like(x) if want_dx else None
    dw = np.empty((Co, 3, 3, Ci), np.float64)
    db = np.empty(Co, np.float64)
    _chk(lib.orc_conv3x3_bwd(_p(x), N, H, W, Ci, _p(w), _p(dy), Co, dil, _p(dx), _p(dw), _p(db)), "orc_conv3x3_bwd")
    return dx, dw, db


def disp_mipmaps(disp, nlevels):
    """CreateDispMipMaps: float levels of a disparity map.  disp: (H, W, 4) uint16
    EDisparityId pixels (Unknown marker 0) or (H, W) int16 plain map (marker -1)."""
    disp = np.ascontiguousarray(disp)
    H, W = disp.shape[:2]
    lvl = np.empty((H, W), np.float32)
    if disp.ndim == 3:
        _chk(lib.orc_mip_level0_ref(_p(np.ascontiguousarray(disp, np.uint16)), W, H, _p(lvl)), "orc_mip_level0_ref")
        unknown = 0.0
    else:
        _chk(lib.orc_mip_level0_plain(_p(np.ascontiguousarray(disp, np.int16)), W, H, _p(lvl)), "orc_mip_level0_plain")
        unknown = -1.0
    out = [lvl]
    for _ in range(nlevels - 1):
        h, w = out[-1].shape
        nxt = np.empty(((h - 1) // 2, (w - 1) // 2), np.float32)
        _chk(lib.orc_mip_next(_p(out[-1]), w, h, unknown, _p(nxt)), "orc_mip_next")
        out.append(nxt)
    return out


def disp_denoise(disp, P=5, support_ratio=0.5, mean_cutoff=1.0, fill=True):
    """Statistics::DispDenoise: (H, W, 4) uint16 EDisparityId pixels -> (H, W) uint16 ids,
    or (H, W) int16 d (-1 invalid) -> int16."""
    disp = np.ascontiguousarray(disp)
    H, W = disp.shape[:2]
    ref = disp.ndim == 3
    out = np.empty((H, W), np.uint16 if ref else np.int16)
    _chk(lib.orc_disp_denoise(1 if ref else 0, _p(np.ascontiguousarray(disp, np.uint16 if ref else np.int16)), W, H, P,
                              support_ratio, mean_cutoff, 1 if fill else 0, _p(out)), "orc_disp_denoise")
    return out


def bm1(L, R, P=9, sim_type=0, dmin=0, dmax=64, offset=0, ltor=True, min_delta=0.0, sim_thresh=0.0,
        grad_thresh=0.0, init=None, fill=0):
    """BlockMatcher1 emulated block by block: (H, W, 4) uint16 pixels (EDisparityId,
    similarity max, prev delta, next delta).  Pixels the reference never writes keep
    `fill` (every channel)."""
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    H, W = L.shape
    out = np.full((H, W, 4), fill, np.uint16)
    if init is not None:
        init = np.ascontiguousarray(init, np.uint16)
        iH, iW = init.shape[:2]
        ip = _p(init)
    else:
        iH = iW = 0
        ip = None
    _chk(lib.orc_bm1(_p(L), _p(R), W, H, P, sim_type, dmin, dmax, offset, 1 if ltor else 0, min_delta, sim_thresh,
                     grad_thresh, ip, iW, iH, _p(out)), "orc_bm1")
    return out


def _init_args(init):
    if init is None:
        return None, 0, 0, None
    init = np.ascontiguousarray(init, np.uint16)
    iH, iW = init.shape[:2]
    return init, iW, iH, _p(init)


def ncc_ref(L, R, P=11, dmin=0, dmax=64, offset=0, ltor=True, min_delta=0, ncc_thresh=0.0, grad_thresh=0.0,
            init=None, fill=0):
    """The reference's BlockMatcherNCC emulated block by block: (H, W, 4) uint16 pixels
    (EDisparityId, byte NCC maximum, prev delta, next delta).  Pixels the reference never
    writes keep `fill` (every channel)."""
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    H, W = L.shape
    out = np.full((H, W, 4), fill, np.uint16)
    init, iW, iH, ip = _init_args(init)
    _chk(lib.orc_ncc_ref(_p(L), _p(R), W, H, P, dmin, dmax, offset, 1 if ltor else 0, int(min_delta), ncc_thresh,
                         grad_thresh, ip, iW, iH, _p(out)), "orc_ncc_ref")
    return out


def aw1_ref(L, R, sub=5, sim_type=0, dmin=0, dmax=64, offset=0, ltor=True, min_delta=0.0, sim_thresh=0.0,
            grad_thresh=0.0, init=None, fill=0):
    """The reference's BlockMatcherAW1 emulated block by block (sub-patch size 5: Patch_15x15,
    3: Patch_9x9): (H, W, 4) uint16 pixels as bm1."""
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    H, W = L.shape
    out = np.full((H, W, 4), fill, np.uint16)
    init, iW, iH, ip = _init_args(init)
    _chk(lib.orc_aw1_ref(_p(L), _p(R), W, H, sub, sim_type, dmin, dmax, offset, 1 if ltor else 0, min_delta,
                         sim_thresh, grad_thresh, ip, iW, iH, _p(out)), "orc_aw1_ref")
    return out


RECTIFY_PARAMS = ("src_fx", "src_fy", "src_cx", "src_cy", "k1", "k2", "p1", "p2", "k3", "rect_fx", "rect_fy",
                  "rect_cx", "rect_cy")


def rectify_params(src=(1.0, 1.0, 0.0, 0.0), dist=(0.0, 0.0, 0.0, 0.0, 0.0), rect=None, R=None):
    """The 22 floats of the rectification: source fx, fy, cx, cy; k1, k2, p1, p2, k3;
    rectified fx, fy, cx, cy (default: the source's); R (rectified to source)."""
    rect = src if rect is None else rect
    R = np.eye(3) if R is None else np.asarray(R)
    return np.array(list(src) + list(dist) + list(rect) + list(np.asarray(R, np.float64).ravel()), np.float32)


def rectify(src, params, W=None, H=None):
    """RectifyImage: uint8 source (sH, sW) -> uint8 (H, W)."""
    src = np.ascontiguousarray(src, np.uint8)
    sH, sW = src.shape
    W = sW if W is None else W
    H = sH if H is None else H
    params = np.ascontiguousarray(params, np.float32)
    assert params.size == 22
    out = np.empty((H, W), np.uint8)
    _chk(lib.orc_rectify(_p(src), sW, sH, _p(params), W, H, _p(out)), "orc_rectify")
    return out


DEFAULT_COLORS = np.array([[10, 100, 180, 255], [0, 150, 50, 255], [0, 200, 200, 255], [200, 0, 200, 255],
                           [200, 50, 50, 255], [240, 50, 150, 255], [100, 10, 10, 255]], np.uint8)


def color_disparity(disp, style, dmin=1, dmax=0xFFEF, inf_offset=0, colors=None):
    """ColorDisparityImage: disp (H, W, 4) uint16, (H, W) uint16 or (H, W) float32 ids;
    styles 0 Analysis / 2 Color -> (H, W, 4) uint8, 1 Gray / 3 DispAbs / 4 DispRel -> (H, W) uint16."""
    disp = np.ascontiguousarray(disp)
    H, W = disp.shape[:2]
    fmt = 0 if disp.ndim == 3 else (2 if disp.dtype == np.float32 else 1)
    colors = np.ascontiguousarray(DEFAULT_COLORS if colors is None else colors, np.uint8)
    out = np.empty((H, W, 4), np.uint8) if style in (0, 2) else np.empty((H, W), np.uint16)
    _chk(lib.orc_color_disparity(_p(disp), fmt, W, H, style, dmin, dmax, inf_offset, _p(colors), _p(out)),
         "orc_color_disparity")
    return out
