"""ctypes binding of the gfx950 stereo engine's C ABI (include/clutec_stereo.h).

The engine is the HIP library lib/libclutec_stereo.so next to this file; there is
no Python or CPU compute path.  Importing this module on a machine without the
built library raises immediately, and every compute call goes through the HIP
kernels.

    import clutec_stereo as cs
    ctx = cs.Context(0)
    m = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_CENSUS, window=(9, 9), census=(9, 7), num_disp=256)
    disp = m.run(left_u8, right_u8)          # numpy in, numpy int16 out
"""
import ctypes as C
import json
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libclutec_stereo.so")

OK = 0
ERROR_INVALID_ARGUMENT = 1
ERROR_INVALID_CONFIG = 2
ERROR_UNSUPPORTED_PIXEL_TYPE = 3
ERROR_FORMAT_MISMATCH = 4
ERROR_INVALID_DEVICE = 5
ERROR_INVALID_FORMAT = 6
ERROR_HIP = 7
ERROR_OUT_OF_MEMORY = 8
ERROR_NOT_READY = 9

PIXEL_LUM_U8 = 1
PIXEL_LUM_U16 = 2
PIXEL_LUM_I16 = 3
PIXEL_LUM_U32 = 4
PIXEL_LUM_U64 = 5
PIXEL_LUM_F32 = 6
PIXEL_RGBA_U16 = 7
PIXEL_RGBA_U8 = 8

ALGO_SAD = 0
ALGO_CENSUS = 1
ALGO_NCC = 2
ALGO_AW1 = 3
ALGO_AW2 = 4

_NP_OF = {
    PIXEL_LUM_U8: (np.uint8, 1),
    PIXEL_LUM_U16: (np.uint16, 1),
    PIXEL_LUM_I16: (np.int16, 1),
    PIXEL_LUM_U32: (np.uint32, 1),
    PIXEL_LUM_U64: (np.uint64, 1),
    PIXEL_LUM_F32: (np.float32, 1),
    PIXEL_RGBA_U16: (np.uint16, 4),
    PIXEL_RGBA_U8: (np.uint8, 4),
}


class ImageFormat(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("pixel_type", C.c_int32)]


class Image(C.Structure):
    _fields_ = [("data", C.c_void_p), ("pitch", C.c_size_t), ("format", ImageFormat)]


class CensusParams(C.Structure):
    _fields_ = [("window_w", C.c_int32), ("window_h", C.c_int32), ("reference_3x3", C.c_int32)]


class MatcherParams(C.Structure):
    _fields_ = [("algorithm", C.c_int32), ("window_w", C.c_int32), ("window_h", C.c_int32),
                ("census_w", C.c_int32), ("census_h", C.c_int32), ("num_disparities", C.c_int32),
                ("lr_tolerance", C.c_int32), ("use_disp_input", C.c_int32), ("reserved", C.c_int32 * 4)]


class CrossCheckParams(C.Structure):
    _fields_ = [("disp_delta_thresh", C.c_int32), ("left_to_right", C.c_int32),
                ("disp_offset", C.c_int32), ("reserved", C.c_int32)]


class Bm1Params(C.Structure):
    _fields_ = [("patch", C.c_int32), ("sim_type", C.c_int32), ("disp_min", C.c_int32), ("disp_max", C.c_int32),
                ("disp_offset", C.c_int32), ("min_delta_thresh", C.c_float), ("sim_thresh", C.c_float),
                ("grad_thresh", C.c_float), ("use_disp_input", C.c_int32), ("left_to_right", C.c_int32),
                ("reserved", C.c_int32 * 4)]


class NccParams(C.Structure):
    _fields_ = [("patch", C.c_int32), ("disp_min", C.c_int32), ("disp_max", C.c_int32), ("disp_offset", C.c_int32),
                ("min_delta_thresh", C.c_int32), ("ncc_thresh", C.c_float), ("grad_thresh", C.c_float),
                ("use_disp_input", C.c_int32), ("left_to_right", C.c_int32), ("reserved", C.c_int32 * 4)]


class Aw1Params(C.Structure):
    _fields_ = [("patch", C.c_int32), ("sim_type", C.c_int32), ("disp_min", C.c_int32), ("disp_max", C.c_int32),
                ("disp_offset", C.c_int32), ("min_delta_thresh", C.c_float), ("sim_thresh", C.c_float),
                ("grad_thresh", C.c_float), ("use_disp_input", C.c_int32), ("left_to_right", C.c_int32),
                ("reserved", C.c_int32 * 4)]


class RectifyParams(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("src_fx", "src_fy", "src_cx", "src_cy", "k1", "k2", "p1", "p2", "k3",
                                         "rect_fx", "rect_fy", "rect_cx", "rect_cy")] + \
               [("R", C.c_float * 9), ("reserved", C.c_int32 * 4)]


class ColorParams(C.Structure):
    _fields_ = [("disp_min", C.c_uint16), ("disp_max", C.c_uint16), ("disp_inf_offset", C.c_uint16),
                ("pad", C.c_uint16), ("colors", (C.c_uint8 * 4) * 7), ("reserved", C.c_int32 * 4)]


COLOR_ANALYSIS, COLOR_GRAY, COLOR_COLOR, COLOR_DISP_ABS, COLOR_DISP_REL = range(5)
# ColorDisparityImage's default palette: Unknown, CannotEvaluate, Saturated, NotSpecific,
# NotFound, NotUnique, Inconsistent
DEFAULT_COLORS = ((10, 100, 180, 255), (0, 150, 50, 255), (0, 200, 200, 255), (200, 0, 200, 255),
                  (200, 50, 50, 255), (240, 50, 150, 255), (100, 10, 10, 255))


class DenoiseParams(C.Structure):
    _fields_ = [("patch", C.c_int32), ("support_ratio", C.c_float), ("mean_cutoff", C.c_float), ("fill", C.c_int32),
                ("reserved", C.c_int32 * 4)]


class BatchTuning(C.Structure):
    _fields_ = [("chunk", C.c_int32), ("grid", C.c_int32), ("order", C.c_int32), ("use_graph", C.c_int32)]

    def as_dict(self):
        return {"chunk": self.chunk, "grid": self.grid, "order": self.order, "use_graph": self.use_graph}


class LearnedParams(C.Structure):
    _fields_ = [("matcher", MatcherParams), ("coarse_disparities", C.c_int32), ("band", C.c_int32),
                ("chunk", C.c_int32), ("logit_thresh", C.c_float), ("centre_residual", C.c_int32),
                ("graph_replay", C.c_int32), ("reserved", C.c_int32 * 2)]


CNN_PARAMS = 7554
# (cin, cout, dilation) of the proposal network's layers
CNN_LAYERS = [(2, 16, 1), (16, 16, 2), (16, 16, 4), (16, 16, 2), (16, 2, 1)]


# every exported entry point: (name, restype, argtypes)
_P = C.c_void_p
_PROTOS = [
    ("stereo_abi_version", C.c_int, []),
    ("stereo_create", C.c_int, [C.POINTER(_P), C.c_int]),
    ("stereo_destroy", None, [_P]),
    ("stereo_last_error", C.c_char_p, [_P]),
    ("stereo_status_string", C.c_char_p, [C.c_int]),
    ("stereo_set_timing", C.c_int, [_P, C.c_int]),
    ("stereo_synchronize", C.c_int, [_P, _P]),
    ("stereo_pixel_size", C.c_size_t, [C.c_int32]),
    ("stereo_image_alloc", C.c_int, [_P, C.POINTER(Image), C.c_int32, C.c_int32, C.c_int32]),
    ("stereo_image_free", C.c_int, [_P, C.POINTER(Image)]),
    ("stereo_image_upload", C.c_int, [_P, C.POINTER(Image), _P, C.c_size_t, _P]),
    ("stereo_image_download", C.c_int, [_P, C.POINTER(Image), _P, C.c_size_t, _P]),
    ("stereo_device_alloc", C.c_int, [_P, C.POINTER(_P), C.c_size_t]),
    ("stereo_device_free", C.c_int, [_P, _P]),
    ("stereo_memcpy", C.c_int, [_P, _P, _P, C.c_size_t, C.c_int32]),
    ("stereo_driver_destroy", C.c_int, [_P]),
    ("stereo_driver_name", C.c_char_p, [_P]),
    ("stereo_driver_last_time_ms", C.c_int, [_P, C.POINTER(C.c_double)]),
    ("stereo_census_configure", C.c_int, [_P, C.POINTER(_P), C.POINTER(ImageFormat), C.POINTER(CensusParams)]),
    ("stereo_census_process", C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), _P]),
    ("stereo_matcher_validate", C.c_int, [C.POINTER(ImageFormat), C.POINTER(MatcherParams), C.c_char_p, C.c_size_t]),
    ("stereo_matcher_configure", C.c_int, [_P, C.POINTER(_P), C.POINTER(ImageFormat), C.POINTER(MatcherParams)]),
    ("stereo_matcher_process_noalloc", C.c_int, [_P, C.POINTER(MatcherParams), C.POINTER(Image), C.POINTER(Image),
                                                 C.POINTER(Image), C.POINTER(Image), C.POINTER(Image), _P]),
    ("stereo_matcher_process", C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), C.POINTER(Image),
                                          C.POINTER(Image), C.POINTER(Image), _P]),
    ("stereo_matcher_process_init", C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), C.POINTER(Image),
                                               C.POINTER(Image), C.POINTER(Image), C.POINTER(Image), _P]),
    ("stereo_disp_mipmaps", C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), C.c_int32, _P]),
    ("stereo_denoise_configure", C.c_int, [_P, C.POINTER(_P), C.POINTER(ImageFormat), C.POINTER(DenoiseParams)]),
    ("stereo_bm1_configure", C.c_int, [_P, C.POINTER(_P), C.POINTER(ImageFormat), C.POINTER(Bm1Params)]),
    ("stereo_rectify_configure", C.c_int, [_P, C.POINTER(_P), C.POINTER(ImageFormat), C.POINTER(RectifyParams)]),
    ("stereo_rectify_process", C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), _P]),
    ("stereo_color_configure", C.c_int, [_P, C.POINTER(_P), C.POINTER(ImageFormat), C.POINTER(ColorParams)]),
    ("stereo_color_process", C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), C.c_int32, _P]),
    ("stereo_bm1_process", C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), C.POINTER(Image), C.POINTER(Image), _P]),
    ("stereo_ncc_configure", C.c_int, [_P, C.POINTER(_P), C.POINTER(ImageFormat), C.POINTER(NccParams)]),
    ("stereo_ncc_process", C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), C.POINTER(Image), C.POINTER(Image), _P]),
    ("stereo_aw1_configure", C.c_int, [_P, C.POINTER(_P), C.POINTER(ImageFormat), C.POINTER(Aw1Params)]),
    ("stereo_aw1_process", C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), C.POINTER(Image), C.POINTER(Image), _P]),
    ("stereo_denoise_process", C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), _P]),
    ("stereo_matcher_scores", C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), C.POINTER(Image), _P]),
    ("stereo_matcher_stage_times", C.c_int, [_P, C.POINTER(C.c_double), C.c_int32]),
    ("stereo_crosscheck_configure", C.c_int, [_P, C.POINTER(_P), C.POINTER(ImageFormat), C.POINTER(CrossCheckParams)]),
    ("stereo_crosscheck_process", C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), C.POINTER(Image), _P]),
    ("stereo_crosscheck_range", C.c_int, [_P, C.POINTER(C.c_uint16), C.POINTER(C.c_uint16)]),
    ("stereo_batch_configure", C.c_int, [_P, C.POINTER(_P), C.c_int32, C.POINTER(ImageFormat), C.POINTER(C.c_int32),
                                          C.POINTER(MatcherParams), C.POINTER(BatchTuning)]),
    ("stereo_batch_process", C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), C.POINTER(Image), C.POINTER(Image),
                                        C.POINTER(Image), _P]),
    ("stereo_batch_autotune", C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), C.POINTER(BatchTuning)]),
    ("stereo_batch_get_tuning", C.c_int, [_P, C.POINTER(BatchTuning)]),
    ("stereo_batch_tuning_table", C.c_int, [_P, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("stereo_batch_apply_tuning_table", C.c_int, [_P, C.c_char_p]),
    ("stereo_cnn_create", C.c_int, [_P, C.POINTER(_P), C.c_int32, C.c_int32, C.c_int32]),
    ("stereo_cnn_destroy", None, [_P]),
    ("stereo_cnn_set_params", C.c_int, [_P, _P, C.c_size_t]),
    ("stereo_cnn_get_params", C.c_int, [_P, _P, C.c_size_t]),
    ("stereo_cnn_get_grads", C.c_int, [_P, _P, C.c_size_t]),
    ("stereo_cnn_state_size", C.c_size_t, []),
    ("stereo_cnn_save_state", C.c_int, [_P, _P, C.c_size_t]),
    ("stereo_cnn_load_state", C.c_int, [_P, _P, C.c_size_t]),
    ("stereo_cnn_set_input", C.c_int, [_P, C.c_int32, C.POINTER(Image), C.POINTER(Image), C.POINTER(Image), _P]),
    ("stereo_cnn_set_labels", C.c_int, [_P, C.c_int32, C.POINTER(Image), _P]),
    ("stereo_cnn_set_conf_labels", C.c_int, [_P, C.c_int32, C.POINTER(Image), _P]),
    ("stereo_cnn_forward", C.c_int, [_P, C.c_int32, _P, _P]),
    ("stereo_cnn_train_step", C.c_int, [_P, C.c_int32, C.c_float, _P, _P]),
    ("stereo_conv3x3_bf16", C.c_int, [_P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, _P, C.c_int32,
                                       C.c_int32, C.c_int32, _P, _P]),
    ("stereo_conv3x3_bwd_bf16", C.c_int, [_P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, _P, C.c_int32,
                                           C.c_int32, _P, _P, _P]),
    ("stereo_learned_configure", C.c_int, [_P, C.POINTER(_P), C.POINTER(ImageFormat), C.POINTER(LearnedParams), _P,
                                            C.c_size_t]),
    ("stereo_learned_process", C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), C.POINTER(Image), C.POINTER(Image),
                                          C.POINTER(Image), _P]),
    ("stereo_learned_stats", C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_double), C.c_int32]),
    ("stereo_learned_tile_ranges", C.c_int, [_P, C.POINTER(C.c_int32), C.c_int32]),
    ("stereo_learned_tile_masks", C.c_int, [_P, C.POINTER(C.c_uint32), C.c_int32]),
    ("stereo_valu_peak", C.c_int, [_P, C.c_int32, C.POINTER(C.c_double)]),
    ("stereo_bf16_rounding_check", C.c_int, [_P, C.POINTER(C.c_uint64)]),
    ("stereo_synth_pair", C.c_int, [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    _P, _P, _P]),
]

EXPORTED = [p[0] for p in _PROTOS]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run module.clutec.cuda_amd/build.py "
                          "(the engine has no fallback path)")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in _PROTOS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class StereoError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"[{status}] {message}")
        self.status = status


def _check(st, ctx=None):
    if st != OK:
        msg = lib.stereo_last_error(ctx.handle if ctx is not None else None)
        raise StereoError(st, (msg or b"").decode() or lib.stereo_status_string(st).decode())


def synth_pair(seed, W, H, kind=0, nregions=8, dmax=63, dbg=8, with_gt=False):
    """Seeded synthetic pair (host generator of the engine)."""
    L = np.empty((H, W), np.uint8)
    R = np.empty((H, W), np.uint8)
    gt = np.empty((H, W), np.int16) if with_gt else None
    rc = lib.stereo_synth_pair(seed, W, H, kind, nregions, dmax, dbg, L.ctypes.data, R.ctypes.data,
                               gt.ctypes.data if with_gt else None)
    if rc != 0:
        raise ValueError(f"stereo_synth_pair failed ({rc})")
    return (L, R, gt) if with_gt else (L, R)


def validate_matcher(W, H, params):
    fmt = ImageFormat(W, H, PIXEL_LUM_U8)
    buf = C.create_string_buffer(256)
    st = lib.stereo_matcher_validate(C.byref(fmt), C.byref(params), buf, 256)
    return st, buf.value.decode()


def matcher_params(algorithm, window=(9, 9), census=(9, 7), num_disp=64, tol=1, use_disp_input=False):
    p = MatcherParams()
    p.algorithm = algorithm
    p.window_w, p.window_h = window
    p.census_w, p.census_h = census
    p.num_disparities = num_disp
    p.lr_tolerance = tol
    p.use_disp_input = 1 if use_disp_input else 0
    return p


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        st = lib.stereo_create(C.byref(h), device)
        if st != OK:
            raise StereoError(st, lib.stereo_last_error(None).decode())
        self.handle = h

    def close(self):
        if self.handle:
            lib.stereo_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timing(self, on=True):
        _check(lib.stereo_set_timing(self.handle, 1 if on else 0), self)

    def synchronize(self, stream=None):
        _check(lib.stereo_synchronize(self.handle, stream), self)

    def valu_peak(self, op=0):
        """Measured INT32 VALU rate in tera lane-ops/s (op 0: v_add_u32, 1: v_bcnt_u32_b32)."""
        v = C.c_double()
        _check(lib.stereo_valu_peak(self.handle, op, C.byref(v)), self)
        return v.value


class DeviceImage:
    """Owning pitched device image (CDeviceImage)."""

    def __init__(self, ctx, W, H, pixel_type):
        self.ctx = ctx
        self.img = Image()
        _check(lib.stereo_image_alloc(ctx.handle, C.byref(self.img), W, H, pixel_type), ctx)

    @property
    def ref(self):
        return C.byref(self.img)

    def free(self):
        if self.img.data:
            lib.stereo_image_free(self.ctx.handle, C.byref(self.img))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def upload(self, arr, stream=None):
        dt, ch = _NP_OF[self.img.format.pixel_type]
        a = np.ascontiguousarray(arr, dtype=dt)
        assert a.shape[0] == self.img.format.height and a.shape[1] == self.img.format.width
        _check(lib.stereo_image_upload(self.ctx.handle, C.byref(self.img), a.ctypes.data, a.strides[0], stream), self.ctx)

    def download(self, stream=None):
        dt, ch = _NP_OF[self.img.format.pixel_type]
        shape = (self.img.format.height, self.img.format.width) + ((ch,) if ch > 1 else ())
        a = np.empty(shape, dt)
        _check(lib.stereo_image_download(self.ctx.handle, C.byref(self.img), a.ctypes.data, a.strides[0], stream), self.ctx)
        return a

    @classmethod
    def from_array(cls, ctx, arr, pixel_type):
        img = cls(ctx, arr.shape[1], arr.shape[0], pixel_type)
        img.upload(arr)
        return img


class DeviceBuffer:
    """Raw device allocation (stereo_device_alloc) holding a numpy array's bytes."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        _check(lib.stereo_device_alloc(ctx.handle, C.byref(self.ptr), self.nbytes), ctx)

    @classmethod
    def from_array(cls, ctx, arr):
        a = np.ascontiguousarray(arr)
        b = cls(ctx, a.nbytes)
        _check(lib.stereo_memcpy(ctx.handle, b.ptr, a.ctypes.data, a.nbytes, 0), ctx)
        return b

    def download(self, dtype, shape):
        a = np.empty(shape, dtype)
        assert a.nbytes <= self.nbytes
        _check(lib.stereo_memcpy(self.ctx.handle, a.ctypes.data, self.ptr, a.nbytes, 1), self.ctx)
        return a

    def free(self):
        if self.ptr:
            lib.stereo_device_free(self.ctx.handle, self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _Driver:
    def __init__(self, ctx):
        self.ctx = ctx
        self.handle = C.c_void_p()

    def close(self):
        if self.handle:
            lib.stereo_driver_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_time_ms(self):
        v = C.c_double()
        _check(lib.stereo_driver_last_time_ms(self.handle, C.byref(v)), self.ctx)
        return v.value

    @property
    def name(self):
        return lib.stereo_driver_name(self.handle).decode()


class Census(_Driver):
    def __init__(self, ctx, W, H, window=(9, 7), reference_3x3=False):
        super().__init__(ctx)
        self.W, self.H = W, H
        self.ref3 = bool(reference_3x3)
        fmt = ImageFormat(W, H, PIXEL_LUM_U8)
        p = CensusParams(window[0], window[1], 1 if reference_3x3 else 0)
        _check(lib.stereo_census_configure(ctx.handle, C.byref(self.handle), C.byref(fmt), C.byref(p)), ctx)

    def process(self, trg, src, stream=None):
        _check(lib.stereo_census_process(self.handle, trg.ref, src.ref, stream), self.ctx)

    def run(self, img):
        src = DeviceImage.from_array(self.ctx, img, PIXEL_LUM_U8)
        trg = DeviceImage(self.ctx, self.W, self.H, PIXEL_LUM_U8 if self.ref3 else PIXEL_LUM_U64)
        self.process(trg, src)
        return trg.download()


class Matcher(_Driver):
    def __init__(self, ctx, W, H, algorithm=ALGO_CENSUS, window=(9, 9), census=(9, 7), num_disp=64, tol=1,
                 use_disp_input=False):
        super().__init__(ctx)
        self.W, self.H = W, H
        fmt = ImageFormat(W, H, PIXEL_LUM_U8)
        self.params = matcher_params(algorithm, window, census, num_disp, tol, use_disp_input)
        _check(lib.stereo_matcher_configure(ctx.handle, C.byref(self.handle), C.byref(fmt), C.byref(self.params)), ctx)

    def process(self, disp, disp_lr, disp_rl, left, right, stream=None):
        _check(lib.stereo_matcher_process(self.handle, disp.ref if disp else None, disp_lr.ref if disp_lr else None,
                                          disp_rl.ref if disp_rl else None, left.ref, right.ref, stream), self.ctx)

    def stage_times_ms(self):
        v = (C.c_double * 3)()
        _check(lib.stereo_matcher_stage_times(self.handle, v, 3), self.ctx)
        return list(v)

    def process_init(self, disp, disp_lr, disp_rl, left, right, init, stream=None):
        """Band search around the initial disparity map `init` (stereo_matcher_process_init)."""
        _check(lib.stereo_matcher_process_init(self.handle, disp.ref if disp else None,
                                               disp_lr.ref if disp_lr else None, disp_rl.ref if disp_rl else None,
                                               left.ref, right.ref, init.ref, stream), self.ctx)

    def run_init(self, L, R, init, maps=False, scores=False):
        """numpy in, numpy out: the band search around `init` (int16 map, -1 invalid, or
        (h, w, 4) uint16 EDisparityId pixels); for NCC also the winning scores."""
        dl = DeviceImage.from_array(self.ctx, L, PIXEL_LUM_U8)
        dr = DeviceImage.from_array(self.ctx, R, PIXEL_LUM_U8)
        init = np.asarray(init)
        di = DeviceImage.from_array(self.ctx, init, PIXEL_RGBA_U16 if init.ndim == 3 else PIXEL_LUM_I16)
        out = DeviceImage(self.ctx, self.W, self.H, PIXEL_LUM_I16)
        lr = DeviceImage(self.ctx, self.W, self.H, PIXEL_LUM_I16) if maps else None
        rl = DeviceImage(self.ctx, self.W, self.H, PIXEL_LUM_I16) if maps else None
        self.process_init(out, lr, rl, dl, dr, di)
        res = [out.download()]
        if maps:
            res += [lr.download(), rl.download()]
        if scores:
            sc = DeviceImage(self.ctx, self.W, self.H, PIXEL_LUM_F32)
            _check(lib.stereo_matcher_scores(self.handle, sc.ref, dl.ref, dr.ref, None), self.ctx)
            res.append(sc.download())
        return res[0] if len(res) == 1 else tuple(res)

    def run(self, L, R, maps=False, dispid=False, scores=False):
        """numpy in, numpy out: the checked map (and the two one-sided maps, and for
        NCC the winning scores)."""
        dl = DeviceImage.from_array(self.ctx, L, PIXEL_LUM_U8)
        dr = DeviceImage.from_array(self.ctx, R, PIXEL_LUM_U8)
        out = DeviceImage(self.ctx, self.W, self.H, PIXEL_RGBA_U16 if dispid else PIXEL_LUM_I16)
        lr = DeviceImage(self.ctx, self.W, self.H, PIXEL_LUM_I16) if maps else None
        rl = DeviceImage(self.ctx, self.W, self.H, PIXEL_LUM_I16) if maps else None
        self.process(out, lr, rl, dl, dr)
        res = [out.download()]
        if maps:
            res += [lr.download(), rl.download()]
        if scores:
            sc = DeviceImage(self.ctx, self.W, self.H, PIXEL_LUM_F32)
            _check(lib.stereo_matcher_scores(self.handle, sc.ref, dl.ref, dr.ref, None), self.ctx)
            res.append(sc.download())
        return res[0] if len(res) == 1 else tuple(res)


def disp_mipmaps(ctx, disp, nlevels, stream=None):
    """CreateDispMipMaps: LUM_F32 levels of a LUM_I16 or RGBA_U16 disparity image
    (stereo_disp_mipmaps); level i+1 is ((w-1)/2) x ((h-1)/2) of level i."""
    w, h = disp.img.format.width, disp.img.format.height
    levels = []
    for i in range(nlevels):
        if i:
            w, h = (w - 1) // 2, (h - 1) // 2
        # a level that does not fit is left to the library to reject
        levels.append(DeviceImage(ctx, max(1, w), max(1, h), PIXEL_LUM_F32))
    arr = (Image * nlevels)(*[lv.img for lv in levels])
    _check(lib.stereo_disp_mipmaps(ctx.handle, disp.ref, arr, nlevels, stream), ctx)
    return levels


class BlockMatcher1(_Driver):
    """BlockMatcher1 with the reference's block-tiled results (stereo_bm1_*): an
    (h, w, 4) uint16 map of EDisparityId and three similarity channels."""

    def __init__(self, ctx, W, H, patch=9, sim_type=0, disp_min=0, disp_max=64, disp_offset=0,
                 min_delta_thresh=0.0, sim_thresh=0.0, grad_thresh=0.0, use_disp_input=False, left_to_right=True):
        super().__init__(ctx)
        self.W, self.H = W, H
        p = Bm1Params(patch, sim_type, disp_min, disp_max, disp_offset, min_delta_thresh, sim_thresh, grad_thresh,
                      1 if use_disp_input else 0, 1 if left_to_right else 0)
        fmt = ImageFormat(W, H, PIXEL_LUM_U8)
        _check(lib.stereo_bm1_configure(ctx.handle, C.byref(self.handle), C.byref(fmt), C.byref(p)), ctx)

    def process(self, disp, left, right, init=None, stream=None):
        _check(lib.stereo_bm1_process(self.handle, disp.ref, left.ref, right.ref, init.ref if init else None, stream),
               self.ctx)

    def run(self, L, R, init=None, fill=0):
        """Host images in, the (H, W, 4) map out; pixels the matcher leaves are `fill`."""
        dl = DeviceImage.from_array(self.ctx, L, PIXEL_LUM_U8)
        dr = DeviceImage.from_array(self.ctx, R, PIXEL_LUM_U8)
        out = DeviceImage.from_array(self.ctx, np.full((self.H, self.W, 4), fill, np.uint16), PIXEL_RGBA_U16)
        di = DeviceImage.from_array(self.ctx, init, PIXEL_RGBA_U16) if init is not None else None
        self.process(out, dl, dr, di)
        return out.download()


class _RefMatcher(_Driver):
    def run(self, L, R, init=None, fill=0):
        """Host images in, the (H, W, 4) map out; pixels the matcher leaves are `fill`."""
        dl = DeviceImage.from_array(self.ctx, L, PIXEL_LUM_U8)
        dr = DeviceImage.from_array(self.ctx, R, PIXEL_LUM_U8)
        out = DeviceImage.from_array(self.ctx, np.full((self.H, self.W, 4), fill, np.uint16), PIXEL_RGBA_U16)
        di = DeviceImage.from_array(self.ctx, init, PIXEL_RGBA_U16) if init is not None else None
        self.process(out, dl, dr, di)
        return out.download()


class BlockMatcherNCC(_RefMatcher):
    """The reference's BlockMatcherNCC operator (stereo_ncc_*): block-tiled, one direction,
    byte scores of the reference's float formula; an (h, w, 4) uint16 map of EDisparityId,
    the byte maximum and the two byte drops around it."""

    def __init__(self, ctx, W, H, patch=11, disp_min=0, disp_max=64, disp_offset=0, min_delta_thresh=0,
                 ncc_thresh=0.0, grad_thresh=0.0, use_disp_input=False, left_to_right=True):
        super().__init__(ctx)
        self.W, self.H = W, H
        p = NccParams(patch, disp_min, disp_max, disp_offset, int(min_delta_thresh), ncc_thresh, grad_thresh,
                      1 if use_disp_input else 0, 1 if left_to_right else 0)
        fmt = ImageFormat(W, H, PIXEL_LUM_U8)
        _check(lib.stereo_ncc_configure(ctx.handle, C.byref(self.handle), C.byref(fmt), C.byref(p)), ctx)

    def process(self, disp, left, right, init=None, stream=None):
        _check(lib.stereo_ncc_process(self.handle, disp.ref, left.ref, right.ref, init.ref if init else None, stream),
               self.ctx)


class BlockMatcherAW1(_RefMatcher):
    """The reference's BlockMatcherAW1 operator (stereo_aw1_*): block-tiled, one direction;
    an (h, w, 4) uint16 map of EDisparityId and three similarity channels."""

    def __init__(self, ctx, W, H, patch=15, sim_type=0, disp_min=0, disp_max=64, disp_offset=0,
                 min_delta_thresh=0.0, sim_thresh=0.0, grad_thresh=0.0, use_disp_input=False, left_to_right=True):
        super().__init__(ctx)
        self.W, self.H = W, H
        p = Aw1Params(patch, sim_type, disp_min, disp_max, disp_offset, min_delta_thresh, sim_thresh, grad_thresh,
                      1 if use_disp_input else 0, 1 if left_to_right else 0)
        fmt = ImageFormat(W, H, PIXEL_LUM_U8)
        _check(lib.stereo_aw1_configure(ctx.handle, C.byref(self.handle), C.byref(fmt), C.byref(p)), ctx)

    def process(self, disp, left, right, init=None, stream=None):
        _check(lib.stereo_aw1_process(self.handle, disp.ref, left.ref, right.ref, init.ref if init else None, stream),
               self.ctx)


class Rectify(_Driver):
    """RectifyImage (stereo_rectify_*): params is the 22-value camera model (source fx,
    fy, cx, cy; k1, k2, p1, p2, k3; rectified fx, fy, cx, cy; R row-major)."""

    def __init__(self, ctx, W, H, params):
        super().__init__(ctx)
        self.W, self.H = W, H
        p = RectifyParams(*[float(v) for v in params[:13]])
        for i in range(9):
            p.R[i] = float(params[13 + i])
        fmt = ImageFormat(W, H, PIXEL_LUM_U8)
        _check(lib.stereo_rectify_configure(ctx.handle, C.byref(self.handle), C.byref(fmt), C.byref(p)), ctx)

    def process(self, target, source, stream=None):
        _check(lib.stereo_rectify_process(self.handle, target.ref, source.ref, stream), self.ctx)

    def run(self, src):
        s = DeviceImage.from_array(self.ctx, src, PIXEL_LUM_U8)
        out = DeviceImage(self.ctx, self.W, self.H, PIXEL_LUM_U8)
        self.process(out, s)
        return out.download()


class ColorDisparity(_Driver):
    """ColorDisparityImage (stereo_color_*): disparity images of RGBA_U16, LUM_U16 or
    LUM_F32 ids to RGBA_U8 (Analysis, Color) or LUM_U16 (Gray, DispAbs, DispRel)."""

    def __init__(self, ctx, W, H, pixel_type=PIXEL_RGBA_U16, disp_min=1, disp_max=0xFFEF, inf_offset=0,
                 colors=DEFAULT_COLORS):
        super().__init__(ctx)
        self.W, self.H, self.pixel_type = W, H, pixel_type
        p = ColorParams(disp_min, disp_max, inf_offset, 0)
        for k in range(7):
            for c in range(4):
                p.colors[k][c] = int(colors[k][c])
        fmt = ImageFormat(W, H, pixel_type)
        _check(lib.stereo_color_configure(ctx.handle, C.byref(self.handle), C.byref(fmt), C.byref(p)), ctx)

    def process(self, out, disp, style, stream=None):
        _check(lib.stereo_color_process(self.handle, out.ref, disp.ref, style, stream), self.ctx)

    def run(self, disp, style):
        d = DeviceImage.from_array(self.ctx, disp, self.pixel_type)
        out = DeviceImage(self.ctx, self.W, self.H, PIXEL_RGBA_U8 if style in (0, 2) else PIXEL_LUM_U16)
        self.process(out, d, style)
        return out.download()


class Denoise(_Driver):
    """Statistics::DispDenoise (stereo_denoise_*): ref_format maps are (h, w, 4) uint16
    EDisparityId pixels in and uint16 ids out; plain maps int16 (-1 invalid) in and out."""

    def __init__(self, ctx, W, H, patch=5, support_ratio=0.5, mean_cutoff=1.0, fill=True, ref_format=False):
        super().__init__(ctx)
        self.W, self.H, self.ref_format = W, H, ref_format
        fmt = ImageFormat(W, H, PIXEL_RGBA_U16 if ref_format else PIXEL_LUM_I16)
        p = DenoiseParams(patch, support_ratio, mean_cutoff, 1 if fill else 0)
        _check(lib.stereo_denoise_configure(ctx.handle, C.byref(self.handle), C.byref(fmt), C.byref(p)), ctx)

    def process(self, out, disp, stream=None):
        _check(lib.stereo_denoise_process(self.handle, out.ref, disp.ref, stream), self.ctx)

    def run(self, disp):
        src = DeviceImage.from_array(self.ctx, disp, PIXEL_RGBA_U16 if self.ref_format else PIXEL_LUM_I16)
        out = DeviceImage(self.ctx, self.W, self.H, PIXEL_LUM_U16 if self.ref_format else PIXEL_LUM_I16)
        self.process(out, src)
        return out.download()


class CrossCheck(_Driver):
    def __init__(self, ctx, W, H, thresh=1, left_to_right=True, offset=0):
        super().__init__(ctx)
        self.W, self.H = W, H
        fmt = ImageFormat(W, H, PIXEL_RGBA_U16)
        p = CrossCheckParams(thresh, 1 if left_to_right else 0, offset, 0)
        _check(lib.stereo_crosscheck_configure(ctx.handle, C.byref(self.handle), C.byref(fmt), C.byref(p)), ctx)

    def process(self, result, lr, rl, stream=None):
        _check(lib.stereo_crosscheck_process(self.handle, result.ref, lr.ref, rl.ref, stream), self.ctx)

    def range(self):
        a, b = C.c_uint16(), C.c_uint16()
        _check(lib.stereo_crosscheck_range(self.handle, C.byref(a), C.byref(b)), self.ctx)
        return a.value, b.value

    def run(self, lr, rl):
        dlr = DeviceImage.from_array(self.ctx, lr, PIXEL_RGBA_U16)
        drl = DeviceImage.from_array(self.ctx, rl, PIXEL_RGBA_U16)
        out = DeviceImage(self.ctx, self.W, self.H, PIXEL_RGBA_U16)
        self.process(out, dlr, drl)
        return out.download()


class BatchMatcher(_Driver):
    """Many pairs of uneven size in one persistent launch (stereo_batch_*)."""

    def __init__(self, ctx, sizes, num_disps, algorithm=ALGO_CENSUS, window=(9, 9), census=(9, 7), tol=1,
                 tuning=None):
        super().__init__(ctx)
        n = len(sizes)
        self.sizes = list(sizes)
        fmts = (ImageFormat * n)(*[ImageFormat(w, h, PIXEL_LUM_U8) for (w, h) in sizes])
        ds = (C.c_int32 * n)(*num_disps)
        self.params = matcher_params(algorithm, window, census, 8, tol)
        t = None
        if tuning is not None:
            t = BatchTuning(tuning.get("chunk", 0), tuning.get("grid", 0), tuning.get("order", 0),
                            tuning.get("use_graph", 1))
        _check(lib.stereo_batch_configure(ctx.handle, C.byref(self.handle), n, fmts, ds, C.byref(self.params),
                                          C.byref(t) if t is not None else None), ctx)
        self.dev = None

    def _images(self, pairs):
        n = len(self.sizes)
        self.left = [DeviceImage.from_array(self.ctx, L, PIXEL_LUM_U8) for (L, R) in pairs]
        self.right = [DeviceImage.from_array(self.ctx, R, PIXEL_LUM_U8) for (L, R) in pairs]
        self.outs = [DeviceImage(self.ctx, w, h, PIXEL_LUM_I16) for (w, h) in self.sizes]
        self.lrs = [DeviceImage(self.ctx, w, h, PIXEL_LUM_I16) for (w, h) in self.sizes]
        self.rls = [DeviceImage(self.ctx, w, h, PIXEL_LUM_I16) for (w, h) in self.sizes]
        arr = lambda imgs: (Image * n)(*[i.img for i in imgs])
        self.dev = (arr(self.outs), arr(self.lrs), arr(self.rls), arr(self.left), arr(self.right))

    def upload(self, pairs):
        self._images(pairs)

    def refill(self, pairs):
        """New content in the same device images (the captured graph stays valid)."""
        for (L, R), dl, dr in zip(pairs, self.left, self.right):
            dl.upload(L)
            dr.upload(R)

    def process(self, stream=None):
        o, a, b, l, r = self.dev
        _check(lib.stereo_batch_process(self.handle, o, a, b, l, r, stream), self.ctx)

    def autotune(self, table_path=None):
        """Times the candidate tunings on the uploaded pairs and keeps the fastest; with
        `table_path`, also writes the per-class table (tuning_table()) there as JSON."""
        t = BatchTuning()
        _check(lib.stereo_batch_autotune(self.handle, self.dev[3], self.dev[4], C.byref(t)), self.ctx)
        if table_path is not None:
            with open(table_path, "w") as f:
                json.dump(self.tuning_table(), f, indent=1)
        return t.as_dict()

    def tuning_table(self):
        """{"matcher", "window", "grid", "order", "use_graph", "chunk", "classes": [...]},
        one class per disparity count and tile count with its work-item length."""
        need = C.c_size_t(0)
        _check(lib.stereo_batch_tuning_table(self.handle, None, 0, C.byref(need)), self.ctx)
        buf = C.create_string_buffer(need.value)
        _check(lib.stereo_batch_tuning_table(self.handle, buf, need.value, None), self.ctx)
        return json.loads(buf.value.decode())

    def apply_tuning_table(self, table):
        """A table from tuning_table() or a JSON file written by autotune(table_path)."""
        if isinstance(table, (str, os.PathLike)):
            with open(table) as f:
                table = json.load(f)
        _check(lib.stereo_batch_apply_tuning_table(self.handle, json.dumps(table).encode()), self.ctx)

    def tuning(self):
        t = BatchTuning()
        _check(lib.stereo_batch_get_tuning(self.handle, C.byref(t)), self.ctx)
        return t.as_dict()

    def results(self):
        return [(o.download(), a.download(), b.download()) for o, a, b in zip(self.outs, self.lrs, self.rls)]

    def run(self, pairs):
        self.upload(pairs)
        self.process()
        return self.results()


class ProposalNet:
    """The learned proposal network (stereo_cnn_*): parameters, training, forward."""

    def __init__(self, ctx, W4, H4, max_frames=1):
        self.ctx = ctx
        self.W4, self.H4, self.maxn = W4, H4, max_frames
        self.handle = C.c_void_p()
        _check(lib.stereo_cnn_create(ctx.handle, C.byref(self.handle), W4, H4, max_frames), ctx)

    def close(self):
        if self.handle:
            lib.stereo_cnn_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib.stereo_cnn_set_params(self.handle, p.ctypes.data, p.size), self.ctx)

    def params(self):
        p = np.empty(CNN_PARAMS, np.float32)
        _check(lib.stereo_cnn_get_params(self.handle, p.ctypes.data, p.size), self.ctx)
        return p

    def grads(self):
        g = np.empty(CNN_PARAMS, np.float32)
        _check(lib.stereo_cnn_get_grads(self.handle, g.ctypes.data, g.size), self.ctx)
        return g

    def save_state(self):
        b = np.empty(lib.stereo_cnn_state_size(), np.uint8)
        _check(lib.stereo_cnn_save_state(self.handle, b.ctypes.data, b.size), self.ctx)
        return b

    def load_state(self, blob):
        b = np.ascontiguousarray(blob, np.uint8)
        _check(lib.stereo_cnn_load_state(self.handle, b.ctypes.data, b.size), self.ctx)

    def set_input(self, index, L4, R4, coarse, stream=None):
        _check(lib.stereo_cnn_set_input(self.handle, index, L4.ref, R4.ref, coarse.ref, stream), self.ctx)

    def set_labels(self, index, disp, stream=None):
        _check(lib.stereo_cnn_set_labels(self.handle, index, disp.ref, stream), self.ctx)

    def set_conf_labels(self, index, conf, stream=None):
        """Confidence targets of slot `index`: LUM_U8 quarter-resolution image, 1 where
        the band search agrees with the full range, 0 where not, other values unset."""
        _check(lib.stereo_cnn_set_conf_labels(self.handle, index, conf.ref, stream), self.ctx)

    def forward_ptr(self, n, out_ptr, stream=None):
        _check(lib.stereo_cnn_forward(self.handle, n, out_ptr, stream), self.ctx)

    def train_step(self, n, lr):
        loss = np.zeros(4, np.float32)
        _check(lib.stereo_cnn_train_step(self.handle, n, lr, loss.ctypes.data, None), self.ctx)
        return loss


def cnn_init_params(seed=0):
    """He-normal weights, zero biases, in the public layer-by-layer layout."""
    rng = np.random.default_rng(seed)
    out = []
    for cin, cout, _ in CNN_LAYERS:
        std = np.sqrt(2.0 / (9 * cin))
        out.append((rng.standard_normal((cout, 3, 3, cin)) * std).astype(np.float32).ravel())
        out.append(np.zeros(cout, np.float32))
    p = np.concatenate(out)
    assert p.size == CNN_PARAMS
    return p


class LearnedMatcher(_Driver):
    """Census matcher with the learned band proposal (stereo_learned_*)."""

    def __init__(self, ctx, W, H, cnn_params, window=(9, 9), census=(9, 7), num_disp=256, tol=1, band=8, chunk=64,
                 logit_thresh=0.0, coarse_disp=0, centre_residual=False, graph_replay=False):
        super().__init__(ctx)
        self.W, self.H = W, H
        fmt = ImageFormat(W, H, PIXEL_LUM_U8)
        lp = LearnedParams()
        lp.matcher = matcher_params(ALGO_CENSUS, window, census, num_disp, tol)
        lp.coarse_disparities = coarse_disp
        lp.band = band
        lp.chunk = chunk
        lp.logit_thresh = logit_thresh
        lp.centre_residual = 1 if centre_residual else 0
        lp.graph_replay = 1 if graph_replay else 0
        p = np.ascontiguousarray(cnn_params, np.float32)
        _check(lib.stereo_learned_configure(ctx.handle, C.byref(self.handle), C.byref(fmt), C.byref(lp),
                                            p.ctypes.data, p.size), ctx)

    def process(self, disp, disp_lr, disp_rl, left, right, stream=None):
        _check(lib.stereo_learned_process(self.handle, disp.ref if disp else None, disp_lr.ref if disp_lr else None,
                                          disp_rl.ref if disp_rl else None, left.ref, right.ref, stream), self.ctx)

    def stats(self, timed=False):
        ev = C.c_int64()
        nt = C.c_int32()
        fb = C.c_int32()
        st = (C.c_double * 5)()
        _check(lib.stereo_learned_stats(self.handle, C.byref(ev), C.byref(nt), C.byref(fb), st if timed else None,
                                        5 if timed else 0), self.ctx)
        d = {"evals": ev.value, "tiles": nt.value, "fallback_tiles": fb.value}
        if timed:
            d["stage_ms"] = list(st)
        return d

    def tile_ranges(self):
        ntx = (self.W + 63) // 64
        nty = (self.H + 55) // 56
        buf = (C.c_int32 * (2 * ntx * nty))()
        _check(lib.stereo_learned_tile_ranges(self.handle, buf, len(buf)), self.ctx)
        return np.array(buf, np.int32).reshape(nty, ntx, 2)

    def tile_masks(self):
        ntx = (self.W + 63) // 64
        nty = (self.H + 55) // 56
        buf = (C.c_uint32 * (ntx * nty))()
        _check(lib.stereo_learned_tile_masks(self.handle, buf, len(buf)), self.ctx)
        return np.array(buf, np.uint32).reshape(nty, ntx)

    def run(self, L, R, maps=False):
        dl = DeviceImage.from_array(self.ctx, L, PIXEL_LUM_U8)
        dr = DeviceImage.from_array(self.ctx, R, PIXEL_LUM_U8)
        out = DeviceImage(self.ctx, self.W, self.H, PIXEL_LUM_I16)
        lr = DeviceImage(self.ctx, self.W, self.H, PIXEL_LUM_I16) if maps else None
        rl = DeviceImage(self.ctx, self.W, self.H, PIXEL_LUM_I16) if maps else None
        self.process(out, lr, rl, dl, dr)
        if maps:
            return out.download(), lr.download(), rl.download()
        return out.download()
