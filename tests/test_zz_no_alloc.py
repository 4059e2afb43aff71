"""stereo_matcher_process_noalloc through the C ABI: the SAD and census matchers with their
workspace in the library's scratch array, equal to stereo_matcher_process of a configured
driver on two calls in a row, the second with a smaller frame and range (the array keeps
the first call's contents), and a frame whose workspace exceeds the array is refused."""
import ctypes as C

import numpy as np
import pytest

import clutec_stereo as cs

pytestmark = pytest.mark.gpu


def _noalloc(ctx, params, L, R, dispid=False):
    H, W = L.shape
    dl = cs.DeviceImage.from_array(ctx, L, cs.PIXEL_LUM_U8)
    dr = cs.DeviceImage.from_array(ctx, R, cs.PIXEL_LUM_U8)
    out = cs.DeviceImage(ctx, W, H, cs.PIXEL_RGBA_U16 if dispid else cs.PIXEL_LUM_I16)
    lr = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_I16)
    rl = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_I16)
    st = cs.lib.stereo_matcher_process_noalloc(ctx.handle, C.byref(params), out.ref, lr.ref, rl.ref, dl.ref, dr.ref,
                                               None)
    assert st == 0, cs.lib.stereo_last_error(ctx.handle)
    ctx.synchronize()
    return out.download(), lr.download(), rl.download()


def _configured(ctx, algo, window, census, D, L, R, dispid=False):
    H, W = L.shape
    m = cs.Matcher(ctx, W, H, algorithm=algo, window=window, census=census, num_disp=D, tol=1)
    try:
        dl = cs.DeviceImage.from_array(ctx, L, cs.PIXEL_LUM_U8)
        dr = cs.DeviceImage.from_array(ctx, R, cs.PIXEL_LUM_U8)
        out = cs.DeviceImage(ctx, W, H, cs.PIXEL_RGBA_U16 if dispid else cs.PIXEL_LUM_I16)
        lr = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_I16)
        rl = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_I16)
        m.process(out, lr, rl, dl, dr)
        ctx.synchronize()
        return out.download(), lr.download(), rl.download()
    finally:
        m.close()


@pytest.mark.parametrize("algo,window,census", [(cs.ALGO_CENSUS, (9, 9), (9, 7)), (cs.ALGO_SAD, (9, 9), (1, 1))])
def test_noalloc_equals_configured_driver_on_two_calls(ctx, algo, window, census):
    for k, (W, H, D, seed, dispid) in enumerate([(1920, 1080, 256, 11, False), (333, 197, 64, 12, True)]):
        print(f"noalloc algo {algo} call {k}: {W}x{H} D={D}", flush=True)
        L, R = cs.synth_pair(seed, W, H, 0, 8, D - 1, 8)
        params = cs.matcher_params(algo, window, census, D, 1, False)
        got = _noalloc(ctx, params, L, R, dispid)
        want = _configured(ctx, algo, window, census, D, L, R, dispid)
        for a, b, what in zip(got, want, ("checked", "left", "right")):
            assert np.array_equal(a, b), f"call {k}: {what} map differs"
        assert (want[1] >= 0).mean() > 0.5


def test_noalloc_refuses_what_does_not_fit(ctx):
    W, H, D = 2560, 1440, 256
    L = np.zeros((H, W), np.uint8)
    params = cs.matcher_params(cs.ALGO_CENSUS, (9, 9), (9, 7), D, 1, False)
    dl = cs.DeviceImage.from_array(ctx, L, cs.PIXEL_LUM_U8)
    out = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_I16)
    st = cs.lib.stereo_matcher_process_noalloc(ctx.handle, C.byref(params), out.ref, None, None, dl.ref, dl.ref, None)
    assert st == cs.ERROR_INVALID_FORMAT
    assert b"scratch array" in cs.lib.stereo_last_error(ctx.handle)
    bad = cs.matcher_params(cs.ALGO_NCC, (9, 9), (1, 1), 64, 1, False)
    small = cs.DeviceImage.from_array(ctx, np.zeros((64, 128), np.uint8), cs.PIXEL_LUM_U8)
    o2 = cs.DeviceImage(ctx, 128, 64, cs.PIXEL_LUM_I16)
    st = cs.lib.stereo_matcher_process_noalloc(ctx.handle, C.byref(bad), o2.ref, None, None, small.ref, small.ref, None)
    assert st == cs.ERROR_INVALID_CONFIG
