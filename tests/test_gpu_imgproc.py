"""RectifyImage and ColorDisparityImage on the GPU against the CPU oracle, bit-exact."""
import numpy as np
import pytest

import clutec_stereo as cs
import orc
from test_oracle_imgproc import _id_map

pytestmark = pytest.mark.gpu


def _rot(ax, ay, az):
    cx, cy, cz = np.cos([ax, ay, az])
    sx, sy, sz = np.sin([ax, ay, az])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    return Rz @ Ry @ Rx


CAMERAS = [
    # source (fx, fy, cx, cy), distortion, rectified camera, rotation angles
    ((1.0, 1.0, 0.0, 0.0), (0, 0, 0, 0, 0), None, (0, 0, 0)),
    ((700.0, 690.0, 300.5, 200.25), (-0.28, 0.09, 0.001, -0.0007, -0.012), (650.0, 650.0, 310.0, 195.0),
     (0.01, -0.02, 0.015)),
    ((400.0, 400.0, 320.0, 240.0), (0.05, -0.01, 0.0, 0.0, 0.0), (420.0, 420.0, 300.0, 250.0), (0.0, 0.0, 0.3)),
    ((350.0, 352.0, 100.0, 80.0), (0, 0, 0, 0, 0), (350.0, 352.0, 100.0, 80.0), (0.0, 0.2, 0.0)),
]


@pytest.mark.parametrize("cam", range(len(CAMERAS)))
def test_rectify_matches_oracle(ctx, cam):
    src_cam, dist, rect_cam, ang = CAMERAS[cam]
    sW, sH, W, H = 640, 400, 600, 420
    print(f"rectify camera {cam} {sW}x{sH} -> {W}x{H}", flush=True)
    src = cs.synth_pair(cam, sW, sH)[0]
    p = orc.rectify_params(src_cam, dist, rect_cam, _rot(*ang))
    got = cs.Rectify(ctx, W, H, p).run(src)
    want = orc.rectify(src, p, W, H)
    assert np.array_equal(got, want), int((got != want).sum())
    assert (got > 0).mean() > 0.2


def test_rectify_identity_1080p(ctx):
    W, H = 1920, 1080
    print("rectify identity 1080p", flush=True)
    L = cs.synth_pair(1, W, H)[0]
    p = orc.rectify_params(src=(1450.0, 1450.0, 959.5, 539.5))
    assert np.array_equal(cs.Rectify(ctx, W, H, p).run(L), L)


@pytest.mark.parametrize("style", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("fmt", ["rgba", "u16", "f32"])
def test_color_matches_oracle(ctx, style, fmt):
    W, H = 333, 121
    print(f"colour style {style} {fmt}", flush=True)
    ids = _id_map(10 + style, W, H)
    if fmt == "rgba":
        disp = np.zeros(ids.shape + (4,), np.uint16)
        disp[..., 0] = ids
        pt = cs.PIXEL_RGBA_U16
    elif fmt == "u16":
        disp, pt = ids, cs.PIXEL_LUM_U16
    else:
        disp = ids.astype(np.float32)
        disp[0, :5] = [1.5, 17.25, 0.5, 65519.5, 299.9]
        pt = cs.PIXEL_LUM_F32
    colors = [list(c) for c in cs.DEFAULT_COLORS]
    colors[6] = [1, 2, 3, 4]
    for dmin, dmax, off in [(1, 0xFFEF, 0), (20, 260, 7), (1, 300, 40000)]:
        got = cs.ColorDisparity(ctx, W, H, pt, dmin, dmax, off, colors).run(disp, style)
        want = orc.color_disparity(disp, style, dmin, dmax, off, np.array(colors, np.uint8))
        assert np.array_equal(got, want), (dmin, dmax, off)


def test_imgproc_configuration_errors(ctx):
    with pytest.raises(cs.StereoError) as e:
        cs.Rectify(ctx, 64, 64, orc.rectify_params(src=(0.0, 1.0, 0.0, 0.0)))
    assert e.value.status == cs.ERROR_INVALID_CONFIG
    with pytest.raises(cs.StereoError) as e:
        cs.ColorDisparity(ctx, 64, 64, cs.PIXEL_LUM_I16)
    assert e.value.status == cs.ERROR_UNSUPPORTED_PIXEL_TYPE
    with pytest.raises(cs.StereoError) as e:
        cs.ColorDisparity(ctx, 64, 64, cs.PIXEL_RGBA_U16, 10, 10)
    assert e.value.status == cs.ERROR_INVALID_CONFIG
    c = cs.ColorDisparity(ctx, 64, 32, cs.PIXEL_LUM_U16)
    d = cs.DeviceImage.from_array(ctx, np.ones((32, 64), np.uint16), cs.PIXEL_LUM_U16)
    wrong = cs.DeviceImage(ctx, 64, 32, cs.PIXEL_LUM_U16)
    with pytest.raises(cs.StereoError):
        c.process(wrong, d, cs.COLOR_ANALYSIS)  # Analysis writes RGBA_U8
    with pytest.raises(cs.StereoError):
        c.process(wrong, d, 7)
