"""Error paths and checkpoint / resume on the GPU:
  - STEREO_FAIL_AT=n makes the n-th guarded operation of a process (an allocation or a
    launch sequence) fail with STEREO_ERROR_HIP; the caller sees the status and message,
    nothing is launched, and the next call works (run in a child process, since the
    variable is read once per process);
  - the network's training state survives stereo_cnn_save_state / load_state: parameters,
    Adam moments and step count, so that a resumed run continues exactly like the
    uninterrupted one."""
import os
import subprocess
import sys

import numpy as np
import pytest

import clutec_stereo as cs
from test_gpu_cnn import _frame

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
sys.path[:0] = [{pkg!r}]
import numpy as np
import clutec_stereo as cs
ctx = cs.Context(0)
L, R = cs.synth_pair(3, 160, 96, 0, 8, 31, 8)
want = None
for attempt in range(4):
    try:
        m = cs.Matcher(ctx, 160, 96, algorithm=cs.ALGO_SAD, window=(7, 7), num_disp=32, tol=1)
        out = m.run(L, R)
    except cs.StereoError as e:
        print("attempt", attempt, "status", e.status, str(e), flush=True)
        continue
    print("attempt", attempt, "ok", int((out >= 0).sum()), flush=True)
"""


@pytest.mark.parametrize("fail_at", [1, 2])
def test_injected_failure_is_reported_and_recovers(fail_at):
    """Guard 1 is the matcher's configure (its workspace allocations), guard 2 the first
    device image of its run."""
    env = dict(os.environ, STEREO_FAIL_AT=str(fail_at))
    code = CHILD.format(pkg=os.path.join(ROOT, "module.clutec.cuda_amd"))
    print(f"child with STEREO_FAIL_AT={fail_at}", flush=True)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = r.stdout.strip().splitlines()
    print("\n".join(lines), flush=True)
    first = lines[0].split()
    assert first[:2] == ["attempt", "0"] and first[2] == "status" and int(first[3]) == cs.ERROR_HIP
    assert "injected failure" in lines[0]
    oks = [l for l in lines[1:] if " ok " in l]
    assert len(oks) == 3 and len({l.split()[-1] for l in oks}) == 1 and int(oks[0].split()[-1]) > 0


PLANES_CHILD = r"""
import sys, hashlib
sys.path[:0] = [{pkg!r}]
import clutec_stereo as cs
ctx = cs.Context(0)
pairs = [cs.synth_pair(6, 320, 240, 1, 5, 63, 8), cs.synth_pair(5, 320, 240, 0, 8, 63, 8)]
for algo, win, cen in ((cs.ALGO_CENSUS, (9, 9), (9, 7)), (cs.ALGO_SAD, (7, 7), (1, 1)), (cs.ALGO_AW1, (5, 5), (1, 1))):
    m = None
    for _ in range(2):  # an injected failure may hit the configure
        try:
            m = cs.Matcher(ctx, 320, 240, algorithm=algo, window=win, census=cen, num_disp=64, tol=1)
            break
        except cs.StereoError as e:
            print(algo, "configure error", str(e), flush=True)
    for attempt in range(3):
        L, R = pairs[min(attempt, 1)]  # frame 0: another pair than frames 1 and 2
        try:
            out = m.run(L, R, maps=True)
        except cs.StereoError as e:
            print(algo, "attempt", attempt, "error", str(e), flush=True)
            continue
        h = hashlib.md5(b"".join(o.tobytes() for o in out)).hexdigest()
        print(algo, "attempt", attempt, "ok", h, flush=True)
"""


def test_frame_after_an_interrupted_merge_is_clean():
    """The census, SAD and adaptive-window matchers fold their keys into two planes that
    every frame's merge clears.  A frame whose merge does not run (a failure injected
    between the matcher and the merge launch) leaves its keys there; the next frame, on
    another pair, must clear them and equal that pair's frame in a run without failures
    (stale keys of the other pair would win some of its atomicMax and change the maps)."""
    code = PLANES_CHILD.format(pkg=os.path.join(ROOT, "module.clutec.cuda_amd"))

    def run(env_extra):
        env = dict(os.environ, **env_extra)
        r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr[-2000:]
        return r.stdout.strip().splitlines()

    clean = {l.split()[0]: l.split()[-1] for l in run({}) if " ok " in l and " attempt 1 " in l}
    assert len(clean) == 3
    checked = set()
    for fail_at in range(1, 60):
        print(f"child with STEREO_FAIL_AT={fail_at}", flush=True)
        lines = run({"STEREO_FAIL_AT": str(fail_at)})
        for l in lines:
            if "between the matcher and the merge" not in l:
                continue
            algo, _, k = l.split()[:3]
            nxt = [m for m in lines if m.startswith(f"{algo} attempt {int(k) + 1} ")]
            if int(k) == 0:
                # the frame after the interrupted one, on the other pair
                assert nxt and nxt[0].split()[3] == "ok" and nxt[0].split()[-1] == clean[algo], lines
                checked.add(algo)
        if len(checked) == 3:
            break
    assert len(checked) == 3, checked


def test_cnn_state_round_trip_resumes_training(ctx):
    W, H, D = 256, 128, 64
    frames = [_frame(ctx, 300 + i, W, H, D)[0] for i in range(2)]

    def net_with_data():
        n = cs.ProposalNet(ctx, W // 4, H // 4, max_frames=2)
        for i, imgs in enumerate(frames):
            n.set_input(i, *imgs[:3])
            n.set_labels(i, imgs[3])
        return n

    a = net_with_data()
    a.set_params(cs.cnn_init_params(7))
    for _ in range(5):
        a.train_step(2, 2e-3)
    blob = a.save_state()
    assert blob.size == cs.lib.stereo_cnn_state_size()
    assert bytes(blob[:8]) == b"CLSTCNN1"
    b = net_with_data()
    b.set_params(cs.cnn_init_params(8))  # different parameters, overwritten by the load
    b.load_state(blob)
    assert np.array_equal(a.params(), b.params())
    assert np.array_equal(a.save_state(), b.save_state())
    # the resumed network continues exactly as the uninterrupted one
    for _ in range(3):
        la = a.train_step(2, 2e-3)
        lb = b.train_step(2, 2e-3)
        assert np.array_equal(la, lb)
    assert np.array_equal(a.params(), b.params())
    # a truncated or foreign blob is refused
    with pytest.raises(cs.StereoError):
        b.load_state(blob[:-4])
    bad = blob.copy()
    bad[:8] = np.frombuffer(b"NOTACNN!", np.uint8)
    with pytest.raises(cs.StereoError):
        b.load_state(bad)
    a.close()
    b.close()
