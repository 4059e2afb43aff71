"""Writes tests/golden/synth_sha256.json: SHA-256 of generated pairs (L, R, ground
truth) for a few small configurations, pinning the synthetic generator byte for
byte.  Regenerate only when the generator's recipe changes on purpose."""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "module.clutec.cuda_amd"))
import clutec_stereo as cs  # noqa: E402

CASES = [
    dict(seed=0, W=64, H=48, kind=0, nregions=8, dmax=63, dbg=8),
    dict(seed=1, W=96, H=54, kind=0, nregions=24, dmax=255, dbg=8),
    dict(seed=2, W=128, H=72, kind=1, nregions=6, dmax=127, dbg=8),
    dict(seed=10000, W=96, H=54, kind=2, nregions=8, dmax=63, dbg=8),
]

out = []
for c in CASES:
    L, R, gt = cs.synth_pair(c["seed"], c["W"], c["H"], c["kind"], c["nregions"], c["dmax"], c["dbg"], with_gt=True)
    out.append(dict(c, sha256=hashlib.sha256(L.tobytes() + R.tobytes() + gt.tobytes()).hexdigest()))
json.dump(out, open(os.path.join(HERE, "synth_sha256.json"), "w"), indent=1)
print(json.dumps(out, indent=1))
