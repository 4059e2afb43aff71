"""Build recipe of the gfx950 stereo engine: every HIP translation unit is compiled
with hipcc for gfx950 and linked, with the C ABI and the host generator, into
module.clutec.cuda_amd/lib/libclutec_stereo.so.  Always rebuilds everything.

    python module.clutec.cuda_amd/build.py [-j N] [--verbose]
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
EXAMPLES = os.path.join(PKG, "examples")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")
LIBNAME = "libclutec_stereo.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 VALU forms are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    c = sorted(f for f in os.listdir(CSRC) if f.endswith(".c"))
    return hip, cpp, c


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(jobs=8, verbose=False):
    if os.path.isdir(BUILD):
        shutil.rmtree(BUILD)
    os.makedirs(BUILD)
    os.makedirs(LIB, exist_ok=True)
    hip, cpp, c = sources()
    cmds = []
    for f in hip:
        cmds.append([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, f), "-o", os.path.join(BUILD, f + ".o")])
    for f in cpp:
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, f), "-o", os.path.join(BUILD, f + ".o")])
    for f in c:
        cmds.append(["gcc", "-O2", "-fPIC", "-std=gnu11", "-I", CSRC, "-c", os.path.join(CSRC, f),
                     "-o", os.path.join(BUILD, f + ".o")])
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(lambda cmd: _run(cmd, verbose), cmds))
    objs = [os.path.join(BUILD, f + ".o") for f in hip + cpp + c]
    tmp = os.path.join(LIB, LIBNAME + ".tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lm"], verbose)
    os.replace(tmp, os.path.join(LIB, LIBNAME))
    # host programs over the C++ driver layer (plain g++, linked against the library)
    for f in sorted(os.listdir(EXAMPLES)):
        if f.endswith(".cpp"):
            exe = os.path.join(LIB, f[:-4])
            _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), os.path.join(EXAMPLES, f),
                  "-o", exe + ".tmp", "-L", LIB, "-lclutec_stereo", "-Wl,-rpath,$ORIGIN"], verbose)
            os.replace(exe + ".tmp", exe)
    return os.path.join(LIB, LIBNAME)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    path = build(a.j, a.verbose)
    print(path)


if __name__ == "__main__":
    sys.exit(main())
