"""The reference-format NCC kernel divides window sums by the window size N = P * P with
a reciprocal product and one fused multiply-add residual step instead of an IEEE
division (csrc/refmatch.hip, div_window).  This checks, for every integer numerator
0 <= x <= N * 65025 (a window sum of bytes or of products of two bytes) and P = 5, 7, 9,
11, 15, that the two agree bit for bit, with gcc's IEEE fmaf and division on the host."""
import os
import shutil
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdio.h>
int main(void) {
    const int Ps[] = {5, 7, 9, 11, 15};
    long bad = 0, n = 0;
    for (int pi = 0; pi < 5; ++pi) {
        const int N = Ps[pi] * Ps[pi];
        const float R = 1.0f / (float)N, Nf = (float)N;
        for (long x = 0; x <= (long)N * 65025; ++x, ++n) {
            const float xf = (float)x;
            volatile float want = xf / Nf;
            const float q0 = xf * R;
            const float q1 = fmaf(fmaf(-q0, Nf, xf), R, q0);
            bad += q1 != want;
        }
    }
    printf("%ld %ld\n", n, bad);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_window_division_is_exact(tmp_path):
    c = tmp_path / "div.c"
    exe = tmp_path / "div"
    c.write_text(SRC)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(c), "-lm"], check=True)
    n, bad = (int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split())
    assert n == sum(p * p * 65025 + 1 for p in (5, 7, 9, 11, 15))
    assert bad == 0


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_window_division_check_fails_on_a_plain_product(tmp_path):
    """The check has teeth: the reciprocal product alone, without the residual step, is
    not correctly rounded for some numerators."""
    c = tmp_path / "div.c"
    exe = tmp_path / "div"
    c.write_text(SRC.replace("fmaf(fmaf(-q0, Nf, xf), R, q0)", "q0"))
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(c), "-lm"], check=True)
    n, bad = (int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split())
    assert bad > 0
