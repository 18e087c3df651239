"""The neighbour build's exact pass divides several numerators by one divisor (rx, ry, rz by |r|
in spiky_kernel_derivative; |r| by h in cubic_kernel, scripts/sim/sph_kernel.py:8-26, 52-61).  It
forms y = 1/b once, correctly rounded, and each quotient as q = RN(x y) corrected once with the
exact residual x - b q (a fused multiply-add): q' = RN(q + (x - b q) y).  Markstein's theorem
(y within half an ulp of 1/b, q within an ulp of x/b, no overflow or underflow) makes q' the
correctly rounded x/b, so density, Aii and the gradients stay bit-exact.

This test compiles a C restatement of that sequence (ivp_device.h, div_rn) and compares it with
IEEE division: every divisor significand of one binade with numerators at the quotient's
critical points (x = b, x just below b, x at binade edges, random x in [-b, b]), random divisors
over [1e-6, 1] with unit-vector numerators, and |r| / h for the shipped smoothing lengths.
Planted mistakes must show mismatches: no correction step (hundreds of thousands), and a
reciprocal one ulp off (a few in the exhaustive binade: a faithful hardware reciprocal would
not do, so y is the correctly rounded division 1/b).
"""

import shutil
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
static uint64_t s = 0x9E3779B97F4A7C15ull;
static inline uint64_t nx(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static inline double u01(void) { return (nx() >> 11) * (1.0 / 9007199254740992.0); }
static int MODE = 0;   /* 0: the sequence; 1: no correction; 2: reciprocal one ulp off */
static inline float div_rn(float x, float b, float y) {
    if (MODE == 2) y = nextafterf(y, 2.0f * y);
    float q = x * y;
    if (MODE == 1) return q;
    return fmaf(fmaf(-b, q, x), y, q);
}
static long n = 0, bad = 0;
static inline void check(float x, float b) {
    volatile float bv = b;
    float y = 1.0f / bv;
    float ref = x / bv;
    float got = div_rn(x, b, y);
    n++;
    if (memcmp(&ref, &got, 4) != 0) bad++;
}
int main(int argc, char** argv) {
    MODE = atoi(argv[1]);
    long R = atol(argv[2]);
    /* every significand of [1, 2), scaled into the range of |r| */
    for (uint32_t m = 0; m < (1u << 23); m += (MODE == 1 ? 97 : 1)) {
        uint32_t bits = 0x3f800000u | m;
        float b1; memcpy(&b1, &bits, 4);
        float b = b1 * 0.0078125f;
        check(b, b);
        check(nextafterf(b, 0.0f), b);
        check(-nextafterf(b, 0.0f), b);
        check(b * 0.5f, b);
        check(nextafterf(b * 0.5f, 1.0f), b);
        check((float)(b * u01()), b);
        check((float)(-b * u01()), b);
    }
    /* random divisors over [1e-6, 1], unit-vector components as numerators */
    for (long t = 0; t < R; ++t) {
        float b = (float)exp(log(1e-6) + (log(1.0) - log(1e-6)) * u01());
        double z = 2 * u01() - 1, ph = 6.283185307179586 * u01(), rr = sqrt(1 - z * z);
        check((float)(b * rr * cos(ph)), b);
        check((float)(b * rr * sin(ph)), b);
        check((float)(b * z), b);
    }
    /* |r| / h for the shipped smoothing lengths (h = 4 r, r = 0.0025, 0.005, 0.01) */
    const float hs[3] = {0.01f, 0.02f, 0.04f};
    for (int k = 0; k < 3; ++k)
        for (long t = 0; t < R / 3; ++t) check((float)(1.001 * hs[k] * u01()), hs[k]);
    printf("%ld %ld\n", n, bad);
    return 0;
}
"""


def _run(tmp_path, mode, samples):
    src = tmp_path / "div.c"
    exe = tmp_path / "div"
    if not exe.exists():
        src.write_text(SRC)
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True)
    out = subprocess.run([str(exe), str(mode), str(samples)], check=True, capture_output=True, text=True).stdout
    n, bad = (int(v) for v in out.split())
    return n, bad


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_corrected_quotient_is_correctly_rounded(tmp_path):
    n, bad = _run(tmp_path, 0, 6_000_000)
    assert n > 70_000_000 and bad == 0


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
@pytest.mark.parametrize("mode", [1, 2], ids=["no_correction", "reciprocal_off_by_one_ulp"])
def test_planted_mistakes_are_caught(tmp_path, mode):
    n, bad = _run(tmp_path, mode, 300_000)
    assert bad > (1000 if mode == 1 else 0)
