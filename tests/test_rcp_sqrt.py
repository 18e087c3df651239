"""The neighbour build's exact pass forms |r| and 1/|r| from the hardware root and reciprocal
(v_sqrt_f32, v_rcp_f32: within one ulp) with one correction each (ivp_device.h, sqrt_rn and
rcp_rn), in place of the compiler's IEEE sequences; density, Aii and the stored gradients must
stay bitwise those of sqrtf and 1.0f / b (spiky_kernel_derivative and cubic_kernel,
scripts/sim/sph_kernel.py:8-26, 52-61).

A C restatement of both corrections runs over every float of one period of the scaling
(sqrt: [1, 4); the reciprocal: every significand of [1, 2)) with every hardware result the error
bound allows (the correctly rounded value and its neighbours), and compares with IEEE sqrtf and
division.  The sqrt correction is also checked at the smallest argument it is exact for (2^-96).
Planted mistakes must show: a reciprocal without the all-ones-significand case, and a sqrt that
moves to only one neighbour."""

import shutil
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
static inline float fb(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t bf(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static int MODE = 0;   /* 0: as shipped; 1: no all-ones case; 2: sqrt moves down only */
static float sqrt_rn(float x, float s) {
    float sm = fb(bf(s) - 1u), sp = fb(bf(s) + 1u);
    float rm = fmaf(-sm, s, x), rp = fmaf(-sp, s, x);
    float r = rm <= 0.0f ? sm : s;
    if (MODE == 2) return r;
    return rp > 0.0f ? sp : r;
}
static float rcp_rn(float b, float y0) {
    float y1 = fmaf(fmaf(-b, y0, 1.0f), y0, y0);
    uint32_t u = bf(b);
    if (MODE == 1) return y1;
    return (u & 0x7fffffu) == 0x7fffffu ? fb((0x7e800000u - (u & 0x7f800000u)) | 1u) : y1;
}
int main(int argc, char** argv) {
    MODE = atoi(argv[1]);
    long n = 0, bad = 0;
    /* reciprocal: b over every significand, scaled into the build's |r| range [1e-6, 0.05] */
    const int ex[3] = {-19, -8, -5};
    for (int e = 0; e < 3; ++e)
        for (uint32_t m = 0; m < (1u << 23); ++m) {
            float b = ldexpf(fb(0x3f800000u | m), ex[e]);
            volatile float bv = b;
            float ref = 1.0f / bv;
            double exact = 1.0 / (double)b;
            for (int d = -2; d <= 2; ++d) {
                float y0 = fb(bf(ref) + d);
                double ulp = (double)fb(bf(y0) + 1u) - (double)y0;
                if (fabs((double)y0 - exact) >= ulp) continue;   /* beyond the one-ulp bound */
                n++;
                if (bf(rcp_rn(b, y0)) != bf(ref)) bad++;
            }
        }
    long ns = 0, bads = 0;
    /* sqrt: every float of [1, 4) and of [2^-96, 2^-94), candidates within one ulp */
    for (int sc = 0; sc < 2; ++sc)
        for (uint32_t m = 0; m < (1u << 24); ++m) {
            float x = ldexpf(fb(0x3f800000u + m), sc ? -96 : 0);
            float ref = sqrtf(x);
            for (int d = -1; d <= 1; ++d) {
                float s = fb(bf(ref) + d);
                double ulp = (double)fb(bf(s) + 1u) - (double)s;
                if (fabs((double)s - sqrt((double)x)) >= ulp) continue;
                ns++;
                if (bf(sqrt_rn(x, s)) != bf(ref)) bads++;
            }
        }
    printf("%ld %ld %ld %ld\n", n, bad, ns, bads);
    return 0;
}
"""


def _run(tmp_path, mode):
    src = tmp_path / "rs.c"
    exe = tmp_path / "rs"
    src.write_text(SRC)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True)
    out = subprocess.run([str(exe), str(mode)], check=True, capture_output=True, text=True).stdout
    return [int(v) for v in out.split()]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_corrected_root_and_reciprocal_are_correctly_rounded(tmp_path):
    n, bad, ns, bads = _run(tmp_path, 0)
    assert n > 3 * (1 << 23) and ns > 2 * (1 << 24)
    assert bad == 0 and bads == 0


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
@pytest.mark.parametrize("mode", [1, 2], ids=["no_all_ones_case", "sqrt_one_neighbour"])
def test_planted_mistakes_are_caught(tmp_path, mode):
    n, bad, ns, bads = _run(tmp_path, mode)
    assert (bad if mode == 1 else bads) > 0
