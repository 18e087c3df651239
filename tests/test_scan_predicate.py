"""The candidate scan's neighbour predicate (ivp_grid.hip, scan_group) decides on d2 formed with
fused multiply-adds, with the correctly rounded root only in a band of relative width 2^-20
around h^2.  The reference's predicate is sqrtf(rx*rx + ry*ry + rz*rz) < h on the unfused
float32 expression (scripts/sim/neighbour_search.py:43-44, `wp.length(r) < support_radius`).

This test compiles a C restatement of both decisions and compares them on random pairs
concentrated within a few parts per million of the support radius, for random and for the
shipped smoothing lengths (h = 4r with r = 0.0025, 0.005, 0.01).  A plausible mistake (margins
too narrow for the fused rounding, a band taken on the wrong side, the root of the fused d2)
shows up as mismatches within a few million samples.
"""

import shutil
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
static uint64_t s = 88172645463325252ull;
static inline uint64_t nx(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static inline double u01(void) { return (nx() >> 11) * (1.0 / 9007199254740992.0); }
int main(int argc, char** argv) {
    long N = atol(argv[1]);
    long bad = 0, band = 0, inside = 0;
    for (long t = 0; t < N; ++t) {
        float h = (float)(0.003 + 0.3 * u01());
        if (t % 7 == 0) h = 0.02f; else if (t % 7 == 1) h = 0.04f; else if (t % 7 == 2) h = 0.01f;
        /* the host's band limits (grid_build) */
        double hh = (double)h * (double)h;
        float lo = nextafterf((float)(hh * (1.0 - 1.0 / 1048576.0)), 0.0f);
        float hi = nextafterf((float)(hh * (1.0 + 1.0 / 1048576.0)), 1.0f);
        double z = 2 * u01() - 1, ph = 6.283185307179586 * u01(), rr = sqrt(1 - z * z);
        double R = (t & 15) == 0 ? 1.5 * h * u01() : h * (1.0 + 8e-6 * (u01() - 0.5));
        float xi = (float)(4.0 * u01()), yi = (float)(u01()), zi = (float)(1.5 * u01());
        float xj = (float)(xi - R * rr * cos(ph)), yj = (float)(yi - R * rr * sin(ph)), zj = (float)(zi - R * z);
        float rx = xi - xj, ry = yi - yj, rz = zi - zj;
        int ref = sqrtf(rx * rx + ry * ry + rz * rz) < h;          /* the reference's test */
        float d2 = fmaf(rz, rz, fmaf(ry, ry, rx * rx));            /* the scan's fused d2 */
        int in = d2 < lo;
        if (!in && d2 < hi) { band++; in = sqrtf(rx * rx + ry * ry + rz * rz) < h; }
        inside += ref;
        bad += in != ref;
    }
    printf("%ld %ld %ld\n", inside, band, bad);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_fused_band_predicate_matches_reference(tmp_path):
    src = tmp_path / "pred.c"
    exe = tmp_path / "pred"
    src.write_text(SRC)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True)
    n = 20_000_000
    out = subprocess.run([str(exe), str(n)], check=True, capture_output=True, text=True).stdout.split()
    inside, band, bad = (int(v) for v in out)
    assert bad == 0
    # the samples straddle the boundary and exercise the band path
    assert 0.4 * n < inside < 0.6 * n
    assert band > n // 20
