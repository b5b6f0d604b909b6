"""Range check of the fixed-point terms (csrc/fixed_point.hpp) compiled with
the host compiler: finite terms below 2^52 convert, NaN, infinities and
terms at or beyond the limit do not, and 2^11 terms at the limit cannot wrap
a signed 64-bit sum."""
import os
import subprocess
import tempfile

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "lbfs-ibm-ade-sp_amd", "csrc")

PROGRAM = r"""
#include <cmath>
#include <cstdio>
#include <limits>
#include "fixed_point.hpp"
using namespace lbfs;
int main()
{
    const double inf = std::numeric_limits<double>::infinity();
    const double nan = std::numeric_limits<double>::quiet_NaN();
    static_assert(fix_in_range(0.0), "zero");
    static_assert(fix_in_range(-1.5e15), "small negative");
    int bad = 0;
    bad += !fix_in_range(std::nextafter(FIX_TERM_LIMIT, 0.0));
    bad += !fix_in_range(-std::nextafter(FIX_TERM_LIMIT, 0.0));
    bad += fix_in_range(FIX_TERM_LIMIT);
    bad += fix_in_range(-FIX_TERM_LIMIT);
    bad += fix_in_range(inf);
    bad += fix_in_range(-inf);
    bad += fix_in_range(nan);
    bad += fix_in_range(1e300);
    // 2^11 terms just below the limit stay inside int64
    const double t = std::nextafter(FIX_TERM_LIMIT, 0.0);
    long double sum = 0.0L;
    for (int k = 0; k < 2048; k++) sum += (long double)(long long)t;
    bad += !(sum < 9223372036854775807.0L);
    // the IBM scale 2^40: |F w| = 4095 converts, 4096 does not
    bad += !fix_in_range(4095.0 * 1099511627776.0);
    bad += fix_in_range(4096.0 * 1099511627776.0);
    std::printf("%d\n", bad);
    return bad != 0;
}
"""


def test_fixed_point_range_host():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t.cpp")
        exe = os.path.join(d, "t")
        with open(src, "w") as f:
            f.write(PROGRAM)
        r = subprocess.run(["g++", "-std=c++17", "-O2", "-I", CSRC, src, "-o", exe], capture_output=True, text=True)
        if r.returncode != 0:
            pytest.fail("g++ failed: " + r.stderr)
        out = subprocess.run([exe], capture_output=True, text=True)
        assert out.returncode == 0 and out.stdout.strip() == "0", out.stdout + out.stderr
