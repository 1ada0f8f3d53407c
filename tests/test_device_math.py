"""Arithmetic identities the device code relies on, checked on the CPU with
the expression taken verbatim from the device header (IEEE FP32 on both
sides; fma is correctly rounded on the host and on gfx950)."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "simpler_amd", "csrc", "device_math.h")

C_MAIN = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#define __device__
#define __forceinline__ inline
#define __builtin_fmaf fmaf
%s
int main(void) {
    const uint32_t lo = %du, hi = 0x7f800000u;  /* every float in [1e-30, FLT_MAX] */
    unsigned long long bad = 0;
    #pragma omp parallel for reduction(+:bad) schedule(static, 1 << 20)
    for (long long b = lo; b < (long long)hi; ++b) {
        const uint32_t u = (uint32_t)b;
        float x, y;
        memcpy(&x, &u, 4);
        if (div10(x) != x / 10.0f) ++bad;
        y = -x;
        if (div10(y) != y / 10.0f) ++bad;
    }
    printf("%%llu\n", bad);
    return 0;
}
"""


def test_div10_equals_ieee_division_by_ten(tmp_path):
    """div10 (the ER step refinement h / 10 of scene.cpp:410) returns the
    correctly rounded FP32 quotient for every |x| >= 1e-30: the refinement
    chain of step sizes is the reference's, bit for bit."""
    src = open(HDR).read()
    m = re.search(r"__device__ __forceinline__ float div10\(float x\) \{.*?\n\}", src, re.S)
    assert m, "div10 not found in device_math.h"
    import struct
    lo = struct.unpack("<I", struct.pack("<f", 1e-30))[0]
    c = tmp_path / "div10.c"
    c.write_text(C_MAIN % (m.group(0), lo))
    exe = str(tmp_path / "div10")
    subprocess.run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", str(c), "-o", exe, "-lm"], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=600,
                         env=dict(os.environ, OMP_NUM_THREADS=str(min(8, os.cpu_count() or 1)))).stdout
    assert int(out) == 0

