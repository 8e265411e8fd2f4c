"""CPU: the device kernels' quotient a / d from d's correctly rounded
reciprocal y (q0 = a y, one correction, then Markstein's correction step;
amr3.hip div_rn) equals the IEEE quotient.  Checked exhaustively over every
float significand of a binade, for the divisors the batched engine uses
(d = 1 / n for every size of the 512-instance draw at its three levels) and
at the two ends of the range where the product form is used; FP64 on
random significands.  The same arithmetic in C (fmaf / fma are correctly
rounded fused multiply-adds, as the GPU's)."""

import ctypes
import shutil
import subprocess

import numpy as np
import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <string.h>
static inline float q32(float a, float d, float y) {
    float q0 = a * y;
    float r0 = fmaf(-q0, d, a);
    float q1 = fmaf(r0, y, q0);
    float r1 = fmaf(-q1, d, a);
    return fmaf(r1, y, q1);
}
static inline double q64(double a, double d, double y) {
    double q0 = a * y;
    double r0 = fma(-q0, d, a);
    double q1 = fma(r0, y, q0);
    double r1 = fma(-q1, d, a);
    return fma(r1, y, q1);
}
/* mismatches over every significand of the binade [2^ex, 2^(ex+1)) */
long long check32(const int* ns, int count, int ex, int stride) {
    long long bad = 0;
#pragma omp parallel for reduction(+:bad) schedule(dynamic, 1)
    for (int k = 0; k < count; ++k) {
        const float d = (float)(1.0 / ns[k]);
        const float y = (float)(1.0 / (double)d);
        for (uint32_t m = 0; m < (1u << 23); m += (uint32_t)stride) {
            uint32_t bits = ((uint32_t)(127 + ex) << 23) | m;
            float a, mq, r;
            memcpy(&a, &bits, 4);
            mq = q32(a, d, y); r = a / d;
            if (memcmp(&mq, &r, 4) != 0) ++bad;
            a = -a; mq = q32(a, d, y); r = a / d;
            if (memcmp(&mq, &r, 4) != 0) ++bad;
        }
    }
    return bad;
}
long long check64(const int* ns, int count, int ex, long per) {
    long long bad = 0;
#pragma omp parallel for reduction(+:bad) schedule(dynamic, 1)
    for (int k = 0; k < count; ++k) {
        const double d = 1.0 / ns[k];
        const double y = 1.0 / d;
        uint64_t s = 0x9e3779b97f4a7c15ull * (uint64_t)(ns[k] + 1);
        for (long j = 0; j < per; ++j) {
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            uint64_t bits = ((uint64_t)(1023 + ex) << 52) | (s & 0x000fffffffffffffull);
            double a, mq, r;
            memcpy(&a, &bits, 8);
            mq = q64(a, d, y); r = a / d;
            if (memcmp(&mq, &r, 8) != 0) ++bad;
        }
    }
    return bad;
}
"""


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    d = tmp_path_factory.mktemp("divchk")
    (d / "divchk.c").write_text(SRC)
    flags = ["-O2", "-ffp-contract=off", "-fPIC", "-shared", "-fopenmp"]
    if "fma" in open("/proc/cpuinfo").read().split():
        flags.append("-mfma")
    subprocess.run([cc, *flags, "-o", str(d / "divchk.so"), str(d / "divchk.c"), "-lm"], check=True)
    so = ctypes.CDLL(str(d / "divchk.so"))
    so.check32.restype = so.check64.restype = ctypes.c_longlong
    so.check32.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    so.check64.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_long]
    return so


def _sizes():
    from subsetix_amd import workloads

    ns = np.array([i["n"] for i in workloads.batched_instances(count=512, seed=4)], np.int64)
    return np.unique(np.concatenate([ns, 2 * ns, 4 * ns, np.arange(1, 130)])).astype(np.int32)


def test_float_quotient_every_significand(lib):
    ns = _sizes()
    for ex in (0, -64, 63):  # the middle and both ends of the product form's range
        bad = lib.check32(ns.ctypes.data, ns.size, ex, 1 if ex == 0 else 7)
        assert bad == 0, f"binade 2^{ex}: {bad} mismatches"


def test_double_quotient_random(lib):
    ns = _sizes()
    for ex in (0, -800, 799):
        assert lib.check64(ns.ctypes.data, ns.size, ex, 20000) == 0
