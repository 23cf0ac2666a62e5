"""The rank-sum test's normal approximation as the device evaluates it
(gsnp_amd/csrc/nearest.h: exp in double-double rounded to nearest, sqrt
checked for correct rounding), built for the host and checked against the
host libm's formula (cpp/normal_dis.cc:3-24, cpp/rank_sum.cc:44-49) and the
oracle: every value not flagged `hard` must be the libm's double, bit for
bit, and few may be flagged."""
from __future__ import annotations

import ctypes
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]

HARNESS = r"""
#include "nearest.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

// the reference's statement with the host libm
static double libm_value(double z)
{
    if (z > 6.0 || z < -6.0) return 0.0;
    double b1 = 0.31938153, b2 = -0.356563782, b3 = 1.781477937, b4 = -1.821255978, b5 = 1.330274429;
    double pp = 0.2316419, c2 = 0.39894228;
    double a = fabs(z);
    double t = 1.0 / (1.0 + a * pp);
    double b = c2 * exp((-z) * (z / 2.0));
    double n = ((((b5 * t + b4) * t + b3) * t + b2) * t + b1) * t;
    n = 1.0 - b * n;
    if (z < 0.0) n = 1.0 - n;
    return n > 0.5 ? 1 - n : n;
}
static double libm_test(int n1, int n2, double T1, double T2)
{
    double u1 = (T1 - n1 * (n1 + n2 + 1) / 2) / sqrt(n1 * n2 * (n1 + n2 + 1) / (double)12);
    double u2 = (T2 - n2 * (n1 + n2 + 1) / 2) / sqrt(n1 * n2 * (n1 + n2 + 1) / (double)12);
    return libm_value(fabs(u1) > fabs(u2) ? u1 : u2);
}

static uint64_t rng(uint64_t* s)
{
    uint64_t x = *s += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static double unif(uint64_t* s) { return (double)(rng(s) >> 11) * 0x1p-53; }

extern "C" int near_test(int n1, int n2, double T1, double T2, double* out)
{
    bool hard = false;
    *out = gsnp::normal_test_nearest(n1, n2, T1, T2, hard);
    return hard ? 1 : 0;
}

extern "C" int near_exp(double x, double* out)
{
    bool hard = false;
    *out = gsnp::exp_nearest(x, hard);
    return hard ? 1 : 0;
}

// exp over [lo, 0]: stats = cases, flagged, unflagged results that differ
// from the libm's
extern "C" void sweep_exp(uint64_t seed, int64_t n, double lo, int64_t* stats)
{
    memset(stats, 0, 3 * sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) {
        const double x = lo * unif(&seed);
        bool hard = false;
        const double v = gsnp::exp_nearest(x, hard);
        stats[0]++;
        if (hard) {
            stats[1]++;
            continue;
        }
        const double w = exp(x);
        if (memcmp(&v, &w, 8) != 0) stats[2]++;
    }
}

// rank-sum cases with n1 + n2 >= 64: half the rank sums drawn near their
// mean (|z| up to about 7), half over their whole range
extern "C" void sweep_test(uint64_t seed, int64_t n, int64_t* stats)
{
    memset(stats, 0, 3 * sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) {
        const int n1 = 1 + (int)(rng(&seed) % 255);
        const int lo2 = n1 >= 64 ? 1 : 64 - n1;
        const int n2 = lo2 + (int)(rng(&seed) % (uint64_t)(256 - lo2));
        const int N = n1 + n2;
        const double tmin = n1 * (n1 + 1) / 2.0, tmax = n1 * (2.0 * N - n1 + 1) / 2.0;
        double T1;
        if (i & 1) {
            T1 = tmin + floor(2 * (tmax - tmin) * unif(&seed)) / 2;
        } else {
            const double mean = n1 * (N + 1) / 2.0, sd = sqrt(n1 * (double)n2 * (N + 1) / 12.0);
            T1 = floor(2 * (mean + sd * 14.0 * (unif(&seed) - 0.5))) / 2;
            T1 = T1 < tmin ? tmin : T1 > tmax ? tmax : T1;
        }
        const double T2 = N * (N + 1) / 2.0 - T1;
        bool hard = false;
        const double v = gsnp::normal_test_nearest(n1, n2, T1, T2, hard);
        stats[0]++;
        if (hard) {
            stats[1]++;
            continue;
        }
        const double w = libm_test(n1, n2, T1, T2);
        if (memcmp(&v, &w, 8) != 0) stats[2]++;
    }
}
"""


@pytest.fixture(scope="module")
def near(tmp_path_factory):
    d = tmp_path_factory.mktemp("nearest")
    src, lib = d / "near.cpp", d / "libnear.so"
    src.write_text(HARNESS)
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
           f"-I{ROOT / 'gsnp_amd' / 'csrc'}", str(src), "-o", str(lib), "-lm"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    L = ctypes.CDLL(str(lib))
    pd = ctypes.POINTER(ctypes.c_double)
    L.near_test.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, pd]
    L.near_exp.argtypes = [ctypes.c_double, pd]
    L.sweep_exp.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(ctypes.c_int64)]
    L.sweep_test.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    return L


def test_exp_nearest_matches_libm(near):
    st = (ctypes.c_int64 * 3)()
    near.sweep_exp(7, 2_000_000, -18.5, st)
    cases, hard, bad = st
    print(f"exp: {cases} cases, {hard} flagged, {bad} unflagged differ", file=sys.stderr)
    assert bad == 0
    assert hard < 0.06 * cases


def test_exp_nearest_special_points(near):
    """exp(0) = 1 exactly; arguments at multiples of ln 2 (the reduction's
    k changes) and the range ends."""
    import math

    out = ctypes.c_double()
    for x in [0.0, -0.0, -math.log(2), -2 * math.log(2), -0.5 * math.log(2), -18.0, -18.5, -1e-300, -5e-17]:
        hard = near.near_exp(x, ctypes.byref(out))
        if not hard:
            assert out.value == math.exp(x), (x, out.value, math.exp(x))
    assert near.near_exp(0.0, ctypes.byref(out)) == 0 and out.value == 1.0


def test_normal_test_matches_libm(near):
    st = (ctypes.c_int64 * 3)()
    near.sweep_test(11, 2_000_000, st)
    cases, hard, bad = st
    print(f"normal test: {cases} cases, {hard} flagged, {bad} unflagged differ", file=sys.stderr)
    assert bad == 0
    assert hard < 0.06 * cases


def test_normal_test_matches_oracle(near):
    """Against the oracle's statement (orc_normal_test) on seeded cases, among
    them values of z near 0, near +-6 (the cut-off) and exact ties of |u1|
    and |u2|."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib

    o = oracle_lib.orc()
    rng = np.random.default_rng(5)
    out = ctypes.c_double()
    checked = 0
    for _ in range(20000):
        n1 = int(rng.integers(1, 256))
        n2 = int(rng.integers(max(1, 64 - n1), 256))
        N = n1 + n2
        mean = n1 * (N + 1) / 2
        sd = (n1 * n2 * (N + 1) / 12) ** 0.5
        T1 = float(np.floor(2 * (mean + sd * rng.uniform(-7, 7))) / 2)
        T1 = min(max(T1, n1 * (n1 + 1) / 2), n1 * (2 * N - n1 + 1) / 2)
        T2 = N * (N + 1) / 2 - T1
        if near.near_test(n1, n2, T1, T2, ctypes.byref(out)):
            continue
        assert out.value == o.orc_normal_test(n1, n2, T1, T2), (n1, n2, T1, T2)
        checked += 1
    assert checked > 17000
