// nearest.h -- the rank-sum test's normal approximation (cpp/normal_dis.cc:3-24,
// cpp/rank_sum.cc:44-49) with exp and sqrt evaluated to the double a host libm
// returns: exp in double-double, rounded to nearest, and an exactness check of
// sqrt.  Where a value lies too close to a rounding boundary to be sure of
// that, `hard` is set and the caller finishes the site on the host.  Plain C++
// for the host compiler as well, so that it can be checked against the host
// libm without a device.
#pragma once

#include <math.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define GSNP_NEAR __host__ __device__ inline
#else
#define GSNP_NEAR inline
#endif

namespace gsnp {

// Double-double arithmetic (Knuth's two-sum, Dekker's product through fma);
// values carry about 106 bits.
struct DD {
    double hi, lo;
};
GSNP_NEAR DD dd_two_sum(double a, double b)
{
    const double s = a + b, bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
GSNP_NEAR DD dd_fast_sum(double a, double b)  // |a| >= |b|
{
    const double s = a + b;
    return {s, b - (s - a)};
}
GSNP_NEAR DD dd_add(DD a, DD b)
{
    DD s = dd_two_sum(a.hi, b.hi);
    s.lo += a.lo + b.lo;
    return dd_fast_sum(s.hi, s.lo);
}
GSNP_NEAR DD dd_mul(DD a, DD b)
{
    const double p = a.hi * b.hi;
    double e = fma(a.hi, b.hi, -p);
    e += a.hi * b.lo + a.lo * b.hi;
    return dd_fast_sum(p, e);
}
GSNP_NEAR DD dd_div(DD a, double b)
{
    const double q1 = a.hi / b;
    const double p = q1 * b, pe = fma(q1, b, -p);
    const double r = ((a.hi - p) - pe) + a.lo;
    return dd_fast_sum(q1, r / b);
}

// True when a normalized double-double (hi is its value rounded to nearest)
// lies within 2 % of an ulp of the midpoint between hi and its neighbour
// towards lo, where a libm whose error may exceed half an ulp could round the
// other way (glibc's exp is documented to stay within 0.511 ulp, so it
// returns hi everywhere else).
GSNP_NEAR bool dd_near_midpoint(DD v)
{
    if (v.lo == 0.0) return false;
    const double gap = fabs(nextafter(v.hi, v.lo > 0.0 ? INFINITY : -INFINITY) - v.hi);
    return fabs(v.lo) > 0.48 * gap;
}

// exp(x) for x in [-18.5, 0] (the normal density's argument for |z| <= 6),
// rounded to nearest: x = k ln 2 + r with |r| <= ln 2 / 2 in double-double
// (ln 2 to 106 bits), exp(r) by its Taylor series to r^22/22! in Horner form
// (the first omitted term is below 1e-31 of the sum), then the exact scaling
// by 2^k.
GSNP_NEAR double exp_nearest(double x, bool& hard)
{
    const double ln2_hi = 0x1.62e42fefa39efp-1, ln2_lo = 0x1.abc9e3b39803fp-56;
    const double k = rint(x * 0x1.71547652b82fep0);
    const double p = k * ln2_hi, pe = fma(k, ln2_hi, -p);  // k ln2_hi exactly
    const double q = k * ln2_lo, qe = fma(k, ln2_lo, -q);
    DD r = dd_two_sum(x, -p);
    r = dd_add(r, DD{-pe, 0.0});
    r = dd_add(r, DD{-q, -qe});
    DD t{1.0, 0.0};
    for (int n = 22; n >= 1; n--) t = dd_add(DD{1.0, 0.0}, dd_div(dd_mul(r, t), (double)n));
    const int ki = (int)k;
    const DD e = dd_fast_sum(ldexp(t.hi, ki), ldexp(t.lo, ki));
    hard |= dd_near_midpoint(e);
    return e.hi;
}

// sqrt(x) for x > 0, checked to be the correctly rounded root (what an IEEE
// sqrt returns): with s the computed root, x - s^2 (through fma, one
// rounding) must stay below s times the ulp towards the true root.
GSNP_NEAR double sqrt_nearest(double x, bool& hard)
{
    const double s = sqrt(x);
    const double r = fma(-s, s, x);  // the true root is s + r / (2 s) to first order
    const double gap = fabs(nextafter(s, r > 0.0 ? INFINITY : 0.0) - s);
    hard |= !(fabs(r) < 0.999 * s * gap);
    return s;
}

// cpp/normal_dis.cc:3-24 (Abramowitz & Stegun 26.2.17), every operation as
// written (built without contraction)
GSNP_NEAR double normal_value_nearest(double z, bool& hard)
{
    if (z > 6.0 || z < -6.0) return 0.0;
    const double b1 = 0.31938153, b2 = -0.356563782, b3 = 1.781477937, b4 = -1.821255978, b5 = 1.330274429;
    const double pp = 0.2316419, c2 = 0.39894228;
    const double a = fabs(z);
    const double t = 1.0 / (1.0 + a * pp);
    const double b = c2 * exp_nearest((-z) * (z / 2.0), hard);
    double n = ((((b5 * t + b4) * t + b3) * t + b2) * t + b1) * t;
    n = 1.0 - b * n;
    if (z < 0.0) n = 1.0 - n;
    return n > 0.5 ? 1 - n : n;
}

// cpp/rank_sum.cc:44-49 (n1 (n1 + n2 + 1) / 2 is an integer division)
GSNP_NEAR double normal_test_nearest(int n1, int n2, double T1, double T2, bool& hard)
{
    const double sd = sqrt_nearest(n1 * n2 * (n1 + n2 + 1) / (double)12, hard);
    const double u1 = (T1 - n1 * (n1 + n2 + 1) / 2) / sd;
    const double u2 = (T2 - n2 * (n1 + n2 + 1) / 2) / sd;
    return normal_value_nearest(fabs(u1) > fabs(u2) ? u1 : u2, hard);
}

}  // namespace gsnp
