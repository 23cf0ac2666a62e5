// gsnp_common.h -- definitions shared by the host code and the HIP kernels:
// window geometry, the observation-record codec, genotype order, and the exact
// decimal formatter used to write .cns rows on the device.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GSNP_HD __host__ __device__ __forceinline__

namespace gsnp {

// Window geometry and limits (cpp/common.h:14-16, :71-81).
constexpr uint32_t kWin = 262144;     // sites per window (32*8192)
constexpr uint32_t kSlots = 256;      // record slots per site (MAX_NUM_NON_ZERO)
constexpr uint32_t kMaxDepUni = 255;  // dep_uni cap (kernel/update_kernel.cu:288)
constexpr uint32_t kScoreRows = 256 * 256 * 4;  // (q_adj, cycle, o_base)
constexpr uint32_t kScoreRowStride = 16;  // doubles per score row: 10 used, padded to one 128-byte line
constexpr uint32_t kPrank = 64 * 64 * 2048;

// Observation record (kernel/likelihood_kernel.cu:24-42):
//   bits 24:23 observed base (A0 C1 T2 G3), 22:17 64-q, 16:9 cycle, 8 strand, 7:0 occurrence.
// Ascending u32 order is base up, quality down, cycle up, strand up.
GSNP_HD uint32_t rec_encode(uint32_t cycle, uint32_t strand, uint32_t o, uint32_t q, uint32_t occ)
{
    return (o << 23) | ((64u - q) << 17) | (cycle << 9) | (strand << 8) | occ;
}
GSNP_HD uint32_t rec_base(uint32_t w) { return (w & 0x01800000u) >> 23; }
GSNP_HD uint32_t rec_q(uint32_t w) { return 64u - ((w & 0x007e0000u) >> 17); }
GSNP_HD uint32_t rec_cycle(uint32_t w) { return (w & 0x0001fe00u) >> 9; }
GSNP_HD uint32_t rec_strand(uint32_t w) { return (w & 0x00000100u) >> 8; }
GSNP_HD uint32_t rec_occ(uint32_t w) { return w & 0xffu; }
GSNP_HD uint32_t rec_key(uint32_t w) { return (w >> 8) & 0x1ffu; }  // (cycle, strand)

// Packed read base byte: (q << 2) | base, with 0 meaning "not observed"
// (N, non-ACGT, or quality below q_min).  q is in [1, 63].
GSNP_HD uint32_t pb_base(uint32_t b) { return b & 3u; }
GSNP_HD uint32_t pb_q(uint32_t b) { return b >> 2; }

// Read metadata word: bits 15:0 hit, 24:16 length (1..256), 31 strand ('-').
GSNP_HD uint32_t rm_hit(uint32_t m) { return m & 0xffffu; }
GSNP_HD uint32_t rm_len(uint32_t m) { return (m >> 16) & 0x1ffu; }
GSNP_HD uint32_t rm_strand(uint32_t m) { return m >> 31; }

// The 10 genotypes a1<=a2 in score-table order AA AC AT AG CC CT CG TT TG GG,
// and their 4-bit codes a1<<2|a2 (kernel/likelihood_kernel.cu:145-156).
GSNP_HD uint32_t geno_code(int g)
{
    return g < 4 ? (uint32_t)g : g < 7 ? (uint32_t)(g + 1) : g < 9 ? (uint32_t)(g + 3) : 15u;
}

// (int)x with x86-64 cvttsd2si semantics: NaN and out-of-range give INT_MIN.
GSNP_HD int trunc_int_x86(double x)
{
    if (!(x > -2147483649.0 && x < 2147483648.0)) return (int)0x80000000;
    return (int)x;
}

// ---------------------------------------------------------------------------
// Decimal output.  fmt_u32 / fmt_i32 write base-10 integers; fmt_g writes the
// C "%#g" conversion (6 significant digits, trailing zeros and point kept),
// correctly rounded with ties to even from the exact binary value, as glibc's
// printf does.  All return the number of characters written.

GSNP_HD int fmt_u32(uint32_t v, char* o)
{
    char t[10];
    int n = 0;
    do {
        t[n++] = (char)('0' + v % 10u);
        v /= 10u;
    } while (v);
    for (int i = 0; i < n; i++) o[i] = t[n - 1 - i];
    return n;
}
GSNP_HD int fmt_i32(int v, char* o)
{
    if (v < 0) {
        o[0] = '-';
        return 1 + fmt_u32(0u - (uint32_t)v, o + 1);
    }
    return fmt_u32((uint32_t)v, o);
}
GSNP_HD int len_u32(uint32_t v)
{
    int n = 1;
    while (v >= 10u) v /= 10u, n++;
    return n;
}
GSNP_HD int len_i32(int v) { return v < 0 ? 1 + len_u32(0u - (uint32_t)v) : len_u32((uint32_t)v); }

// Rounds v * 10^k (v = m * 2^e exactly, m < 2^53) to an integer, ties to even
// (or truncates when trunc is set).
// Fast path: 0 <= k <= 27 with 128-bit arithmetic; otherwise a 1280-bit
// multiword path.  Returns the rounded value saturated to 2^63.
struct BigNum {
    static constexpr int kLimbs = 40;
    uint32_t d[kLimbs];
    int n;
};

GSNP_HD void big_set_shl(BigNum& b, uint64_t m, int sh)
{
    for (int i = 0; i < BigNum::kLimbs; i++) b.d[i] = 0;
    int limb = sh >> 5, bit = sh & 31;
    uint64_t lo = m << bit;
    uint64_t hi = bit ? (m >> (64 - bit)) : 0;
    if (limb < BigNum::kLimbs) b.d[limb] = (uint32_t)lo;
    if (limb + 1 < BigNum::kLimbs) b.d[limb + 1] = (uint32_t)(lo >> 32);
    if (limb + 2 < BigNum::kLimbs) b.d[limb + 2] = (uint32_t)hi;
    b.n = BigNum::kLimbs;
}
GSNP_HD void big_mul_small(BigNum& b, uint32_t f)
{
    uint64_t c = 0;
    for (int i = 0; i < BigNum::kLimbs; i++) {
        uint64_t x = (uint64_t)b.d[i] * f + c;
        b.d[i] = (uint32_t)x;
        c = x >> 32;
    }
}
GSNP_HD uint32_t big_div_small(BigNum& b, uint32_t f)
{
    uint64_t r = 0;
    for (int i = BigNum::kLimbs - 1; i >= 0; i--) {
        uint64_t x = (r << 32) | b.d[i];
        b.d[i] = (uint32_t)(x / f);
        r = x % f;
    }
    return (uint32_t)r;
}
GSNP_HD bool big_bit(const BigNum& b, int i) { return i >= 0 && i < 32 * BigNum::kLimbs && ((b.d[i >> 5] >> (i & 31)) & 1u); }
GSNP_HD bool big_any_below(const BigNum& b, int i)
{
    for (int j = 0; j < BigNum::kLimbs && 32 * j < i; j++) {
        uint32_t w = b.d[j];
        int top = i - 32 * j;
        if (top < 32) w &= (1u << top) - 1u;
        if (w) return true;
    }
    return false;
}
GSNP_HD uint64_t big_shr_u64(const BigNum& b, int s)
{
    uint64_t r = 0;
    for (int i = 0; i < 64; i++)
        if (big_bit(b, s + i)) r |= (1ull << i);
    return r;
}

// Multiword path of round_scaled for scalings outside 10^0..10^27 (values
// below 1e-22 or from 1e6 up); kept out of line so the common path stays in
// registers.
__host__ __device__ __noinline__ inline uint64_t round_scaled_wide(uint64_t m, int e, int k, bool trunc)
{
    // general path: v = A / 2^s with A = m << max(e,0), s = max(-e,0)
    BigNum b;
    int s = e < 0 ? -e : 0;
    big_set_shl(b, m, e > 0 ? e : 0);
    if (k >= 0) {
        for (int i = 0; i < k; i++) big_mul_small(b, 10u);
        uint64_t q = big_shr_u64(b, s);
        bool guard = s > 0 && big_bit(b, s - 1);
        bool sticky = s > 1 && big_any_below(b, s - 1);
        if (!trunc && guard && (sticky || (q & 1))) q += 1;
        return q;
    }
    // k < 0: integer part I = A >> s with fraction F, then divide by 10^|k|
    bool frac = s > 0 && big_any_below(b, s);
    BigNum I;
    for (int i = 0; i < BigNum::kLimbs; i++) I.d[i] = 0;
    I.n = BigNum::kLimbs;
    for (int i = 0; i + s < 32 * BigNum::kLimbs; i++)
        if (big_bit(b, i + s)) I.d[i >> 5] |= 1u << (i & 31);
    uint32_t top = 0;
    bool sticky = frac;
    for (int i = 0; i < -k; i++) {
        if (top) sticky = true;
        top = big_div_small(I, 10u);
    }
    uint64_t q = big_shr_u64(I, 0);
    if (!trunc && (top > 5 || (top == 5 && (sticky || (q & 1))))) q += 1;
    return q;
}

GSNP_HD uint64_t round_scaled(uint64_t m, int e, int k, bool trunc)
{
    if (k >= 0 && k <= 27) {
        unsigned __int128 p5 = 1;
        for (int i = 0; i < k; i++) p5 *= 5u;
        unsigned __int128 N = (unsigned __int128)m * p5;  // < 2^116
        int s = -(e + k);
        if (s <= 0) {  // integral result; saturate, callers only compare it with 10^5..10^6
            if (-s >= 11) return (1ull << 63);
            const unsigned __int128 v = N << (-s);
            return (v >> 63) ? (1ull << 63) : (uint64_t)v;
        }
        if (s >= 128) return 0;
        unsigned __int128 q = N >> s;
        unsigned __int128 rem = N - (q << s);
        unsigned __int128 half = (unsigned __int128)1 << (s - 1);
        if (!trunc && (rem > half || (rem == half && (q & 1)))) q += 1;
        return (uint64_t)q;
    }
    return round_scaled_wide(m, e, k, trunc);
}

// Six significant digits D (10^5 <= D < 10^6) with decimal exponent X, as
// %#g lays them out.
GSNP_HD int fmt_g_digits(uint64_t D, int X, char* o, int n)
{
    char dg[6];
    for (int i = 5; i >= 0; i--) {
        dg[i] = (char)('0' + D % 10u);
        D /= 10u;
    }
    if (X >= -4 && X < 6) {
        if (X >= 0) {
            for (int i = 0; i <= X; i++) o[n++] = dg[i];
            o[n++] = '.';
            for (int i = X + 1; i < 6; i++) o[n++] = dg[i];
        } else {
            o[n++] = '0';
            o[n++] = '.';
            for (int i = 0; i < -X - 1; i++) o[n++] = '0';
            for (int i = 0; i < 6; i++) o[n++] = dg[i];
        }
    } else {
        o[n++] = dg[0];
        o[n++] = '.';
        for (int i = 1; i < 6; i++) o[n++] = dg[i];
        o[n++] = 'e';
        int ax = X;
        if (ax < 0) {
            o[n++] = '-';
            ax = -ax;
        } else {
            o[n++] = '+';
        }
        if (ax < 10) o[n++] = '0';
        n += fmt_u32((uint32_t)ax, o + n);
    }
    return n;
}

// Exact %#g of any double: the decimal exponent and the rounded digits come
// from integer arithmetic on the binary value (kept out of line).
__host__ __device__ __noinline__ inline int fmt_g_exact(double v, char* o)
{
    int n = 0;
    uint64_t bits;
    __builtin_memcpy(&bits, &v, 8);
    if (bits >> 63) o[n++] = '-';
    int be = (int)((bits >> 52) & 0x7ff);
    uint64_t m = bits & ((1ull << 52) - 1);
    if (be == 0x7ff) {  // inf / nan as glibc prints them
        const char* s = m ? "nan" : "inf";
        for (int i = 0; i < 3; i++) o[n++] = s[i];
        return n;
    }
    if (be == 0 && m == 0) {
        const char* s = "0.00000";
        for (int i = 0; i < 7; i++) o[n++] = s[i];
        return n;
    }
    int e;
    if (be == 0) {
        e = -1074;
    } else {
        m |= (1ull << 52);
        e = be - 1075;
    }
    // estimate X = floor(log10(v)) from the binary exponent, then correct
    int bl = 0;
    for (uint64_t t = m; t; t >>= 1) bl++;
    int l2 = e + bl - 1;  // floor(log2 v)
    int X = (l2 * 1233) >> 12;  // floor(l2 * log10 2), arithmetic shift floors negatives
    // exact X = floor(log10 v) from the truncated scaling, then round
    for (int iter = 0; iter < 4; iter++) {
        uint64_t Dt = round_scaled(m, e, 5 - X, true);
        if (Dt >= 1000000ull)
            X += 1;
        else if (Dt < 100000ull)
            X -= 1;
        else
            break;
    }
    uint64_t D = round_scaled(m, e, 5 - X, false);
    if (D >= 1000000ull) {
        if (X == 5) {
            // glibc keeps the fixed-style precision (0 decimals) when rounding
            // carries 999999.5.. up to 10^6, printing "1.e+06"
            const char* s = "1.e+06";
            for (int i = 0; i < 6; i++) o[n++] = s[i];
            return n;
        }
        X += 1;
        D = 100000ull;
    }
    return fmt_g_digits(D, X, o, n);
}

// %#g as glibc's printf writes it (6 significant digits, trailing zeros and
// the point kept, round half to even on the exact binary value).  Values in
// [1, 10^6) -- copy numbers -- take an FP64 path (fmt_g_fast): X from exact
// powers of ten, y = v * 10^(5-X) rounded once (error below 2^-33 since
// y < 2^20), and the digit rounding is exact unless y sits within that error
// of a half, which goes to the exact routine.
// The cases settled without the exact routine; -1 for the others.
// The six rounded digits D and the exponent X (0..5) of v's %#g text when the
// FP64 path settles them; false for the others.
GSNP_HD bool g_fast_digits(double v, uint32_t& D, int& X)
{
    if (!(v >= 1.0 && v < 1e6)) return false;
    X = v < 10.0 ? 0 : v < 100.0 ? 1 : v < 1e3 ? 2 : v < 1e4 ? 3 : v < 1e5 ? 4 : 5;
    const double scale = X == 0 ? 1e5 : X == 1 ? 1e4 : X == 2 ? 1e3 : X == 3 ? 1e2 : X == 4 ? 1e1 : 1.0;
    const double y = v * scale;
    const double fl = floor(y);
    const double frac = y - fl;  // exact: y < 2^20
    if (!(fabs(frac - 0.5) > 2.5e-10)) return false;
    const uint64_t d = (uint64_t)fl + (frac > 0.5 ? 1u : 0u);
    if (d >= 1000000ull) return false;
    D = (uint32_t)d;
    return true;
}

GSNP_HD int fmt_g_fast(double v, char* o)
{
    uint32_t D;
    int X;
    return g_fast_digits(v, D, X) ? fmt_g_digits(D, X, o, 0) : -1;
}

GSNP_HD int fmt_g(double v, char* o)
{
    const int n = fmt_g_fast(v, o);
    return n >= 0 ? n : fmt_g_exact(v, o);
}

// The %#g text of a row's rank-sum and copy-number columns when either needs
// the exact routine (written by a pre-pass, so the row kernels stay small).
// Byte 15 of each field holds its length.
struct GText {
    char rank[16];
    char copy[16];
    uint32_t site;  // the row's site
};

}  // namespace gsnp
