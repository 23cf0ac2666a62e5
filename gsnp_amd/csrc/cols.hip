// cols.hip -- GSNP's -C column files, encoded on the device (Column_* and
// extract_*, cpp/compress.h:33-1312; codecs kernel/compress_kernel.cu:85-393
// and cpp/dict_compression.h:47-296).  Every reference window of a called
// span becomes one chunk in each of the five column files; the device builds
// those chunks byte for byte and the host only writes them out.
//
// Stages over a span cut into windows of GSNP_WIN sites (window w of every
// per-window array starts at w*GSNP_WIN):
//   k_col_count   per 256-site block: 16 counters (the sparse lists, the N
//                 lists of columns 3 and 10, the run heads of the six
//                 run-length columns, the sites that are not N rows)
//   k_col_scan    per window: exclusive scan of the block counters
//   k_col_emit    per site: the sparse lists in site order, run values and
//                 starts, the 2-bit words of columns 3 and 10
//   k_col_runlen  run lengths from the run starts
//   k_dict_setbits, k_dict_rank, k_dict_pack
//                 the sorted-unique dictionary and packed ranks of the 14
//                 integer arrays, over several workgroups per array: a
//                 presence bitmap of the values, its prefix popcounts, then
//                 each element's rank
//   k_col_dict    per (window, array), one workgroup: the same for the
//                 copy numbers (doubles) and any integer array the bitmap
//                 does not cover (a hash set gives the distinct values)
//   k_col_sizes   chunk sizes and their offsets in each file
//   k_col_write   the chunks
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/gsnp_abi.h"
#include "cols.h"
#include "cwriter.h"
#include "engine.h"

namespace gsnp {

namespace {

constexpr uint32_t kWin = GSNP_WIN;
constexpr uint32_t kCB = 256;               // sites per block of the per-site kernels
constexpr uint32_t kWinBlocks = kWin / kCB;  // 1024
constexpr int kNC = 16;                      // counters per block
constexpr int kNDict = 15;                   // dictionary-coded arrays per window
constexpr uint32_t kTabStride = 2 * kWin;    // hash slots per (window, array)
constexpr uint32_t kDictThreads = 1024;
constexpr int kLdsLg = 13;
constexpr uint32_t kLdsSlots = 1u << kLdsLg;  // LDS hash set of a dictionary: up to kLdsSlots/2 distinct values
constexpr uint32_t kBmpWords = kLdsSlots;    // presence bitmap of small-range integer arrays, in the hash set's LDS
constexpr uint32_t kBmpBits = 32 * kBmpWords;  // 262,144 values: any window's
constexpr uint32_t kSubWindows = 32;         // windows per encoder pass
constexpr uint64_t kEmpty = ~0ull;           // never a key: int keys use 32 bits, f64 keys are not NaN

// counters: sparse lists, N lists, non-N sites, then the run heads of
// columns 5, 7, 8, 9, 14, 16
enum : int { kC4 = 0, kC6, kCN3, kCN10, kCCom, kCX12, kCX13, kC15, kC17, kCNotN, kCRun0 };

// dictionary-coded arrays of a window: 0-4 run values of columns 5, 7, 8, 9,
// 14 (int32), 5 run values of column 16 (f64), 6-11 the six run lengths,
// 12-14 columns 11, 12, 13 of the common list

__constant__ char kColACTGN[8] = {'A', 'C', 'T', 'G', 'N', 'N', 'N', 'N'};
__constant__ char kColAbbv[17] = {'A', 'M', 'W', 'R', 'M', 'C', 'Y', 'S', 'W',
                                  'Y', 'T', 'K', 'R', 'S', 'K', 'G', 'N'};
__constant__ uint32_t kColCode2[4] = {0, 1, 3, 2};  // 2-bit code (A0 C1 G2 T3) of "ACTG"[i]

struct ColIn {
    const uint8_t* ref;  // reference codes of site 0
    SiteArrays a;
    Results res;
    uint32_t n;
};

struct ColSite {
    bool nflag, nrow, n3, n10, com, x12, x13, f15, f17;
    uint32_t c3, c4, c6, code3, code10;
    int32_t v[5];
    double cn;
    int32_t a, b, d;
    double rank;
};

__device__ __forceinline__ int pick(int4 v, uint32_t b) { return b == 0 ? v.x : b == 1 ? v.y : b == 2 ? v.z : v.w; }

// The column values of site j: extract_3_4_6 (cpp/compress.h:87-185),
// extract_5/7/8/9/14/16 (:413-500), extract_10 (:899-913) and
// extract_11_12_13 (:915-944), with the N row of call_cns2's -C branch.
__device__ ColSite col_site(const ColIn& in, uint32_t j)
{
    ColSite s;
    const uint32_t ori = in.ref[j] & 0xF;
    const int32_t depth = in.a.depth[j];
    const uint32_t b1 = in.res.base1[j], b2 = in.res.base2[j];
    s.nflag = (ori & 4) && depth == 0;
    s.nrow = s.nflag || b1 >= 4;
    s.c3 = s.nrow ? 'N' : (uint32_t)kColACTGN[ori & 7];
    s.c4 = s.nrow ? 'N' : (uint32_t)kColAbbv[min((uint32_t)in.res.type1[j], 16u)];
    s.c6 = s.nrow ? 'N' : (uint32_t)kColACTGN[b1 & 7];
    s.n3 = s.c3 == 'N';
    s.code3 = s.n3 ? 0u : kColCode2[ori & 3];
    s.n10 = s.nrow || b2 >= 4;
    s.code10 = s.n10 ? 0u : kColCode2[b2 & 3];
    const int4 qs = in.a.qs[j], cu = in.a.cu[j], ca = in.a.ca[j];
    if (!s.nrow) {
        const int qsb = pick(qs, b1), cub = pick(cu, b1);
        s.v[0] = in.res.q_cns[j];
        s.v[1] = (qsb == 0 || cub == 0) ? 0 : qsb / cub;
        s.v[2] = cub;
        s.v[3] = pick(ca, b1);
        s.v[4] = depth;
        s.cn = depth == 0 ? 255.0 : (double)in.a.repeat[j] / (double)depth;
    } else {
        s.v[0] = s.v[1] = s.v[2] = s.v[3] = s.v[4] = 0;
        s.cn = 255.0;
    }
    s.a = s.b = s.d = 0;
    if (!s.n10) {
        const int qsb = pick(qs, b2);
        s.b = pick(cu, b2);
        s.a = (qsb == 0 || s.b == 0) ? 0 : qsb / s.b;
        s.d = pick(ca, b2);
    }
    s.com = s.a != 0;
    s.x12 = !s.com && s.a != s.b;
    s.x13 = !s.com && !s.x12 && s.a != s.d;
    s.rank = in.res.rank[j];
    s.f15 = !s.nrow && s.rank != 1.0;
    s.f17 = !s.nrow && (ori & 8);
    return s;
}

// Site j's counter bits (bit c for counter c); run heads compare with the
// previous site of the window through LDS.  Every thread of the block calls it.
__device__ uint32_t col_flags(const ColIn& in, uint32_t j, bool ok, ColSite& s, int32_t (*sv)[kCB], double* sd)
{
    const uint32_t t = threadIdx.x;
    uint32_t f = 0;
    if (ok) {
        s = col_site(in, j);
        f |= (uint32_t)(s.c4 != s.c3) << kC4;
        f |= (uint32_t)(s.c6 != s.c3) << kC6;
        f |= (uint32_t)s.n3 << kCN3;
        f |= (uint32_t)s.n10 << kCN10;
        f |= (uint32_t)s.com << kCCom;
        f |= (uint32_t)s.x12 << kCX12;
        f |= (uint32_t)s.x13 << kCX13;
        f |= (uint32_t)s.f15 << kC15;
        f |= (uint32_t)s.f17 << kC17;
        f |= (uint32_t)!s.nflag << kCNotN;
        for (int c = 0; c < 5; c++) sv[c][t] = s.v[c];
        sd[t] = s.cn;
    }
    __syncthreads();
    if (ok) {
        const uint32_t i = j % kWin;
        if (i == 0) {
            f |= 0x3Fu << kCRun0;
        } else {
            int32_t pv[5];
            double pc;
            if (t > 0) {
                for (int c = 0; c < 5; c++) pv[c] = sv[c][t - 1];
                pc = sd[t - 1];
            } else {
                const ColSite p = col_site(in, j - 1);
                for (int c = 0; c < 5; c++) pv[c] = p.v[c];
                pc = p.cn;
            }
            for (int c = 0; c < 5; c++) f |= (uint32_t)(s.v[c] != pv[c]) << (kCRun0 + c);
            f |= (uint32_t)(s.cn != pc) << (kCRun0 + 5);
        }
    }
    return f;
}

__global__ __launch_bounds__(256) void k_col_count(ColIn in, uint32_t* cnt)
{
    __shared__ int32_t sv[5][kCB];
    __shared__ double sd[kCB];
    __shared__ uint32_t wc[4][kNC];
    const uint32_t t = threadIdx.x, j = blockIdx.x * kCB + t;
    ColSite s;
    const uint32_t f = col_flags(in, j, j < in.n, s, sv, sd);
    for (int c = 0; c < kNC; c++) {
        const uint64_t b = __ballot((f >> c) & 1u);
        if ((t & 63) == 0) wc[t >> 6][c] = (uint32_t)__popcll(b);
    }
    __syncthreads();
    if (t < kNC) cnt[(size_t)blockIdx.x * kNC + t] = wc[0][t] + wc[1][t] + wc[2][t] + wc[3][t];
}

// per window: exclusive scan of each counter over the window's blocks
__global__ __launch_bounds__(256) void k_col_scan(const uint32_t* cnt, uint32_t n, uint32_t* boff, uint32_t* tot)
{
    __shared__ uint32_t wsum[4];
    const uint32_t w = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const uint32_t C = min(kWin, n - w * kWin);
    const uint32_t nb = (C + kCB - 1) / kCB;
    const size_t b0 = (size_t)w * kWinBlocks;
    for (int c = 0; c < kNC; c++) {
        uint32_t v[4], sum = 0;
        for (int k = 0; k < 4; k++) {
            const uint32_t b = t * 4 + k;
            v[k] = b < nb ? cnt[(b0 + b) * kNC + c] : 0u;
            sum += v[k];
        }
        uint32_t incl = sum;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off);
            if ((int)lane >= off) incl += y;
        }
        if (lane == 63) wsum[wv] = incl;
        __syncthreads();
        uint32_t run = incl - sum;
        for (uint32_t k = 0; k < wv; k++) run += wsum[k];
        for (int k = 0; k < 4; k++) {
            const uint32_t b = t * 4 + k;
            if (b < nb) boff[(b0 + b) * kNC + c] = run;
            run += v[k];
        }
        if (t == 255) tot[(size_t)w * kNC + c] = run;
        __syncthreads();
    }
}

struct EmitOut {
    int32_t *id4, *id6, *idn, *nidx, *x12i, *x12v, *x13i, *x13v, *cid, *id15, *id17;
    uint8_t *v4, *v6;
    uint32_t *code3, *code10, *rstart, *in32;
    uint64_t* in64;
    double* val15;
    size_t stride;  // elements per array of rstart / in32
};

__global__ __launch_bounds__(256) void k_col_emit(ColIn in, const uint32_t* boff, EmitOut o)
{
    __shared__ int32_t sv[5][kCB];
    __shared__ double sd[kCB];
    __shared__ uint32_t wc[4][kNC];
    const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const uint32_t j = blockIdx.x * kCB + t;
    const bool ok = j < in.n;
    ColSite s;
    const uint32_t f = col_flags(in, j, ok, s, sv, sd);
    uint64_t bal[kNC];
    for (int c = 0; c < kNC; c++) {
        bal[c] = __ballot((f >> c) & 1u);
        if (lane == 0) wc[wv][c] = (uint32_t)__popcll(bal[c]);
    }
    __syncthreads();
    const uint32_t w = j / kWin, i = j % kWin;
    const size_t seg = (size_t)w * kWin;
    const uint64_t lt = (1ull << lane) - 1ull;
    auto pos = [&](int c) -> size_t {
        uint32_t p = boff[(size_t)blockIdx.x * kNC + c] + (uint32_t)__popcll(bal[c] & lt);
        for (uint32_t k = 0; k < wv; k++) p += wc[k][c];
        return seg + p;
    };
    if (ok) {
        if (f & (1u << kC4)) {
            const size_t p = pos(kC4);
            o.id4[p] = (int32_t)i;
            o.v4[p] = (uint8_t)s.c4;
        }
        if (f & (1u << kC6)) {
            const size_t p = pos(kC6);
            o.id6[p] = (int32_t)i;
            o.v6[p] = (uint8_t)s.c6;
        }
        if (f & (1u << kCN3)) o.idn[pos(kCN3)] = (int32_t)i;
        if (f & (1u << kCN10)) o.nidx[pos(kCN10)] = (int32_t)i;
        if (f & (1u << kCCom)) {
            const size_t p = pos(kCCom);
            o.cid[p] = (int32_t)i;
            o.in32[12 * o.stride + p] = (uint32_t)s.a;
            o.in32[13 * o.stride + p] = (uint32_t)s.b;
            o.in32[14 * o.stride + p] = (uint32_t)s.d;
        }
        if (f & (1u << kCX12)) {
            const size_t p = pos(kCX12);
            o.x12i[p] = (int32_t)i;
            o.x12v[p] = s.b;
        }
        if (f & (1u << kCX13)) {
            const size_t p = pos(kCX13);
            o.x13i[p] = (int32_t)i;
            o.x13v[p] = s.d;
        }
        if (f & (1u << kC15)) {
            const size_t p = pos(kC15);
            o.id15[p] = (int32_t)i;
            o.val15[p] = s.rank;
        }
        if (f & (1u << kC17)) o.id17[pos(kC17)] = (int32_t)i;
        for (int c = 0; c < 6; c++) {
            if (!(f & (1u << (kCRun0 + c)))) continue;
            const size_t p = pos(kCRun0 + c);
            o.rstart[c * o.stride + p] = i;
            if (c < 5)
                o.in32[c * o.stride + p] = (uint32_t)s.v[c];
            else
                o.in64[p] = (uint64_t)__double_as_longlong(s.cn);
        }
    }
    // 2-bit words of columns 3 and 10: 16 sites per word, the first site in
    // the high bits, the window's last word right-aligned
    // (kernel/compress_kernel.cu:104-153)
    const uint32_t C = ok ? min(kWin, in.n - w * kWin) : 0u;
    const uint32_t last = min((i & ~15u) + 15u, C - 1u);
    uint32_t x3 = ok ? s.code3 << (2 * (last - i)) : 0u;
    uint32_t x10 = ok ? s.code10 << (2 * (last - i)) : 0u;
    for (int m = 1; m < 16; m <<= 1) {
        x3 |= __shfl_xor(x3, m);
        x10 |= __shfl_xor(x10, m);
    }
    if (ok && (i & 15) == 0) {
        o.code3[(size_t)w * (kWin / 16) + i / 16] = x3;
        o.code10[(size_t)w * (kWin / 16) + i / 16] = x10;
    }
}

// one thread per run: grid (runs / 256, column, window)
__global__ __launch_bounds__(256) void k_col_runlen(uint32_t n, const uint32_t* tot, const uint32_t* rstart,
                                                    uint32_t* in32, size_t stride)
{
    const uint32_t c = blockIdx.y, w = blockIdx.z;
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nrun = tot[(size_t)w * kNC + kCRun0 + c];
    if (r >= nrun) return;
    const uint32_t C = min(kWin, n - w * kWin);
    const uint32_t* rs = rstart + c * stride + (size_t)w * kWin;
    in32[(6 + c) * stride + (size_t)w * kWin + r] = (r + 1 < nrun ? rs[r + 1] : C) - rs[r];
}

struct DictArgs {
    const uint32_t* tot;
    const uint32_t* in32;
    const uint64_t* in64;
    size_t stride;
    uint64_t* tab;
    uint32_t* trank;
    uint64_t* dkeys;
    uint32_t* dslot;
    uint32_t* eslot;  // the slot of each element
    uint32_t* words;
    uint32_t* dict32;
    uint64_t* dict64;
    uint32_t* nd;
    uint32_t* nw;
    const uint32_t* state;  // [wk][3] of the split integer path: done flags at [2] (null: not run)
    uint32_t lds_max;       // distinct values kept in the LDS set before falling back to HBM (0: HBM only)
    uint32_t bitmap;        // integer arrays of small range through the LDS presence bitmap
    uint8_t pow2bits[32];  // index bits of a dictionary of 2^e entries, as the host's libm gives them
};

// Distinct values above this count are ranked by sorting them (below it, by
// counting the smaller ones: O(distinct^2) compares, cheap for a few hundred).
constexpr uint32_t kRankSortMin = 512;

// unsigned keys in the order of the values: int32 bits (zero-extended), or
// double bits with negatives below positives and their order reversed
__device__ __forceinline__ uint64_t order_key(uint64_t raw, bool dbl)
{
    if (!dbl) return (uint32_t)raw ^ 0x80000000u;
    return (raw >> 63) ? ~raw : (raw | 0x8000000000000000ull);
}
__device__ __forceinline__ uint64_t raw_key(uint64_t k, bool dbl)
{
    if (!dbl) return (uint32_t)k ^ 0x80000000u;
    return (k >> 63) ? (k & ~0x8000000000000000ull) : ~k;
}

__device__ __forceinline__ uint32_t slot_of(uint64_t key, int lg)
{
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - lg));
}

__device__ __forceinline__ uint64_t tab_load(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One workgroup per (window, array): distinct values through an open-addressed
// hash set (load at most 1/2, so every probe sequence meets a free slot), each
// distinct value's rank = the count of distinct values below it, then the
// ranks packed floor(32/bits) per word, MSB first, the last word right-aligned
// (dict_encode, cpp/dict_compression.h:116-166).  The set lives in LDS while
// it holds at most lds_max values (the usual case: a window's arrays have a
// few hundred distinct values among up to 262,144 elements); past that the
// workgroup starts again with a set in HBM sized for every element.
__global__ __launch_bounds__(kDictThreads) void k_col_dict(DictArgs a)
{
    __shared__ uint32_t s_nd, s_over;
    __shared__ unsigned long long ltab[kLdsSlots];
    __shared__ uint32_t lrank[kLdsSlots];
    __shared__ uint64_t tile[1024];
    const uint32_t w = blockIdx.x, k = blockIdx.y, t = threadIdx.x;
    const uint32_t* tw = a.tot + (size_t)w * kNC;
    const bool all_n = tw[kCNotN] == 0;
    const uint32_t m = all_n ? 0u : k < 12 ? tw[kCRun0 + k % 6] : tw[kCCom];
    const size_t wk = (size_t)w * kNDict + k;
    if (a.state && a.state[3 * wk + 2]) return;  // encoded by the split integer path
    if (m == 0) {
        if (t == 0) a.nd[wk] = a.nw[wk] = 0;
        return;
    }
    const bool dbl = k == 5;
    const uint32_t* in = a.in32 + k * a.stride + (size_t)w * kWin;
    const uint64_t* in64 = a.in64 + (size_t)w * kWin;
    auto key_at = [&](uint32_t e) -> uint64_t { return dbl ? in64[e] : (uint64_t)in[e]; };
    if (!dbl && a.bitmap) {
        // integer arrays whose values span fewer than kBmpBits: a presence
        // bitmap over [min, min + kBmpBits) in LDS (ltab's space) gives the
        // distinct values in order; a value's rank is the set bits below it
        // (the word prefix in lrank's space plus a popcount in its word)
        // (every pass over the elements keeps eight loads in flight per thread:
        // one at a time, the passes wait on memory latency)
        constexpr uint32_t U = 8;
        __shared__ int32_t s_lo, s_hi;
        int32_t lo = INT32_MAX, hi = INT32_MIN;
        for (uint32_t e0 = t; e0 < m; e0 += U * kDictThreads) {
            int32_t v[U];
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                const uint32_t e = e0 + u * kDictThreads;
                v[u] = e < m ? (int32_t)in[e] : (int32_t)in[e0];
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                lo = min(lo, v[u]);
                hi = max(hi, v[u]);
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            lo = min(lo, __shfl_xor(lo, off));
            hi = max(hi, __shfl_xor(hi, off));
        }
        if (t == 0) s_lo = INT32_MAX, s_hi = INT32_MIN;
        __syncthreads();
        if ((t & 63) == 0) {
            atomicMin(&s_lo, lo);
            atomicMax(&s_hi, hi);
        }
        __syncthreads();
        const int32_t vmin = s_lo;
        const int64_t range = (int64_t)s_hi - (int64_t)vmin + 1;
        if (range <= (int64_t)kBmpBits) {
            uint32_t* bits = reinterpret_cast<uint32_t*>(ltab);
            uint32_t* wpre = lrank;
            const uint32_t nbw = (uint32_t)((range + 31) / 32);
            for (uint32_t i = t; i < nbw; i += kDictThreads) bits[i] = 0u;
            __syncthreads();
            for (uint32_t e0 = t; e0 < m; e0 += U * kDictThreads) {
                int32_t v[U];
#pragma unroll
                for (uint32_t u = 0; u < U; u++) {
                    const uint32_t e = e0 + u * kDictThreads;
                    v[u] = e < m ? (int32_t)in[e] : (int32_t)in[e0];
                }
#pragma unroll
                for (uint32_t u = 0; u < U; u++) {
                    const uint32_t d = (uint32_t)(v[u] - vmin);
                    const uint32_t b = 1u << (d & 31u);
                    if (!(bits[d >> 5] & b)) atomicOr(&bits[d >> 5], b);  // most values are already present
                }
            }
            __syncthreads();
            // exclusive prefix of the words' popcounts: kBmpWords / kDictThreads
            // words per thread, then a block scan of the thread sums
            constexpr uint32_t PW = kBmpWords / kDictThreads;
            uint32_t c[PW], sum = 0;
#pragma unroll
            for (uint32_t j = 0; j < PW; j++) {
                const uint32_t i = t * PW + j;
                c[j] = i < nbw ? (uint32_t)__popc(bits[i]) : 0u;
                sum += c[j];
            }
            uint32_t incl = sum;
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t y = __shfl_up(incl, off);
                if ((int)(t & 63) >= off) incl += y;
            }
            uint32_t* wsum = reinterpret_cast<uint32_t*>(tile);
            if ((t & 63) == 63) wsum[t >> 6] = incl;
            __syncthreads();
            uint32_t run = incl - sum;
            for (uint32_t q = 0; q < (t >> 6); q++) run += wsum[q];
            if (t == kDictThreads - 1) s_nd = run + sum;
#pragma unroll
            for (uint32_t j = 0; j < PW; j++) {
                const uint32_t i = t * PW + j;
                if (i < nbw) {
                    wpre[i] = run;
                    uint32_t x = bits[i], r = run;
                    while (x) {  // the dictionary, in order
                        a.dict32[wk * kWin + r++] = (uint32_t)(vmin + (int32_t)(i * 32u + (uint32_t)(__ffs(x) - 1)));
                        x &= x - 1u;
                    }
                }
                run += c[j];
            }
            __syncthreads();
            const uint32_t nd = s_nd;
            const int bits_per = (nd & (nd - 1)) == 0 ? (int)a.pow2bits[__ffs(nd) - 1] : 32 - __clz(nd - 1);
            const uint32_t per = 32u / (uint32_t)bits_per;
            const uint32_t nwords = (m + per - 1) / per;
            uint32_t* wo = a.words + wk * kWin;
            // each thread packs a contiguous run of whole words, sixteen
            // elements in flight per step (a word at a time left only as many
            // loads in flight as a word holds elements)
            constexpr uint32_t PK = 16;
            const uint32_t wpt = (nwords + kDictThreads - 1) / kDictThreads;
            const uint32_t q0 = t * wpt, q1 = min(q0 + wpt, nwords);
            if (q0 < q1) {
                const uint32_t eb0 = q0 * per, eb1 = min(q1 * per, m);
                uint32_t acc = 0, k = 0, q = q0;  // k: ranks in acc
                for (uint32_t eb = eb0; eb < eb1; eb += PK) {
                    int32_t v[PK];
#pragma unroll
                    for (uint32_t u = 0; u < PK; u++) v[u] = eb + u < eb1 ? (int32_t)in[eb + u] : vmin;
#pragma unroll
                    for (uint32_t u = 0; u < PK; u++) {
                        if (eb + u >= eb1) break;
                        const uint32_t d = (uint32_t)(v[u] - vmin);
                        const uint32_t r = wpre[d >> 5] + (uint32_t)__popc(bits[d >> 5] & ((1u << (d & 31u)) - 1u));
                        acc = bits_per >= 32 ? 0u : (acc << bits_per);
                        acc |= r;
                        if (++k == per || eb + u + 1 == m) {  // a full word, or the window's last (right-aligned)
                            wo[q++] = acc;
                            acc = 0;
                            k = 0;
                        }
                    }
                }
            }
            if (t == 0) {
                a.nd[wk] = nd;
                a.nw[wk] = nwords;
            }
            return;
        }
        __syncthreads();  // s_lo / s_hi read by every thread before the hash path reuses LDS
    }
    uint32_t cap = 64;
    while (cap < 2 * m) cap <<= 1;
    const int lg = __ffs(cap) - 1;
    uint64_t* tab = a.tab + wk * kTabStride;
    uint32_t* trank = a.trank + wk * kTabStride;
    uint64_t* dk = a.dkeys + wk * kWin;
    uint32_t* ds = a.dslot + wk * kWin;
    uint32_t* es = a.eslot + wk * kWin;
    bool lds = a.lds_max > 0;
    if (lds) {
        const uint32_t limit = min(a.lds_max, kLdsSlots / 2);
        if (t == 0) s_nd = s_over = 0;
        for (uint32_t h = t; h < kLdsSlots; h += kDictThreads) ltab[h] = kEmpty;
        __syncthreads();
        const volatile uint32_t* over = &s_over;
        // four elements per step, their keys loaded before any is inserted
        for (uint32_t e0 = t; e0 < m && !*over; e0 += 4 * kDictThreads) {
            uint64_t keys[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t e = e0 + u * kDictThreads;
                keys[u] = e < m ? key_at(e) : 0ull;
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t e = e0 + u * kDictThreads;
                if (e >= m) break;
                const uint64_t key = keys[u];
                uint32_t h = slot_of(key, kLdsLg);
                for (uint32_t probe = 0; probe < kLdsSlots; probe++) {
                    unsigned long long cur = ltab[h];
                    if (cur == kEmpty) {
                        cur = atomicCAS(&ltab[h], (unsigned long long)kEmpty, (unsigned long long)key);
                        if (cur == kEmpty) {
                            const uint32_t p = atomicAdd(&s_nd, 1u);
                            if (p < limit) {
                                dk[p] = key;
                                ds[p] = h;
                            } else {
                                s_over = 1;
                            }
                            break;
                        }
                    }
                    if (cur == key) break;
                    h = (h + 1) & (kLdsSlots - 1);
                }
                es[e] = h;
            }
        }
        __syncthreads();
        lds = s_over == 0;
    }
    if (!lds) {
        __syncthreads();  // every thread has read s_over
        if (t == 0) s_nd = 0;
        for (uint32_t h = t; h < cap; h += kDictThreads) tab[h] = kEmpty;
        __syncthreads();
    }
    // (four keys loaded before any is inserted, as in the LDS set)
    for (uint32_t e0 = t; e0 < m && !lds; e0 += 4 * kDictThreads) {
        uint64_t keys[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t e = e0 + u * kDictThreads;
            keys[u] = e < m ? key_at(e) : 0ull;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t e = e0 + u * kDictThreads;
            if (e >= m) break;
            const uint64_t key = keys[u];
            uint32_t h = slot_of(key, lg);
            for (uint32_t probe = 0; probe < cap; probe++) {
                uint64_t cur = tab_load(&tab[h]);
                if (cur == kEmpty) {
                    cur = atomicCAS(reinterpret_cast<unsigned long long*>(&tab[h]), (unsigned long long)kEmpty,
                                    (unsigned long long)key);
                    if (cur == kEmpty) {
                        const uint32_t p = atomicAdd(&s_nd, 1u);
                        dk[p] = key;
                        ds[p] = h;
                        break;
                    }
                }
                if (cur == key) break;
                h = (h + 1) & (cap - 1);
            }
            es[e] = h;
        }
    }
    __syncthreads();
    const uint32_t nd = s_nd;
    if (nd > kRankSortMin) {
        // many distinct values: sort (order key, slot) pairs of the distinct
        // values in place (a bitonic network over the padded power of two,
        // one block barrier per step); a value's rank is its sorted position
        uint32_t npad = 1;
        while (npad < nd) npad <<= 1;
        // in LDS when the set was (the hash keys are no longer needed: the
        // first half of their space takes the order keys, the second the
        // slots), else in the distinct-key arrays in HBM
        uint64_t* sk = lds ? reinterpret_cast<uint64_t*>(ltab) : dk;
        uint32_t* ss = lds ? reinterpret_cast<uint32_t*>(ltab + kLdsSlots / 2) : ds;
        for (uint32_t d = t; d < npad; d += kDictThreads) {
            sk[d] = d < nd ? order_key(dk[d], dbl) : ~0ull;
            ss[d] = d < nd ? ds[d] : 0u;
        }
        __syncthreads();
        for (uint32_t kk = 2; kk <= npad; kk <<= 1)
            for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
                for (uint32_t i = t; i < npad; i += kDictThreads) {
                    const uint32_t l = i ^ j;
                    if (l <= i) continue;
                    const uint64_t x = sk[i], y = sk[l];
                    if ((i & kk) == 0 ? x > y : x < y) {
                        sk[i] = y;
                        sk[l] = x;
                        const uint32_t si = ss[i];
                        ss[i] = ss[l];
                        ss[l] = si;
                    }
                }
                __syncthreads();
            }
        for (uint32_t r = t; r < nd; r += kDictThreads) {
            const uint64_t x = raw_key(sk[r], dbl);
            if (lds)
                lrank[ss[r]] = r;
            else
                trank[ss[r]] = r;
            if (dbl)
                a.dict64[(size_t)w * kWin + r] = x;
            else
                a.dict32[wk * kWin + r] = (uint32_t)x;
        }
    }
    for (uint32_t base = 0; base < nd && nd <= kRankSortMin; base += kDictThreads) {
        const uint32_t d = base + t;
        const uint64_t x = d < nd ? dk[d] : 0ull;
        uint32_t r = 0;
        for (uint32_t t0 = 0; t0 < nd; t0 += 1024) {
            const uint32_t len = min(1024u, nd - t0);
            for (uint32_t u = t; u < len; u += kDictThreads) tile[u] = dk[t0 + u];
            __syncthreads();
            if (d < nd) {
                if (dbl) {
                    const double xd = __longlong_as_double((long long)x);
                    for (uint32_t u = 0; u < len; u++) r += __longlong_as_double((long long)tile[u]) < xd;
                } else {
                    const int32_t xi = (int32_t)(uint32_t)x;
                    for (uint32_t u = 0; u < len; u++) r += (int32_t)(uint32_t)tile[u] < xi;
                }
            }
            __syncthreads();
        }
        if (d < nd) {
            if (lds)
                lrank[ds[d]] = r;
            else
                trank[ds[d]] = r;
            if (dbl)
                a.dict64[(size_t)w * kWin + r] = x;
            else
                a.dict32[wk * kWin + r] = (uint32_t)x;
        }
    }
    __syncthreads();
    const int bits = (nd & (nd - 1)) == 0 ? (int)a.pow2bits[__ffs(nd) - 1] : 32 - __clz(nd - 1);
    const uint32_t per = 32u / (uint32_t)bits;
    const uint32_t nwords = (m + per - 1) / per;
    uint32_t* wo = a.words + wk * kWin;
    // a contiguous run of whole words per thread, sixteen slots in flight
    constexpr uint32_t PK = 16;
    const uint32_t wpt = (nwords + kDictThreads - 1) / kDictThreads;
    const uint32_t q0 = t * wpt, q1 = min(q0 + wpt, nwords);
    if (q0 < q1) {
        const uint32_t eb0 = q0 * per, eb1 = min(q1 * per, m);
        uint32_t acc = 0, k = 0, q = q0;  // k: ranks in acc
        for (uint32_t eb = eb0; eb < eb1; eb += PK) {
            uint32_t sl[PK];
#pragma unroll
            for (uint32_t u = 0; u < PK; u++) sl[u] = eb + u < eb1 ? es[eb + u] : 0u;
#pragma unroll
            for (uint32_t u = 0; u < PK; u++) {
                if (eb + u >= eb1) break;
                acc = bits >= 32 ? 0u : (acc << bits);
                acc |= lds ? lrank[sl[u]] : trank[sl[u]];
                if (++k == per || eb + u + 1 == m) {  // a full word, or the window's last (right-aligned)
                    wo[q++] = acc;
                    acc = 0;
                    k = 0;
                }
            }
        }
    }
    if (t == 0) {
        a.nd[wk] = nd;
        a.nw[wk] = nwords;
    }
}

// The dictionary-coded arrays are encoded by several workgroups each:
//
// Integer arrays whose values all lie in [0, kBmpBits) (every one on real
// data), in three launches:
//  k_dict_setbits  kDictParts workgroups per array set the values' bits in an
//                  LDS bitmap of their slice, then OR its non-zero words into
//                  the array's bitmap in HBM, with the largest value and an
//                  out-of-range flag;
//  k_dict_rank     one workgroup per array scans the words' popcounts, writes
//                  the dictionary (the set bits in order), the word prefix and
//                  a copy of the bits, and clears the bitmap and the flags for
//                  the next pass;
//  k_dict_pack     workgroups per array pack tiles of whole words: a tile's
//                  elements are loaded coalesced, 16 per thread in flight, and
//                  ranked (the word prefix plus the set bits below the value
//                  in its word) into LDS, then each thread assembles words.
// The copy numbers (doubles), with at most kHashSortMax distinct values per
// window:
//  k_hash_insert   workgroups per array insert the values into an open-
//                  addressed set in HBM (compare-and-swap), each element
//                  keeping its slot and each new value appended to a list;
//  k_hash_rank     one workgroup per array ranks the listed values (up to
//                  kHashRankMax the count of smaller ones; above that the
//                  position after a bitonic sort of (order key, list index)
//                  pairs in LDS: the keys are distinct, so the two agree),
//                  writes the dictionary and each slot's rank, and empties
//                  the used slots for the next pass;
//  k_dict_pack     as above, with an element's rank from its slot.
// An array these do not cover is left (done flag 0) to k_col_dict.
constexpr uint32_t kDictParts = 16;
constexpr uint32_t kDictPartThreads = 512;
constexpr uint32_t kIntDicts = kNDict - 1;  // all but the copy numbers (doubles, array 5)
constexpr uint32_t kPackK = 16;             // elements per thread of a pack tile
constexpr uint32_t kPackTile = kPackK * kDictPartThreads;
constexpr uint32_t kPackLds = 2048;         // bitmap words a pack workgroup keeps in LDS
constexpr uint32_t kHashRankMax = 1024;     // distinct copy numbers ranked by counting
constexpr uint32_t kHashSortMax = 8192;     // distinct copy numbers ranked by k_hash_rank (sorted above kHashRankMax)
constexpr uint32_t kHashLg = 19;            // slots of a window's set: 2 kWin (load at most 1/2)
constexpr uint32_t kHashSlots = 1u << kHashLg;
static_assert(kHashSlots >= 2 * kWin, "the set holds a window's values at load 1/2");

struct SplitArgs {
    const uint32_t* tot;
    const uint32_t* in32;
    const uint64_t* in64;
    size_t stride;
    uint32_t* gbits;  // [wk][kBmpWords], zero between passes
    uint32_t* rbits;  // [wk][kBmpWords]
    uint32_t* wpre;   // [wk][kBmpWords]
    uint32_t* state;  // [wk][3]: largest value (doubles: distinct count), out of range, done;
                      // the first two zero between passes
    uint64_t* htab;   // [w][kHashSlots], all empty between passes
    uint32_t* hrank;  // [w][kHashSlots]
    uint64_t* hkeys;  // [w][kWin] distinct values in insertion order
    uint32_t* hslot;  // [w][kWin] their slots
    uint32_t* eslot;  // [wk][kWin] each element's slot
    uint32_t* dict32;
    uint64_t* dict64;
    uint32_t* words;
    uint32_t* nd;
    uint32_t* nw;
    uint8_t pow2bits[32];
};

// array index of blockIdx.y in the integer launches (array 5 is skipped)
__device__ __forceinline__ uint32_t int_dict(uint32_t y) { return y < 5 ? y : y + 1; }

__device__ __forceinline__ uint32_t dict_elems(const uint32_t* tot, uint32_t w, uint32_t k)
{
    const uint32_t* tw = tot + (size_t)w * kNC;
    return tw[kCNotN] == 0 ? 0u : k < 12 ? tw[kCRun0 + k % 6] : tw[kCCom];
}

__device__ __forceinline__ int dict_index_bits(uint32_t nd, const uint8_t* pow2bits)
{
    return (nd & (nd - 1)) == 0 ? (int)pow2bits[__ffs(nd) - 1] : 32 - __clz(nd - 1);
}

// grid (window, integer array, part)
__global__ __launch_bounds__(kDictPartThreads) void k_dict_setbits(SplitArgs a)
{
    __shared__ uint32_t bits[kBmpWords];
    __shared__ uint32_t s_max, s_over;
    const uint32_t w = blockIdx.x, k = int_dict(blockIdx.y), p = blockIdx.z, t = threadIdx.x;
    const uint32_t m = dict_elems(a.tot, w, k);
    const uint32_t e0 = (uint32_t)((uint64_t)m * p / kDictParts), e1 = (uint32_t)((uint64_t)m * (p + 1) / kDictParts);
    if (e0 == e1) return;
    const size_t wk = (size_t)w * kNDict + k;
    for (uint32_t i = t; i < kBmpWords; i += kDictPartThreads) bits[i] = 0u;
    if (t == 0) s_max = 0u, s_over = 0u;
    __syncthreads();
    const uint32_t* in = a.in32 + k * a.stride + (size_t)w * kWin;
    uint32_t hi = 0, over = 0;
    constexpr uint32_t U = 16;  // loads in flight per thread
    for (uint32_t e = e0 + t; e < e1; e += U * kDictPartThreads) {
        uint32_t v[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t ee = e + u * kDictPartThreads;
            v[u] = in[ee < e1 ? ee : e];  // a repeat of element e sets its bit again
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            if (v[u] >= kBmpBits) {  // negative ones too, as unsigned
                over = 1u;
                continue;
            }
            hi = max(hi, v[u]);
            const uint32_t b = 1u << (v[u] & 31u);
            if (!(bits[v[u] >> 5] & b)) atomicOr(&bits[v[u] >> 5], b);  // most values are already present
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        hi = max(hi, (uint32_t)__shfl_xor((int)hi, off));
        over |= (uint32_t)__shfl_xor((int)over, off);
    }
    if ((t & 63) == 0) {
        atomicMax(&s_max, hi);
        if (over) atomicOr(&s_over, 1u);
    }
    __syncthreads();
    const uint32_t nbw = s_max / 32u + 1u;
    uint32_t* gb = a.gbits + wk * kBmpWords;
    for (uint32_t i = t; i < nbw; i += kDictPartThreads) {
        const uint32_t x = bits[i];
        if (x) atomicOr(&gb[i], x);
    }
    if (t == 0) {
        atomicMax(&a.state[3 * wk], s_max);
        if (s_over) atomicOr(&a.state[3 * wk + 1], 1u);
    }
}

// block-wide exclusive scan of one value per thread (kDictThreads threads);
// `total` receives the sum
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum, uint32_t* total)
{
    const uint32_t t = threadIdx.x;
    uint32_t incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)incl, off);
        if ((int)(t & 63) >= off) incl += y;
    }
    if ((t & 63) == 63) wsum[t >> 6] = incl;
    __syncthreads();
    uint32_t run = incl - v;
    for (uint32_t q = 0; q < (t >> 6); q++) run += wsum[q];
    if (t == kDictThreads - 1) *total = run + v;
    return run;
}

// grid (window, integer array)
__global__ __launch_bounds__(kDictThreads) void k_dict_rank(SplitArgs a)
{
    __shared__ uint32_t s_nbw, s_over, s_nd;
    __shared__ uint32_t wsum[kDictThreads / 64];
    const uint32_t w = blockIdx.x, k = int_dict(blockIdx.y), t = threadIdx.x;
    const uint32_t m = dict_elems(a.tot, w, k);
    const size_t wk = (size_t)w * kNDict + k;
    if (m == 0) {  // no part touched the bitmap or the flags
        if (t == 0) a.nd[wk] = a.nw[wk] = 0, a.state[3 * wk + 2] = 1u;
        return;
    }
    if (t == 0) {
        s_nbw = a.state[3 * wk] / 32u + 1u;  // every set bit lies below
        s_over = a.state[3 * wk + 1];
    }
    __syncthreads();
    const uint32_t nbw = s_nbw;
    uint32_t* gb = a.gbits + wk * kBmpWords;
    if (s_over) {  // left to k_col_dict; the bitmap is cleared for the next pass
        for (uint32_t i = t; i < nbw; i += kDictThreads) gb[i] = 0u;
        if (t == 0) a.state[3 * wk] = a.state[3 * wk + 1] = a.state[3 * wk + 2] = 0u;
        return;
    }
    uint32_t* rb = a.rbits + wk * kBmpWords;
    uint32_t* wp = a.wpre + wk * kBmpWords;
    // kBmpWords / kDictThreads words per thread, then a block scan of the
    // thread sums
    constexpr uint32_t PW = kBmpWords / kDictThreads;
    uint32_t x[PW], sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < PW; j++) {
        const uint32_t i = t * PW + j;
        x[j] = i < nbw ? gb[i] : 0u;
        sum += (uint32_t)__popc(x[j]);
    }
    uint32_t run = block_excl_scan(sum, wsum, &s_nd);
    uint32_t* dict = a.dict32 + wk * kWin;
#pragma unroll
    for (uint32_t j = 0; j < PW; j++) {
        const uint32_t i = t * PW + j;
        if (i < nbw) {
            rb[i] = x[j];
            wp[i] = run;
            gb[i] = 0u;
            uint32_t y = x[j], r = run;
            while (y) {  // the dictionary, in order
                dict[r++] = i * 32u + (uint32_t)(__ffs(y) - 1);
                y &= y - 1u;
            }
        }
        run += (uint32_t)__popc(x[j]);
    }
    __syncthreads();
    if (t == 0) {
        const uint32_t nd = s_nd;
        const uint32_t per = 32u / (uint32_t)dict_index_bits(nd, a.pow2bits);
        a.nd[wk] = nd;
        a.nw[wk] = (m + per - 1) / per;
        a.state[3 * wk] = a.state[3 * wk + 1] = 0u;
        a.state[3 * wk + 2] = 1u;
    }
}

__device__ __forceinline__ uint32_t hash_slot(uint64_t key)
{
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - kHashLg));
}

// the slot of `key` in the window's set, inserting it (and listing it as
// new) when absent; load at most 1/2, so every probe sequence meets an empty
// slot
__device__ __forceinline__ uint32_t hash_insert(uint64_t* tab, uint64_t key, uint32_t* count, uint64_t* keys,
                                                uint32_t* slots)
{
    uint32_t h = hash_slot(key);
    for (uint32_t probe = 0; probe < kHashSlots; probe++) {
        uint64_t cur = tab_load(&tab[h]);
        if (cur == kEmpty) {
            cur = atomicCAS(reinterpret_cast<unsigned long long*>(&tab[h]), (unsigned long long)kEmpty,
                            (unsigned long long)key);
            if (cur == kEmpty) {
                const uint32_t d = atomicAdd(count, 1u);
                keys[d] = key;
                slots[d] = h;
                return h;
            }
        }
        if (cur == key) return h;
        h = (h + 1u) & (kHashSlots - 1u);
    }
    return h;  // not reached
}

constexpr uint32_t kHashParts = kWin / (8 * kDictPartThreads);  // eight elements per thread: 64
static_assert(kHashParts * 8 * kDictPartThreads == kWin, "the parts cover a window's array");
constexpr uint32_t kHashLdsLg = 11;                             // a workgroup's own set: 2048 slots
constexpr uint32_t kHashLds = 1u << kHashLdsLg;
constexpr uint32_t kHashLdsProbes = 64;  // past this an element goes to the window's set directly

// grid (window, part): the copy numbers (array 5) into the window's set.
// A workgroup first gathers its slice's distinct values in a set of its own
// in LDS, so that only those (a few hundred) reach the window's set: with
// every element inserted there, the few hot values took a compare-and-swap
// per element.
__global__ __launch_bounds__(kDictPartThreads) void k_hash_insert(SplitArgs a)
{
    __shared__ unsigned long long lset[kHashLds];
    __shared__ uint32_t lglob[kHashLds];
    const uint32_t w = blockIdx.x, p = blockIdx.y, t = threadIdx.x;
    const uint32_t m = dict_elems(a.tot, w, 5);
    const uint32_t e0 = p * 8u * kDictPartThreads;
    if (e0 >= m) return;
    const size_t wk = (size_t)w * kNDict + 5;
    for (uint32_t i = t; i < kHashLds; i += kDictPartThreads) lset[i] = kEmpty;
    __syncthreads();
    const uint64_t* in = a.in64 + (size_t)w * kWin;
    uint64_t* tab = a.htab + (size_t)w * kHashSlots;
    uint32_t* es = a.eslot + wk * kWin;
    uint32_t* count = &a.state[3 * wk];
    uint64_t* keys = a.hkeys + (size_t)w * kWin;
    uint32_t* slots = a.hslot + (size_t)w * kWin;
    uint64_t key[8];
    uint32_t ls[8];  // the element's slot in the workgroup's set, or kNoSlot
    constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t u = 0; u < 8; u++) {
        const uint32_t e = e0 + t + u * kDictPartThreads;
        key[u] = e < m ? in[e] : kEmpty;
    }
#pragma unroll
    for (uint32_t u = 0; u < 8; u++) {
        ls[u] = kNoSlot;
        if (key[u] == kEmpty) continue;  // past the array's end (values are never NaN)
        uint32_t h = (uint32_t)((key[u] * 0x9E3779B97F4A7C15ull) >> (64 - kHashLdsLg));
        for (uint32_t probe = 0; probe < kHashLdsProbes; probe++) {
            unsigned long long cur = lset[h];
            if (cur == kEmpty) cur = atomicCAS(&lset[h], (unsigned long long)kEmpty, (unsigned long long)key[u]);
            if (cur == kEmpty || cur == key[u]) {
                ls[u] = h;
                break;
            }
            h = (h + 1u) & (kHashLds - 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = t; i < kHashLds; i += kDictPartThreads) {
        const uint64_t k = lset[i];
        if (k != kEmpty) lglob[i] = hash_insert(tab, k, count, keys, slots);
    }
    __syncthreads();
#pragma unroll
    for (uint32_t u = 0; u < 8; u++) {
        const uint32_t e = e0 + t + u * kDictPartThreads;
        if (e >= m) break;
        es[e] = ls[u] != kNoSlot ? lglob[ls[u]] : hash_insert(tab, key[u], count, keys, slots);
    }
}

// grid (window): rank of each listed copy number, the dictionary in order
__global__ __launch_bounds__(kDictThreads) void k_hash_rank(SplitArgs a)
{
    __shared__ uint64_t ok[kHashSortMax];
    __shared__ uint16_t ix[kHashSortMax];
    const uint32_t w = blockIdx.x, t = threadIdx.x;
    const uint32_t m = dict_elems(a.tot, w, 5);
    const size_t wk = (size_t)w * kNDict + 5;
    uint64_t* tab = a.htab + (size_t)w * kHashSlots;
    const uint64_t* hk = a.hkeys + (size_t)w * kWin;
    const uint32_t* hs = a.hslot + (size_t)w * kWin;
    const uint32_t nd = a.state[3 * wk];  // written by the previous launch only
    if (m == 0 || nd > kHashSortMax) {  // none, or left to k_col_dict: the set is emptied
        for (uint32_t d = t; d < nd; d += kDictThreads) tab[hs[d]] = kEmpty;
        __syncthreads();  // every thread has read the count
        if (t == 0) {
            if (m == 0) a.nd[wk] = a.nw[wk] = 0;
            a.state[3 * wk] = 0u;
            a.state[3 * wk + 2] = m == 0 ? 1u : 0u;
        }
        return;
    }
    uint32_t* rk = a.hrank + (size_t)w * kHashSlots;
    if (nd <= kHashRankMax) {
        for (uint32_t d = t; d < nd; d += kDictThreads) ok[d] = order_key(hk[d], true);
        __syncthreads();
        for (uint32_t d = t; d < nd; d += kDictThreads) {
            const uint64_t x = ok[d];
            uint32_t r = 0;
            for (uint32_t u = 0; u < nd; u++) r += ok[u] < x;  // all lanes read one word: a broadcast
            rk[hs[d]] = r;
            a.dict64[(size_t)w * kWin + r] = hk[d];
            tab[hs[d]] = kEmpty;
        }
    } else {
        // (order key, list index) pairs over the next power of two, padded
        // with keys above every value (no copy number is NaN), sorted
        // ascending; the position of a value is its rank
        uint32_t np = kHashRankMax;
        while (np < nd) np <<= 1;  // <= kHashSortMax
        for (uint32_t d = t; d < np; d += kDictThreads) {
            ok[d] = d < nd ? order_key(hk[d], true) : ~0ull;
            ix[d] = (uint16_t)d;
        }
        __syncthreads();
        for (uint32_t k = 2; k <= np; k <<= 1)
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t q = t; q < np / 2; q += kDictThreads) {
                    const uint32_t i = 2u * j * (q / j) + q % j, l = i + j;  // pair (i, i + j), i & j == 0
                    const uint64_t x = ok[i], y = ok[l];
                    if ((x > y) == ((i & k) == 0)) {
                        ok[i] = y;
                        ok[l] = x;
                        const uint16_t u = ix[i];
                        ix[i] = ix[l];
                        ix[l] = u;
                    }
                }
                __syncthreads();
            }
        for (uint32_t r = t; r < nd; r += kDictThreads) {
            const uint32_t d = ix[r];
            rk[hs[d]] = r;
            a.dict64[(size_t)w * kWin + r] = hk[d];
            tab[hs[d]] = kEmpty;
        }
    }
    __syncthreads();
    if (t == 0) {
        const uint32_t per = 32u / (uint32_t)dict_index_bits(nd, a.pow2bits);
        a.nd[wk] = nd;
        a.nw[wk] = (m + per - 1) / per;
        a.state[3 * wk] = 0u;
        a.state[3 * wk + 2] = 1u;
    }
}

// grid (window, array, part); array 5 takes its ranks from the slots of
// k_hash_insert, the others from their bitmaps
__global__ __launch_bounds__(kDictPartThreads) void k_dict_pack(SplitArgs a)
{
    __shared__ uint32_t lb[kPackLds], lp[kPackLds], rk[kPackTile];
    const uint32_t w = blockIdx.x, k = blockIdx.y, p = blockIdx.z, t = threadIdx.x;
    const uint32_t m = dict_elems(a.tot, w, k);
    const size_t wk = (size_t)w * kNDict + k;
    if (m == 0 || !a.state[3 * wk + 2]) return;
    const uint32_t nd = a.nd[wk], nwords = a.nw[wk];
    const int bits = dict_index_bits(nd, a.pow2bits);
    const uint32_t per = 32u / (uint32_t)bits;
    const uint32_t tw = kPackTile / per;  // whole words per tile
    const uint32_t ntiles = (nwords + tw - 1) / tw;
    const uint32_t t0 = ntiles * p / kDictParts, t1 = ntiles * (p + 1) / kDictParts;
    if (t0 == t1) return;
    const bool dbl = k == 5;
    const uint32_t* gb = a.rbits + wk * kBmpWords;
    const uint32_t* gp = a.wpre + wk * kBmpWords;
    const uint32_t nbw = dbl ? 0u : a.dict32[wk * kWin + nd - 1] / 32u + 1u;  // the largest value's word
    const bool inl = nbw <= kPackLds;
    if (inl)
        for (uint32_t i = t; i < nbw; i += kDictPartThreads) {
            lb[i] = gb[i];
            lp[i] = gp[i];
        }
    const uint32_t* in = a.in32 + k * a.stride + (size_t)w * kWin;
    const uint32_t* es = a.eslot + wk * kWin;
    const uint32_t* hr = a.hrank + (size_t)w * kHashSlots;
    uint32_t* wo = a.words + wk * kWin;
    for (uint32_t tile = t0; tile < t1; tile++) {
        const uint32_t q0 = tile * tw, q1 = min(q0 + tw, nwords);
        const uint32_t e0 = q0 * per, e1 = min(q1 * per, m);
        uint32_t v[kPackK];
#pragma unroll
        for (uint32_t i = 0; i < kPackK; i++) {
            const uint32_t e = e0 + t + i * kDictPartThreads;
            v[i] = e < e1 ? (dbl ? es[e] : in[e]) : 0u;
        }
        __syncthreads();  // the bitmap copy is in; the previous tile's ranks are used
#pragma unroll
        for (uint32_t i = 0; i < kPackK; i++) {
            const uint32_t e = e0 + t + i * kDictPartThreads;
            if (e >= e1) break;
            uint32_t r;
            if (dbl) {
                r = hr[v[i]];
            } else {
                const uint32_t wd = v[i] >> 5, below = (1u << (v[i] & 31u)) - 1u;
                r = inl ? lp[wd] + (uint32_t)__popc(lb[wd] & below) : gp[wd] + (uint32_t)__popc(gb[wd] & below);
            }
            rk[e - e0] = r;
        }
        __syncthreads();
        // a word per thread and step, ranks MSB first; the array's last word
        // holds the remainder, right-aligned
        for (uint32_t q = q0 + t; q < q1; q += kDictPartThreads) {
            const uint32_t eb = (q - q0) * per, ee = min(eb + per, e1 - e0);
            uint32_t acc = 0;
            for (uint32_t e = eb; e < ee; e++) acc = bits >= 32 ? rk[e] : ((acc << bits) | rk[e]);
            wo[q] = acc;
        }
    }
}

struct WriteArgs {
    uint32_t n;
    const uint32_t* tot;
    const uint32_t *nd, *nw;
    const int32_t *id4, *id6, *idn, *nidx, *x12i, *x12v, *x13i, *x13v, *cid, *id15, *id17;
    const uint8_t *v4, *v6;
    const uint32_t *code3, *code10, *words, *dict32;
    const uint64_t* dict64;
    const double* val15;
    uint64_t* foff;  // [window][file]
    uint64_t* ftot;  // [file]
    uint8_t* out;    // [file][cap]
    uint64_t cap;
};

// The pieces of window w's chunk in file f, in file order (Column_3_4_6
// cpp/compress.h:187-409, Column_5_7_8_9_14_16 :501-783, Column_10 :786-897,
// Column_11_12_13 :951-1175, Column_15_17 :1178-1312).
template <class Sink>
__device__ void chunk_pieces(const WriteArgs& a, uint32_t w, int f, Sink& s)
{
    const uint32_t* tw = a.tot + (size_t)w * kNC;
    const uint64_t C = min(kWin, a.n - w * kWin);
    const bool all_n = tw[kCNotN] == 0;
    const size_t seg = (size_t)w * kWin;
    const uint64_t nind = all_n ? 1u : 0u;
    const uint64_t ncode = (C + 15) / 16;
    if (f == 0) {  // .3_4_6
        if (all_n) {
            const uint64_t p[7] = {1, 0, C, 0, 0, 0, 0};
            s.vals(p, 7, 56);
            return;
        }
        const uint64_t n4 = tw[kC4], n6 = tw[kC6], nN = tw[kCN3];
        const uint64_t p[7] = {0, 5 * n4 + 5 * n6 + 4 * nN + 4 * ncode, C, ncode, nN, n4, n6};
        s.vals(p, 7, 56);
        s.bytes(a.id4 + seg, 4 * n4);
        s.bytes(a.v4 + seg, n4);
        s.bytes(a.id6 + seg, 4 * n6);
        s.bytes(a.v6 + seg, n6);
        s.bytes(a.idn + seg, 4 * nN);
        s.bytes(a.code3 + (size_t)w * (kWin / 16), 4 * ncode);
        return;
    }
    s.vals(&nind, 1, 4);
    if (all_n) return;
    const size_t wk0 = (size_t)w * kNDict;
    if (f == 1) {  // .5_7_8_9_14_16
        for (int c = 0; c < 6; c++) {
            const size_t kv = wk0 + c, kl = wk0 + 6 + c;
            const uint64_t sz = c == 5 ? 8 : 4;
            const uint64_t nwv = a.nw[kv], ndv = a.nd[kv], nwl = a.nw[kl], ndl = a.nd[kl];
            const uint64_t p[6] = {tw[kCRun0 + c], nwv, nwl, ndv, ndl, 4 * nwv + 4 * nwl + sz * ndv + 4 * ndl};
            s.vals(p, 6, 48);
            s.bytes(a.words + kv * kWin, 4 * nwv);
            s.bytes(a.words + kl * kWin, 4 * nwl);
            if (c == 5)
                s.bytes(a.dict64 + seg, 8 * ndv);
            else
                s.bytes(a.dict32 + kv * kWin, 4 * ndv);
            s.bytes(a.dict32 + kl * kWin, 4 * ndl);
        }
    } else if (f == 2) {  // .10
        const uint64_t nN = tw[kCN10];
        const uint64_t p[3] = {nN, ncode, C};
        s.vals(p, 3, 24);
        s.bytes(a.code10 + (size_t)w * (kWin / 16), 4 * ncode);
        s.bytes(a.nidx + seg, 4 * nN);
    } else if (f == 3) {  // .11_12_13
        const uint64_t ncom = tw[kCCom], n12 = tw[kCX12], n13 = tw[kCX13];
        const uint64_t p[4] = {C, ncom, n12, n13};
        s.vals(p, 4, 32);
        s.bytes(a.x12i + seg, 4 * n12);
        s.bytes(a.x12v + seg, 4 * n12);
        s.bytes(a.x13i + seg, 4 * n13);
        s.bytes(a.x13v + seg, 4 * n13);
        s.bytes(a.cid + seg, 4 * ncom);
        for (int k = 12; k < 15; k++) {
            const size_t kk = wk0 + k;
            const uint64_t dp[2] = {a.nw[kk], a.nd[kk]};
            s.vals(dp, 2, 16);
            s.bytes(a.words + kk * kWin, 4 * dp[0]);
            s.bytes(a.dict32 + kk * kWin, 4 * dp[1]);
        }
    } else {  // .15_17
        const uint64_t n15 = tw[kC15], n17 = tw[kC17];
        s.vals(&n15, 1, 4);
        s.bytes(a.id15 + seg, 4 * n15);
        s.bytes(a.val15 + seg, 8 * n15);
        s.vals(&n17, 1, 4);
        s.bytes(a.id17 + seg, 4 * n17);
    }
}

struct SizeSink {
    uint64_t n = 0;
    __device__ void vals(const uint64_t*, int, uint64_t nbytes) { n += nbytes; }
    __device__ void bytes(const void*, uint64_t nbytes) { n += nbytes; }
};

// one thread per (window, file): chunk sizes, then their offsets per file
__global__ __launch_bounds__(256) void k_col_sizes(WriteArgs a, uint32_t nwin)
{
    __shared__ uint64_t sz[kSubWindows * 5];
    const uint32_t t = threadIdx.x;
    if (t < nwin * 5) {
        SizeSink s;
        chunk_pieces(a, t / 5, (int)(t % 5), s);
        sz[t] = s.n;
    }
    __syncthreads();
    if (t < 5) {
        uint64_t run = 0;
        for (uint32_t w = 0; w < nwin; w++) {
            a.foff[w * 5 + t] = run;
            run += sz[w * 5 + t];
        }
        a.ftot[t] = run;
    }
}

constexpr uint32_t kWriteParts = 16;  // blocks per chunk

// Block `part` of the chunk's kWriteParts blocks copies its slice of every
// piece (slices start on 4-byte boundaries); part 0 also writes the headers.
struct CopySink {
    uint8_t* dst;
    uint32_t part;
    uint64_t* stage;  // LDS, 8 values
    __device__ void vals(const uint64_t* v, int count, uint64_t nbytes)
    {
        if (part == 0) {
            if (threadIdx.x < (uint32_t)count) stage[threadIdx.x] = v[threadIdx.x];
            __syncthreads();
            const uint8_t* src = reinterpret_cast<const uint8_t*>(stage);
            for (uint32_t b = threadIdx.x; b < nbytes; b += blockDim.x) dst[b] = src[b];
            __syncthreads();
        }
        dst += nbytes;
    }
    __device__ void bytes(const void* p, uint64_t nbytes)
    {
        const uint8_t* src = static_cast<const uint8_t*>(p);
        const uint64_t lo = (nbytes * part / kWriteParts) & ~3ull;
        const uint64_t hi = part + 1 == kWriteParts ? nbytes : (nbytes * (part + 1) / kWriteParts) & ~3ull;
        if (((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 3) == 0) {
            const uint64_t q0 = lo / 4, q1 = hi / 4;
            for (uint64_t q = q0 + threadIdx.x; q < q1; q += blockDim.x)
                reinterpret_cast<uint32_t*>(dst)[q] = reinterpret_cast<const uint32_t*>(src)[q];
            for (uint64_t b = q1 * 4 + threadIdx.x; b < hi; b += blockDim.x) dst[b] = src[b];
        } else {
            for (uint64_t b = lo + threadIdx.x; b < hi; b += blockDim.x) dst[b] = src[b];
        }
        dst += nbytes;
    }
};

// grid (window, file, part)
__global__ __launch_bounds__(256) void k_col_write(WriteArgs a)
{
    __shared__ uint64_t stage[8];
    const uint32_t w = blockIdx.x;
    const int f = (int)blockIdx.y;
    CopySink s;
    s.dst = a.out + (size_t)f * a.cap + a.foff[w * 5 + f];
    s.part = blockIdx.z;
    s.stage = stage;
    chunk_pieces(a, w, f, s);
}

}  // namespace

#define HIPCHK(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return fail(GSNP_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

static int alloc_cols(ColBufs& b, uint32_t win)
{
    if (b.win >= win) return GSNP_OK;
    const size_t S = (size_t)win * kWin;
    const size_t nb = (size_t)win * kWinBlocks;
    const size_t wk = (size_t)win * kNDict;
    b.win = 0;
    HIPCHK(b.cnt.alloc(nb * kNC));
    HIPCHK(b.boff.alloc(nb * kNC));
    HIPCHK(b.tot.alloc((size_t)win * kNC));
    for (DevBuf<int32_t>* p : {&b.id4, &b.id6, &b.idn, &b.nidx, &b.x12i, &b.x12v, &b.x13i, &b.x13v, &b.cid, &b.id15,
                               &b.id17})
        HIPCHK(p->alloc(S));
    HIPCHK(b.v4.alloc(S));
    HIPCHK(b.v6.alloc(S));
    HIPCHK(b.code3.alloc(S / 16));
    HIPCHK(b.code10.alloc(S / 16));
    HIPCHK(b.val15.alloc(S));
    HIPCHK(b.rstart.alloc(6 * S));
    HIPCHK(b.in32.alloc(kNDict * S));
    HIPCHK(b.in64.alloc(S));
    HIPCHK(b.tab.alloc(wk * kTabStride));
    HIPCHK(b.trank.alloc(wk * kTabStride));
    HIPCHK(b.dkeys.alloc(wk * kWin));
    HIPCHK(b.dslot.alloc(wk * kWin));
    HIPCHK(b.eslot.alloc(wk * kWin));
    HIPCHK(b.words.alloc(wk * kWin));
    HIPCHK(b.dict32.alloc(wk * kWin));
    HIPCHK(b.dict64.alloc(S));
    HIPCHK(b.nd.alloc(wk));
    HIPCHK(b.nw.alloc(wk));
    HIPCHK(b.foff.alloc((size_t)win * 5));
    HIPCHK(b.ftot.alloc(5));
    HIPCHK(b.gbits.alloc(wk * kBmpWords));
    HIPCHK(b.rbits.alloc(wk * kBmpWords));
    HIPCHK(b.wpre.alloc(wk * kBmpWords));
    HIPCHK(b.gstate.alloc(3 * wk));
    HIPCHK(b.htab.alloc((size_t)win * kHashSlots));
    HIPCHK(b.hrank.alloc((size_t)win * kHashSlots));
    HIPCHK(b.hkeys.alloc(S));
    HIPCHK(b.hslot.alloc(S));
    b.win = win;
    return GSNP_OK;
}

int encode_columns(Ctx& c, uint32_t offset, uint32_t ws_abs, uint32_t n, const ColumnSink& sink)
{
    if (!n || ws_abs % kWin) return fail(GSNP_E_ARG, "column chunks start on a window boundary and hold sites");
    hipStream_t st = c.stream;
    ColBufs& b = c.cols;
    const uint32_t want = std::min<uint32_t>(kSubWindows, (n + kWin - 1) / kWin);
    if (b.win < want) {
        // a queued pass reads the encoder arrays until its writes are done
        int rc = finish_columns(c);
        if (rc) {
            drop_columns(c);
            return rc;
        }
        HIPCHK(hipStreamSynchronize(st));
        rc = alloc_cols(b, want);
        if (rc) return rc;
        // the split dictionaries' bitmap and flags start at zero; each pass
        // leaves them so
        const size_t wk = (size_t)b.win * kNDict;
        launch_fill_u32(st, b.gbits.p, (uint32_t)(wk * kBmpWords), 0u);
        launch_fill_u32(st, b.gstate.p, (uint32_t)(3 * wk), 0u);
        launch_fill_u32(st, b.htab.p, (uint32_t)(2 * (size_t)b.win * kHashSlots), 0xFFFFFFFFu);  // kEmpty
        HIPCHK(hipGetLastError());
    }
    if (!b.copy) {
        HIPCHK(hipStreamCreateWithFlags(&b.copy, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&b.ev_written, hipEventDisableTiming));
        for (hipEvent_t& e : b.ev_copied) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    DictArgs da{};
    for (int e = 0; e < 32; e++) da.pow2bits[e] = (uint8_t)dict_bits(1ull << e);
    // GSNP_DICT_LDS_MAX: distinct values a dictionary keeps in LDS (at most
    // 2048, the default; 0 builds every dictionary in HBM)
    da.lds_max = kLdsSlots / 2;
    if (const char* v = std::getenv("GSNP_DICT_LDS_MAX")) da.lds_max = (uint32_t)std::strtoul(v, nullptr, 10);
    // GSNP_DICT_BITMAP=0 sends every integer array through the hash set too
    da.bitmap = 1;
    if (const char* v = std::getenv("GSNP_DICT_BITMAP")) da.bitmap = (uint32_t)std::strtoul(v, nullptr, 10);
    const size_t S = (size_t)b.win * kWin;
    int& cur = b.cur;
    for (uint32_t s0 = 0; s0 < n; s0 += kSubWindows * kWin) {
        const uint32_t ns = std::min<uint32_t>(n - s0, kSubWindows * kWin);
        const uint32_t nwin = (ns + kWin - 1) / kWin;
        const uint32_t nblk = (ns + kCB - 1) / kCB;
        ColIn in;
        in.ref = c.ref.p + ws_abs + s0;
        in.a = c.site_view(offset + s0);
        in.res = c.result_view(offset + s0);
        in.n = ns;
        c.timer.begin(kStColumns, st);
        hipLaunchKernelGGL(k_col_count, dim3(nblk), dim3(kCB), 0, st, in, b.cnt.p);
        hipLaunchKernelGGL(k_col_scan, dim3(nwin), dim3(256), 0, st, b.cnt.p, ns, b.boff.p, b.tot.p);
        EmitOut eo;
        eo.id4 = b.id4.p, eo.id6 = b.id6.p, eo.idn = b.idn.p, eo.nidx = b.nidx.p;
        eo.x12i = b.x12i.p, eo.x12v = b.x12v.p, eo.x13i = b.x13i.p, eo.x13v = b.x13v.p;
        eo.cid = b.cid.p, eo.id15 = b.id15.p, eo.id17 = b.id17.p;
        eo.v4 = b.v4.p, eo.v6 = b.v6.p, eo.code3 = b.code3.p, eo.code10 = b.code10.p;
        eo.rstart = b.rstart.p, eo.in32 = b.in32.p, eo.in64 = b.in64.p, eo.val15 = b.val15.p;
        eo.stride = S;
        hipLaunchKernelGGL(k_col_emit, dim3(nblk), dim3(kCB), 0, st, in, b.boff.p, eo);
        hipLaunchKernelGGL(k_col_runlen, dim3(kWinBlocks, 6, nwin), dim3(256), 0, st, ns, b.tot.p, b.rstart.p,
                           b.in32.p, S);
        da.tot = b.tot.p, da.in32 = b.in32.p, da.in64 = b.in64.p, da.stride = S;
        da.tab = b.tab.p, da.trank = b.trank.p, da.dkeys = b.dkeys.p, da.dslot = b.dslot.p, da.eslot = b.eslot.p;
        da.words = b.words.p, da.dict32 = b.dict32.p, da.dict64 = b.dict64.p, da.nd = b.nd.p, da.nw = b.nw.p;
        da.state = nullptr;
        if (da.bitmap) {
            SplitArgs sa;
            sa.tot = b.tot.p, sa.in32 = b.in32.p, sa.in64 = b.in64.p, sa.stride = S;
            sa.gbits = b.gbits.p, sa.rbits = b.rbits.p, sa.wpre = b.wpre.p, sa.state = b.gstate.p;
            sa.htab = b.htab.p, sa.hrank = b.hrank.p, sa.hkeys = b.hkeys.p, sa.hslot = b.hslot.p;
            sa.eslot = b.eslot.p, sa.dict32 = b.dict32.p, sa.dict64 = b.dict64.p, sa.words = b.words.p;
            sa.nd = b.nd.p, sa.nw = b.nw.p;
            std::copy(da.pow2bits, da.pow2bits + 32, sa.pow2bits);
            hipLaunchKernelGGL(k_dict_setbits, dim3(nwin, kIntDicts, kDictParts), dim3(kDictPartThreads), 0, st, sa);
            hipLaunchKernelGGL(k_hash_insert, dim3(nwin, kHashParts), dim3(kDictPartThreads), 0, st, sa);
            hipLaunchKernelGGL(k_dict_rank, dim3(nwin, kIntDicts), dim3(kDictThreads), 0, st, sa);
            hipLaunchKernelGGL(k_hash_rank, dim3(nwin), dim3(kDictThreads), 0, st, sa);
            hipLaunchKernelGGL(k_dict_pack, dim3(nwin, kNDict, kDictParts), dim3(kDictPartThreads), 0, st, sa);
            da.state = b.gstate.p;
        }
        hipLaunchKernelGGL(k_col_dict, dim3(nwin, kNDict), dim3(kDictThreads), 0, st, da);
        WriteArgs wa;
        wa.n = ns, wa.tot = b.tot.p, wa.nd = b.nd.p, wa.nw = b.nw.p;
        wa.id4 = b.id4.p, wa.id6 = b.id6.p, wa.idn = b.idn.p, wa.nidx = b.nidx.p;
        wa.x12i = b.x12i.p, wa.x12v = b.x12v.p, wa.x13i = b.x13i.p, wa.x13v = b.x13v.p;
        wa.cid = b.cid.p, wa.id15 = b.id15.p, wa.id17 = b.id17.p, wa.v4 = b.v4.p, wa.v6 = b.v6.p;
        wa.code3 = b.code3.p, wa.code10 = b.code10.p, wa.words = b.words.p, wa.dict32 = b.dict32.p;
        wa.dict64 = b.dict64.p, wa.val15 = b.val15.p, wa.foff = b.foff.p, wa.ftot = b.ftot.p;
        wa.out = nullptr, wa.cap = 0;
        hipLaunchKernelGGL(k_col_sizes, dim3(1), dim3(256), 0, st, wa, nwin);
        c.timer.end(kStColumns, st);
        HIPCHK(hipGetLastError());
        uint64_t tot[5];
        std::vector<uint64_t> offsets((size_t)nwin * 5);
        HIPCHK(hipMemcpyAsync(tot, b.ftot.p, sizeof tot, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(offsets.data(), b.foff.p, 8ull * nwin * 5, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        // buffer `cur` was last copied from two passes ago (this call's or an
        // earlier one's), and that pass has been handed over, so it is free
        const uint64_t need = (*std::max_element(tot, tot + 5) + 255) & ~255ull;  // files stay 256-byte aligned
        if (need > b.stream_cap[cur]) {
            HIPCHK(b.stream[cur].alloc(5 * need));
            b.stream_cap[cur] = need;
        }
        wa.out = b.stream[cur].p, wa.cap = b.stream_cap[cur];
        c.timer.begin(kStColumns, st);
        hipLaunchKernelGGL(k_col_write, dim3(nwin, 5, kWriteParts), dim3(256), 0, st, wa);
        c.timer.end(kStColumns, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(b.ev_written, st));
        HIPCHK(hipStreamWaitEvent(b.copy, b.ev_written, 0));
        const uint64_t all = tot[0] + tot[1] + tot[2] + tot[3] + tot[4];
        if (all > b.host_cap[cur]) {
            if (b.host[cur]) HIPCHK(hipHostFree(b.host[cur]));
            b.host[cur] = nullptr;
            b.host_cap[cur] = 0;
            HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&b.host[cur]), all, hipHostMallocDefault));
            b.host_cap[cur] = all;
        }
        ColumnPending d;
        d.buf = cur;
        d.sink = sink;
        d.offsets.swap(offsets);
        uint64_t at = 0;
        for (int f = 0; f < 5; f++) {
            if (tot[f])
                HIPCHK(hipMemcpyAsync(b.host[cur] + at, b.stream[cur].p + (size_t)f * b.stream_cap[cur], tot[f],
                                      hipMemcpyDeviceToHost, b.copy));
            d.pass.file[f] = b.host[cur] + at;
            d.pass.bytes[f] = tot[f];
            at += tot[f];
        }
        HIPCHK(hipEventRecord(b.ev_copied[cur], b.copy));
        for (uint32_t w = 0; w < nwin; w++) {
            d.starts.push_back((uint64_t)ws_abs + s0 + (uint64_t)w * kWin + 1);
            d.sizes.push_back(std::min<uint32_t>(kWin, ns - w * kWin));
        }
        // the pass before this one (of this call or an earlier one) goes to
        // its sink now that this pass is queued
        int rc = finish_columns(c);
        if (rc) {
            drop_columns(c);
            return rc;
        }
        b.prev = std::move(d);
        b.have_prev = true;
        cur ^= 1;
    }
    return GSNP_OK;
}

int finish_columns(Ctx& c)
{
    ColBufs& b = c.cols;
    if (!b.have_prev) return GSNP_OK;
    b.have_prev = false;
    ColumnPending& d = b.prev;
    HIPCHK(hipEventSynchronize(b.ev_copied[d.buf]));
    d.pass.starts = d.starts.data();
    d.pass.sizes = d.sizes.data();
    d.pass.offsets = d.offsets.data();
    d.pass.n_windows = (uint32_t)d.sizes.size();
    const int rc = d.sink(d.pass);
    d.sink = nullptr;
    return rc;
}

void drop_columns(Ctx& c)
{
    ColBufs& b = c.cols;
    if (b.copy) (void)hipStreamSynchronize(b.copy);
    b.have_prev = false;
    b.prev.sink = nullptr;
}

}  // namespace gsnp
