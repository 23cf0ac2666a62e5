// dict.hip -- device dictionary encoding of 32-bit integer columns
// (cuda_dictCompression, kernel/compress_kernel.cu:185-334): the dictionary is
// the sorted set of distinct values, each element is replaced by its rank and
// the ranks are packed MSB-first, floor(32/bits) per word, the last word
// holding the remainder right-aligned.
//
// The distinct set comes from a presence bitmap over [min, max] (at most 2^32
// bits = 512 MiB, so every int32 column fits): set bits, prefix-popcount the
// bitmap words, and a value's rank is the popcount of the set bits below it.
//
// The char, uint32 and double overloads (cpp/dict_compression.h:26-37) map
// each value to a 64-bit unsigned key with the same order, sort the keys
// (a bitonic network over the power-of-two padded array, one launch per
// step), keep the distinct ones, and rank every element by binary search in
// that dictionary.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <type_traits>

#include <algorithm>
#include <climits>
#include <string>
#include <vector>

#include "../../include/gsnp_abi.h"
#include "cwriter.h"
#include "engine.h"

namespace gsnp {

namespace {

__global__ void k_minmax(const int32_t* v, uint64_t n, int32_t* mm)
{
    int lo = INT_MAX, hi = INT_MIN;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        lo = min(lo, v[i]);
        hi = max(hi, v[i]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, __shfl_xor(lo, off));
        hi = max(hi, __shfl_xor(hi, off));
    }
    if (__lane_id() == 0) {
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
    }
}

__global__ void k_setbits(const int32_t* v, uint64_t n, int32_t vmin, uint32_t* bits)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t d = (uint32_t)((int64_t)v[i] - vmin);
        atomicOr(&bits[d >> 5], 1u << (d & 31));
    }
}

// per-block popcount sums of 256-word chunks
__global__ __launch_bounds__(256) void k_popblock(const uint32_t* bits, uint64_t nw, uint64_t* bsum)
{
    __shared__ uint64_t red[4];
    const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
    uint64_t c = i < nw ? (uint64_t)__popc(bits[i]) : 0;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (__lane_id() == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// word prefix (exclusive) from the scanned block sums; emits the dictionary
__global__ __launch_bounds__(256) void k_wordprefix(const uint32_t* bits, uint64_t nw, const uint64_t* boff,
                                                    int32_t vmin, uint32_t* wpre, int32_t* dict)
{
    __shared__ uint32_t wsum[4];
    const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
    const uint32_t w = i < nw ? bits[i] : 0u;
    const uint32_t c = (uint32_t)__popc(w);
    uint32_t incl = c;
    const uint32_t lane = __lane_id(), wv = threadIdx.x >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off);
        if ((int)lane >= off) incl += y;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t k = 0; k < wv; k++) before += wsum[k];
    const uint64_t pre = boff[blockIdx.x] + before + incl - c;
    if (i >= nw) return;
    wpre[i] = (uint32_t)pre;
    uint32_t x = w, r = 0;
    while (x) {
        const int b = __ffs(x) - 1;
        dict[pre + r] = (int32_t)((int64_t)vmin + (int64_t)(i * 32 + b));
        x &= x - 1;
        r++;
    }
}

__global__ void k_pack(const int32_t* v, uint64_t n, int32_t vmin, const uint32_t* bits, const uint32_t* wpre,
                       int nbits, uint32_t per, uint64_t nwords, uint32_t* out)
{
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < nwords; w += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t e0 = w * per;
        const uint64_t e1 = e0 + per < n ? e0 + per : n;
        uint32_t acc = 0;
        for (uint64_t e = e0; e < e1; e++) {
            const uint32_t d = (uint32_t)((int64_t)v[e] - vmin);
            const uint32_t word = bits[d >> 5];
            const uint32_t rank = wpre[d >> 5] + (uint32_t)__popc(word & ((1u << (d & 31)) - 1u));
            acc = nbits >= 32 ? 0u : (acc << nbits);
            acc |= rank;
        }
        out[w] = acc;
    }
}


// Order-preserving unsigned keys.  char is signed (as on the reference's
// hosts); double keys put negatives below positives with their order
// reversed, and both zeros map to the key of +0.0 (the reference's
// dictionary compares with ==, so they are one value).
template <class T>
struct KeyOf;
template <>
struct KeyOf<char> {
    using K = uint64_t;
    __host__ __device__ static K key(char v) { return (uint8_t)v ^ 0x80u; }
    __host__ __device__ static char value(K k) { return (char)(uint8_t)(k ^ 0x80u); }
};
template <>
struct KeyOf<uint32_t> {
    using K = uint64_t;
    __host__ __device__ static K key(uint32_t v) { return v; }
    __host__ __device__ static uint32_t value(K k) { return (uint32_t)k; }
};
template <>
struct KeyOf<double> {
    using K = uint64_t;
    __host__ __device__ static K key(double v)
    {
        uint64_t b;
        __builtin_memcpy(&b, &v, 8);
        if (b == 0x8000000000000000ull) b = 0;  // -0.0 == +0.0
        return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
    }
    __host__ __device__ static double value(K k)
    {
        const uint64_t b = (k >> 63) ? (k & ~0x8000000000000000ull) : ~k;
        double v;
        __builtin_memcpy(&v, &b, 8);
        return v;
    }
};

template <class T>
__global__ void k_tokey(const T* in, uint64_t n, typename KeyOf<T>::K* keys)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        keys[i] = KeyOf<T>::key(in[i]);
}

// each element's rank (lower bound in the sorted distinct keys), packed as
// k_pack does: per ranks MSB-first in a word, the last word right-aligned
template <class T>
__global__ void k_rank_pack(const T* in, uint64_t n, const typename KeyOf<T>::K* dict, uint64_t nd, int nbits,
                            uint32_t per, uint64_t nwords, uint32_t* out)
{
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < nwords; w += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t e0 = w * per;
        const uint64_t e1 = e0 + per < n ? e0 + per : n;
        uint32_t acc = 0;
        for (uint64_t e = e0; e < e1; e++) {
            const typename KeyOf<T>::K k = KeyOf<T>::key(in[e]);
            uint64_t lo = 0, hi = nd;
            while (lo < hi) {
                const uint64_t mid = (lo + hi) / 2;
                if (dict[mid] < k)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            acc = nbits >= 32 ? 0u : (acc << nbits);
            acc |= (uint32_t)lo;
        }
        out[w] = acc;
    }
}

template <class T>
__global__ void k_fromkey(const typename KeyOf<T>::K* keys, uint64_t n, T* out)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = KeyOf<T>::value(keys[i]);
}

// one compare-exchange step (distance j, merge size k) of an ascending
// bitonic sort over n (a power of two) keys
__global__ void k_bitonic_step(uint64_t* v, uint64_t n, uint64_t j, uint64_t k)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t l = i ^ j;
        if (l <= i) continue;
        const uint64_t a = v[i], b = v[l];
        const bool up = (i & k) == 0;
        if (up ? a > b : a < b) {
            v[i] = b;
            v[l] = a;
        }
    }
}

// distinct keys of a sorted array: per-256 block head counts, then the heads
// written at their exclusive offsets (block offsets scanned in between)
__global__ __launch_bounds__(256) void k_head_count(const uint64_t* v, uint64_t n, uint64_t* bsum)
{
    __shared__ uint64_t red[4];
    const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
    uint64_t c = i < n && (i == 0 || v[i] != v[i - 1]) ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (__lane_id() == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_head_write(const uint64_t* v, uint64_t n, const uint64_t* boff, uint64_t* out)
{
    __shared__ uint32_t wsum[4];
    const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
    const uint32_t h = i < n && (i == 0 || v[i] != v[i - 1]) ? 1u : 0u;
    uint32_t incl = h;
    const uint32_t lane = __lane_id(), wv = threadIdx.x >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off);
        if ((int)lane >= off) incl += y;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t k = 0; k < wv; k++) before += wsum[k];
    if (h) out[boff[blockIdx.x] + before + incl - 1] = v[i];
}

}  // namespace

#define HIPCHK(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return fail(GSNP_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Encodes n device values; dictionary and words land in device buffers owned
// by the context; counts are returned.
int dict_encode_device(Ctx& c, const int32_t* d_in, uint64_t n, uint64_t* n_words, uint64_t* n_dict)
{
    hipStream_t st = c.stream;
    if (n == 0) {
        *n_words = *n_dict = 0;
        return GSNP_OK;
    }
    if (c.dict_mm.n < 2) HIPCHK(c.dict_mm.alloc(2));
    const int32_t init[2] = {INT_MAX, INT_MIN};
    HIPCHK(hipMemcpyAsync(c.dict_mm.p, init, 8, hipMemcpyHostToDevice, st));
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_minmax, dim3(grid), dim3(256), 0, st, d_in, n, c.dict_mm.p);
    int32_t mm[2];
    HIPCHK(hipMemcpyAsync(mm, c.dict_mm.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const uint64_t range = (uint64_t)((int64_t)mm[1] - mm[0]) + 1;
    const uint64_t nw = (range + 31) / 32;
    if (c.dict_bits.n < nw) HIPCHK(c.dict_bits.alloc(nw));
    if (c.dict_wpre.n < nw) HIPCHK(c.dict_wpre.alloc(nw));
    const uint64_t nb = (nw + 255) / 256;
    if (c.dict_bsum.n < nb + 1) HIPCHK(c.dict_bsum.alloc(nb + 1));
    if (c.dict_out.n < n) HIPCHK(c.dict_out.alloc(n));
    if (c.dict_vals.n < n) HIPCHK(c.dict_vals.alloc(n));
    if (c.total.n < 1) HIPCHK(c.total.alloc(1));
    HIPCHK(hipMemsetAsync(c.dict_bits.p, 0, nw * 4, st));
    hipLaunchKernelGGL(k_setbits, dim3(grid), dim3(256), 0, st, d_in, n, mm[0], c.dict_bits.p);
    hipLaunchKernelGGL(k_popblock, dim3((uint32_t)nb), dim3(256), 0, st, c.dict_bits.p, nw, c.dict_bsum.p);
    launch_scan_blocks(st, c.dict_bsum.p, (uint32_t)nb, c.total.p);
    hipLaunchKernelGGL(k_wordprefix, dim3((uint32_t)nb), dim3(256), 0, st, c.dict_bits.p, nw, c.dict_bsum.p, mm[0],
                       c.dict_wpre.p, c.dict_vals.p);
    uint64_t nd = 0;
    HIPCHK(hipMemcpyAsync(&nd, c.total.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const int bits = dict_bits(nd);
    const uint32_t per = (uint32_t)(32 / bits);
    const uint64_t nwords = (n + per - 1) / per;
    const uint32_t g2 = (uint32_t)std::min<uint64_t>((nwords + 255) / 256, 4096);
    hipLaunchKernelGGL(k_pack, dim3(g2), dim3(256), 0, st, d_in, n, mm[0], c.dict_bits.p, c.dict_wpre.p, bits, per,
                       nwords, c.dict_out.p);
    HIPCHK(hipGetLastError());
    *n_words = nwords;
    *n_dict = nd;
    return GSNP_OK;
}


// cuda_dictCompression for char, uint32 and double columns of host values:
// dictionary and words to host memory.  With no +0.0 among the values the
// dictionary's zero is -0.0 (the value the input holds).  NaN has no rank
// (the reference's search for it fails) and is rejected.
template <class T>
int dict_encode_sorted(Ctx& c, const T* in, uint64_t n, uint32_t* words, uint64_t* n_words, T* dict, uint64_t* n_dict)
{
    using K = typename KeyOf<T>::K;
    *n_words = *n_dict = 0;
    if (n == 0) return GSNP_OK;
    bool neg_zero_only = false;
    if constexpr (std::is_same<T, double>::value) {
        bool pos0 = false, neg0 = false;
        for (uint64_t i = 0; i < n; i++) {
            if (std::isnan(in[i])) return fail(GSNP_E_ARG, "NaN in a dictionary-coded column");
            if (in[i] == 0.0) (std::signbit(in[i]) ? neg0 : pos0) = true;
        }
        neg_zero_only = neg0 && !pos0;
    }
    hipStream_t st = c.stream;
    DevBuf<T> d_in, d_dict;
    DevBuf<K> keys, uniq;
    DevBuf<uint64_t> bsum, tot;
    DevBuf<uint32_t> d_words;
    uint64_t npad = 1;
    while (npad < n) npad <<= 1;
    const uint64_t nb = (n + 255) / 256;
    HIPCHK(d_in.alloc(n));
    HIPCHK(keys.alloc(npad));
    HIPCHK(uniq.alloc(n));
    HIPCHK(bsum.alloc(nb + 1));
    HIPCHK(tot.alloc(1));
    HIPCHK(hipMemcpyAsync(d_in.p, in, sizeof(T) * n, hipMemcpyHostToDevice, st));
    const uint32_t grid = (uint32_t)std::min<uint64_t>((npad + 255) / 256, 4096);
    launch_fill_u32(st, reinterpret_cast<uint32_t*>(keys.p), (uint32_t)(2 * npad), 0xFFFFFFFFu);  // padding sorts last
    hipLaunchKernelGGL(k_tokey<T>, dim3(grid), dim3(256), 0, st, d_in.p, n, keys.p);
    for (uint64_t k = 2; k <= npad; k <<= 1)
        for (uint64_t j = k >> 1; j > 0; j >>= 1)
            hipLaunchKernelGGL(k_bitonic_step, dim3(grid), dim3(256), 0, st, keys.p, npad, j, k);
    hipLaunchKernelGGL(k_head_count, dim3((uint32_t)nb), dim3(256), 0, st, keys.p, n, bsum.p);
    launch_scan_blocks(st, bsum.p, (uint32_t)nb, tot.p);
    hipLaunchKernelGGL(k_head_write, dim3((uint32_t)nb), dim3(256), 0, st, keys.p, n, bsum.p, uniq.p);
    HIPCHK(hipGetLastError());
    uint64_t nd = 0;
    HIPCHK(hipMemcpyAsync(&nd, tot.p, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const int bits = dict_bits(nd);
    const uint32_t per = (uint32_t)(32 / bits);
    const uint64_t nwords = (n + per - 1) / per;
    HIPCHK(d_words.alloc(nwords));
    HIPCHK(d_dict.alloc(nd));
    const uint32_t g2 = (uint32_t)std::min<uint64_t>((nwords + 255) / 256, 4096);
    hipLaunchKernelGGL(k_rank_pack<T>, dim3(g2), dim3(256), 0, st, d_in.p, n, uniq.p, nd, bits, per, nwords,
                       d_words.p);
    hipLaunchKernelGGL(k_fromkey<T>, dim3((uint32_t)std::min<uint64_t>((nd + 255) / 256, 4096)), dim3(256), 0, st,
                       uniq.p, nd, d_dict.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(words, d_words.p, 4 * nwords, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(dict, d_dict.p, sizeof(T) * nd, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if constexpr (std::is_same<T, double>::value) {
        if (neg_zero_only)
            for (uint64_t i = 0; i < nd; i++)
                if (dict[i] == 0.0) dict[i] = -0.0;
    }
    *n_words = nwords;
    *n_dict = nd;
    return GSNP_OK;
}

}  // namespace gsnp

extern "C" int gsnp_dict_compress_i32(gsnp_ctx* c, const int32_t* in, uint64_t n, uint32_t* words, uint64_t* n_words,
                                      int32_t* dict, uint64_t* n_dict)
{
    using namespace gsnp;
    if (!c || (!in && n) || !words || !n_words || !dict || !n_dict) return fail(GSNP_E_ARG, "null argument");
    DevBuf<int32_t> d_in;
    HIPCHK(d_in.alloc(n));
    if (n) HIPCHK(hipMemcpyAsync(d_in.p, in, 4 * n, hipMemcpyHostToDevice, c->stream));
    int rc = dict_encode_device(*c, d_in.p, n, n_words, n_dict);
    if (rc) return rc;
    if (*n_words) HIPCHK(hipMemcpyAsync(words, c->dict_out.p, 4 * *n_words, hipMemcpyDeviceToHost, c->stream));
    if (*n_dict) HIPCHK(hipMemcpyAsync(dict, c->dict_vals.p, 4 * *n_dict, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return GSNP_OK;
}

template <class T>
static int dict_compress_entry(gsnp_ctx* c, const T* in, uint64_t n, uint32_t* words, uint64_t* n_words, T* dict,
                               uint64_t* n_dict)
{
    using namespace gsnp;
    if (!c || (!in && n) || !words || !n_words || !dict || !n_dict) return fail(GSNP_E_ARG, "null argument");
    if (n > (1ull << 30)) return fail(GSNP_E_ARG, "more than 2^30 values");
    return guarded([&] { return dict_encode_sorted<T>(*c, in, n, words, n_words, dict, n_dict); });
}

extern "C" int gsnp_dict_compress_char(gsnp_ctx* c, const char* in, uint64_t n, uint32_t* words, uint64_t* n_words,
                                       char* dict, uint64_t* n_dict)
{
    return dict_compress_entry<char>(c, in, n, words, n_words, dict, n_dict);
}

extern "C" int gsnp_dict_compress_u32(gsnp_ctx* c, const uint32_t* in, uint64_t n, uint32_t* words, uint64_t* n_words,
                                      uint32_t* dict, uint64_t* n_dict)
{
    return dict_compress_entry<uint32_t>(c, in, n, words, n_words, dict, n_dict);
}

extern "C" int gsnp_dict_compress_f64(gsnp_ctx* c, const double* in, uint64_t n, uint32_t* words, uint64_t* n_words,
                                      double* dict, uint64_t* n_dict)
{
    return dict_compress_entry<double>(c, in, n, words, n_words, dict, n_dict);
}
