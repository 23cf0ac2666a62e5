// cdec.hip -- -C column files back to .cns text on the device
// (decompressAll, cpp/decompression.cpp:130-255, with the column layouts of
// cpp/compress.h:151-1312 and the codecs of cpp/dict_compression.h:47-166 and
// kernel/compress_kernel.cu:85-153).
//
// The host reads and checks each window's chunks (ColumnReader::read_chunk:
// every count and index against the window size), packs a batch of windows'
// raw arrays into one upload, and the device rebuilds the 17 columns -- 2-bit
// words, dictionary indices, run lengths turned into run starts, sparse lists
// scattered -- then sizes and writes every row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/gsnp_abi.h"
#include "cwriter.h"
#include "engine.h"
#include "gsnp_common.h"
#include "kernels.h"

namespace gsnp {

namespace {

// The 17 columns of the batch's sites (columns 1 and 2 come from the window).
struct DecCols {
    char *c3, *c4, *c6, *c10;
    int32_t *c5, *c7, *c8, *c9, *c11, *c12, *c13, *c14, *c17;
    double *c15, *c16;
};

enum : uint32_t { kColC3 = 0, kColC4, kColC6, kColC10, kColC5, kColC7, kColC8, kColC9, kColC11, kColC12, kColC13,
                  kColC14, kColC17, kColC15, kColC16 };

__device__ __forceinline__ char* char_col(const DecCols& c, uint32_t k)
{
    return k == kColC3 ? c.c3 : k == kColC4 ? c.c4 : k == kColC6 ? c.c6 : c.c10;
}
__device__ __forceinline__ int32_t* int_col(const DecCols& c, uint32_t k)
{
    switch (k) {
    case kColC5: return c.c5;
    case kColC7: return c.c7;
    case kColC8: return c.c8;
    case kColC9: return c.c9;
    case kColC11: return c.c11;
    case kColC12: return c.c12;
    case kColC13: return c.c13;
    case kColC14: return c.c14;
    default: return c.c17;
    }
}

// Work items of the batch, each over `n` elements.  Offsets are in bytes of
// the uploaded blob or in elements of the scratch arrays.
struct Unpack {  // dictionary indices -> values (cpp/dict_compression.h:68-88)
    uint64_t words, dict, out;  // blob, blob, scratch element offset
    uint32_t n, nd, bits, per, f64;
};
struct Scatter {  // ids[i] of a window's sparse list -> column
    uint64_t ids, vals, site0;  // blob offsets; first site of the window
    uint32_t n, col, kind;      // kind: 0 constant, 1 char values, 2 i32 values, 3 f64 values, 4 i32 scratch values
    int32_t cval;
};
struct TwoBit {  // 2-bit words -> bases A0 C1 G2 T3 (kernel/compress_kernel.cu:133-153)
    uint64_t words, site0;
    uint32_t n, col;
};
struct Expand {  // runs (values, exclusive starts) -> a column's sites
    uint64_t vals, starts, site0;  // scratch element offsets
    uint32_t nrun, n, col, f64;
};

// index of the item whose [pre[j], pre[j+1]) holds element e
__device__ __forceinline__ uint32_t find_item(const uint64_t* pre, uint32_t items, uint64_t e)
{
    uint32_t lo = 0, hi = items;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (pre[mid] <= e)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_dec_init(DecCols c, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u) {
        c.c3[i] = c.c4[i] = c.c6[i] = c.c10[i] = 'N';
        c.c5[i] = c.c7[i] = c.c8[i] = c.c9[i] = c.c11[i] = c.c12[i] = c.c13[i] = c.c14[i] = c.c17[i] = 0;
        c.c15[i] = 1.0;
        c.c16[i] = 255.0;
    }
}

__global__ __launch_bounds__(256) void k_dec_2bit(DecCols c, const uint8_t* blob, const TwoBit* items,
                                                  const uint64_t* pre, uint32_t n_items)
{
    const uint64_t total = pre[n_items];
    for (uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x; e < total; e += (uint64_t)gridDim.x * 256u) {
        const uint32_t j = find_item(pre, n_items, e);
        const TwoBit& t = items[j];
        const uint32_t i = (uint32_t)(e - pre[j]);
        const uint32_t* w = reinterpret_cast<const uint32_t*>(blob + t.words);
        const uint32_t nw = (t.n + 15u) / 16u, wi = i / 16u;
        const uint32_t in_word = wi == nw - 1u ? t.n - 16u * wi : 16u;  // the last word is right-aligned
        const uint32_t code = (w[wi] >> (2u * (in_word - (i % 16u) - 1u))) & 3u;
        char_col(c, t.col)[t.site0 + i] = code == 0 ? 'A' : code == 1 ? 'C' : code == 2 ? 'G' : 'T';
    }
}

__global__ __launch_bounds__(256) void k_dec_scatter(DecCols c, const uint8_t* blob, const int32_t* scratch32,
                                                     const Scatter* items, const uint64_t* pre, uint32_t n_items)
{
    const uint64_t total = pre[n_items];
    for (uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x; e < total; e += (uint64_t)gridDim.x * 256u) {
        const uint32_t j = find_item(pre, n_items, e);
        const Scatter& s = items[j];
        const uint32_t i = (uint32_t)(e - pre[j]);
        const uint64_t site = s.site0 + (uint32_t)reinterpret_cast<const int32_t*>(blob + s.ids)[i];
        switch (s.kind) {
        case 0:
            if (s.col <= kColC10)
                char_col(c, s.col)[site] = (char)s.cval;
            else
                int_col(c, s.col)[site] = s.cval;
            break;
        case 1: char_col(c, s.col)[site] = reinterpret_cast<const char*>(blob + s.vals)[i]; break;
        case 2: int_col(c, s.col)[site] = reinterpret_cast<const int32_t*>(blob + s.vals)[i]; break;
        case 3: c.c15[site] = reinterpret_cast<const double*>(blob + s.vals)[i]; break;
        default: int_col(c, s.col)[site] = scratch32[s.vals + i]; break;
        }
    }
}

// columns 4 and 6 start as column 3 (Column_3_4_6::decompress, cpp/compress.h:288-318)
__global__ __launch_bounds__(256) void k_dec_copy3(DecCols c, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u)
        c.c4[i] = c.c6[i] = c.c3[i];
}

__global__ __launch_bounds__(256) void k_dec_unpack(const uint8_t* blob, int32_t* scratch32, double* scratch64,
                                                    const Unpack* items, const uint64_t* pre, uint32_t n_items)
{
    const uint64_t total = pre[n_items];
    for (uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x; e < total; e += (uint64_t)gridDim.x * 256u) {
        const uint32_t j = find_item(pre, n_items, e);
        const Unpack& u = items[j];
        const uint32_t i = (uint32_t)(e - pre[j]);
        const uint32_t* w = reinterpret_cast<const uint32_t*>(blob + u.words);
        const uint32_t nw = (u.n + u.per - 1u) / u.per, wi = i / u.per;
        const uint32_t in_word = wi == nw - 1u ? u.n - u.per * wi : u.per;  // the last word is right-aligned
        const uint32_t mask = u.bits >= 32u ? 0xffffffffu : ((1u << u.bits) - 1u);
        const uint32_t idx = (w[wi] >> (u.bits * (in_word - (i % u.per) - 1u))) & mask;  // < nd (checked on the host)
        if (u.f64)
            scratch64[u.out + i] = reinterpret_cast<const double*>(blob + u.dict)[idx];
        else
            scratch32[u.out + i] = reinterpret_cast<const int32_t*>(blob + u.dict)[idx];
    }
}

// run lengths -> exclusive run starts, one workgroup per column of a window
__global__ __launch_bounds__(256) void k_dec_starts(int32_t* scratch32, const Expand* items)
{
    __shared__ uint32_t wsum[4];
    const Expand& x = items[blockIdx.x];
    uint32_t* v = reinterpret_cast<uint32_t*>(scratch32) + x.starts;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < x.nrun; base += 256u) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t len = i < x.nrun ? v[i] : 0u;
        uint32_t incl = len;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off);
            if ((int)lane >= off) incl += y;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        uint32_t before = 0, tile = 0;
        for (uint32_t k = 0; k < 4; k++) {
            if (k < w) before += wsum[k];
            tile += wsum[k];
        }
        if (i < x.nrun) v[i] = carry + before + incl - len;
        carry += tile;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_dec_expand(DecCols c, const int32_t* scratch32, const double* scratch64,
                                                    const Expand* items, const uint64_t* pre, uint32_t n_items)
{
    const uint64_t total = pre[n_items];
    for (uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x; e < total; e += (uint64_t)gridDim.x * 256u) {
        const uint32_t j = find_item(pre, n_items, e);
        const Expand& x = items[j];
        const uint32_t i = (uint32_t)(e - pre[j]);
        const uint32_t* st = reinterpret_cast<const uint32_t*>(scratch32) + x.starts;
        uint32_t lo = 0, hi = x.nrun;  // the last run starting at or before site i
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (st[mid] <= i)
                lo = mid;
            else
                hi = mid;
        }
        if (x.f64)
            c.c16[x.site0 + i] = scratch64[x.vals + lo];
        else
            int_col(c, x.col)[x.site0 + i] = scratch32[x.vals + lo];
    }
}

// A window of the batch: its sites, 1-based start and chromosome name.
struct DecWin {
    uint64_t site0, start;
    uint32_t n, name_off, name_len, pad;
};

template <class Put>
__device__ __forceinline__ int put_dec(uint64_t v, int n, Put& put)
{
    int k = 1;
    for (uint64_t t = v / 10u; t; t /= 10u) k++;
    for (int p = n + k - 1; p >= n; p--) {
        put(p, (char)('0' + v % 10u));
        v /= 10u;
    }
    return n + k;
}

template <class Put>
__device__ __forceinline__ int put_int(int32_t v, int n, Put& put)
{
    if (v < 0) {
        put(n, '-');
        return put_dec((uint64_t)(0u - (uint32_t)v), n + 1, put);
    }
    return put_dec((uint64_t)v, n, put);
}

template <class Put>
__device__ __forceinline__ int put_gtext(double v, int n, Put& put)
{
    char t[32];
    const int k = fmt_g(v, t);
    for (int i = 0; i < k; i++) put(n + i, t[i]);
    return n + k;
}

__constant__ char kDecNRow[] = "\tN\tN\t0\tN\t0\t0\t0\tN\t0\t0\t0\t0\t1.000\t255.000\t0\n";

// One row as decompressAll prints it (cpp/decompression.cpp:202-221): the N
// literal where columns 3, 4 and 6 are all 'N', else the 17 columns with
// columns 15 and 16 as %#g.
template <class Put>
__device__ int emit_dec_row(const DecCols& c, uint64_t s, const char* name, uint32_t name_len, uint64_t pos, Put put)
{
    int n = 0;
    for (uint32_t i = 0; i < name_len; i++) put(n++, name[i]);
    put(n++, '\t');
    n = put_dec(pos, n, put);
    if (c.c3[s] == 'N' && c.c4[s] == 'N' && c.c6[s] == 'N') {
        for (int i = 0; i < (int)sizeof(kDecNRow) - 1; i++) put(n + i, kDecNRow[i]);
        return n + (int)sizeof(kDecNRow) - 1;
    }
    const char ch[4] = {c.c3[s], c.c4[s], c.c6[s], c.c10[s]};
    put(n++, '\t');
    put(n++, ch[0]);
    put(n++, '\t');
    put(n++, ch[1]);
    put(n++, '\t');
    n = put_int(c.c5[s], n, put);
    put(n++, '\t');
    put(n++, ch[2]);
    put(n++, '\t');
    n = put_int(c.c7[s], n, put);
    put(n++, '\t');
    n = put_int(c.c8[s], n, put);
    put(n++, '\t');
    n = put_int(c.c9[s], n, put);
    put(n++, '\t');
    put(n++, ch[3]);
    put(n++, '\t');
    n = put_int(c.c11[s], n, put);
    put(n++, '\t');
    n = put_int(c.c12[s], n, put);
    put(n++, '\t');
    n = put_int(c.c13[s], n, put);
    put(n++, '\t');
    n = put_int(c.c14[s], n, put);
    put(n++, '\t');
    n = put_gtext(c.c15[s], n, put);
    put(n++, '\t');
    n = put_gtext(c.c16[s], n, put);
    put(n++, '\t');
    n = put_int(c.c17[s], n, put);
    put(n++, '\n');
    return n;
}

__device__ __forceinline__ uint32_t find_win(const DecWin* w, uint32_t nw, uint64_t s)
{
    uint32_t lo = 0, hi = nw;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (w[mid].site0 <= s)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_dec_rowlen(DecCols c, const uint8_t* blob, const DecWin* wins, uint32_t nw,
                                                    uint64_t n, uint32_t* rowlen, uint64_t* block_sum)
{
    __shared__ uint64_t red[4];
    const uint64_t s = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    uint32_t len = 0;
    if (s < n) {
        const DecWin& w = wins[find_win(wins, nw, s)];
        len = (uint32_t)emit_dec_row(c, s, reinterpret_cast<const char*>(blob + w.name_off), w.name_len,
                                     w.start + (s - w.site0), [](int, char) {});
        rowlen[s] = len;
    }
    uint64_t v = len;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) block_sum[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_dec_format(DecCols c, const uint8_t* blob, const DecWin* wins, uint32_t nw,
                                                    uint64_t n, const uint32_t* rowlen, const uint64_t* block_off,
                                                    char* out)
{
    __shared__ uint32_t wsum[4];
    const uint64_t s = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t len = s < n ? rowlen[s] : 0u;
    uint32_t incl = len;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off);
        if ((int)lane >= off) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t k = 0; k < w; k++) before += wsum[k];
    if (s >= n) return;
    char* o = out + block_off[blockIdx.x] + before + incl - len;
    const DecWin& wn = wins[find_win(wins, nw, s)];
    emit_dec_row(c, s, reinterpret_cast<const char*>(blob + wn.name_off), wn.name_len, wn.start + (s - wn.site0),
                 [&](int i, char ch) { o[i] = ch; });
}

uint32_t grid_for(uint64_t n)
{
    const uint64_t b = (n + 255) / 256;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 8192));
}

#define DECCHK(x)                                                                                  \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return fail(GSNP_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Host staging of one batch of windows.
struct Batch {
    std::vector<uint8_t> blob;
    std::vector<DecWin> wins;
    std::vector<TwoBit> twobit;
    std::vector<Scatter> scat1, scat2, scat3;  // N lists; column 4/6 differences; columns 11-13, 15, 17
    std::vector<Unpack> unpack;
    std::vector<Expand> expand;
    uint64_t sites = 0, n32 = 0, n64 = 0;  // scratch elements

    template <class T>
    uint64_t put(const std::vector<T>& v)
    {
        const uint64_t off = (blob.size() + 15) & ~uint64_t(15);
        blob.resize(off + v.size() * sizeof(T));
        if (!v.empty()) std::memcpy(blob.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
    uint64_t put_bytes(const void* p, size_t n)
    {
        const uint64_t off = (blob.size() + 15) & ~uint64_t(15);
        blob.resize(off + n);
        if (n) std::memcpy(blob.data() + off, p, n);
        return off;
    }
    void unpack_item(const std::vector<uint32_t>& words, uint64_t dict_off, uint64_t n, uint64_t nd, bool f64,
                     uint64_t out)
    {
        if (!n) return;
        Unpack u{};
        u.words = put(words);
        u.dict = dict_off;
        u.out = out;
        u.n = (uint32_t)n;
        u.nd = (uint32_t)nd;
        u.bits = (uint32_t)dict_bits(nd);
        u.per = (uint32_t)std::floor(32 / (double)u.bits);
        u.f64 = f64;
        unpack.push_back(u);
    }
    void scatter(std::vector<Scatter>& list, const std::vector<int32_t>& ids, uint64_t vals, uint32_t col, uint32_t kind,
                 int32_t cval)
    {
        if (ids.empty()) return;
        Scatter sc{};
        sc.ids = put(ids);
        sc.vals = vals;
        sc.site0 = sites;
        sc.n = (uint32_t)ids.size();
        sc.col = col;
        sc.kind = kind;
        sc.cval = cval;
        list.push_back(sc);
    }

    void add(const RawChunk& c)
    {
        DecWin w{};
        w.site0 = sites;
        w.start = c.start;
        w.n = (uint32_t)c.n;
        w.name_off = (uint32_t)put_bytes(c.chr.data(), c.chr.size());
        w.name_len = (uint32_t)c.chr.size();
        wins.push_back(w);
        if (!c.n346 && c.n) {
            twobit.push_back({put(c.code3), sites, (uint32_t)c.n, kColC3});
            scatter(scat1, c.idn, 0, kColC3, 0, 'N');
            scatter(scat2, c.id4, put(c.v4), kColC4, 1, 0);
            scatter(scat2, c.id6, put(c.v6), kColC6, 1, 0);
        }
        if (!c.n5) {
            static const uint32_t cols[6] = {kColC5, kColC7, kColC8, kColC9, kColC14, kColC16};
            for (int k = 0; k < 6; k++) {
                const RawRle& r = c.rle[k];
                const bool f64 = k == 5;
                Expand x{};
                x.vals = f64 ? n64 : n32;
                unpack_item(r.wv, f64 ? put(r.dv64) : put(r.dv32), r.nrun, r.nd_val, f64, x.vals);
                if (f64)
                    n64 += r.nrun;
                else
                    n32 += r.nrun;
                x.starts = n32;
                unpack_item(r.wl, put(r.dl), r.nrun, r.nd_len, false, x.starts);
                n32 += r.nrun;
                x.site0 = sites;
                x.nrun = (uint32_t)r.nrun;
                x.n = (uint32_t)c.n;
                x.col = cols[k];
                x.f64 = f64;
                expand.push_back(x);
            }
        }
        if (!c.n10 && c.n) {
            twobit.push_back({put(c.code10), sites, (uint32_t)c.n, kColC10});
            std::vector<int32_t> ni(c.ni10.begin(), c.ni10.end());
            scatter(scat1, ni, 0, kColC10, 0, 'N');
        }
        if (!c.n11) {
            const uint64_t nc = c.cid.size();
            static const uint32_t cols[3] = {kColC11, kColC12, kColC13};
            for (int k = 0; k < 3; k++) {
                const uint64_t out = n32;
                unpack_item(c.w11[k], put(c.d11[k]), nc, c.nd11[k], false, out);
                n32 += nc;
                scatter(scat2, c.cid, out, cols[k], 4, 0);  // the commons first ...
            }
            scatter(scat3, c.x12i, put(c.x12v), kColC12, 2, 0);  // ... then the extra lists
            scatter(scat3, c.x13i, put(c.x13v), kColC13, 2, 0);
        }
        if (!c.n15) {
            scatter(scat3, c.id15, put(c.v15), kColC15, 3, 0);
            scatter(scat3, c.id17, 0, kColC17, 0, 1);
        }
        sites += c.n;
    }
};

template <class T>
std::vector<uint64_t> prefix_of(const std::vector<T>& items)
{
    std::vector<uint64_t> pre(items.size() + 1, 0);
    for (size_t i = 0; i < items.size(); i++) pre[i + 1] = pre[i] + items[i].n;
    return pre;
}

struct DevMem {
    std::vector<void*> ptrs;
    ~DevMem()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    hipError_t alloc(T** p, size_t count)
    {
        void* v = nullptr;
        hipError_t e = hipMalloc(&v, std::max<size_t>(1, count) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(v);
        *p = static_cast<T*>(v);
        return e;
    }
};

// The device work of one batch; the rows go to fo.
int decode_batch(hipStream_t st, Batch& b, FILE* fo, const std::string& out)
{
    if (!b.sites) return GSNP_OK;
    DevMem m;
    uint8_t* blob;
    DecCols c;
    DECCHK(m.alloc(&blob, b.blob.size()));
    DECCHK(hipMemcpyAsync(blob, b.blob.data(), b.blob.size(), hipMemcpyHostToDevice, st));
    const uint64_t n = b.sites;
    DECCHK(m.alloc(&c.c3, n));
    DECCHK(m.alloc(&c.c4, n));
    DECCHK(m.alloc(&c.c6, n));
    DECCHK(m.alloc(&c.c10, n));
    for (int32_t** p : {&c.c5, &c.c7, &c.c8, &c.c9, &c.c11, &c.c12, &c.c13, &c.c14, &c.c17}) DECCHK(m.alloc(p, n));
    DECCHK(m.alloc(&c.c15, n));
    DECCHK(m.alloc(&c.c16, n));
    int32_t* s32;
    double* s64;
    DECCHK(m.alloc(&s32, b.n32));
    DECCHK(m.alloc(&s64, b.n64));
    // item lists and their element prefixes
    auto upload = [&](const auto& items, auto** d_items, uint64_t** d_pre) -> int {
        using T = typename std::decay_t<decltype(items)>::value_type;
        const std::vector<uint64_t> pre = prefix_of(items);
        DECCHK(m.alloc(d_items, items.size()));
        DECCHK(m.alloc(d_pre, pre.size()));
        if (!items.empty()) DECCHK(hipMemcpyAsync(*d_items, items.data(), items.size() * sizeof(T), hipMemcpyHostToDevice, st));
        DECCHK(hipMemcpyAsync(*d_pre, pre.data(), pre.size() * 8, hipMemcpyHostToDevice, st));
        return GSNP_OK;
    };
    TwoBit* d_tb;
    Scatter *d_s1, *d_s2, *d_s3;
    Unpack* d_up;
    Expand* d_ex;
    uint64_t *p_tb, *p_s1, *p_s2, *p_s3, *p_up, *p_ex;
    int rc = upload(b.twobit, &d_tb, &p_tb);
    if (!rc) rc = upload(b.scat1, &d_s1, &p_s1);
    if (!rc) rc = upload(b.scat2, &d_s2, &p_s2);
    if (!rc) rc = upload(b.scat3, &d_s3, &p_s3);
    if (!rc) rc = upload(b.unpack, &d_up, &p_up);
    if (!rc) rc = upload(b.expand, &d_ex, &p_ex);
    if (rc) return rc;
    auto total = [](const auto& items) {
        uint64_t t = 0;
        for (const auto& x : items) t += x.n;
        return t;
    };
    hipLaunchKernelGGL(k_dec_init, dim3(grid_for(n)), dim3(256), 0, st, c, n);
    if (!b.twobit.empty())
        hipLaunchKernelGGL(k_dec_2bit, dim3(grid_for(total(b.twobit))), dim3(256), 0, st, c, blob, d_tb, p_tb,
                           (uint32_t)b.twobit.size());
    if (!b.scat1.empty())
        hipLaunchKernelGGL(k_dec_scatter, dim3(grid_for(total(b.scat1))), dim3(256), 0, st, c, blob, s32, d_s1, p_s1,
                           (uint32_t)b.scat1.size());
    hipLaunchKernelGGL(k_dec_copy3, dim3(grid_for(n)), dim3(256), 0, st, c, n);
    if (!b.unpack.empty())
        hipLaunchKernelGGL(k_dec_unpack, dim3(grid_for(total(b.unpack))), dim3(256), 0, st, blob, s32, s64, d_up, p_up,
                           (uint32_t)b.unpack.size());
    if (!b.scat2.empty())
        hipLaunchKernelGGL(k_dec_scatter, dim3(grid_for(total(b.scat2))), dim3(256), 0, st, c, blob, s32, d_s2, p_s2,
                           (uint32_t)b.scat2.size());
    if (!b.scat3.empty())
        hipLaunchKernelGGL(k_dec_scatter, dim3(grid_for(total(b.scat3))), dim3(256), 0, st, c, blob, s32, d_s3, p_s3,
                           (uint32_t)b.scat3.size());
    if (!b.expand.empty()) {
        hipLaunchKernelGGL(k_dec_starts, dim3((uint32_t)b.expand.size()), dim3(256), 0, st, s32, d_ex);
        hipLaunchKernelGGL(k_dec_expand, dim3(grid_for(total(b.expand))), dim3(256), 0, st, c, s32, s64, d_ex, p_ex,
                           (uint32_t)b.expand.size());
    }
    // rows: lengths, block offsets, text
    DecWin* d_w;
    DECCHK(m.alloc(&d_w, b.wins.size()));
    DECCHK(hipMemcpyAsync(d_w, b.wins.data(), b.wins.size() * sizeof(DecWin), hipMemcpyHostToDevice, st));
    const uint64_t nblk = (n + 255) / 256;
    uint32_t* rowlen;
    uint64_t *bsum, *tot;
    DECCHK(m.alloc(&rowlen, n));
    DECCHK(m.alloc(&bsum, nblk));
    DECCHK(m.alloc(&tot, 1));
    hipLaunchKernelGGL(k_dec_rowlen, dim3((uint32_t)nblk), dim3(256), 0, st, c, blob, d_w, (uint32_t)b.wins.size(), n,
                       rowlen, bsum);
    launch_scan_blocks(st, bsum, (uint32_t)nblk, tot);
    uint64_t bytes = 0;
    DECCHK(hipMemcpyAsync(&bytes, tot, 8, hipMemcpyDeviceToHost, st));
    DECCHK(hipStreamSynchronize(st));
    char* text;
    DECCHK(m.alloc(&text, bytes));
    hipLaunchKernelGGL(k_dec_format, dim3((uint32_t)nblk), dim3(256), 0, st, c, blob, d_w, (uint32_t)b.wins.size(), n,
                       rowlen, bsum, text);
    DECCHK(hipGetLastError());
    std::unique_ptr<char[]> host(new char[bytes ? bytes : 1]);
    DECCHK(hipMemcpyAsync(host.get(), text, bytes, hipMemcpyDeviceToHost, st));
    DECCHK(hipStreamSynchronize(st));
    if (bytes && std::fwrite(host.get(), 1, bytes, fo) != bytes) return fail(GSNP_E_IO, "cannot write " + out);
    return GSNP_OK;
}

}  // namespace

// -C files to .cns text with the columns rebuilt and the rows written on the
// device, a batch of windows (up to 16, or 4 Mi sites) per upload.
int decompress_device(const std::string& prefix, const std::string& out)
{
    ColumnReader rd;
    std::string err;
    int rc = rd.open(prefix, err);
    if (rc) return fail(rc == 1 ? GSNP_E_IO : GSNP_E_FORMAT, err);
    FILE* fo = std::fopen(out.c_str(), "w");
    if (!fo) return fail(GSNP_E_IO, "cannot create " + out);
    hipStream_t st = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) {
        std::fclose(fo);
        return fail(GSNP_E_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    Batch b;
    RawChunk c;
    bool got = true;
    rc = GSNP_OK;
    while (!rc) {
        int r = rd.read_chunk(c, got, err);
        if (r) {
            rc = fail(GSNP_E_FORMAT, err);
            break;
        }
        if (got) b.add(c);
        if (!got || b.wins.size() >= 16 || b.sites >= (4u << 20)) {
            rc = decode_batch(st, b, fo, out);
            b = Batch();
        }
        if (!got) break;
    }
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    if (std::fclose(fo) != 0 && !rc) rc = fail(GSNP_E_IO, "cannot write " + out);
    return rc;
}

}  // namespace gsnp

extern "C" int gsnp_decompress_device(const char* prefix, const char* out)
{
    if (!prefix || !out) return gsnp::fail(GSNP_E_ARG, "null argument");
    return gsnp::guarded([&] { return gsnp::decompress_device(prefix, out); });
}
