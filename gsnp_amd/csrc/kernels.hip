// kernels.hip -- CDNA4 (gfx950) kernels of the SOAPsnp calling path.
//
// Compiled with -ffp-contract=off: the likelihood is a sequential FP64 sum in
// sorted-record order and every comparison downstream depends on its exact
// rounding, so no FMA contraction anywhere.
#include <hip/hip_runtime.h>

#include "gsnp_common.h"
#include "kernels.h"
#include "nearest.h"
#include "walk.h"

namespace gsnp {

namespace {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// lower_bound / upper_bound over positions[lo, hi)
__device__ uint64_t lower_pos(const int32_t* pos, uint64_t lo, uint64_t hi, int64_t v)
{
    while (lo < hi) {
        uint64_t m = (lo + hi) >> 1;
        if ((int64_t)pos[m] < v)
            lo = m + 1;
        else
            hi = m;
    }
    return lo;
}

}  // namespace

// ---------------------------------------------------------------------------
// (a) site matrix.  One thread per site, 64 consecutive sites per block (one
// wave: a read covers more of a narrow block's lanes).  The
// block walks, in input order, every read of [r0, r1) that can cover one of
// its sites.  The reads' positions, metadata and packed bytes are first
// staged into LDS with wide coalesced loads (in chunks when a pileup is
// deeper than the stage), so the walk itself runs at LDS latency; read fields
// are block-uniform broadcasts and the lanes of a covering read take
// consecutive bytes of it.
// Per-site semantics follow kernel/update_kernel.cu:233-326: depth and
// repeat for every covering read; for an observed base (ACGT, quality >=
// q_min) while dep_uni < 255: unique hits append a record and bump
// count_uni/q_sum/dep_uni; every hit bumps count_all.  Records are buffered
// four at a time in registers (a shift register, newest last) and stored as
// 16-byte words.
// Sites below `carry` continue from the values already in memory (the
// recycled read-length tail of the previous window); the others start at 0.
// Only reads starting in [lo_min, hi_max] take part; `range` holds each
// block's read interval (k_read_ranges).  With head_qs set, a site on a
// window boundary (absolute position a multiple of the window) also records
// its q_sum before the reads that start on it, which is what the per-window
// path holds in its first tail site when the window is called.

constexpr uint32_t kBuildSites = 64;     // sites per block (one wave)
constexpr uint32_t kStageBytes = 8192;   // packed read bytes staged per chunk
constexpr uint32_t kStageReads = 128;

// read interval [b, e) of each 64-site block: reads starting in
// [ws + s0 - (L-1), ws + s0 + 63] clipped to [lo_min, hi_max]
__global__ __launch_bounds__(256) void k_read_ranges(const int32_t* pos, uint64_t r0, uint64_t r1, uint32_t ws,
                                                     uint32_t n_blocks, uint32_t L, int64_t lo_min, int64_t hi_max,
                                                     uint64_t* range)
{
    const uint32_t blk = blockIdx.x * 256u + threadIdx.x;
    if (blk >= n_blocks) return;
    const uint32_t s0 = blk * kBuildSites;
    int64_t lo = (int64_t)ws + (int64_t)s0 - ((int64_t)L - 1);
    if (lo < lo_min) lo = lo_min;
    int64_t hi = (int64_t)ws + (int64_t)s0 + (kBuildSites - 1);
    if (hi > hi_max) hi = hi_max;
    const uint64_t b = lower_pos(pos, r0, r1, lo);
    uint64_t e = lower_pos(pos, b, r1, hi + 1);
    if (e < b) e = b;
    range[2 * blk] = b;
    range[2 * blk + 1] = e;
}

__global__ __launch_bounds__(kBuildSites) void k_site_build(ReadView r, const uint64_t* range, uint32_t ws, uint32_t n_sites,
                                                    uint32_t carry, SiteArrays a, int4* head_qs)
{
    __shared__ __align__(16) uint8_t sseq[kStageBytes + 32];
    __shared__ int32_t spos[kStageReads];
    __shared__ uint32_t smeta[kStageReads];
    __shared__ int sca[4 * kBuildSites];  // count_all, one LDS add per observed base
    const uint32_t s = blockIdx.x * kBuildSites + threadIdx.x;
    const uint32_t pitch = r.pitch;
    const uint32_t kr = kStageBytes / pitch < kStageReads ? kStageBytes / pitch : kStageReads;
    // block-uniform bounds: keep the read loop on scalar registers
    const uint64_t rb = range[2 * blockIdx.x], re = range[2 * blockIdx.x + 1];
    const uint64_t b = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(rb >> 32)) << 32) |
                       (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)rb);
    const uint64_t e = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(re >> 32)) << 32) |
                       (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)re);
    const bool live = s < n_sites;
    int depth = 0, repeat = 0, nrec = 0;
    uint32_t cu = 0;   // 4 x 8-bit count_uni
    uint64_t qs = 0;   // 4 x 16-bit q_sum
    int ca0 = 0, ca1 = 0, ca2 = 0, ca3 = 0;
    uint32_t* row = a.rec + (size_t)s * kSlots;
    // records of slots [nrec & ~3, nrec), right-aligned: the newest in pb3
    uint32_t pb0 = 0, pb1 = 0, pb2 = 0, pb3 = 0;
    if (live && s < carry) {
        depth = a.depth[s];
        repeat = a.repeat[s];
        nrec = a.nrec[s];
        int4 c = a.cu[s], q = a.qs[s], al = a.ca[s];
        cu = (uint32_t)c.x | ((uint32_t)c.y << 8) | ((uint32_t)c.z << 16) | ((uint32_t)c.w << 24);
        qs = (uint64_t)(uint32_t)q.x | ((uint64_t)(uint32_t)q.y << 16) | ((uint64_t)(uint32_t)q.z << 32) |
             ((uint64_t)(uint32_t)q.w << 48);
        ca0 = al.x, ca1 = al.y, ca2 = al.z, ca3 = al.w;
        const int k = nrec & 3;
        if (k) {
            const uint4 v = *reinterpret_cast<const uint4*>(row + (nrec & ~3));
            pb1 = v.x;
            pb2 = k == 2 ? v.x : v.y;
            pb3 = k == 1 ? v.x : k == 2 ? v.y : v.z;
        }
    }
    // count_all per base and lane, base-major (word o * 64 + lane): a lane
    // stays in its own bank whatever base it adds to
    sca[0 * kBuildSites + threadIdx.x] = ca0;
    sca[1 * kBuildSites + threadIdx.x] = ca1;
    sca[2 * kBuildSites + threadIdx.x] = ca2;
    sca[3 * kBuildSites + threadIdx.x] = ca3;
    const int64_t p = (int64_t)ws + s;
    const int64_t p0 = (int64_t)ws + (int64_t)blockIdx.x * kBuildSites;  // the block's first site
    const int32_t lane = (int32_t)threadIdx.x;
    const bool head = head_qs && live && (p % kWin) == 0;
    const bool any_head = __ballot(head) != 0;  // block-uniform: the snapshot test leaves the read loop of other blocks
    bool snapped = false;
    for (uint64_t c0 = b; c0 < e; c0 += kr) {
        const uint32_t cn = (uint32_t)(e - c0 < kr ? e - c0 : kr);
        __syncthreads();  // the previous chunk has been consumed
        // read starts relative to the block's first site (they lie within
        // [-(L-1), 63]): the read loop works in 32-bit scalar registers
        for (uint32_t i = threadIdx.x; i < cn; i += kBuildSites) {
            spos[i] = (int32_t)((int64_t)r.pos[c0 + i] - p0);
            smeta[i] = r.meta[c0 + i];
        }
        // bytes [c0*pitch, (c0+cn)*pitch) from the 16-byte boundary below;
        // the read store is padded so the last vector stays in the allocation
        const uint64_t g0 = c0 * pitch;
        const uint64_t a0 = g0 & ~15ull;
        const uint32_t delta = (uint32_t)(g0 - a0);
        const uint32_t nvec = (uint32_t)((g0 + (uint64_t)cn * pitch - a0 + 15) >> 4);
        const uint4* src = reinterpret_cast<const uint4*>(r.seq + a0);
        uint4* dst = reinterpret_cast<uint4*>(sseq);
        for (uint32_t v = threadIdx.x; v < nvec; v += kBuildSites) dst[v] = src[v];
        __syncthreads();
        for (uint32_t j = 0; j < cn; j++) {
            const int32_t start = __builtin_amdgcn_readfirstlane(spos[j]);  // block-uniform
            const uint32_t m = __builtin_amdgcn_readfirstlane(smeta[j]);
            if (any_head && head && !snapped && start >= lane) {
                head_qs[p / kWin - ws / kWin] = make_int4((int)(qs & 0xffff), (int)((qs >> 16) & 0xffff),
                                                          (int)((qs >> 32) & 0xffff), (int)(qs >> 48));
                snapped = true;
            }
            const int32_t coord = lane - start;
            const uint32_t len = rm_len(m);
            if (!live || coord < 0 || coord >= (int32_t)len) continue;
            const uint32_t hit = rm_hit(m);
            depth += 1;
            repeat += (int)hit;
            const uint32_t byte = sseq[delta + j * pitch + (uint32_t)coord];
            if (byte == 0 || nrec >= (int)kMaxDepUni) continue;
            const uint32_t o = pb_base(byte);
            if (hit == 1) {
                const uint32_t q = pb_q(byte);
                const uint32_t strand = rm_strand(m);
                const uint32_t cyc = strand ? (uint32_t)(len - 1 - coord) : (uint32_t)coord;
                const uint32_t w = rec_encode(cyc, strand, o, q, 1);
                pb0 = pb1;
                pb1 = pb2;
                pb2 = pb3;
                pb3 = w;
                nrec += 1;
                if ((nrec & 3) == 0) *reinterpret_cast<uint4*>(row + (nrec - 4)) = make_uint4(pb0, pb1, pb2, pb3);
                cu += 1u << (8 * o);
                qs += (uint64_t)q << (16 * o);
            }
            atomicAdd(&sca[o * kBuildSites + threadIdx.x], 1);
        }
    }
    if (!live) return;
    if (const int k = nrec & 3) {  // the last k records, left-aligned
        const uint32_t x = k == 1 ? pb3 : k == 2 ? pb2 : pb1;
        *reinterpret_cast<uint4*>(row + (nrec & ~3)) = make_uint4(x, k == 2 ? pb3 : pb2, pb3, 0u);
    }
    if (head && !snapped)
        head_qs[p / kWin - ws / kWin] =
            make_int4((int)(qs & 0xffff), (int)((qs >> 16) & 0xffff), (int)((qs >> 32) & 0xffff), (int)(qs >> 48));
    a.depth[s] = depth;
    a.repeat[s] = repeat;
    a.nrec[s] = nrec;
    a.cu[s] = make_int4((int)(cu & 0xff), (int)((cu >> 8) & 0xff), (int)((cu >> 16) & 0xff), (int)(cu >> 24));
    a.qs[s] = make_int4((int)(qs & 0xffff), (int)((qs >> 16) & 0xffff), (int)((qs >> 32) & 0xffff), (int)(qs >> 48));
    a.ca[s] = make_int4(sca[0 * kBuildSites + threadIdx.x], sca[1 * kBuildSites + threadIdx.x],
                        sca[2 * kBuildSites + threadIdx.x], sca[3 * kBuildSites + threadIdx.x]);
}

void launch_site_build(hipStream_t st, const ReadView& r, uint64_t r0, uint64_t r1, uint32_t ws, uint32_t n_sites,
                       int64_t lo_min, int64_t hi_max, uint32_t carry, const SiteArrays& a, int4* head_qs,
                       uint64_t* range)
{
    const uint32_t blocks = (n_sites + kBuildSites - 1) / kBuildSites;
    if (!blocks) return;
    hipLaunchKernelGGL(k_read_ranges, dim3((blocks + 255) / 256), dim3(256), 0, st, r.pos, r0, r1, ws, blocks, r.pitch,
                       lo_min, hi_max, range);
    hipLaunchKernelGGL(k_site_build, dim3(blocks), dim3(kBuildSites), 0, st, r, range, ws, n_sites, carry, a, head_qs);
}

// ---------------------------------------------------------------------------
// (b) segmented sort.  Sites are bucketed by record count (<=8, 16, 32, 64,
// 128, 256; 0 or 1 record needs no sort) and each bucket runs a bitonic
// network in registers: for widths up to 64 a site occupies `W` lanes of a
// wave, exchanging partners with ds_swizzle (xor < 32) or ds_bpermute
// (xor 32); widths 128/256 keep 2/4 records per lane and do the xor >= 64
// steps inside the lane.  Padding lanes hold 0xFFFFFFFF, so ascending u32
// order (base, quality descending, cycle, strand) comes out in the first n
// slots, as kernel/bitonic_kernel.cu:119-176 produces.

// word fill: clears counters and flags with a kernel, so that the sequence
// can be captured into a graph (memset nodes are avoided there)
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill_u32(hipStream_t st, void* p, uint32_t n_words, uint32_t v)
{
    if (n_words)
        hipLaunchKernelGGL(k_fill_u32, dim3((n_words + 255) / 256), dim3(256), 0, st, static_cast<uint32_t*>(p),
                           n_words, v);
}

// Bucketing: each block classifies 4096 sites (16 per thread) with
// wave-aggregated LDS counters, reserves its slice of every list with one
// global atomic per class, then scatters the site indices.  With `all` set,
// sites of 0 or 1 record go to the first class too (the fused path calls
// every site); the records of all sites are summed into n_records.
constexpr uint32_t kBucketPer = 16;

__global__ __launch_bounds__(256) void k_bucket(const int32_t* nrec, uint32_t n_sites, uint32_t* lists,
                                                uint32_t pitch, uint32_t* counts, int all, int min_class,
                                                unsigned long long* n_records)
{
    __shared__ uint32_t cnt[6], base[6];
    __shared__ unsigned long long rsum[4];
    if (threadIdx.x < 6) cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t lane = lane_id();
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t s0 = blockIdx.x * 256u * kBucketPer;
    int cls[kBucketPer];
    uint32_t loc[kBucketPer];
    unsigned long long recs = 0;
    // places within the wave's share of each class (wave-uniform running
    // counts), then one LDS reservation per class per wave
    uint32_t run[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (uint32_t it = 0; it < kBucketPer; it++) {
        const uint32_t s = s0 + it * 256u + threadIdx.x;
        int b = -1;
        if (s < n_sites) {
            const int n = nrec[s];
            recs += (unsigned long long)n;
            b = n <= 1 ? (all ? 0 : -1) : n <= 8 ? 0 : n <= 16 ? 1 : n <= 32 ? 2 : n <= 64 ? 3 : n <= 128 ? 4 : 5;
            if (b >= 0 && b < min_class) b = min_class;
        }
        cls[it] = b;
        loc[it] = 0;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const uint64_t mask = __ballot(b == k);
            if (b == k) loc[it] = run[k] + (uint32_t)__popcll(mask & below);
            run[k] += (uint32_t)__popcll(mask);
        }
    }
    uint32_t wbase[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        uint32_t off = 0;
        if (lane == 0 && run[k]) off = atomicAdd(&cnt[k], run[k]);
        wbase[k] = (uint32_t)__shfl((int)off, 0);
    }
#pragma unroll
    for (uint32_t it = 0; it < kBucketPer; it++) {
        const int b = cls[it];
        uint32_t wb = 0;
#pragma unroll
        for (int k = 0; k < 6; k++) wb = b == k ? wbase[k] : wb;
        loc[it] += wb;
    }
    for (int off = 32; off > 0; off >>= 1) recs += __shfl_xor(recs, off);
    if (lane == 0) rsum[threadIdx.x >> 6] = recs;
    __syncthreads();
    if (threadIdx.x < 6) base[threadIdx.x] = cnt[threadIdx.x] ? atomicAdd(&counts[threadIdx.x], cnt[threadIdx.x]) : 0u;
    if (threadIdx.x == 0 && n_records) {
        const unsigned long long t = rsum[0] + rsum[1] + rsum[2] + rsum[3];
        if (t) atomicAdd(n_records, t);
    }
    __syncthreads();
#pragma unroll
    for (uint32_t it = 0; it < kBucketPer; it++) {
        const int b = cls[it];
        if (b >= 0) lists[(size_t)b * pitch + base[b] + loc[it]] = s0 + it * 256u + threadIdx.x;
    }
}

void launch_bucket(hipStream_t st, const int32_t* nrec, uint32_t n_sites, uint32_t* lists, uint32_t pitch,
                   uint32_t* counts, int all, unsigned long long* n_records, int min_class)
{
    launch_fill_u32(st, counts, 6, 0u);
    const uint32_t blocks = (n_sites + 256 * kBucketPer - 1) / (256 * kBucketPer);
    if (blocks)
        hipLaunchKernelGGL(k_bucket, dim3(blocks), dim3(256), 0, st, nrec, n_sites, lists, pitch, counts, all,
                           min_class, n_records);
}

template <int J>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v)
{
    if constexpr (J == 1) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    } else if constexpr (J == 2) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    } else if constexpr (J < 32) {
        return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (J << 10));  // bitmask mode, xor J
    } else {
        return (uint32_t)__shfl_xor((int)v, J);
    }
}

// median of three: with c = 0 the smaller of a and b, with c = ~0 the
// larger (one v_med3_u32; c depends on the lane and the stage only)
__device__ __forceinline__ uint32_t umed3(uint32_t a, uint32_t b, uint32_t c)
{
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    const uint32_t t = hi < c ? hi : c;
    return lo > t ? lo : t;
}

// all ones where the lane keeps the larger value of a compare-exchange at
// distance J within merge K (ascending where (e & K) == 0)
template <int K, int J>
__device__ __forceinline__ uint32_t keeps_max(uint32_t e)
{
    return ((e & K) == 0) == ((e & J) == 0) ? 0u : 0xFFFFFFFFu;
}

template <int K, int J>
__device__ __forceinline__ void bitonic_step(uint32_t& v, uint32_t e)
{
    const uint32_t w = xor_lane<J>(v);
    v = umed3(v, w, keeps_max<K, J>(e));
}

template <int K, int J>
struct Stage {
    __device__ __forceinline__ static void run(uint32_t& v, uint32_t e)
    {
        bitonic_step<K, J>(v, e);
        if constexpr (J > 1) Stage<K, J / 2>::run(v, e);
    }
};

template <int K, int S>
__device__ __forceinline__ void bitonic_all(uint32_t& v, uint32_t e)
{
    Stage<K, K / 2>::run(v, e);
    if constexpr (K < S) bitonic_all<K * 2, S>(v, e);
}

// Two independent networks step by step, so that each cross-lane exchange of
// one overlaps the other's.
template <int K, int J>
__device__ __forceinline__ void bitonic_step2(uint32_t& v0, uint32_t& v1, uint32_t e)
{
    const uint32_t w0 = xor_lane<J>(v0);
    const uint32_t w1 = xor_lane<J>(v1);
    const uint32_t c = keeps_max<K, J>(e);
    v0 = umed3(v0, w0, c);
    v1 = umed3(v1, w1, c);
}

template <int K, int J>
struct Stage2 {
    __device__ __forceinline__ static void run(uint32_t& v0, uint32_t& v1, uint32_t e)
    {
        bitonic_step2<K, J>(v0, v1, e);
        if constexpr (J > 1) Stage2<K, J / 2>::run(v0, v1, e);
    }
};

template <int K, int S>
__device__ __forceinline__ void bitonic_all2(uint32_t& v0, uint32_t& v1, uint32_t e)
{
    Stage2<K, K / 2>::run(v0, v1, e);
    if constexpr (K < S) bitonic_all2<K * 2, S>(v0, v1, e);
}

// widths 8..64: 64/S sites per wave, one record per lane
template <int S>
__global__ __launch_bounds__(256) void k_sort_lanes(uint32_t* rec, const int32_t* nrec, const uint32_t* list,
                                                    const uint32_t* count_ptr)
{
    const uint32_t count = *count_ptr;
    const uint32_t lane = lane_id();
    const uint32_t sub = lane / S, e = lane % S;
    constexpr uint32_t kPerWave = 64 / S;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwave = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t g = wave * kPerWave; g < count; g += nwave * kPerWave) {
        const uint32_t idx = g + sub;
        const bool ok = idx < count;
        uint32_t site = ok ? list[idx] : 0;
        const uint32_t n = ok ? (uint32_t)nrec[site] : 0;
        uint32_t* row = rec + (size_t)site * kSlots;
        uint32_t v = (e < n) ? row[e] : 0xFFFFFFFFu;
        bitonic_all<2, S>(v, e);
        if (e < n) row[e] = v;
    }
}

// one bitonic merge stage of width K >= 128 over R registers per lane
// (element e = r*64 + lane): in-lane compare-exchanges for j = K/2 .. 64,
// then cross-lane j = 32 .. 1; ascending where (e & K) == 0.
template <int R, int K, int JR>
__device__ __forceinline__ void regs_inlane(uint32_t* v, uint32_t lane)
{
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (r & JR) continue;
        const uint32_t e = (uint32_t)r * 64u + lane;
        const bool up = (e & (uint32_t)K) == 0;
        const uint32_t a = v[r], b = v[r | JR];
        const uint32_t mn = a < b ? a : b, mx = a < b ? b : a;
        v[r] = up ? mn : mx;
        v[r | JR] = up ? mx : mn;
    }
    if constexpr (JR > 1) regs_inlane<R, K, JR / 2>(v, lane);
}

template <int R, int K>
__device__ __forceinline__ void regs_merge(uint32_t* v, uint32_t lane)
{
    regs_inlane<R, K, K / 128>(v, lane);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t e = (uint32_t)r * 64u + lane;
        const bool up = (e & (uint32_t)K) == 0;
        uint32_t x = v[r];
#define GSNP_XSTEP(J)                                          \
    {                                                          \
        const uint32_t w = xor_lane<J>(x);                     \
        const bool low = (e & (J)) == 0;                       \
        const uint32_t mn = x < w ? x : w, mx = x < w ? w : x; \
        x = (up == low) ? mn : mx;                             \
    }
        GSNP_XSTEP(32) GSNP_XSTEP(16) GSNP_XSTEP(8) GSNP_XSTEP(4) GSNP_XSTEP(2) GSNP_XSTEP(1)
#undef GSNP_XSTEP
        v[r] = x;
    }
}

// widths 128/256: one site per wave, R = S/64 records per lane at e = r*64+lane
template <int S>
__global__ __launch_bounds__(256) void k_sort_regs(uint32_t* rec, const int32_t* nrec, const uint32_t* list,
                                                   const uint32_t* count_ptr)
{
    constexpr int R = S / 64;
    const uint32_t count = *count_ptr;
    const uint32_t lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwave = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t g = wave; g < count; g += nwave) {
        const uint32_t site = list[g];
        const uint32_t n = (uint32_t)nrec[site];
        uint32_t* row = rec + (size_t)site * kSlots;
        uint32_t v[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint32_t e = r * 64 + lane;
            v[r] = e < n ? row[e] : 0xFFFFFFFFu;
        }
        // k = 2..64: every step is cross-lane with e = r*64+lane, (e & k) from lane bits
#pragma unroll
        for (int r = 0; r < R; r++) bitonic_all<2, 64>(v[r], (uint32_t)(r * 64) + lane);
        // k = 128 (and 256): in-lane steps for j >= 64, then cross-lane j = 32..1
        regs_merge<R, 128>(v, lane);
        if constexpr (S == 256) regs_merge<R, 256>(v, lane);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint32_t e = r * 64 + lane;
            if (e < n) row[e] = v[r];
        }
    }
}

void launch_sort(hipStream_t st, uint32_t* rec, const int32_t* nrec, const uint32_t* lists, uint32_t pitch,
                 const uint32_t* counts, uint32_t max_sites)
{
    if (!max_sites) return;
    auto grid = [&](uint32_t per_wave) {
        uint32_t waves = (max_sites + per_wave - 1) / per_wave;
        uint32_t blocks = (waves + 3) / 4;
        return blocks < 2048u ? (blocks ? blocks : 1u) : 2048u;
    };
    hipLaunchKernelGGL(k_sort_lanes<8>, dim3(grid(8)), dim3(256), 0, st, rec, nrec, lists + 0 * (size_t)pitch, counts + 0);
    hipLaunchKernelGGL(k_sort_lanes<16>, dim3(grid(4)), dim3(256), 0, st, rec, nrec, lists + 1 * (size_t)pitch, counts + 1);
    hipLaunchKernelGGL(k_sort_lanes<32>, dim3(grid(2)), dim3(256), 0, st, rec, nrec, lists + 2 * (size_t)pitch, counts + 2);
    hipLaunchKernelGGL(k_sort_lanes<64>, dim3(grid(1)), dim3(256), 0, st, rec, nrec, lists + 3 * (size_t)pitch, counts + 3);
    hipLaunchKernelGGL(k_sort_regs<128>, dim3(grid(1)), dim3(256), 0, st, rec, nrec, lists + 4 * (size_t)pitch, counts + 4);
    hipLaunchKernelGGL(k_sort_regs<256>, dim3(grid(1)), dim3(256), 0, st, rec, nrec, lists + 5 * (size_t)pitch, counts + 5);
}

// ---------------------------------------------------------------------------
// (c) likelihood.  One thread per site over its sorted records
// (kernel/likelihood_kernel.cu:168-256): walk_adjusted gives each record's
// dependency-adjusted quality in record order and the 10 genotype scores of
// its row (one 128-byte line of the score table) are added in that order.
__global__ __launch_bounds__(256) void k_likelihood(Tables t, SiteArrays a, uint32_t n_sites, double* lk,
                                                    uint32_t* err)
{
    __shared__ uint32_t seen_all[256 * 17];
    uint32_t* seen = seen_all + threadIdx.x * 17;
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= n_sites) return;
    const uint32_t n = (uint32_t)a.nrec[s];
    const uint32_t* row = a.rec + (size_t)s * kSlots;
    double acc[10];
#pragma unroll
    for (int g = 0; g < 10; g++) acc[g] = 0.0;
    const uint32_t bad = walk_adjusted(row, n, t, seen, [&](uint32_t qa, uint32_t cyc, uint32_t o) {
        const double2* sr =
            reinterpret_cast<const double2*>(t.score + (size_t)((qa << 10) | (cyc << 2) | o) * kScoreRowStride);
#pragma unroll
        for (int h = 0; h < 5; h++) {
            const double2 v = sr[h];
            acc[2 * h] += v.x;
            acc[2 * h + 1] += v.y;
        }
    });
    double2* out = reinterpret_cast<double2*>(lk + (size_t)s * 10);
#pragma unroll
    for (int h = 0; h < 5; h++) out[h] = make_double2(acc[2 * h], acc[2 * h + 1]);
    if (bad) atomicOr(err, bad);
}

void launch_likelihood(hipStream_t st, const Tables& t, const SiteArrays& a, uint32_t n_sites, double* lk, uint32_t* err)
{
    const uint32_t blocks = (n_sites + 255) / 256;
    if (blocks) hipLaunchKernelGGL(k_likelihood, dim3(blocks), dim3(256), 0, st, t, a, n_sites, lk, err);
}

// ---------------------------------------------------------------------------
// (d) posterior (kernel/posterior_kernel.cu:635-786 and the q_cns rules of
// :833-864).  One thread per site.  Priors come as host-computed log10 rows
// (reference code rows, or a known-SNP row found by binary search).  A het
// call under the rank-sum mode is handed to k_rank through a job list.

__device__ __forceinline__ int get4(const int4& v, uint32_t i)
{
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}

// Calls site s from its ten log-likelihoods lkv(0..9).
template <class LK>
__device__ __forceinline__ void posterior_site(const Tables& t, const RefView& ref, uint32_t ws, const SiteArrays& a,
                                               uint32_t s, LK lkv, const int4* head_qs, const Results& out,
                                               RankJob* jobs, uint32_t* job_count)
{
    const uint32_t ori = ref.code[ws + s] & 0xF;
    const int depth = a.depth[s];
    if ((ori & 4) && depth == 0) return;  // N row, not called (nflag)
    const int dep_uni = a.nrec[s];
    const int4 qs = a.qs[s];
    const int4 ca = a.ca[s];
    // base1/base2 and qual1..3 (>= keeps the later base on ties)
    uint32_t base1 = 0, base2 = 0;
    int qual1 = -1, qual2 = -2, qual3 = -3;
    const int4 key = dep_uni ? qs : ca;
    int k1 = dep_uni ? -1 : 0, k2 = dep_uni ? -2 : 0, k3 = dep_uni ? -3 : 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) {
        const int v = get4(key, i);
        if (v >= k1) {
            k3 = k2;
            base2 = base1;
            k2 = k1;
            base1 = i;
            k1 = v;
        } else if (v >= k2) {
            k3 = k2;
            base2 = i;
            k2 = v;
        } else if (v >= k3) {
            k3 = v;
        }
    }
    if (k1 == 0)
        base1 = ori & 7;
    else if (k2 == 0 && base1 != (ori & 7))
        base2 = ori & 7;
    if (dep_uni) {
        qual1 = k1;
        qual2 = k2;
        qual3 = k3;
    }
    // log10 priors
    const double* lp = t.lprior + (ori & 7) * 10;
    if ((ori & 8) && t.refine && ref.n_snp) {
        // the known SNPs of the site's 256-site block, then a search among them
        const uint32_t p = ws + s;
        uint32_t lo = ref.snp_first[p >> 8], hi = ref.snp_first[(p >> 8) + 1];
        while (lo < hi) {
            uint32_t m = (lo + hi) >> 1;
            if (ref.snp_pos[m] < p)
                lo = m + 1;
            else
                hi = m;
        }
        if (lo < ref.n_snp && ref.snp_pos[lo] == p) lp = ref.snp_lp + (size_t)lo * 10;
    }
    uint32_t type1 = 16, type2 = 16;
    double tp1 = 0.0, tp2 = 0.0;
#pragma unroll
    for (int g = 0; g < 10; g++) {
        const uint32_t gt = geno_code(g);
        if (t.monoploid && (gt >> 2) != (gt & 3)) continue;
        const double v = lkv(g) + lp[g];
        if (v >= tp1 || type1 == 16) {
            type2 = type1;
            tp2 = tp1;
            type1 = gt;
            tp1 = v;
        } else if (v >= tp2 || type2 == 16) {
            type2 = gt;
            tp2 = v;
        }
    }
    const double dtp = tp1 - tp2;
    const bool hom = (type1 & 3) == ((type1 >> 2) & 3);
    // HET cap inputs; q_sum is flat over sites, so base codes 4..7 read the next
    // site -- for the last site of a window, as it stands before the reads
    // that start on the window boundary (head_qs in the batched span)
    bool het_ok = false;
    if (!hom) {
        const uint32_t het = base1 < base2 ? (base1 << 2 | base2) : (base2 << 2 | base1);
        int4 nq = qs;
        if (base1 >= 4 || base2 >= 4) {
            const uint32_t nx = ws + s + 1;
            nq = (head_qs && nx % kWin == 0) ? head_qs[nx / kWin - ws / kWin] : a.qs[s + 1];
        }
        const int q_b1 = base1 < 4 ? get4(qs, base1) : get4(nq, base1 & 3);
        const int q_b2 = base2 < 4 ? get4(qs, base2) : get4(nq, base2 & 3);
        het_ok = q_b1 > 0 && q_b2 > 0 && type1 == het;
    }
    double rank = 1.0;
    int q_cns;
    if (t.rank_mode && !hom) {
        const int4 cu = a.cu[s];
        if (get4(cu, type1 & 3) == 0 || get4(cu, (type1 >> 2) & 3) == 0) {
            rank = 0.0;
        } else {
            const uint32_t j = atomicAdd(job_count, 1u);
            // written as three words (a struct copy with byte fields would go through scratch)
            uint32_t* jw = reinterpret_cast<uint32_t*>(jobs + j);
            jw[0] = s;
            jw[1] = (type1 & 0xffu) | ((het_ok ? 1u : 0u) << 8);
            jw[2] = (uint32_t)(qual2 - qual3);
            jobs[j].dtp = dtp;
            out.base1[s] = (uint8_t)base1;
            out.base2[s] = (uint8_t)base2;
            out.type1[s] = (uint8_t)type1;
            return;  // rank and q_cns come from k_rank
        }
    }
    if (rank == 0.0)
        q_cns = 0;
    else
        q_cns = trunc_int_x86(10 * dtp + 0.0);  // 10*log10(1) == 0
    if (hom) {
        if (qual1 > 0 && base1 != (type1 & 3))
            q_cns = 0;
        else if (q_cns > qual1 - qual2)
            q_cns = qual1 - qual2;
    } else {
        if (het_ok) {
            if (q_cns > qual2 - qual3) q_cns = qual2 - qual3;
        } else {
            q_cns = 0;
        }
    }
    if (q_cns > 99) q_cns = 99;
    if (q_cns < 0) q_cns = 0;
    out.base1[s] = (uint8_t)base1;
    out.base2[s] = (uint8_t)base2;
    out.type1[s] = (uint8_t)type1;
    out.q_cns[s] = q_cns;
    out.rank[s] = rank;
}

__global__ __launch_bounds__(256) void k_posterior(Tables t, RefView ref, uint32_t ws, SiteArrays a, uint32_t n_sites,
                                                   const int4* head_qs, Results out, RankJob* jobs,
                                                   uint32_t* job_count)
{
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= n_sites) return;
    const double* lk = out.lk + (size_t)s * 10;
    posterior_site(t, ref, ws, a, s, [&](int g) { return lk[g]; }, head_qs, out, jobs, job_count);
}

void launch_posterior(hipStream_t st, const Tables& t, const RefView& ref, uint32_t ws, const SiteArrays& a,
                      uint32_t n_sites, const int4* head_qs, Results& out, RankJob* jobs, uint32_t* job_count)
{
    const uint32_t blocks = (n_sites + 255) / 256;
    if (blocks)
        hipLaunchKernelGGL(k_posterior, dim3(blocks), dim3(256), 0, st, t, ref, ws, a, n_sites, head_qs, out, jobs,
                           job_count);
}

// ---------------------------------------------------------------------------
// Batched form: sort + likelihood + posterior fused per depth class (the
// per-window path runs the same launch without the posterior).  A wave takes
// a group of G sites of one class (record count <= S) from the bucket lists;
// lane i owns site i of the group.
//  1. Sort: the record slots of a chunk of passes are loaded first (the loads
//     are in flight together), then for S <= 64 lane groups of S sort 64/S
//     sites per pass with the k_sort_lanes networks, for S = 128/256 one site
//     per pass with S/64 records per lane.  Each sorted record then learns,
//     in parallel (kernel/likelihood_kernel.cu:168-256 walks them in order):
//       t - 1: how many records before it share its (base, strand, cycle)
//              key -- the same base means the same segment;
//       g:     how many records of its base segment up to and including it
//              are first occurrences (t = 1), minus one.
//     One ballot per key bit gives each lane the lanes of its site holding
//     its key (all eleven bits) and its base (the two base bits), and one
//     more ballot the first occurrences; the wide classes, with S/64 records
//     per lane, take the ballots of every register.  (q, t - 1, g) indexes
//     the adjusted-quality table (copied into LDS by the classes up to 32
//     records when it is small) and the record is parked in LDS ([G][RS]
//     words) as its score row (q_adj << 10 | cycle << 2 | base).
//  2. Five lanes per site, two genotypes each, add the site's score values
//     in record order, eight records in flight: each load instruction reads
//     whole 80-byte rows of twelve sites (five contiguous lanes per row, one
//     line) instead of one 16-byte piece of 64 rows, which divides the
//     address work of the gathers by five.  Records are read once, with
//     non-temporal loads, so they do not push the score table out of L2.
//  3. The ten sums of a site go through its LDS row to the lane that owns
//     the site; posterior_site calls it from registers.

// G: sites per wave group.  The 32- and 64-record classes take 32: with
// 16-pass chunks and eight gathers in flight they fit in about 100 VGPRs, and
// their smaller LDS rows then let 16 waves share a CU; the wide classes hold
// few sites, so their groups are small to spread them over more waves.
template <int S>
constexpr uint32_t fused_group()
{
    return S <= 16 ? 64u : S <= 64 ? 32u : S == 128 ? 16u : 8u;
}

// LDS words per site row: odd, so that row walks of the 64 lanes are
// conflict-free, and at least 20 so that a row can hand over the ten sums
template <int S>
constexpr uint32_t fused_row_stride()
{
    return S + 1 < 21 ? 21u : (uint32_t)S + 1u;
}

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// (q, t - 1, g) table index << 10 | cycle << 2 | base of a sorted record
__device__ __forceinline__ uint32_t qadj_slot(uint32_t w, uint32_t t, uint32_t g, uint32_t td, uint32_t gd,
                                              uint32_t& bad)
{
    uint32_t q = rec_q(w);
    if (q > 63) {
        bad |= kErrQadjRange;
        q = 63;
    }
    const uint32_t tm1 = t < td ? t : td - 1;
    const uint32_t gq = g < gd ? g : gd - 1;
    return (((q * td + tm1) * gd + gq) << 10) | (rec_cycle(w) << 2) | rec_base(w);  // index < 2^22
}

// score row of a qadj_slot value
__device__ __forceinline__ uint32_t score_row(const uint8_t* qadj, uint32_t slot, uint32_t& bad)
{
    uint32_t qa = qadj[slot >> 10];
    if (qa == 0) {
        bad |= kErrQadjRange;
        qa = 1;
    }
    return (qa << 10) | (slot & 1023u);
}

// The same for a site sorted over R registers (record e = r * 64 + lane):
// ballots of every register per key bit, so that record (r, lane) counts the
// matching lanes of registers r' < r whole and of register r below it.
template <int R>
__device__ __forceinline__ void wide_slots(const uint32_t* v, uint64_t below, uint64_t upto, uint32_t td, uint32_t gd,
                                           uint32_t& bad, uint32_t* slot)
{
    uint64_t live[R], eq[R][R], ob[R][R];  // [mine][theirs], theirs <= mine
#pragma unroll
    for (int r = 0; r < R; r++) live[r] = __ballot(v[r] != 0xFFFFFFFFu);
    // lanes whose key bit differs from mine, per 32-bit half: one signed
    // one-bit extract and one ballot per register and bit, one three-input
    // operation per (mine, theirs) pair and half; base bits first, so that
    // the base masks are the key masks' start
    uint32_t dlo[R][R], dhi[R][R], olo[R][R], ohi[R][R];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int q = 0; q <= r; q++) dlo[r][q] = dhi[r][q] = olo[r][q] = ohi[r][q] = 0u;
#pragma unroll
    for (int b = 0; b < 11; b++) {
        const uint32_t bit = b < 2 ? 23u + (uint32_t)b : b == 2 ? 8u : (uint32_t)(6 + b);  // 23, 24, 8, 9..16
        uint32_t m[R];
        uint64_t bal[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            m[r] = (uint32_t)__builtin_amdgcn_sbfe((int)v[r], bit, 1u);
            bal[r] = __ballot((int)m[r] < 0);
        }
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int q = 0; q <= r; q++) {
                dlo[r][q] = __builtin_amdgcn_bitop3_b32(dlo[r][q], (uint32_t)bal[q], m[r], 0xF6);
                dhi[r][q] = __builtin_amdgcn_bitop3_b32(dhi[r][q], (uint32_t)(bal[q] >> 32), m[r], 0xF6);
                if (b == 1) {
                    olo[r][q] = dlo[r][q];
                    ohi[r][q] = dhi[r][q];
                }
            }
    }
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int q = 0; q <= r; q++) {
            eq[r][q] = live[q] & ~(((uint64_t)dhi[r][q] << 32) | dlo[r][q]);
            ob[r][q] = live[q] & ~(((uint64_t)ohi[r][q] << 32) | olo[r][q]);
        }
    uint32_t t[R];
    uint64_t first[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        uint32_t c = (uint32_t)__popcll(eq[r][r] & below);
#pragma unroll
        for (int q = 0; q < r; q++) c += (uint32_t)__popcll(eq[r][q]);
        t[r] = c;
        first[r] = __ballot(v[r] != 0xFFFFFFFFu && c == 0);
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        uint32_t g = (uint32_t)__popcll(ob[r][r] & first[r] & upto);
#pragma unroll
        for (int q = 0; q < r; q++) g += (uint32_t)__popcll(ob[r][q] & first[q]);
        uint32_t b2 = 0;
        slot[r] = v[r] != 0xFFFFFFFFu ? qadj_slot(v[r], t[r], g - 1u, td, gd, b2) : 0xFFFFFFFFu;
        bad |= b2;
    }
}

// Adjusted-quality tables up to this size are copied into LDS by the
// classes of S <= 32 (QL), whose rows leave room for them.
constexpr uint32_t kQadjLdsMax = 20480;

__host__ __device__ constexpr uint32_t qadj_lds_words(uint32_t td, uint32_t gd)
{
    return ((64u * td * gd + 15u) / 16u) * 4u;  // whole 16-byte units
}

// Up to 128 records the kernel is held to 128 VGPRs, four waves per SIMD
// (the 64-record class would take 148 and three waves; with four, the
// stage is about 3 % faster); the 256 class keeps its registers.
template <int S, bool QL = false, bool POST = true>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(S <= 128 ? 4 : 1, 8))) void k_fused(FusedArgs f, const uint32_t* list, const uint32_t* count_ptr)
{
    constexpr uint32_t G = fused_group<S>();
    constexpr uint32_t RS = fused_row_stride<S>();
    extern __shared__ uint32_t lds[];
    const uint32_t lane = lane_id();
    const uint32_t wib = threadIdx.x >> 6;
    const uint8_t* qtab = f.t.qadj;
    uint32_t qwords = 0;
    if constexpr (QL) {
        // the block copies the table once (words, then the last bytes)
        const uint32_t qbytes = 64u * (uint32_t)f.t.td * (uint32_t)f.t.gd, nw = qbytes / 4u;
        const uint32_t* src = reinterpret_cast<const uint32_t*>(f.t.qadj);
        for (uint32_t i = threadIdx.x; i < nw; i += blockDim.x) lds[i] = src[i];
        if (threadIdx.x < (qbytes & 3u))
            reinterpret_cast<uint8_t*>(lds)[4u * nw + threadIdx.x] = f.t.qadj[4u * nw + threadIdx.x];
        __syncthreads();
        qtab = reinterpret_cast<const uint8_t*>(lds);
        qwords = qadj_lds_words((uint32_t)f.t.td, (uint32_t)f.t.gd);
    }
    uint32_t* srt = lds + qwords + wib * (G * RS);
    const uint32_t count = *count_ptr;
    const uint32_t ngroups = (count + G - 1) / G;
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + wib;
    const uint32_t nwave = gridDim.x * (blockDim.x >> 6);
    // table extents are at most 256 (the masks let the index use 24-bit multiplies)
    const uint32_t td = (uint32_t)f.t.td & 511u, gd = (uint32_t)f.t.gd & 511u;
    // phase-2 roles: lane = 5 * site slot + 16-byte piece of the row (lanes 60..63 idle)
    const uint32_t sl = lane / 5u, cl = lane - 5u * sl;
    const char* score = reinterpret_cast<const char*>(f.t.score);
    uint32_t bad = 0;
    for (uint32_t grp = wave; grp < ngroups; grp += nwave) {
        const uint32_t g0 = grp * G;
        const uint32_t gcount = count - g0 < G ? count - g0 : G;
        // lane gi owns site gi of the group
        const bool act = lane < gcount;
        const uint32_t site = act ? list[g0 + lane] : 0u;
        const uint32_t n = act ? (uint32_t)f.a.nrec[site] : 0u;
        wave_lds_sync();  // the previous group's rows have been read
        // phase 1: sort, dependency counts, score rows parked in LDS
        if constexpr (S <= 64) {
            constexpr uint32_t P = 64 / S, NP = G / P;
            const uint32_t sub = lane / S, e = lane % S, gbase = lane & ~(uint32_t)(S - 1);
            const uint64_t gmask = (S == 64 ? ~0ull : ((1ull << S) - 1ull)) << gbase;  // lanes of my group
            const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
            const uint64_t upto = below | (1ull << lane);
            uint32_t qf_min = 63u, qa_min = 255u;  // out-of-table qualities: either reaches 0
            // in chunks of CH passes (bounds the registers)
            constexpr uint32_t CH = NP < 16 ? NP : S == 64 ? 8 : 16;  // 64: fewer registers
            static_assert(G % P == 0 && NP % CH == 0 && CH % 2 == 0, "whole chunks of paired passes");
#pragma unroll
            for (uint32_t k0 = 0; k0 < NP; k0 += CH) {
                uint32_t v[CH];
                uint32_t pad = 0;  // bit k: my slot of pass k is padding
                // every lane loads (a padding lane's slot lies in its site's
                // row, an idle group's in site 0's, and on cache lines the
                // site's records touch anyway), so that no load waits behind
                // a branch; padding slots are set to all ones afterwards
                uint32_t nk[CH];
#pragma unroll
                for (uint32_t k = 0; k < CH; k++) {
                    const uint32_t gi = (k0 + k) * P + sub;
                    uint32_t st;
                    if constexpr (P == 1) {  // one site per pass: its index and count in scalar registers
                        st = (uint32_t)__builtin_amdgcn_readlane((int)site, (int)(k0 + k));
                        nk[k] = (uint32_t)__builtin_amdgcn_readlane((int)n, (int)(k0 + k));
                    } else {
                        st = (uint32_t)__shfl((int)site, (int)gi);
                        nk[k] = (uint32_t)__shfl((int)n, (int)gi);
                    }
                    v[k] = __builtin_nontemporal_load(f.a.rec + (size_t)st * kSlots + e);
                }
#pragma unroll
                for (uint32_t k = 0; k < CH; k++) v[k] = e < nk[k] ? v[k] : 0xFFFFFFFFu;
#pragma unroll
                for (uint32_t k = 0; k < CH; k += 2) bitonic_all2<2, S>(v[k], v[k + 1], e);  // CH is even
#pragma unroll
                for (uint32_t k = 0; k < CH; k++) {
                    // the lanes of my site that agree with me on every key bit
                    // (one ballot per bit, 32-bit halves), below me; and those
                    // with my base, up to me; padding lanes (all ones: a record's
                    // occurrence byte is 1) take no part.  Per bit: a signed
                    // one-bit extract (all ones where my bit is set), the compare
                    // that makes the ballot, and per half one three-input
                    // operation that adds the lanes whose bit differs from mine
                    // (d | (ballot ^ m)) to the difference mask.  The two base
                    // bits go first, so that the base mask is the key mask's
                    // start.
                    const uint32_t w = v[k];
                    const bool my = w != 0xFFFFFFFFu;
                    const uint64_t live = __ballot(my) & gmask;
                    uint32_t dlo = 0, dhi = 0, olo = 0, ohi = 0;
#pragma unroll
                    for (int b = 0; b < 11; b++) {
                        const uint32_t bit = b < 2 ? 23u + (uint32_t)b : b == 2 ? 8u : (uint32_t)(6 + b);  // 23, 24, 8, 9..16
                        const uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int)w, bit, 1u);
                        const uint64_t bal = __ballot((int)m < 0);
                        dlo = __builtin_amdgcn_bitop3_b32(dlo, (uint32_t)bal, m, 0xF6);
                        dhi = __builtin_amdgcn_bitop3_b32(dhi, (uint32_t)(bal >> 32), m, 0xF6);
                        if (b == 1) {
                            olo = (uint32_t)(live & upto) & ~dlo;  // my base, up to me
                            ohi = (uint32_t)((live & upto) >> 32) & ~dhi;
                        }
                    }
                    const uint32_t plo = (uint32_t)(live & below) & ~dlo, phi = (uint32_t)((live & below) >> 32) & ~dhi;
                    const uint32_t t = (uint32_t)(__popc(plo) + __popc(phi));
                    // first occurrences; olo/ohi hold live lanes only
                    const uint64_t first = __ballot(t == 0);
                    const uint32_t g =
                        (uint32_t)(__popc(olo & (uint32_t)first) + __popc(ohi & (uint32_t)(first >> 32))) - 1u;
                    // (q, t - 1, g) table index << 10 | cycle << 2 | base;
                    // padding lanes take a q 63 record's
                    // (an index inside the table; their row is replaced below).
                    // A quality field of 0 (q = 64, beyond the table) is
                    // remembered in qf_min and clamped to q 63.
                    const uint32_t x = my ? w : 0x00020001u;
                    const uint32_t qf = (x >> 17) & 63u;  // 64 - q
                    qf_min = qf < qf_min ? qf : qf_min;
                    const uint32_t q = 64u - (qf > 1u ? qf : 1u);
                    const uint32_t tm1 = t < td ? t : td - 1u;
                    const uint32_t gq = g < gd ? g : gd - 1u;
                    const uint32_t idx = (q * td + tm1) * gd + gq;  // < 2^22
                    v[k] = (idx << 10) | ((x >> 7) & 0x3FCu) | ((x >> 23) & 3u);
                    pad |= my ? 0u : 1u << k;
                }
                // score rows into LDS; padding slots [n, S) get row 0, whose
                // ten values are +0.0, so that phase 2 adds whole groups of
                // eight slots with no per-slot test (a sum that starts at +0.0
                // is unchanged by adding +0.0)
#pragma unroll
                for (uint32_t k = 0; k < CH; k++) {
                    const bool my = !((pad >> k) & 1u);
                    const uint32_t qa = qtab[v[k] >> 10];  // 0 marks a quality beyond the table
                    qa_min = my && qa < qa_min ? qa : qa_min;
                    const uint32_t row = ((qa > 1u ? qa : 1u) << 10) | (v[k] & 1023u);
                    srt[((k0 + k) * P + sub) * RS + e] = my ? row : 0u;
                }
            }
            if (qf_min == 0u || qa_min == 0u) bad |= kErrQadjRange;
        } else {
            constexpr int R = S / 64;
            uint32_t v[G][R];
#pragma unroll
            for (uint32_t gi = 0; gi < G; gi++) {
                const uint32_t st = (uint32_t)__builtin_amdgcn_readlane((int)site, (int)gi);
                const uint32_t nn = (uint32_t)__builtin_amdgcn_readlane((int)n, (int)gi);
#pragma unroll
                for (int r = 0; r < R; r++) {  // unconditional, as above: e < S <= kSlots
                    const uint32_t e = (uint32_t)r * 64u + lane;
                    const uint32_t w = f.a.rec[(size_t)st * kSlots + e];
                    v[gi][r] = e < nn ? w : 0xFFFFFFFFu;
                }
            }
            const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
            const uint64_t upto = below | (1ull << lane);
#pragma unroll
            for (uint32_t gi = 0; gi < G; gi++) {
#pragma unroll
                for (int r = 0; r < R; r++) bitonic_all<2, 64>(v[gi][r], (uint32_t)(r * 64) + lane);
                regs_merge<R, 128>(v[gi], lane);
                if constexpr (S == 256) regs_merge<R, 256>(v[gi], lane);
                uint32_t slot[R];
                wide_slots<R>(v[gi], below, upto, td, gd, bad, slot);
#pragma unroll
                for (int r = 0; r < R; r++)  // padding slots: row 0, all +0.0
                    srt[gi * RS + (uint32_t)r * 64u + lane] =
                        slot[r] != 0xFFFFFFFFu ? score_row(f.t.qadj, slot[r], bad) : 0u;
            }
        }
        wave_lds_sync();
        // phase 2: twelve sites at a time, lane (sl, cl) adds genotypes 2cl and
        // 2cl+1 of site r * 12 + sl in record order, eight records in flight; the
        // sums replace words 4cl..4cl+3 of the site's row once every lane is
        // past the row
        constexpr uint32_t U = 8;
#pragma unroll 1
        for (uint32_t r = 0; r < (G + 11) / 12; r++) {
            const uint32_t gi = r * 12u + sl;
            const uint32_t ng = (uint32_t)__shfl((int)n, (int)(gi & 63u));
            const bool on = lane < 60u && gi < gcount;
            const uint32_t m = on ? ng : 0u;
            const uint32_t* rw = srt + (on ? gi : 0u) * RS;
            double a0 = 0.0, a1 = 0.0;
            // i + u < i + U <= S rounded up to U: inside the row (RS > S, RS >= 21);
            // slots past the site's records hold row 0, all +0.0
            for (uint32_t i = 0; i < m; i += U) {
                double2 x[U];
#pragma unroll
                for (uint32_t u = 0; u < U; u++)
                    x[u] = *reinterpret_cast<const double2*>(score + ((rw[i + u] << 7) | (cl << 4)));
#pragma unroll
                for (uint32_t u = 0; u < U; u++) {
                    a0 += x[u].x;
                    a1 += x[u].y;
                }
            }
            wave_lds_sync();
            if (on) {
                const uint64_t b0 = (uint64_t)__double_as_longlong(a0), b1 = (uint64_t)__double_as_longlong(a1);
                uint32_t* d = srt + gi * RS + 4u * cl;
                d[0] = (uint32_t)b0;
                d[1] = (uint32_t)(b0 >> 32);
                d[2] = (uint32_t)b1;
                d[3] = (uint32_t)(b1 >> 32);
            }
        }
        wave_lds_sync();
        if (!act) continue;
        // phase 3: posterior from the sums
        double acc[10];
#pragma unroll
        for (int g = 0; g < 10; g++)
            acc[g] = __longlong_as_double((long long)(((uint64_t)srt[lane * RS + 2 * g + 1] << 32) |
                                                      srt[lane * RS + 2 * g]));
        if (f.write_lk) {
            double2* out = reinterpret_cast<double2*>(f.out.lk + (size_t)site * 10);
#pragma unroll
            for (int h = 0; h < 5; h++) out[h] = make_double2(acc[2 * h], acc[2 * h + 1]);
        }
        if constexpr (POST)
            posterior_site(f.t, f.ref, f.ws, f.a, site, [&](int g) { return acc[g]; }, f.head_qs, f.out, f.jobs,
                           f.job_count);
    }
    if (bad) atomicOr(f.err, bad);
}

template <int S>
static void launch_fused_class(hipStream_t st, const FusedArgs& f, const uint32_t* list, const uint32_t* count,
                               uint32_t max_sites, uint32_t want_wpb, uint32_t max_blocks)
{
    constexpr uint32_t G = fused_group<S>();
    const uint32_t per_wave = G * fused_row_stride<S>() * 4u;
    const bool ql = S <= 32 && 64u * (uint32_t)f.t.td * (uint32_t)f.t.gd <= kQadjLdsMax;
    const uint32_t table = ql ? 4u * qadj_lds_words((uint32_t)f.t.td, (uint32_t)f.t.gd) : 0u;
    // waves per block: eight when the LDS table is shared (one copy per block)
    uint32_t wpb = (ql && table + 8u * per_wave <= 65536u) ? 8u
                 : table + 4u * per_wave <= 65536u         ? 4u
                 : table + 2u * per_wave <= 65536u         ? 2u
                                                            : 1u;
    if (want_wpb && want_wpb < wpb) wpb = want_wpb;
    const uint32_t groups = (max_sites + G - 1) / G;
    uint32_t blocks = (groups + wpb - 1) / wpb;
    if (blocks > max_blocks) blocks = max_blocks;
    if (!blocks) blocks = 1;
    // the batched form's launches (posterior fused) and the per-window path's
    // (likelihoods only) are separate instantiations
    if constexpr (S <= 32) {
        if (ql) {
            if (f.posterior)
                hipLaunchKernelGGL((k_fused<S, true, true>), dim3(blocks), dim3(64 * wpb), table + wpb * per_wave, st,
                                   f, list, count);
            else
                hipLaunchKernelGGL((k_fused<S, true, false>), dim3(blocks), dim3(64 * wpb), table + wpb * per_wave, st,
                                   f, list, count);
            return;
        }
    }
    if (f.posterior)
        hipLaunchKernelGGL((k_fused<S, false, true>), dim3(blocks), dim3(64 * wpb), wpb * per_wave, st, f, list, count);
    else
        hipLaunchKernelGGL((k_fused<S, false, false>), dim3(blocks), dim3(64 * wpb), wpb * per_wave, st, f, list,
                           count);
}

void launch_fused_one(hipStream_t st, const FusedArgs& f, const uint32_t* lists, uint32_t pitch,
                      const uint32_t* counts, uint32_t max_sites, uint32_t cls, uint32_t wpb, uint32_t max_blocks)
{
    if (!max_sites) return;
    const uint32_t* list = lists + (size_t)cls * pitch;
    const uint32_t* count = counts + cls;
    if (!max_blocks) max_blocks = 8192u;
    switch (cls) {
    case 0: launch_fused_class<8>(st, f, list, count, max_sites, wpb, max_blocks); break;
    case 1: launch_fused_class<16>(st, f, list, count, max_sites, wpb, max_blocks); break;
    case 2: launch_fused_class<32>(st, f, list, count, max_sites, wpb, max_blocks); break;
    case 3: launch_fused_class<64>(st, f, list, count, max_sites, wpb, max_blocks); break;
    case 4: launch_fused_class<128>(st, f, list, count, max_sites, wpb, max_blocks); break;
    default: launch_fused_class<256>(st, f, list, count, max_sites, wpb, max_blocks); break;
    }
}

void launch_fused(hipStream_t st, const FusedArgs& f, const uint32_t* lists, uint32_t pitch, const uint32_t* counts,
                  uint32_t max_sites, const FusedTune& tune)
{
    if (!max_sites) return;
    // deepest classes first: their groups are the longest
    for (int cls = 5; cls >= (int)tune.min_class; cls--)
        launch_fused_one(st, f, lists, pitch, counts, max_sites, (uint32_t)cls, tune.wpb[cls], tune.max_blocks[cls]);
}

// ---------------------------------------------------------------------------
// (e) Wilcoxon rank-sum test on het calls (kernel/posterior_kernel.cu:506-572,
// cpp/rank_sum.cc:44-58).  One wave per job: lane q counts the records of the
// two alleles with quality q, a wave prefix sum gives the mid-ranks, and
// rank sums T1/T2 (exact half-integers) are reduced across lanes.  N < 64
// uses the exact table (including its <<16 interpolation index); N >= 64
// takes the normal approximation, whose exp and sqrt are evaluated so that
// they give the host libm's doubles (below).  A site whose exp or sqrt lies
// too close to a rounding boundary to be sure of that, or whose log10 lies
// within 1e-9 of an integer boundary, goes to the host fix-up list.

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__global__ __launch_bounds__(256) void k_rank(Tables t, SiteArrays a, const RankJob* jobs, const uint32_t* job_count,
                                              Results out, FixJob* fixes, uint32_t* fix_count,
                                              unsigned long long* normal_count)
{
    const uint32_t count = *job_count;
    const uint32_t lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwave = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t j = wave; j < count; j += nwave) {
        const RankJob jb = jobs[j];
        const uint32_t s = jb.site;
        const uint32_t A = jb.type1 & 3, B = (jb.type1 >> 2) & 3;  // n1 allele, n2 allele
        const uint32_t n = (uint32_t)a.nrec[s];
        const uint32_t* row = a.rec + (size_t)s * kSlots;
        // same-quality counts: lane q
        int cnt = 0;
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t w = row[i];
            const uint32_t o = rec_base(w);
            if ((o == A || o == B) && rec_q(w) == lane) cnt += (int)rec_occ(w);
        }
        // exclusive prefix over lanes 0..q (only q <= q_max_min+1 take part)
        const bool part = (int)lane <= t.q_max_min + 1;
        int incl = part ? cnt : 0;
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(incl, off);
            if ((int)lane >= off) incl += y;
        }
        const int excl = incl - (part ? cnt : 0);
        const double rank_q = part ? excl + (1 + cnt) / 2.0 : 0.0;
        double ta = 0.0, tb = 0.0;
        for (uint32_t i0 = 0; i0 < n; i0 += 64) {
            const uint32_t i = i0 + lane;
            const uint32_t w = i < n ? row[i] : 0u;
            const uint32_t q = rec_q(w) & 63;
            const double rq = __shfl(rank_q, (int)q);
            if (i < n) {
                const uint32_t o = rec_base(w);
                const double c = rq * rec_occ(w);
                if (o == A) ta += c;
                if (o == B) tb += c;
            }
        }
        ta = wave_sum(ta);
        tb = wave_sum(tb);
        if (lane != 0) continue;
        const int4 cu = a.cu[s];
        const int n1 = get4(cu, A), n2 = get4(cu, B);
        bool need_host = false;
        double rank = 0.0;
        if (n1 + n2 < 64) {
            const double* p = t.prank;
            if (n1 <= n2) {
                const int i1 = (n1 + n2) << 17 | n1 << 11 | (int)(ta);
                const int i2 = (n1 + n2) << 16 | n1 << 11 | (int)(ta + 1);
                rank = p[i1] + (ta - (int)ta) * (p[i2] - p[i1]);
            } else {
                const int i1 = (n1 + n2) << 17 | n2 << 11 | (int)(tb);
                const int i2 = (n1 + n2) << 16 | n2 << 11 | (int)(tb + 1);
                rank = p[i1] + (tb - (int)tb) * (p[i2] - p[i1]);
            }
        } else {
            bool hard = false;
            rank = normal_test_nearest(n1, n2, ta, tb, hard);
            need_host = hard;
            if (!hard) atomicAdd(normal_count, 1ull);
        }
        int q_cns = 0;
        if (!need_host) {
            if (rank == 0.0) {
                q_cns = 0;
            } else {
                const double l = rank == 1.0 ? 0.0 : log10(rank);
                const double y = 10 * jb.dtp + 10 * l;
                if (y == y) {  // NaN gives INT_MIN on both sides, no guard needed
                    const double r = rint(y);
                    const double band = 1e-9 * (1.0 + fabs(y) + fabs(10 * l));
                    if (fabs(y - r) <= band || !(fabs(y) < 2.0e9)) need_host = true;
                }
                q_cns = trunc_int_x86(y);
            }
        }
        if (need_host) {
            const uint32_t f = atomicAdd(fix_count, 1u);
            FixJob fx;
            fx.site = s;
            fx.n1 = n1;
            fx.n2 = n2;
            fx.het_ok = jb.het_ok;
            fx.normal = n1 + n2 >= 64;
            fx.pad0 = fx.pad1 = 0;
            fx.cap = jb.cap;
            fx.T1 = ta;
            fx.T2 = tb;
            fx.dtp = jb.dtp;
            fx.table_rank = rank;
            fixes[f] = fx;
            continue;
        }
        if (jb.het_ok) {
            if (q_cns > jb.cap) q_cns = jb.cap;
        } else {
            q_cns = 0;
        }
        if (q_cns > 99) q_cns = 99;
        if (q_cns < 0) q_cns = 0;
        out.q_cns[s] = q_cns;
        out.rank[s] = rank;
    }
}

void launch_rank(hipStream_t st, const Tables& t, const SiteArrays& a, const RankJob* jobs, const uint32_t* job_count,
                 uint32_t max_jobs, Results& out, FixJob* fixes, uint32_t* fix_count, unsigned long long* normal_count)
{
    uint32_t blocks = (max_jobs + 3) / 4;
    if (blocks > 4096) blocks = 4096;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_rank, dim3(blocks), dim3(256), 0, st, t, a, jobs, job_count, out, fixes, fix_count,
                       normal_count);
}

__global__ void k_apply_fix(const FixJob* fixes, const double* rank, const int32_t* qcns, uint32_t n, Results out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = fixes[i].site;
    out.rank[s] = rank[i];
    out.q_cns[s] = qcns[i];
}

void launch_apply_fix(hipStream_t st, const FixJob* fixes, const double* rank, const int32_t* qcns, uint32_t n,
                      Results& out)
{
    if (n) hipLaunchKernelGGL(k_apply_fix, dim3((n + 255) / 256), dim3(256), 0, st, fixes, rank, qcns, n, out);
}

// ---------------------------------------------------------------------------
// (f) .cns rows (cpp/call_genotype.cc:353-411).  k_rowlen sizes each row and
// sums per block; a one-block scan turns block sums into offsets; k_format
// writes each block's rows into LDS at their block-local offsets and copies
// the contiguous block text out with wide stores (rows of very long
// chromosome names fall back to direct stores).

__device__ __forceinline__ double copy_number(int depth, int repeat)
{
    return depth == 0 ? 255.0 : (double)repeat / depth;
}

struct RowData {
    uint32_t site;
    bool nrow;
    uint32_t ori, type1, b1, b2;
    int q_cns, depth, repeat;
    int avg1, cu1, ca1, avg2, cu2, ca2;
    double rank;
};

__device__ __forceinline__ RowData load_row(const RefView& ref, uint32_t ws, const SiteArrays& a, const Results& r,
                                            uint32_t s)
{
    RowData d;
    d.site = s;
    d.ori = ref.code[ws + s] & 0xF;
    d.depth = a.depth[s];
    d.nrow = (d.ori & 4) && d.depth == 0;
    d.b1 = d.nrow ? 4 : r.base1[s];
    d.nrow = d.nrow || d.b1 >= 4;
    if (d.nrow) return d;
    d.b2 = r.base2[s];
    d.type1 = r.type1[s];
    d.q_cns = r.q_cns[s];
    d.repeat = a.repeat[s];
    d.rank = r.rank[s];
    const int4 qs = a.qs[s], cu = a.cu[s], ca = a.ca[s];
    const int q1 = get4(qs, d.b1);
    d.cu1 = get4(cu, d.b1);
    d.ca1 = get4(ca, d.b1);
    d.avg1 = q1 == 0 ? 0 : q1 / d.cu1;
    if (d.b2 < 4) {
        const int q2 = get4(qs, d.b2);
        d.cu2 = get4(cu, d.b2);
        d.ca2 = get4(ca, d.b2);
        d.avg2 = q2 == 0 ? 0 : q2 / d.cu2;
    }
    return d;
}

__constant__ char kNRow[] = "\tN\tN\t0\tN\t0\t0\t0\tN\t0\t0\t0\t0\t1.000\t255.000\t0\n";
constexpr int kNRowLen = sizeof(kNRow) - 1;
__constant__ char kACTGN[] = "ACTGNNNN";
__constant__ char kAbbv[] = "AMWRMCYSWYTKRSKGN";

// Row output: characters go through put(pos, c), each digit straight to its
// position (a character array read at run-time indexes would compile to
// compare/select chains over the whole array).
__device__ __forceinline__ int digits_u32(uint32_t v)
{
    return v < 10u ? 1 : v < 100u ? 2 : v < 1000u ? 3 : v < 10000u ? 4 : v < 100000u ? 5 : v < 1000000u ? 6
         : v < 10000000u ? 7 : v < 100000000u ? 8 : v < 1000000000u ? 9 : 10;
}

template <class Put>
__device__ __forceinline__ int put_u32(uint32_t v, int n, Put& put)
{
    const int k = digits_u32(v);
    for (int p = n + k - 1; p >= n; p--) {
        put(p, (char)('0' + v % 10u));
        v /= 10u;
    }
    return n + k;
}

template <class Put>
__device__ __forceinline__ int put_i32(int v, int n, Put& put)
{
    if (v < 0) {
        put(n, '-');
        return put_u32(0u - (uint32_t)v, n + 1, put);
    }
    return put_u32((uint32_t)v, n, put);
}

// %#g text: six digits with the point after digit X (fmt_g_digits for
// 0 <= X <= 5), or the pre-pass text of the site (rank field 0, copy number
// 1) when the FP64 path does not settle it
template <class Put>
__device__ __forceinline__ int put_g(double v, const RowCtx& rc, uint32_t site, int field, int n, Put& put)
{
    uint32_t D;
    int X;
    if (g_fast_digits(v, D, X)) {
        for (int i = 5; i >= 0; i--) {
            put(n + i + (i > X ? 1 : 0), (char)('0' + D % 10u));
            D /= 10u;
        }
        put(n + X + 1, '.');
        return n + 7;
    }
    const GText& gt = rc.gtext[rc.gidx[site]];
    const char* t = field == 0 ? gt.rank : gt.copy;
    const int k = t[15];
    for (int i = 0; i < k; i++) put(n + i, t[i]);
    return n + k;
}

// Writes the row through `put(pos, char)`; returns its length.
template <class Put>
__device__ __forceinline__ int emit_row(const RowData& d, const RowCtx& rc, uint32_t pos1, Put put)
{
    int n = 0;
    for (int i = 0; i < rc.name_len; i++) put(n++, rc.name[i]);
    put(n++, '\t');
    n = put_u32(pos1, n, put);
    if (d.nrow) {
#pragma unroll
        for (int i = 0; i < kNRowLen; i++) put(n + i, kNRow[i]);
        return n + kNRowLen;
    }
    put(n++, '\t');
    put(n++, kACTGN[d.ori & 7]);
    put(n++, '\t');
    put(n++, kAbbv[d.type1]);
    put(n++, '\t');
    n = put_i32(d.q_cns, n, put);
    put(n++, '\t');
    put(n++, kACTGN[d.b1]);
    put(n++, '\t');
    n = put_i32(d.avg1, n, put);
    put(n++, '\t');
    n = put_i32(d.cu1, n, put);
    put(n++, '\t');
    n = put_i32(d.ca1, n, put);
    put(n++, '\t');
    if (d.b2 < 4) {
        put(n++, kACTGN[d.b2]);
        put(n++, '\t');
        n = put_i32(d.avg2, n, put);
        put(n++, '\t');
        n = put_i32(d.cu2, n, put);
        put(n++, '\t');
        n = put_i32(d.ca2, n, put);
    } else {
        put(n++, 'N');
        put(n++, '\t');
        put(n++, '0');
        put(n++, '\t');
        put(n++, '0');
        put(n++, '\t');
        put(n++, '0');
    }
    put(n++, '\t');
    n = put_i32(d.depth, n, put);
    put(n++, '\t');
    n = put_g(d.rank, rc, d.site, 0, n, put);
    put(n++, '\t');
    n = put_g(copy_number(d.depth, d.repeat), rc, d.site, 1, n, put);
    put(n++, '\t');
    put(n++, (d.ori & 8) ? '1' : '0');
    put(n++, '\n');
    return n;
}

__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v, uint64_t* red)
{
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) red[w] = v;
    __syncthreads();
    uint64_t t = 0;
    for (uint32_t i = 0; i < blockDim.x / 64; i++) t += red[i];
    return t;
}

// Row lengths and per-block sums.  A row whose rank-sum or copy number the
// fast %#g path cannot settle is listed for k_gtext (gidx: site -> entry,
// entry.site) and sized without those columns; k_gtext adds them.
__global__ __launch_bounds__(256) void k_rowlen(RefView ref, uint32_t ws, SiteArrays a, Results res, uint32_t n_sites,
                                                RowCtx rc, uint32_t* gidx, GText* gtext, uint32_t* gcount,
                                                uint32_t* rowlen, uint64_t* block_sum)
{
    __shared__ uint64_t red[4];
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    uint32_t len = 0;
    if (s < n_sites) {
        const RowData d = load_row(ref, ws, a, res, s);
        if (!d.nrow) {
            uint32_t D;
            int X;
            if (!g_fast_digits(d.rank, D, X) || !g_fast_digits(copy_number(d.depth, d.repeat), D, X)) {
                const uint32_t j = atomicAdd(gcount, 1u);
                gidx[s] = j;
                gtext[j].site = s;
                gtext[j].rank[15] = 0;
                gtext[j].copy[15] = 0;
            }
        }
        len = (uint32_t)emit_row(d, rc, rc.pos_base + s, [](int, char) {});
        rowlen[s] = len;
    }
    const uint64_t tot = block_sum_u64(len, red);
    if (threadIdx.x == 0) block_sum[blockIdx.x] = tot;
}

void launch_rowlen(hipStream_t st, const RefView& ref, uint32_t ws, const SiteArrays& a, const Results& res,
                   uint32_t n_sites, RowCtx rc, uint32_t* gidx, GText* gtext, uint32_t* gcount, uint32_t* rowlen,
                   uint64_t* block_sum)
{
    const uint32_t blocks = (n_sites + kRowBlock - 1) / kRowBlock;
    if (blocks)
        hipLaunchKernelGGL(k_rowlen, dim3(blocks), dim3(kRowBlock), 0, st, ref, ws, a, res, n_sites, rc, gidx, gtext,
                           gcount, rowlen, block_sum);
}

// exclusive scan of block sums in place (one block of 1024 threads, 16
// consecutive values per thread per tile); writes the total
__global__ __launch_bounds__(1024) void k_scan_blocks(uint64_t* v, uint32_t n, uint64_t* total)
{
    constexpr uint32_t K = 16;
    __shared__ uint64_t wsum[16];
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    uint64_t carry = 0;  // the same in every thread
    for (uint32_t base = 0; base < n; base += 1024u * K) {
        const uint32_t i0 = base + threadIdx.x * K;
        uint64_t x[K], sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < K; k++) {
            x[k] = i0 + k < n ? v[i0 + k] : 0;
            sum += x[k];
        }
        uint64_t incl = sum;
        for (int off = 1; off < 64; off <<= 1) {
            const uint64_t y = __shfl_up(incl, off);
            if ((int)lane >= off) incl += y;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        uint64_t before = 0, tile = 0;
        for (uint32_t k = 0; k < 16; k++) {
            if (k < w) before += wsum[k];
            tile += wsum[k];
        }
        uint64_t run = carry + before + incl - sum;
#pragma unroll
        for (uint32_t k = 0; k < K; k++) {
            if (i0 + k < n) v[i0 + k] = run;
            run += x[k];
        }
        carry += tile;
        __syncthreads();  // wsum is rewritten by the next tile
    }
    if (threadIdx.x == 0) *total = carry;
}

void launch_scan_blocks(hipStream_t st, uint64_t* block_sum, uint32_t n_blocks, uint64_t* total)
{
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, block_sum, n_blocks, total);
}

// Exact %#g text of the listed rows (k_rowlen), and the lengths of the
// columns that use it added to the row and its block.
__global__ __launch_bounds__(256) void k_gtext(SiteArrays a, Results res, GText* gtext, const uint32_t* gcount,
                                               uint32_t* rowlen, uint64_t* block_sum)
{
    const uint32_t cnt = *gcount;
    for (uint32_t j = blockIdx.x * 256u + threadIdx.x; j < cnt; j += gridDim.x * 256u) {
        GText& g = gtext[j];
        const uint32_t s = g.site;
        const double rank = res.rank[s], copy = copy_number(a.depth[s], a.repeat[s]);
        const int lr = fmt_g(rank, g.rank), lc = fmt_g(copy, g.copy);
        g.rank[15] = (char)lr;
        g.copy[15] = (char)lc;
        uint32_t D;
        int X;
        const uint32_t add = (g_fast_digits(rank, D, X) ? 0u : (uint32_t)lr) + (g_fast_digits(copy, D, X) ? 0u : (uint32_t)lc);
        rowlen[s] += add;
        atomicAdd(&block_sum[s / kRowBlock], (uint64_t)add);
    }
}

void launch_gtext(hipStream_t st, const SiteArrays& a, const Results& res, uint32_t n_sites, GText* gtext,
                  const uint32_t* gcount, uint32_t* rowlen, uint64_t* block_sum)
{
    uint32_t blocks = (n_sites + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks)
        hipLaunchKernelGGL(k_gtext, dim3(blocks), dim3(256), 0, st, a, res, gtext, gcount, rowlen, block_sum);
}

constexpr uint32_t kFmtLds = 24576;  // 256 rows of up to 96 bytes; larger blocks write directly

__global__ __launch_bounds__(256) void k_format(RefView ref, uint32_t ws, SiteArrays a, Results res, uint32_t n_sites,
                                                RowCtx rc, const uint32_t* rowlen, const uint64_t* block_off, char* out)
{
    __shared__ __align__(16) char stage[kFmtLds + 16];
    __shared__ uint32_t wsum[4];
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint32_t len = s < n_sites ? rowlen[s] : 0;
    uint32_t incl = len;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off);
        if ((int)lane >= off) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (uint32_t k = 0; k < 4; k++) {
        if (k < w) before += wsum[k];
        total += wsum[k];
    }
    const uint32_t my = before + incl - len;
    const uint64_t base = block_off[blockIdx.x];
    if (total <= kFmtLds) {
        if (s < n_sites) {
            const RowData d = load_row(ref, ws, a, res, s);
            emit_row(d, rc, rc.pos_base + s, [&](int i, char c) { stage[my + i] = c; });
        }
        __syncthreads();
        // copy stage[0, total) to out[base, base + total): bytes up to 16-byte
        // alignment of the destination, then 16-byte stores whose four words
        // are joined from five aligned LDS words, then the tail bytes
        const uint32_t mis = (uint32_t)(base & 15u);
        const uint32_t h = mis == 0 ? 0u : (16u - mis < total ? 16u - mis : total);
        for (uint32_t i = threadIdx.x; i < h; i += 256) out[base + i] = stage[i];
        const uint32_t nq = (total - h) / 16u, sh = h & 3u;
        const uint32_t* sw = reinterpret_cast<const uint32_t*>(stage) + (h >> 2);
        uint4* dst = reinterpret_cast<uint4*>(out + base + h);
        for (uint32_t i = threadIdx.x; i < nq; i += 256) {
            const uint32_t* w = sw + 4u * i;  // w[4] lies at most 3 bytes past the text: inside the stage's slack
            const uint32_t w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3], w4 = w[4];
            uint4 q;
            q.x = __builtin_amdgcn_alignbyte(w1, w0, sh);
            q.y = __builtin_amdgcn_alignbyte(w2, w1, sh);
            q.z = __builtin_amdgcn_alignbyte(w3, w2, sh);
            q.w = __builtin_amdgcn_alignbyte(w4, w3, sh);
            dst[i] = q;
        }
        for (uint32_t i = h + 16u * nq + threadIdx.x; i < total; i += 256) out[base + i] = stage[i];
    } else if (s < n_sites) {
        const RowData d = load_row(ref, ws, a, res, s);
        char* o = out + base + my;
        emit_row(d, rc, rc.pos_base + s, [&](int i, char c) { o[i] = c; });
    }
}

void launch_format(hipStream_t st, const RefView& ref, uint32_t ws, const SiteArrays& a, const Results& res,
                   uint32_t n_sites, RowCtx rc, const uint32_t* rowlen, const uint64_t* block_off, char* out)
{
    const uint32_t blocks = (n_sites + 255) / 256;
    if (blocks)
        hipLaunchKernelGGL(k_format, dim3(blocks), dim3(256), 0, st, ref, ws, a, res, n_sites, rc, rowlen, block_off,
                           out);
}

// ---------------------------------------------------------------------------
// window carry-over (kernel/recycle_kernel.cu:91-136): the read-length tail
// [win, win+L) moves to [0, L), records included.
__global__ __launch_bounds__(256) void k_recycle(SiteArrays a, uint32_t win, uint32_t L)
{
    const uint32_t i = blockIdx.x;  // one block per tail site
    if (i >= L) return;
    const uint32_t src = win + i;
    const int n = a.nrec[src];
    for (int k = threadIdx.x; k < n; k += blockDim.x) a.rec[(size_t)i * kSlots + k] = a.rec[(size_t)src * kSlots + k];
    if (threadIdx.x == 0) {
        a.nrec[i] = n;
        a.depth[i] = a.depth[src];
        a.repeat[i] = a.repeat[src];
        a.cu[i] = a.cu[src];
        a.qs[i] = a.qs[src];
        a.ca[i] = a.ca[src];
    }
}

void launch_recycle(hipStream_t st, const SiteArrays& a, uint32_t win, uint32_t L)
{
    if (L) hipLaunchKernelGGL(k_recycle, dim3(L), dim3(256), 0, st, a, win, L);
}

}  // namespace gsnp
