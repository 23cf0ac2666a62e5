// adjoint.hip -- gradient of the summed per-site log marginal likelihood
//
//   l = sum_s ln sum_g prior_s(g) 10^{lk_s(g)},
//   lk_s(g) = sum_r log10(0.5 p[qa_r, c_r, a1(g), o_r] + 0.5 p[qa_r, c_r, a2(g), o_r])
//
// with respect to ln p[q][c][a][o] (q < 64) and ln het_rate, where the records
// r of a site, their adjusted qualities qa_r and the genotype order are those
// of the calling path (kernel/likelihood_kernel.cu:137-256) and the priors
// those of cpp/prior.cc.  No reference counterpart: the reference has no
// gradient.  With w_s(g) = prior_s(g) 10^{lk_s(g)} / sum_g', the gradient is
//
//   dl/d ln p[q,c,a,o] = sum_{records r with key (q,c,o)} sum_g w_s(g) share_g(a)
//   share_g(a) = p_a ([a == a1] + [a == a2]) / (p_a1 + p_a2)
//   dl/d ln het = sum_s sum_g w_s(g) [prior_s(g) carries one factor of het]
//
// Accumulation is deterministic: the records are sorted once by score-row key
// (stable LSD radix sort), and each evaluation sums their contributions in a
// fixed order: each piece of at most 512 records of one key (in key order) by
// one wave, whose lane l computes the contributions of records l, l + 64, ...
// from their sites' weights and adds them in that order before a fixed xor
// tree, then the pieces of a key in order.  No floating-point atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gsnp_abi.h"
#include "engine.h"
#include "gsnp_common.h"
#include "kernels.h"
#include "tables.h"
#include "walk.h"

namespace gsnp {

namespace {

constexpr uint32_t kNoRow = 0xFFFFFFFFu;
constexpr uint32_t kPiece = 512;      // records per reduction piece at most
constexpr uint32_t kKeys = 65536;     // score-row keys with q < 64: q<<10 | cycle<<2 | o
constexpr uint32_t kRadixChunk = 4096;  // items per block of a radix pass (16 tiles of 256)
constexpr uint32_t kErrAdjQ = 4u;     // adjusted quality >= 64 (outside the gradient table)

__device__ __forceinline__ int get4i(const int4& v, uint32_t i)
{
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}

// per site: record count of a called site and its prior row (0..7 reference
// rows, 8+i known SNP i); N rows without coverage are not called
__global__ __launch_bounds__(256) void k_adj_sites(Tables t, RefView ref, uint32_t ws, SiteArrays a, uint32_t n,
                                                   uint32_t* cnt, uint32_t* prow)
{
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= n) return;
    const uint32_t ori = ref.code[ws + s] & 0xF;
    if ((ori & 4) && a.depth[s] == 0) {
        cnt[s] = 0;
        prow[s] = kNoRow;
        return;
    }
    cnt[s] = (uint32_t)a.nrec[s];
    uint32_t row = ori & 7;
    if ((ori & 8) && t.refine && ref.n_snp) {
        uint32_t lo = 0, hi = ref.n_snp;
        const uint32_t p = ws + s;
        while (lo < hi) {
            const uint32_t m = (lo + hi) >> 1;
            if (ref.snp_pos[m] < p)
                lo = m + 1;
            else
                hi = m;
        }
        if (lo < ref.n_snp && ref.snp_pos[lo] == p) row = 8 + lo;
    }
    prow[s] = row;
}

// exclusive scan of u32 counts into u64 offsets: per-block local scan, block
// sums scanned by k_scan_blocks, then the block offsets added
__global__ __launch_bounds__(256) void k_scan_local(const uint32_t* in, uint32_t n, uint64_t* out, uint64_t* bsum)
{
    __shared__ uint64_t ws[4];
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const uint64_t x = i < n ? in[i] : 0;
    uint64_t incl = x;
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t y = __shfl_up(incl, off);
        if ((int)lane >= off) incl += y;
    }
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    uint64_t before = 0;
    for (uint32_t k = 0; k < w; k++) before += ws[k];
    if (i < n) out[i] = before + incl - x;
    if (threadIdx.x == 255) bsum[blockIdx.x] = before + incl;
}

__global__ __launch_bounds__(256) void k_scan_add(uint64_t* out, uint32_t n, const uint64_t* boff, const uint64_t* total)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) out[i] += boff[blockIdx.x];
    if (i == n) out[n] = *total;
}

// adjusted-quality keys of each called site's sorted records, at the site's
// offset (record order)
__global__ __launch_bounds__(256) void k_adj_keys(Tables t, SiteArrays a, uint32_t n, const uint32_t* prow,
                                                  const uint64_t* off, uint32_t* key, uint32_t* err)
{
    __shared__ uint32_t seen_all[256 * 17];
    uint32_t* seen = seen_all + threadIdx.x * 17;
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= n || prow[s] == kNoRow) return;
    const uint32_t* row = a.rec + (size_t)s * kSlots;
    uint32_t* dst = key + off[s];
    uint32_t k = 0, bad = 0;
    bad |= walk_adjusted(row, (uint32_t)a.nrec[s], t, seen, [&](uint32_t qa, uint32_t cyc, uint32_t o) {
        if (qa >= 64) {
            bad |= kErrAdjQ;
            qa = 63;
        }
        dst[k++] = (qa << 10) | (cyc << 2) | o;
    });
    if (bad) atomicOr(err, bad);
}

// ---- stable LSD radix sort of (key, value) by 8-bit digits ----------------

__global__ __launch_bounds__(256) void k_radix_hist(const uint32_t* key, uint32_t n, uint32_t shift, uint32_t nblocks,
                                                    uint32_t* hist)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t c0 = blockIdx.x * kRadixChunk;
    for (uint32_t i = threadIdx.x; i < kRadixChunk; i += 256) {
        const uint32_t idx = c0 + i;
        if (idx < n) atomicAdd(&h[(key[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Each block scatters its chunk in input order, one tile of 256 at a time:
// an item's place is its digit's offset for the block, plus the items of the
// same digit in earlier tiles, earlier waves and earlier lanes of its wave.
__global__ __launch_bounds__(256) void k_radix_scatter(const uint32_t* key, const uint32_t* val, uint32_t n,
                                                       uint32_t shift, uint32_t nblocks, const uint64_t* off,
                                                       uint32_t* key_out, uint32_t* val_out)
{
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[4][256];
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    run[threadIdx.x] = 0;
    const uint32_t c0 = blockIdx.x * kRadixChunk;
    for (uint32_t t0 = 0; t0 < kRadixChunk; t0 += 256) {
        const uint32_t idx = c0 + t0 + threadIdx.x;
        const bool valid = idx < n;
        const uint32_t k = valid ? key[idx] : 0u;
        const uint32_t d = (k >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & below);
        const uint32_t cntw = (uint32_t)__popcll(peers);
        for (uint32_t j = 0; j < 4; j++) wcnt[j][threadIdx.x] = 0;
        __syncthreads();
        if (valid && rank == cntw - 1) wcnt[w][d] = cntw;
        __syncthreads();
        if (valid) {
            uint32_t before = run[d];
            for (uint32_t j = 0; j < w; j++) before += wcnt[j][d];
            const uint64_t pos = off[(size_t)d * nblocks + blockIdx.x] + before + rank;
            key_out[pos] = k;
            val_out[pos] = val[idx];
        }
        __syncthreads();
        run[threadIdx.x] += wcnt[0][threadIdx.x] + wcnt[1][threadIdx.x] + wcnt[2][threadIdx.x] + wcnt[3][threadIdx.x];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_iota(uint32_t* v, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) v[i] = i;
}

// piece heads: a new piece starts at every key change and every kPiece-th item
__global__ __launch_bounds__(256) void k_piece_flags(const uint32_t* skey, uint32_t n, uint32_t* flag)
{
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j < n) flag[j] = (j % kPiece == 0 || skey[j] != skey[j - 1]) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_piece_starts(const uint32_t* flag, const uint64_t* pidx, uint32_t n,
                                                      uint32_t* pstart, uint32_t npieces)
{
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j < n && flag[j]) pstart[pidx[j]] = j;
    if (j == 0) pstart[npieces] = n;
}

// first piece of every key (pieces are in key order)
__global__ __launch_bounds__(256) void k_key_pieces(const uint32_t* skey, const uint32_t* pstart, uint32_t npieces,
                                                    uint32_t* kpiece)
{
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k > kKeys) return;
    uint32_t lo = 0, hi = npieces;
    while (lo < hi) {
        const uint32_t m = (lo + hi) >> 1;
        if (skey[pstart[m]] < k)
            lo = m + 1;
        else
            hi = m;
    }
    kpiece[k] = lo;
}

// ---- evaluation ------------------------------------------------------------

// device score rows for q < 64 from the current p_matrix (device log10; the
// calling path's exact table is host-built)
__global__ __launch_bounds__(256) void k_adj_table(const double* pm, double* T)
{
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;  // q<<10 | c<<2 | o
    if (r >= kKeys) return;
    const uint32_t q = r >> 10, c = (r >> 2) & 255u, o = r & 3u;
    const double* row = pm + ((size_t)q << 12) + ((size_t)c << 4);
    double* out = T + (size_t)r * kScoreRowStride;
    int g = 0;
    for (uint32_t a1 = 0; a1 < 4; a1++)
        for (uint32_t a2 = a1; a2 < 4; a2++) {
            const double h1 = 0.5 * row[(a1 << 2) | o], h2 = 0.5 * row[(a2 << 2) | o];
            out[g++] = log10(h1 + h2);
        }
}

// per site: log-likelihoods in record order, log marginal, genotype weights
// and the het derivative.  Five lanes per site add its score rows (two
// genotypes each, every genotype in record order, eight records in flight):
// one load instruction reads whole 80-byte rows of twelve sites instead of a
// 16-byte piece of 64 rows.  The sums meet in LDS and the site's first lane
// finishes the site.
constexpr uint32_t kAdjSitesPerWave = 12;

__global__ __launch_bounds__(256) void k_adj_site(const double* T, const uint32_t* key, const uint64_t* off,
                                                  const uint32_t* prow, const double* lp, const uint32_t* hmask,
                                                  uint32_t n, int monoploid, double* ll, double* dh, double* wout)
{
    __shared__ double sums[4][kAdjSitesPerWave][10];
    const uint32_t lane = __lane_id(), wib = threadIdx.x >> 6;
    const uint32_t sl = lane / 5u, cl = lane - 5u * sl;
    const uint32_t s = (blockIdx.x * 4u + wib) * kAdjSitesPerWave + sl;
    const bool on = lane < 60u && s < n;
    const uint64_t r0 = on ? off[s] : 0, r1 = on ? off[s + 1] : 0;
    double a0 = 0.0, a1 = 0.0;
    constexpr uint32_t U = 8;
    for (uint64_t r = r0; r < r1; r += U) {
        uint32_t kk[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) kk[u] = key[r + u < r1 ? r + u : r1 - 1];  // inside the site's keys
        double2 x[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++)
            x[u] = *reinterpret_cast<const double2*>(T + (size_t)kk[u] * kScoreRowStride + 2u * cl);
#pragma unroll
        for (uint32_t u = 0; u < U; u++)
            if (r + u < r1) {
                a0 += x[u].x;
                a1 += x[u].y;
            }
    }
    if (lane < 60u) {
        sums[wib][sl][2 * cl] = a0;
        sums[wib][sl][2 * cl + 1] = a1;
    }
    __syncthreads();
    if (!on || cl != 0) return;
    const uint32_t pr = prow[s];
    double* w = wout + (size_t)s * 10;
    if (pr == kNoRow) {
        ll[s] = 0.0;
        dh[s] = 0.0;
        for (int g = 0; g < 10; g++) w[g] = 0.0;
        return;
    }
    double lk[10];
#pragma unroll
    for (int g = 0; g < 10; g++) lk[g] = sums[wib][sl][g];
    const double* row = lp + (size_t)pr * 10;
    double v[10], m = -INFINITY;
#pragma unroll
    for (int g = 0; g < 10; g++) {
        const uint32_t gt = geno_code(g);
        const bool het = (gt >> 2) != (gt & 3);
        v[g] = (monoploid && het) ? -INFINITY : row[g] + lk[g];
        m = v[g] > m ? v[g] : m;
    }
    double z = 0.0;
#pragma unroll
    for (int g = 0; g < 10; g++) {
        v[g] = exp10(v[g] - m);
        z += v[g];
    }
    const uint32_t mask = hmask[pr];
    double d = 0.0;
#pragma unroll
    for (int g = 0; g < 10; g++) {
        w[g] = v[g] / z;
        if ((mask >> g) & 1u) d += w[g];
    }
    ll[s] = m * 2.302585092994045684 + log(z);
    dh[s] = d;
}

// per record: the site that owns it (record order; sites own contiguous runs)
__global__ __launch_bounds__(256) void k_adj_rsite(const uint64_t* off, uint32_t n, uint32_t* rsite)
{
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= n) return;
    for (uint64_t r = off[s]; r < off[s + 1]; r++) rsite[r] = s;
}

// the site of each record in key order (prepare): ssite[j] = rsite[perm[j]]
__global__ __launch_bounds__(256) void k_adj_ssite(const uint32_t* perm, const uint32_t* rsite, uint32_t nr,
                                                   uint32_t* ssite)
{
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j < nr) ssite[j] = rsite[perm[j]];
}

// One wave per piece (at most 512 records of one key, in key order): lane l
// computes the contributions of records l, l + 64, ... of the piece to the
// four ln p entries of the key's row, adds them in that order, then a fixed
// xor tree over the wave (a fixed order: the sums are the same on every run).
// A record's contribution is sum_g w_s(g) share_g(a) from its site's ten
// weights and the key's four p values, one division per genotype pair.
__global__ __launch_bounds__(256) void k_adj_pieces(const double* pm, const uint32_t* skey, const uint32_t* ssite,
                                                    const double* wsite, const uint32_t* pstart, uint32_t npieces,
                                                    double* psum)
{
    const uint32_t p = blockIdx.x * 4u + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
    if (p >= npieces) return;  // wave-uniform
    const uint32_t j0 = pstart[p], j1 = pstart[p + 1];
    const uint32_t k = skey[j0];  // one key per piece
    const uint32_t q = k >> 10, cyc = (k >> 2) & 255u, o = k & 3u;
    const double* prow = pm + ((size_t)q << 12) + ((size_t)cyc << 4) + o;
    double pa[4];
#pragma unroll
    for (int a = 0; a < 4; a++) pa[a] = prow[a << 2];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (uint32_t j = j0 + lane; j < j1; j += 64u) {
        const double* ws = wsite + (size_t)ssite[j] * 10;
        double w[10];
#pragma unroll
        for (int g = 0; g < 10; g++) w[g] = ws[g];
        double c[4] = {0.0, 0.0, 0.0, 0.0};
        int g = 0;
#pragma unroll
        for (int a1 = 0; a1 < 4; a1++)
#pragma unroll
            for (int a2 = a1; a2 < 4; a2++) {
                if (a1 == a2) {
                    c[a1] += w[g];
                } else {
                    const double inv = 1.0 / (pa[a1] + pa[a2]);  // one division per pair
                    c[a1] += w[g] * (pa[a1] * inv);
                    c[a2] += w[g] * (pa[a2] * inv);
                }
                g++;
            }
        s0 += c[0];
        s1 += c[1];
        s2 += c[2];
        s3 += c[3];
    }
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off);
        s1 += __shfl_xor(s1, off);
        s2 += __shfl_xor(s2, off);
        s3 += __shfl_xor(s3, off);
    }
    if (lane == 0) {
        double2* dst = reinterpret_cast<double2*>(psum + (size_t)p * 4);
        dst[0] = make_double2(s0, s1);
        dst[1] = make_double2(s2, s3);
    }
}

// gradient of one key's four entries, p_matrix layout q<<12 | c<<4 | a<<2 | o
__global__ __launch_bounds__(256) void k_adj_grad(const double* psum, const uint32_t* kpiece, double* grad)
{
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= kKeys) return;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t p = kpiece[k]; p < kpiece[k + 1]; p++)
#pragma unroll
        for (int a = 0; a < 4; a++) s[a] += psum[(size_t)p * 4 + a];
    const uint32_t q = k >> 10, cyc = (k >> 2) & 255u, o = k & 3u;
#pragma unroll
    for (int a = 0; a < 4; a++) grad[((size_t)q << 12) | ((size_t)cyc << 4) | ((size_t)a << 2) | o] = s[a];
}

// fixed-order sums of two per-site arrays: 256-site block partials (fixed
// shuffle tree), then one block summing the partials in order
__global__ __launch_bounds__(256) void k_sum2_blocks(const double* x, const double* y, uint32_t n, double* part)
{
    __shared__ double sx[4], sy[4];
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    double a = i < n ? x[i] : 0.0, b = i < n ? y[i] : 0.0;
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    if (__lane_id() == 0) {
        sx[threadIdx.x >> 6] = a;
        sy[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = (sx[0] + sx[1]) + (sx[2] + sx[3]);
        part[2 * blockIdx.x + 1] = (sy[0] + sy[1]) + (sy[2] + sy[3]);
    }
}

__global__ __launch_bounds__(64) void k_sum2_final(const double* part, uint32_t nb, double* out)
{
    double a = 0.0, b = 0.0;
    for (uint32_t i = __lane_id(); i < nb; i += 64) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    if (__lane_id() == 0) {
        out[0] = a;
        out[1] = b;
    }
}

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + 255) / 256); }

}  // namespace

#define HIPCHK(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return fail(GSNP_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// exclusive scan of n u32 counts into out[0, n] (out[n] = total)
static int scan_counts(Ctx& c, const uint32_t* in, uint32_t n, uint64_t* out)
{
    const uint32_t nb = blocks_for(n) ? blocks_for(n) : 1;
    if (c.adj_bsum.n < nb + 1) HIPCHK(c.adj_bsum.alloc(nb + 1));
    if (c.total.n < 1) HIPCHK(c.total.alloc(1));
    hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(256), 0, c.stream, in, n, out, c.adj_bsum.p);
    launch_scan_blocks(c.stream, c.adj_bsum.p, nb, c.total.p);
    hipLaunchKernelGGL(k_scan_add, dim3(blocks_for((uint64_t)n + 1)), dim3(256), 0, c.stream, out, n, c.adj_bsum.p,
                       c.total.p);
    HIPCHK(hipGetLastError());
    return GSNP_OK;
}

// stable sort of (key, value) pairs by the low 16 bits of key, two 8-bit passes
static int radix_sort16(Ctx& c, uint32_t* key, uint32_t* val, uint32_t* key_tmp, uint32_t* val_tmp, uint32_t n)
{
    if (!n) return GSNP_OK;
    const uint32_t nb = (n + kRadixChunk - 1) / kRadixChunk;
    const uint32_t nh = 256 * nb;
    if (c.adj_hist.n < nh) HIPCHK(c.adj_hist.alloc(nh));
    if (c.adj_hoff.n < (size_t)nh + 1) HIPCHK(c.adj_hoff.alloc((size_t)nh + 1));
    uint32_t *ki = key, *vi = val, *ko = key_tmp, *vo = val_tmp;
    for (uint32_t shift = 0; shift < 16; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(256), 0, c.stream, ki, n, shift, nb, c.adj_hist.p);
        int rc = scan_counts(c, c.adj_hist.p, nh, c.adj_hoff.p);
        if (rc) return rc;
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(256), 0, c.stream, ki, vi, n, shift, nb, c.adj_hoff.p, ko,
                           vo);
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    HIPCHK(hipGetLastError());
    return GSNP_OK;  // two passes: the result is back in key/val
}

}  // namespace gsnp

using namespace gsnp;

extern "C" int gsnp_adjoint_prepare(gsnp_ctx* c, uint32_t start, uint32_t n)
{
    if (!c) return fail(GSNP_E_ARG, "null context");
    if ((uint64_t)start + n > c->chr_len) return fail(GSNP_E_ARG, "span outside the chromosome");
    if (!c->n_reads && n) return fail(GSNP_E_ARG, "no reads loaded");
    int rc = c->alloc_window(n + 2);
    if (rc) return rc;
    c->adj_n = 0;
    SiteArrays a = c->site_view(0);
    // site matrix of the span and each site's sorted records
    launch_site_build(c->stream, c->read_view(), 0, c->n_reads, start, n + 1, INT64_MIN, (int64_t)start + n, 0, a,
                      nullptr, c->ranges.p);
    launch_bucket(c->stream, a.nrec, n, c->lists.p, c->win_cap, c->counts.p, 0, nullptr);
    launch_sort(c->stream, a.rec, a.nrec, c->lists.p, c->win_cap, c->counts.p, n);
    HIPCHK(hipGetLastError());
    HIPCHK(c->adj_cnt.alloc(n + 1));
    HIPCHK(c->adj_prow.alloc(n + 1));
    HIPCHK(c->adj_off.alloc((size_t)n + 1));
    if (n)
        hipLaunchKernelGGL(k_adj_sites, dim3(blocks_for(n)), dim3(256), 0, c->stream, c->tables(), c->ref_view(), start,
                           a, n, c->adj_cnt.p, c->adj_prow.p);
    rc = scan_counts(*c, c->adj_cnt.p, n, c->adj_off.p);
    if (rc) return rc;
    uint64_t nrec = 0;
    HIPCHK(hipMemcpyAsync(&nrec, c->adj_off.p + n, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (nrec >= 0xFFFFFFFFull) return fail(GSNP_E_RANGE, "span holds too many records for one adjoint");
    const uint32_t nr = (uint32_t)nrec;
    HIPCHK(c->adj_key.alloc(nr + 1));
    HIPCHK(c->adj_rsite.alloc(nr + 1));
    if (n)
        hipLaunchKernelGGL(k_adj_rsite, dim3(blocks_for(n)), dim3(256), 0, c->stream, c->adj_off.p, n, c->adj_rsite.p);
    HIPCHK(c->adj_skey.alloc(nr + 1));
    HIPCHK(c->adj_perm.alloc(nr + 1));
    HIPCHK(c->adj_tmp.alloc(2 * (size_t)nr + 2));
    HIPCHK(hipMemsetAsync(c->err.p, 0, 4, c->stream));
    if (n)
        hipLaunchKernelGGL(k_adj_keys, dim3(blocks_for(n)), dim3(256), 0, c->stream, c->tables(), a, n, c->adj_prow.p,
                           c->adj_off.p, c->adj_key.p, c->err.p);
    // records in key order (stable): perm, its inverse, and the sorted keys
    HIPCHK(hipMemcpyAsync(c->adj_skey.p, c->adj_key.p, 4ull * nr, hipMemcpyDeviceToDevice, c->stream));
    if (nr) hipLaunchKernelGGL(k_iota, dim3(blocks_for(nr)), dim3(256), 0, c->stream, c->adj_perm.p, nr);
    rc = radix_sort16(*c, c->adj_skey.p, c->adj_perm.p, c->adj_tmp.p, c->adj_tmp.p + nr + 1, nr);
    if (rc) return rc;
    HIPCHK(c->adj_ssite.alloc(nr + 1));
    if (nr)
        hipLaunchKernelGGL(k_adj_ssite, dim3(blocks_for(nr)), dim3(256), 0, c->stream, c->adj_perm.p, c->adj_rsite.p,
                           nr, c->adj_ssite.p);
    // reduction pieces
    uint32_t* flag = c->adj_tmp.p;
    HIPCHK(c->adj_pidx.alloc((size_t)nr + 1));
    if (nr) hipLaunchKernelGGL(k_piece_flags, dim3(blocks_for(nr)), dim3(256), 0, c->stream, c->adj_skey.p, nr, flag);
    rc = scan_counts(*c, flag, nr, c->adj_pidx.p);
    if (rc) return rc;
    uint64_t np = 0;
    HIPCHK(hipMemcpyAsync(&np, c->adj_pidx.p + nr, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const uint32_t npieces = (uint32_t)np;
    HIPCHK(c->adj_pstart.alloc(npieces + 1));
    HIPCHK(c->adj_kpiece.alloc(kKeys + 1));
    hipLaunchKernelGGL(k_piece_starts, dim3(blocks_for(nr ? nr : 1)), dim3(256), 0, c->stream, flag, c->adj_pidx.p, nr,
                       c->adj_pstart.p, npieces);
    hipLaunchKernelGGL(k_key_pieces, dim3(blocks_for(kKeys + 1)), dim3(256), 0, c->stream, c->adj_skey.p,
                       c->adj_pstart.p, npieces, c->adj_kpiece.p);
    HIPCHK(hipGetLastError());
    // evaluation buffers
    HIPCHK(c->adj_T.alloc((size_t)kKeys * kScoreRowStride));
    HIPCHK(c->adj_w.alloc((size_t)n * 10 + 1));
    HIPCHK(c->adj_ll.alloc(n + 1));
    HIPCHK(c->adj_dh.alloc(n + 1));
    HIPCHK(c->adj_psum.alloc(4 * (size_t)npieces + 4));
    HIPCHK(c->adj_part.alloc(2 * (size_t)blocks_for(n) + 2));
    HIPCHK(c->adj_out.alloc(2));
    HIPCHK(c->adj_pm.alloc((size_t)256 * 256 * 16));
    HIPCHK(c->adj_grad.alloc((size_t)64 * 256 * 16));
    uint32_t e = 0;
    HIPCHK(hipMemcpyAsync(&e, c->err.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (e & kErrQadjRange) return fail(GSNP_E_RANGE, "adjusted quality exceeds the score table");
    if (e & kErrAdjQ) return fail(GSNP_E_RANGE, "adjusted quality of 64 or more: outside the gradient table");
    c->adj_start = start;
    c->adj_n = n;
    c->adj_nrec = nr;
    c->adj_npieces = npieces;
    return GSNP_OK;
}

// Prior rows (log10) at het rate exp(log_het): the 8 reference-code rows
// (cpp/prior.cc:3-43), then one row per known SNP (cpp/prior.cc:45-97); and
// per row the genotypes whose prior carries one factor of het.
static void adjoint_priors(const Ctx& c, double log_het, std::vector<double>& lp, std::vector<uint32_t>& mask)
{
    gsnp_params pp = c.p;
    pp.het_novel_r = std::exp(log_het);
    std::vector<double> prior(128);
    prior_gen(pp, prior.data());
    lp = log_prior_rows(prior.data());
    static const int kGenoCodes[10] = {0, 1, 2, 3, 5, 6, 7, 10, 11, 15};
    uint32_t het_bits = 0;
    for (int g = 0; g < 10; g++)
        if ((kGenoCodes[g] >> 2) != (kGenoCodes[g] & 3)) het_bits |= 1u << g;
    mask.assign(8, het_bits);
    const size_t nsnp = c.h_snps.size();
    lp.resize((8 + nsnp) * 10);
    mask.resize(8 + nsnp);
    for (size_t i = 0; i < nsnp; i++) {
        const gsnp_snp& sn = c.h_snps[i];
        snp_log_prior(pp, prior.data(), c.h_snp_ref[i], sn, &lp[(8 + i) * 10]);
        uint32_t m = 0;
        for (int g = 0; g < 10; g++) {
            const int a1 = kGenoCodes[g] >> 2, a2 = kGenoCodes[g] & 3;
            const bool replaced = !sn.indel && sn.freq[a1] > 0 && sn.freq[a2] > 0;
            if (a1 != a2 && !replaced) m |= 1u << g;
        }
        mask[8 + i] = m;
    }
}

extern "C" int gsnp_adjoint_eval(gsnp_ctx* c, const double* p_matrix, double log_het, double* loglik,
                                 double* grad_log_p, double* grad_log_het, double* site_loglik)
{
    if (!c || !p_matrix) return fail(GSNP_E_ARG, "null argument");
    if (!c->adj_n) return fail(GSNP_E_ARG, "gsnp_adjoint_prepare has not been called");
    const uint32_t n = c->adj_n, nr = (uint32_t)c->adj_nrec;
    std::vector<double> lp;
    std::vector<uint32_t> mask;
    adjoint_priors(*c, log_het, lp, mask);
    if (c->adj_lp.n < lp.size()) HIPCHK(c->adj_lp.alloc(lp.size()));
    if (c->adj_mask.n < mask.size()) HIPCHK(c->adj_mask.alloc(mask.size()));
    hipStream_t st = c->stream;
    HIPCHK(hipMemcpyAsync(c->adj_lp.p, lp.data(), 8 * lp.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(c->adj_mask.p, mask.data(), 4 * mask.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(c->adj_pm.p, p_matrix, 8ull * 256 * 256 * 16, hipMemcpyHostToDevice, st));
    c->timer.begin(kStAdjoint, st);
    hipLaunchKernelGGL(k_adj_table, dim3(blocks_for(kKeys)), dim3(256), 0, st, c->adj_pm.p, c->adj_T.p);
    if (n) {
        hipLaunchKernelGGL(k_adj_site, dim3((n + 4 * kAdjSitesPerWave - 1) / (4 * kAdjSitesPerWave)), dim3(256), 0, st,
                           c->adj_T.p, c->adj_key.p, c->adj_off.p,
                           c->adj_prow.p, c->adj_lp.p, c->adj_mask.p, n, c->p.is_monoploid, c->adj_ll.p, c->adj_dh.p,
                           c->adj_w.p);
    }
    if (c->adj_npieces)
        hipLaunchKernelGGL(k_adj_pieces, dim3((c->adj_npieces + 3) / 4), dim3(256), 0, st, c->adj_pm.p,
                           c->adj_skey.p, c->adj_ssite.p, c->adj_w.p, c->adj_pstart.p, c->adj_npieces,
                           c->adj_psum.p);
    hipLaunchKernelGGL(k_adj_grad, dim3(blocks_for(kKeys)), dim3(256), 0, st, c->adj_psum.p, c->adj_kpiece.p,
                       c->adj_grad.p);
    const uint32_t nb = blocks_for(n) ? blocks_for(n) : 1;
    hipLaunchKernelGGL(k_sum2_blocks, dim3(nb), dim3(256), 0, st, c->adj_ll.p, c->adj_dh.p, n, c->adj_part.p);
    hipLaunchKernelGGL(k_sum2_final, dim3(1), dim3(64), 0, st, c->adj_part.p, nb, c->adj_out.p);
    c->timer.end(kStAdjoint, st);
    HIPCHK(hipGetLastError());
    double out2[2];
    HIPCHK(hipMemcpyAsync(out2, c->adj_out.p, 16, hipMemcpyDeviceToHost, st));
    if (grad_log_p) HIPCHK(hipMemcpyAsync(grad_log_p, c->adj_grad.p, 8ull * 64 * 256 * 16, hipMemcpyDeviceToHost, st));
    if (site_loglik && n) HIPCHK(hipMemcpyAsync(site_loglik, c->adj_ll.p, 8ull * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (loglik) *loglik = out2[0];
    if (grad_log_het) *grad_log_het = out2[1];
    return GSNP_OK;
}

extern "C" int gsnp_adjoint_info(gsnp_ctx* c, uint32_t* n_sites, uint64_t* n_records, uint32_t* n_prior_rows)
{
    if (!c) return fail(GSNP_E_ARG, "null context");
    if (n_sites) *n_sites = c->adj_n;
    if (n_records) *n_records = c->adj_nrec;
    if (n_prior_rows) *n_prior_rows = 8 + (uint32_t)c->h_snps.size();
    return GSNP_OK;
}

extern "C" int gsnp_adjoint_export(gsnp_ctx* c, uint32_t* keys, uint64_t* offsets, uint32_t* prior_row)
{
    if (!c) return fail(GSNP_E_ARG, "null context");
    if (!c->adj_n) return fail(GSNP_E_ARG, "gsnp_adjoint_prepare has not been called");
    const uint32_t n = c->adj_n;
    if (keys && c->adj_nrec)
        HIPCHK(hipMemcpyAsync(keys, c->adj_key.p, 4ull * c->adj_nrec, hipMemcpyDeviceToHost, c->stream));
    if (offsets) HIPCHK(hipMemcpyAsync(offsets, c->adj_off.p, 8ull * (n + 1), hipMemcpyDeviceToHost, c->stream));
    if (prior_row) HIPCHK(hipMemcpyAsync(prior_row, c->adj_prow.p, 4ull * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return GSNP_OK;
}

extern "C" int gsnp_adjoint_priors(gsnp_ctx* c, double log_het, double* lp10, uint32_t* het_mask, uint32_t cap_rows)
{
    if (!c) return fail(GSNP_E_ARG, "null context");
    std::vector<double> lp;
    std::vector<uint32_t> mask;
    adjoint_priors(*c, log_het, lp, mask);
    if (cap_rows < mask.size()) return fail(GSNP_E_ARG, "prior row buffer too small");
    if (lp10) std::memcpy(lp10, lp.data(), 8 * lp.size());
    if (het_mask) std::memcpy(het_mask, mask.data(), 4 * mask.size());
    return GSNP_OK;
}
