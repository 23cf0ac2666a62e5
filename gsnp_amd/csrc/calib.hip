// calib.hip -- pass-1 calibration counts on the device (the counting loop of
// Prob_matrix::matrix_gen, cpp/matrix.cc:58-121): every base of a unique-hit
// read that was observed (A, C, G or T at or above the quality floor) and
// lies on a reference base that is neither N nor a known SNP counts once in
// the bin (quality, cycle, reference base, observed base).  The counts are
// exact integers, so the order of the additions does not matter; the matrix
// itself is then derived from them on the host (matrix_from_counts, a few
// thousand rows).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/gsnp_abi.h"
#include "engine.h"
#include "gsnp_common.h"

namespace gsnp {

namespace {

// Each workgroup counts one slice of 16 cycles over a range of reads into a
// private histogram in LDS -- 63 qualities x 16 (reference, observed) pairs x
// 16 cycles of 32-bit counters, 64,512 bytes -- then adds its non-zero bins
// to the device histogram with one 64-bit atomic each.  Counting in LDS
// takes the per-base global atomics (and their contention on the few bins
// of high-quality matches) off the device memory path.
constexpr uint32_t kSliceCycles = 16;
constexpr uint32_t kCalibQ = 63;  // q - q_min in [1, 63] on unique-hit bases
constexpr uint32_t kSliceBins = kCalibQ * kSliceCycles * 16;
constexpr uint32_t kCalibThreads = 1024;
constexpr uint32_t kCalibReads = kCalibThreads / kSliceCycles;  // reads per step of a workgroup
constexpr uint32_t kCalibUnroll = 4;                            // steps whose loads are in flight together

// LDS word of (q - q_min - 1, reference, observed, cycle in the slice).  The
// pair goes in as (ref ^ obs) << 2 | obs, so that the four matches, which
// take most of the counts, are pairs 0-3: with the cycle in the low bits,
// the 16 cycles of a wave's read on the four matching pairs fall on 64
// different banks.
__device__ __forceinline__ uint32_t calib_word(uint32_t qi, uint32_t ref, uint32_t obs, uint32_t j)
{
    return ((qi * 16u + (((ref ^ obs) << 2) | obs)) << 4) | j;
}

// blockIdx.y: the cycle slice [16 y, 16 y + 16); lane 16 r + j of a step
// takes cycle 16 y + j of read r, and kCalibUnroll steps load together.  The
// read store holds (q - q_min) << 2 | base per base, 0 where the base was
// not observed.
__global__ __launch_bounds__(kCalibThreads) void k_calib_count(ReadView r, uint64_t n_reads, const uint8_t* ref,
                                                               unsigned long long* hist)
{
    __shared__ uint32_t bins[kSliceBins];
    for (uint32_t i = threadIdx.x; i < kSliceBins; i += kCalibThreads) bins[i] = 0;
    __syncthreads();
    const uint32_t c0 = blockIdx.y * kSliceCycles;
    const uint32_t j = threadIdx.x & 15u;
    const uint32_t cyc = c0 + j;
    const uint64_t stride = (uint64_t)gridDim.x * kCalibReads * kCalibUnroll;
    for (uint64_t i0 = (uint64_t)blockIdx.x * kCalibReads * kCalibUnroll + (threadIdx.x >> 4); i0 < n_reads;
         i0 += stride) {
        uint32_t m[kCalibUnroll], b[kCalibUnroll], o[kCalibUnroll];
        uint64_t at[kCalibUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kCalibUnroll; u++) {
            const uint64_t i = i0 + u * kCalibReads;
            m[u] = i < n_reads ? r.meta[i] : 0u;  // hit count 0: skipped
        }
#pragma unroll
        for (uint32_t u = 0; u < kCalibUnroll; u++) {
            const uint64_t i = i0 + u * kCalibReads;
            const uint32_t len = rm_len(m[u]);
            const bool ok = rm_hit(m[u]) == 1 && cyc < len;
            const uint32_t k = rm_strand(m[u]) ? len - 1u - cyc : cyc;  // read coordinate of this cycle
            b[u] = ok ? r.seq[i * (uint64_t)r.pitch + k] : 0u;
            at[u] = ok ? (uint64_t)r.pos[i] + k : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < kCalibUnroll; u++) o[u] = b[u] ? ref[at[u]] : 4u;
#pragma unroll
        for (uint32_t u = 0; u < kCalibUnroll; u++)
            if (!(o[u] & 12u))  // observed, and neither N nor a known SNP on the reference
                atomicAdd(&bins[calib_word(pb_q(b[u]) - 1u, o[u] & 3u, pb_base(b[u]), j)], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kSliceBins; i += kCalibThreads) {
        const uint32_t v = bins[i];
        if (!v) continue;
        const uint32_t q = (i >> 8) + 1u, pm = (i >> 4) & 15u, jj = i & 15u;
        const uint32_t obs = pm & 3u, rb = (pm >> 2) ^ obs;
        atomicAdd(&hist[(q << 12) | ((c0 + jj) << 4) | (rb << 2) | obs], (unsigned long long)v);
    }
}

}  // namespace

#define HIPCHK(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return fail(GSNP_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace gsnp

extern "C" int gsnp_calib_count(gsnp_ctx* c, uint64_t* counts)
{
    using namespace gsnp;
    if (!c || !counts) return fail(GSNP_E_ARG, "null argument");
    if (!c->n_reads) return GSNP_OK;
    if (c->calib.n < (size_t)kCalibBins) HIPCHK(c->calib.alloc(kCalibBins));
    launch_fill_u32(c->stream, c->calib.p, 2u * kCalibBins, 0u);
    // a cycle slice per blockIdx.y; two workgroups fit a CU's LDS, so about
    // 512 of them in all fill the device once
    const uint32_t slices = ((uint32_t)c->p.read_length + kSliceCycles - 1) / kSliceCycles;
    const uint64_t per_step = (uint64_t)kCalibReads * kCalibUnroll;
    const uint64_t per_slice =
        std::max<uint64_t>(1, std::min<uint64_t>((c->n_reads + per_step - 1) / per_step, 512 / slices));
    c->timer.begin(kStCalib, c->stream);
    hipLaunchKernelGGL(k_calib_count, dim3((uint32_t)per_slice, slices), dim3(kCalibThreads), 0, c->stream, c->read_view(),
                       c->n_reads, c->ref.p, c->calib.p);
    c->timer.end(kStCalib, c->stream);
    HIPCHK(hipGetLastError());
    std::vector<unsigned long long> h(kCalibBins);
    HIPCHK(hipMemcpyAsync(h.data(), c->calib.p, 8ull * kCalibBins, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < kCalibBins; i++) counts[i] += h[i];
    return GSNP_OK;
}
