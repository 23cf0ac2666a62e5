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
// private histogram in LDS -- 63 qualities x 16 cycles x 16 (reference,
// observed) pairs of 32-bit counters, 64,512 bytes -- then adds its non-zero
// bins to the device histogram with one 64-bit atomic each.  Counting in LDS
// takes the per-base global atomics (and their contention on the few bins
// of high-quality matches) off the device memory path.
constexpr uint32_t kSliceCycles = 16;
constexpr uint32_t kCalibQ = 63;  // q - q_min in [1, 63] on unique-hit bases
constexpr uint32_t kSliceBins = kCalibQ * kSliceCycles * 16;

// blockIdx.y: the cycle slice [16 y, 16 y + 16); a wave takes four reads at a
// time, lane 16 r + j on cycle 16 y + j of read r.  The read store holds
// (q - q_min) << 2 | base per base, 0 where the base was not observed.
__global__ __launch_bounds__(256) void k_calib_count(ReadView r, uint64_t n_reads, const uint8_t* ref,
                                                     unsigned long long* hist)
{
    __shared__ uint32_t bins[kSliceBins];
    for (uint32_t i = threadIdx.x; i < kSliceBins; i += 256) bins[i] = 0;
    __syncthreads();
    const uint32_t c0 = blockIdx.y * kSliceCycles;
    const uint32_t j = threadIdx.x & 15u;
    const uint32_t cyc = c0 + j;
    const uint64_t stride = (uint64_t)gridDim.x * 16u;
    for (uint64_t i = (uint64_t)blockIdx.x * 16u + (threadIdx.x >> 4); i < n_reads; i += stride) {
        const uint32_t m = r.meta[i];
        const uint32_t len = rm_len(m);
        if (rm_hit(m) != 1 || cyc >= len) continue;
        const uint32_t k = rm_strand(m) ? len - 1u - cyc : cyc;  // read coordinate of this cycle
        const uint32_t b = r.seq[i * (uint64_t)r.pitch + k];
        if (!b) continue;
        const uint32_t o = ref[(uint64_t)r.pos[i] + k];
        if (o & 12u) continue;  // N or known SNP on the reference
        atomicAdd(&bins[((pb_q(b) - 1u) * kSliceCycles + j) * 16u + ((o & 3u) << 2) + pb_base(b)], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kSliceBins; i += 256) {
        const uint32_t v = bins[i];
        if (!v) continue;
        const uint32_t q = i / (kSliceCycles * 16u) + 1u, jj = (i / 16u) % kSliceCycles, t = i & 15u;
        atomicAdd(&hist[(q << 12) | ((c0 + jj) << 4) | t], (unsigned long long)v);
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
    // a cycle slice per blockIdx.y; reads spread over about 2048 workgroups in all
    const uint32_t slices = ((uint32_t)c->p.read_length + kSliceCycles - 1) / kSliceCycles;
    const uint64_t per_slice = std::max<uint64_t>(1, std::min<uint64_t>((c->n_reads + 15) / 16, 2048 / slices));
    c->timer.begin(kStCalib, c->stream);
    hipLaunchKernelGGL(k_calib_count, dim3((uint32_t)per_slice, slices), dim3(256), 0, c->stream, c->read_view(),
                       c->n_reads, c->ref.p, c->calib.p);
    c->timer.end(kStCalib, c->stream);
    HIPCHK(hipGetLastError());
    std::vector<unsigned long long> h(kCalibBins);
    HIPCHK(hipMemcpyAsync(h.data(), c->calib.p, 8ull * kCalibBins, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < kCalibBins; i++) counts[i] += h[i];
    return GSNP_OK;
}
