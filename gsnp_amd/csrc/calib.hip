// calib.hip -- pass-1 calibration counts on the device (the counting loop of
// Prob_matrix::matrix_gen, cpp/matrix.cc:58-121): every base of a unique-hit
// read that was observed (A, C, G or T at or above the quality floor) and
// lies on a reference base that is neither N nor a known SNP counts once in
// the bin (quality, cycle, reference base, observed base).  The counts are
// exact integers, so their order does not matter; the matrix itself is then
// derived from them on the host (matrix_from_counts, a few thousand rows).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/gsnp_abi.h"
#include "engine.h"
#include "gsnp_common.h"

namespace gsnp {

namespace {

constexpr uint32_t kCalibCopies = 8;  // sub-histograms (block index mod 8), against atomic contention

// One wave per read, lane k on base k (every 64th base for reads over 64).
// The read store holds (q - q_min) << 2 | base per base, 0 where the base was
// not observed; unique-hit bases hold q - q_min in [1, 63].
__global__ __launch_bounds__(256) void k_calib_count(ReadView r, uint64_t n_reads, const uint8_t* ref,
                                                     unsigned long long* hist)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t nwave = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    unsigned long long* h = hist + (size_t)(blockIdx.x % kCalibCopies) * kCalibBins;
    for (uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n_reads; i += nwave) {
        const uint32_t m = r.meta[i];
        if (rm_hit(m) != 1) continue;
        const uint32_t len = rm_len(m);
        const bool rev = rm_strand(m) != 0;
        const uint8_t* seq = r.seq + i * (uint64_t)r.pitch;
        const uint8_t* rc = ref + r.pos[i];
        for (uint32_t k = lane; k < len; k += 64) {
            const uint32_t b = seq[k];
            if (!b) continue;
            const uint32_t o = rc[k];
            if (o & 12u) continue;  // N or known SNP on the reference
            const uint32_t cyc = rev ? len - 1u - k : k;
            atomicAdd(&h[(pb_q(b) << 12) | (cyc << 4) | ((o & 3u) << 2) | pb_base(b)], 1ull);
        }
    }
}

__global__ __launch_bounds__(256) void k_calib_sum(unsigned long long* hist)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= kCalibBins) return;
    unsigned long long s = 0;
    for (uint32_t c = 0; c < kCalibCopies; c++) s += hist[(size_t)c * kCalibBins + i];
    hist[i] = s;
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
    if (c->calib.n < (size_t)kCalibCopies * kCalibBins) HIPCHK(c->calib.alloc((size_t)kCalibCopies * kCalibBins));
    launch_fill_u32(c->stream, c->calib.p, 2u * kCalibCopies * kCalibBins, 0u);
    const uint64_t blocks = std::min<uint64_t>((c->n_reads + 3) / 4, 8192);
    hipLaunchKernelGGL(k_calib_count, dim3((uint32_t)blocks), dim3(256), 0, c->stream, c->read_view(), c->n_reads,
                       c->ref.p, c->calib.p);
    hipLaunchKernelGGL(k_calib_sum, dim3(kCalibBins / 256), dim3(256), 0, c->stream, c->calib.p);
    HIPCHK(hipGetLastError());
    std::vector<unsigned long long> h(kCalibBins);
    HIPCHK(hipMemcpyAsync(h.data(), c->calib.p, 8ull * kCalibBins, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < kCalibBins; i++) counts[i] += h[i];
    return GSNP_OK;
}
