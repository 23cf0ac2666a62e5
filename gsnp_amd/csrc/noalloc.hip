// noalloc.hip -- the segmented sort + likelihood entry point on caller-owned
// device memory that allocates nothing.  Its only working memory (the depth
// class lists, their counters and an error word) is a fixed device array,
// `scratch`, one copy per GPU for the life of the process.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/gsnp_abi.h"
#include "engine.h"
#include "gsnp_common.h"
#include "kernels.h"

namespace gsnp {

// sites per call; the lists hold one u32 per site and class
constexpr uint32_t kNoallocMaxSites = GSNP_WIN;
constexpr uint32_t kNoallocCounters = 6u * kNoallocMaxSites;  // 8 class counters
constexpr uint32_t kNoallocErr = kNoallocCounters + 8u;        // kernel error flags
constexpr uint32_t kNoallocWords = kNoallocCounters + 64u;

// [6][kNoallocMaxSites] class lists, then the class counters, then the error
// word: 6 MiB + 256 B per GPU.  The name is a host-side handle only; kernels
// get the device address that hipGetSymbolAddress returns for each device.
static __device__ __align__(256) uint32_t scratch[kNoallocWords];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

__global__ __launch_bounds__(256) void k_check_counts(const int32_t* counts, uint32_t n, uint32_t* err)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) {
        const int32_t c = counts[i];
        if (c < 0 || c > (int32_t)kSlots) atomicOr(err, 1u);
    }
}

}  // namespace

// The device address of `scratch` on `device`, looked up once per device with
// that device current.
int noalloc_scratch(int device, uint32_t** out)
{
    if (device < 0 || device >= kMaxDevices) return fail(GSNP_E_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return fail(GSNP_E_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
        if (cur != device) return fail(GSNP_E_ARG, "the context's device is not the current device");
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(GSNP_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch_ptr[device];
    return GSNP_OK;
}

}  // namespace gsnp

using namespace gsnp;

extern "C" int gsnp_sort_likelihood_noalloc(gsnp_ctx* c, uint32_t* records, const int32_t* counts, uint32_t n_sites,
                                            double* lk10)
{
    if (!c || !records || !counts || !lk10) return fail(GSNP_E_ARG, "null argument");
    if (n_sites > kNoallocMaxSites)
        return fail(GSNP_E_ARG, "n_sites above the 262144 sites of the fixed working array");
    if (!c->score.p) return fail(GSNP_E_ARG, "context has no score table");
    uint32_t* ptr = nullptr;
    int rc = noalloc_scratch(c->device, &ptr);
    if (rc) return rc;
    if (!n_sites) return GSNP_OK;
    uint32_t* lists = ptr;
    uint32_t* cls_count = ptr + kNoallocCounters;
    uint32_t* err = ptr + kNoallocErr;
    hipStream_t s = c->stream;
    // the array keeps the previous call's contents: clear the error word (the
    // class counters are cleared by launch_bucket)
    launch_fill_u32(s, err, 1, 0u);
    hipLaunchKernelGGL(k_check_counts, dim3((n_sites + 255) / 256), dim3(256), 0, s, counts, n_sites, err);
    uint32_t e = 0;
    hipError_t he = hipMemcpyAsync(&e, err, 4, hipMemcpyDeviceToHost, s);
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (he != hipSuccess) return fail(GSNP_E_DEVICE, std::string("count check: ") + hipGetErrorString(he));
    if (e) return fail(GSNP_E_ARG, "a record count is outside [0, 256]");
    SiteArrays a;
    a.rec = records;
    a.nrec = const_cast<int32_t*>(counts);  // read only on this path
    launch_bucket(s, a.nrec, n_sites, lists, kNoallocMaxSites, cls_count, 0, nullptr);
    launch_sort(s, records, a.nrec, lists, kNoallocMaxSites, cls_count, n_sites);
    launch_likelihood(s, c->tables(), a, n_sites, lk10, err);
    he = hipGetLastError();
    if (he == hipSuccess) he = hipMemcpyAsync(&e, err, 4, hipMemcpyDeviceToHost, s);
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (he != hipSuccess) return fail(GSNP_E_DEVICE, std::string("sort + likelihood: ") + hipGetErrorString(he));
    if (e & kErrQadjRange) return fail(GSNP_E_RANGE, "adjusted quality exceeds the score table");
    return GSNP_OK;
}
