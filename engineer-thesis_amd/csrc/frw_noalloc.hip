// Walks that allocate nothing: frw_walk's walks with every scratch buffer in
// one fixed device array of the library, for callers that cannot allocate
// device memory (code under graph capture, serving loops, harnesses that
// forbid hipMalloc).  No counterpart in the reference.
//
// The array `scratch` holds, per GPU: the job record of the launch, the
// work-queue counter, the kernel error flag, the hop counter and up to
// FRW_NOALLOC_MAX_RECTS + 2 count bins.  Its name is a host-side stand-in in
// host code: the device address of each GPU's copy comes from
// hipGetSymbolAddress, taken once per device at that device's first call and
// kept in scratch_ptr[].  Every call on one GPU shares that GPU's array, and
// the array keeps the last call's contents: the bins, the hop counter and
// the error flag are cleared at the start of each call (the work counter is
// cleared by the walk's own prologue), so calls that overlap on one GPU race.
#include <chrono>
#include <mutex>

#include "frw_internal.h"

namespace frw {
namespace {

constexpr size_t kJobBytes = 1024;                    // job record
constexpr size_t kWorkOff = kJobBytes;                // work-queue counter (8 B)
constexpr size_t kErrOff = kWorkOff + 8;              // error flag (4 B)
constexpr size_t kHopsOff = kErrOff + 8;              // hop counter (8 B)
constexpr size_t kCountsOff = 2048;                   // count bins (8 B each)
constexpr int64_t kMaxBins = (int64_t)FRW_NOALLOC_MAX_RECTS + 2;
constexpr size_t kScratchBytes = kCountsOff + 8 * (size_t)kMaxBins;  // 8 MiB + 2 KiB
static_assert(sizeof(WalkJob) <= kJobBytes, "job record must fit its slot");

static __device__ __align__(256) unsigned long long scratch[kScratchBytes / 8];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
char* scratch_ptr[kMaxDevices] = {};  // device address of each GPU's array, once known

// Device address of the current device's array (taken on its first use).
int scratch_of(int dev, char** out)
{
    if (dev < 0 || dev >= kMaxDevices) return set_error(FRW_E_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        const int rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (rc != FRW_OK) return rc;
        if (!p) return set_error(FRW_E_DEVICE, "scratch array has no device address");
        scratch_ptr[dev] = static_cast<char*>(p);
    }
    *out = scratch_ptr[dev];
    return FRW_OK;
}

// Zeroes the bins, the hop counter and the error flag (a fill kernel rather
// than a memset, so that a captured call records no memset node).
__global__ void clear_scratch(unsigned long long* counts, int64_t n_bins, unsigned long long* hops, int* err)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_bins) counts[i] = 0ull;
    if (i == 0) {
        *hops = 0ull;
        *err = 0;
    }
}

}  // namespace
}  // namespace frw

using namespace frw;

extern "C" int frw_walk_noalloc(const frw_tree* t, const frw_walk_params* p, void* stream, int64_t* counts_host,
                                frw_walk_summary* summary)
{
    int rc = check_walk_params(t, p);
    if (rc != FRW_OK) return rc;
    if (p->mode != FRW_MODE_EXACT && p->mode != FRW_MODE_BISECT)
        return set_error(FRW_E_INVALID, "frw_walk_noalloc runs exact or bisection walks");
    if (summary && !counts_host) return set_error(FRW_E_INVALID, "the summary needs counts_host");
    if (t->n_rects > FRW_NOALLOC_MAX_RECTS)
        return set_error(FRW_E_RANGE, "frw_walk_noalloc: the tree has " + std::to_string(t->n_rects) +
                                          " rects, the fixed scratch array holds bins for at most " +
                                          std::to_string((int64_t)FRW_NOALLOC_MAX_RECTS));
    int dev = -1;
    if ((rc = check_hip(hipGetDevice(&dev), "hipGetDevice")) != FRW_OK) return rc;
    if (dev != t->device) return set_error(FRW_E_INVALID, "the tree's device must be the current device");
    char* base = nullptr;
    if ((rc = scratch_of(dev, &base)) != FRW_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int64_t nb = t->n_rects + 2;
    auto* counts = reinterpret_cast<unsigned long long*>(base + kCountsOff);
    auto* hops = reinterpret_cast<unsigned long long*>(base + kHopsOff);
    int* err = reinterpret_cast<int*>(base + kErrOff);
    const WalkScratch ws{reinterpret_cast<WalkJob*>(base), reinterpret_cast<unsigned long long*>(base + kWorkOff),
                         err};
    const auto t0 = std::chrono::steady_clock::now();
    clear_scratch<<<(unsigned)((nb + 255) / 256), 256, 0, s>>>(counts, nb, hops, err);
    rc = check_hip(hipGetLastError(), "clear_scratch");
    if (rc == FRW_OK) {
        const frw_walk_outputs o{reinterpret_cast<int64_t*>(counts), reinterpret_cast<int64_t*>(hops), nullptr,
                                 nullptr};
        rc = launch_walk_in(t, p, &o, s, ws);
    }
    int64_t h = 0;
    int e = 0;
    if (rc == FRW_OK && counts_host)
        rc = check_hip(hipMemcpyAsync(counts_host, counts, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, s),
                       "copy counts");
    if (rc == FRW_OK) rc = check_hip(hipMemcpyAsync(&h, hops, sizeof(h), hipMemcpyDeviceToHost, s), "copy hops");
    if (rc == FRW_OK) rc = check_hip(hipMemcpyAsync(&e, err, sizeof(e), hipMemcpyDeviceToHost, s), "copy error flag");
    if (rc == FRW_OK) rc = check_hip(hipStreamSynchronize(s), "frw_walk_noalloc");
    if (rc != FRW_OK) return rc;
    if (e) return set_error(FRW_E_KERNEL, "kernel error flag " + std::to_string(e));
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (summary) fill_summary(t, p, counts_host, h, ms, summary);
    return FRW_OK;
}
