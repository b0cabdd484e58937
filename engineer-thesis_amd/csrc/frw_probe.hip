// Measurement hook: read bandwidth of a buffer small enough to stay in the
// caches (or large enough not to), the ceiling the walk kernel's cached
// record reads are compared with.  No reference counterpart.
#include <algorithm>

#include "frw_internal.h"

namespace frw {
namespace {

// Each wave reads whole 1 KiB chunks (64 lanes x 16 B, coalesced): chunk
// visit k goes to chunk (k * 40503) mod n_chunks (a power of two, so a
// bijection of each run of n_chunks visits), which scatters consecutive
// visits of one wave and the waves of one CU over the buffer, so a buffer
// larger than L1 is served by L2 (or, larger still, by HBM).  Four visits
// per iteration keep four loads in flight per lane.  The xor of everything
// read is stored only if it equals a value it practically never takes, which
// keeps every load live.
__global__ void __launch_bounds__(256) read_sweep(const int4* __restrict__ buf, int64_t n_chunks, int64_t visits,
                                                  int4* __restrict__ sink)
{
    const int64_t mask = n_chunks - 1;
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t n_waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    int4 acc = make_int4(0, 0, 0, 0);
    int64_t k = wave;
    for (; k + 3 * n_waves < visits; k += 4 * n_waves) {
        int4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = buf[(((k + u * n_waves) * 40503) & mask) * 64 + lane];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            acc.x ^= v[u].x;
            acc.y ^= v[u].y;
            acc.z ^= v[u].z;
            acc.w ^= v[u].w;
        }
    }
    for (; k < visits; k += n_waves) {
        const int4 v = buf[((k * 40503) & mask) * 64 + lane];
        acc.x ^= v.x;
        acc.y ^= v.y;
        acc.z ^= v.z;
        acc.w ^= v.w;
    }
    if (acc.x == 0x7f3a5c11 && acc.y == 0x0badf00d && acc.z == 0x5eed5eed && acc.w == 0x1234abcd) sink[0] = acc;
}

__global__ void fill_words(int4* buf, int64_t n16)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n16) buf[i] = make_int4((int)i, (int)(i * 3), (int)(i * 5), (int)(i * 7));
}

}  // namespace
}  // namespace frw

using namespace frw;

extern "C" int frw_probe_read_bandwidth(int64_t bytes, int32_t repeats, double* gbps)
{
    if (bytes < 4096 || bytes > ((int64_t)1 << 34) || repeats < 1 || repeats > (1 << 20) || !gbps)
        return set_error(FRW_E_INVALID, "bytes in [4 KiB, 16 GiB], repeats in [1, 2^20]");
    int64_t n_chunks = 1;  // the largest power of two of 1 KiB chunks in `bytes`
    while (n_chunks * 2 * 1024 <= bytes) n_chunks *= 2;
    const int64_t n16 = n_chunks * 64;
    int4 *buf = nullptr, *sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int dev = 0, cus = 0;
    int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc == FRW_OK) rc = check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CUs");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&buf, (size_t)n16 * 16), "hipMalloc probe buffer");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&sink, 16), "hipMalloc sink");
    if (rc == FRW_OK) {
        fill_words<<<(unsigned)((n16 + 255) / 256), 256>>>(buf, n16);
        rc = check_hip(hipGetLastError(), "fill");
    }
    if (rc == FRW_OK) rc = check_hip(hipEventCreate(&e0), "event");
    if (rc == FRW_OK) rc = check_hip(hipEventCreate(&e1), "event");
    // eight 256-lane blocks per CU
    const unsigned grid = (unsigned)std::max(cus, 1) * 8;
    float ms = 0.f;
    if (rc == FRW_OK) {
        read_sweep<<<grid, 256>>>(buf, n_chunks, n_chunks, sink);  // one sweep warms the caches
        rc = check_hip(hipGetLastError(), "read_sweep");
    }
    if (rc == FRW_OK) rc = check_hip(hipEventRecord(e0, nullptr), "record");
    if (rc == FRW_OK) {
        read_sweep<<<grid, 256>>>(buf, n_chunks, n_chunks * repeats, sink);
        rc = check_hip(hipGetLastError(), "read_sweep");
    }
    if (rc == FRW_OK) rc = check_hip(hipEventRecord(e1, nullptr), "record");
    if (rc == FRW_OK) rc = check_hip(hipEventSynchronize(e1), "sync");
    if (rc == FRW_OK) rc = check_hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    if (rc == FRW_OK) *gbps = (double)n_chunks * 1024.0 * repeats / ((double)ms * 1e-3) / 1e9;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (buf) (void)hipFree(buf);
    if (sink) (void)hipFree(sink);
    return rc;
}
