// Persistent random-walk kernel (replaces randomWalkCuda + randomWalkCudaWrapper,
// randomWalk/Cuda/randomwalk.cu:64-134).
//
// Every lane owns one walker.  A lane whose walk has ended takes the next walk
// id from a device work counter (one atomic per wave per refill), so uneven
// walk lengths never idle a wave while work remains.  Per hop (reference
// RandomWalk, randomWalk/CPU/mainFunctions.cpp:48-90):
//   u     = Philox4x32-10(walk, job, hop/4, 'WALK'; seed)[hop % 4]
//   panel = largest i with T[i] <= u          (Green CDF thresholds in LDS)
//   p     = getPointFromNindex(square, panel) (FP64, reference op order)
//   p outside the strict bounds            -> walk escaped (hop not counted)
//   exact mode:  d* = nearest-rect distance; d* <= 0 -> hit (Q2 for the id),
//                else square = biggest square replayed from d*
//   bisect mode: square = literal bisection (one Q1 per step), hit = Q2(p)
// Tallies: int64 hit counts per original rect id + escaped + stuck bins, and
// the total hop count; optional per-walk outputs for replay checks.
#include "frw_internal.h"

namespace frw {
namespace {

struct WalkArgs {
    const double* surface;    // Gaussian surface of the master (4 doubles)
    const uint32_t* thresholds;
    unsigned long long* work;
    int* err;
    unsigned long long* counts;
    unsigned long long* hops_total;
    int32_t* hit_out;
    int32_t* hops_out;
    unsigned long long* traffic;  // kStats: node records, rect records
    uint64_t seed;
    uint64_t n_walks;
    uint32_t first_walk;
    uint32_t job;
    uint32_t max_hops;
};

__global__ void zero_counter(unsigned long long* work) { *work = 0ull; }

template <int kMode, bool kStats>
__global__ void __launch_bounds__(256) walk_kernel(DevTree t, WalkArgs a)
{
    extern __shared__ int32_t stack_mem[];
    __shared__ uint32_t sT[kNSample];
    for (int i = threadIdx.x; i < kNSample; i += blockDim.x) sT[i] = a.thresholds[i];
    __syncthreads();
    const LaneStack st{stack_mem + threadIdx.x, (int)blockDim.x};
    const int lane = threadIdx.x & (kWaveSize - 1);
    const double s0x = a.surface[0], s0y = a.surface[1], s1x = a.surface[2], s1y = a.surface[3];
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);

    bool active = false, done = false;
    uint64_t w = 0;
    uint32_t hop = 0;
    DBox sq{0, 0, 0, 0};
    uint32_t rnd[4] = {0, 0, 0, 0};
    unsigned long long my_hops = 0;
    Counters cnt{0u, 0u};
    Counters* cp = kStats ? &cnt : nullptr;
    unsigned long long my_nodes = 0, my_rects = 0;

    for (;;) {
        const bool need = !active && !done;
        const unsigned long long mask = __ballot(need);
        if (mask) {
            const int leader = __ffsll((long long)mask) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(a.work, (unsigned long long)__popcll(mask));
            base = __shfl(base, leader);
            if (need) {
                const unsigned long long below = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
                w = base + (unsigned long long)__popcll(below);
                if (w < a.n_walks) {
                    active = true;
                    hop = 0;
                    sq = DBox{s0x, s0y, s1x, s1y};
                } else {
                    done = true;
                }
            }
        }
        if (__ballot(active) == 0ull) break;
        if (!active) continue;

        if ((hop & 3u) == 0u) {
            rnd[0] = a.first_walk + (uint32_t)w;
            rnd[1] = a.job;
            rnd[2] = hop >> 2;
            rnd[3] = kWalkDomain;
            philox4x32_10(rnd, k0, k1);
        }
        const uint32_t u = rnd[hop & 3u];
        const int panel = sample_panel(sT, u);
        double px, py;
        point_from_panel(sq.x0, sq.y0, sq.x1, sq.y1, panel, px, py);
        int result = -3;  // walk continues
        if (!in_bounds_strict(t.bounds, px, py)) {
            result = -1;
        } else {
            if (kMode == kModeExact) {
                const double ds = nearest_rect(t, st, px, py, cp);
                if (ds <= 0.0)
                    result = q2_point(t, st, px, py, cp);
                else
                    sq = biggest_square<false>(t, st, px, py, ds, nullptr, cp, a.err);
            } else {
                sq = biggest_square<true>(t, st, px, py, 0.0, nullptr, cp, a.err);
                const int h = q2_point(t, st, px, py, cp);
                if (h >= 0) result = h;
            }
            if (kStats) {
                my_nodes += cnt.nodes;
                my_rects += cnt.rects;
                cnt.nodes = cnt.rects = 0u;
            }
            hop++;
            if (result == -3 && hop >= a.max_hops) result = -2;
        }
        if (result != -3) {
            const int bin = result >= 0 ? result : (result == -1 ? t.n_rects : t.n_rects + 1);
            atomicAdd(&a.counts[bin], 1ull);
            my_hops += hop;
            if (a.hit_out) a.hit_out[w] = result;
            if (a.hops_out) a.hops_out[w] = (int32_t)hop;
            active = false;
        }
    }
    for (int off = 32; off >= 1; off >>= 1) my_hops += __shfl_xor(my_hops, off);
    if (lane == 0 && my_hops) atomicAdd(a.hops_total, my_hops);
    if (kStats) {
        for (int off = 32; off >= 1; off >>= 1) {
            my_nodes += __shfl_xor(my_nodes, off);
            my_rects += __shfl_xor(my_rects, off);
        }
        if (lane == 0) {
            atomicAdd(&a.traffic[0], my_nodes);
            atomicAdd(&a.traffic[1], my_rects);
        }
    }
}

}  // namespace

int launch_walk(const frw_tree* t, const frw_walk_params* p, const frw_walk_outputs* o,
                hipStream_t s, unsigned long long* traffic)
{
    if (p->n_walks == 0) return FRW_OK;
    const int4 m = t->host_rects[p->master];
    int rc = launch_gauss(t, m, t->d_surface, s);
    if (rc != FRW_OK) return rc;
    zero_counter<<<1, 1, 0, s>>>(t->d_work);
    if ((rc = check_hip(hipGetLastError(), "zero_counter")) != FRW_OK) return rc;

    int block = p->block;
    if (block < 0 || block % kWaveSize != 0 || block > 256)
        return set_error(FRW_E_INVALID, "block must be a multiple of 64 and <= 256");
    if (block == 0) block = block_for_stack(t, 256, kNSample * 4);
    if (block == 0 || (size_t)block * stack_bytes_per_lane(t) + kNSample * 4 > (size_t)lds_per_block(t))
        return set_error(FRW_E_RANGE, "tree too deep for the LDS traversal stack at this block size");
    const size_t lds = (size_t)block * stack_bytes_per_lane(t);
    int grid = p->grid;
    if (grid <= 0) {
        int cus = 0, lds_cu = 0;
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, t->device);
        hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, t->device);
        if (lds_cu <= 0) lds_cu = 64 * 1024;
        const int per_cu_lds = (int)((size_t)lds_cu / (lds + kNSample * 4));
        const int per_cu = std::max(1, std::min(per_cu_lds, 2048 / block));
        grid = std::max(1, cus * per_cu);
        const int64_t need = (p->n_walks + block - 1) / block;
        if (need < grid) grid = (int)need;
    }
    WalkArgs a;
    a.surface = t->d_surface;
    a.thresholds = t->d_thresholds;
    a.work = t->d_work;
    a.err = t->d_err;
    a.counts = (unsigned long long*)o->counts;
    a.hops_total = (unsigned long long*)o->hops;
    a.hit_out = o->hit_out;
    a.hops_out = o->hops_out;
    a.traffic = traffic;
    a.seed = p->seed;
    a.n_walks = (uint64_t)p->n_walks;
    a.first_walk = (uint32_t)p->first_walk;
    a.job = p->job;
    a.max_hops = (uint32_t)(p->max_hops > 0 ? p->max_hops : 10000000);
    const DevTree v = dev_view(t);
    if (traffic) {
        if (p->mode == kModeBisect)
            walk_kernel<kModeBisect, true><<<grid, block, lds, s>>>(v, a);
        else
            walk_kernel<kModeExact, true><<<grid, block, lds, s>>>(v, a);
    } else {
        if (p->mode == kModeBisect)
            walk_kernel<kModeBisect, false><<<grid, block, lds, s>>>(v, a);
        else
            walk_kernel<kModeExact, false><<<grid, block, lds, s>>>(v, a);
    }
    return check_hip(hipGetLastError(), "walk_kernel");
}

}  // namespace frw
