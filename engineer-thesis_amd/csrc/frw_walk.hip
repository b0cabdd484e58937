// Persistent random-walk kernel (replaces randomWalkCuda + randomWalkCudaWrapper,
// randomWalk/Cuda/randomwalk.cu:64-134).
//
// Every lane owns one walker.  A lane whose walk has ended takes the next walk
// index from a device work counter (one atomic per wave per refill), so uneven
// walk lengths never idle a wave while work remains.  A launch runs a table of
// jobs (layout tree, master surface, walk range); one job is the plain
// single-layout walk, many jobs are the batched form, where the global walk
// index is mapped to its job by a binary search over the jobs' walk offsets.
// Per hop (reference RandomWalk, randomWalk/CPU/mainFunctions.cpp:48-90):
//   u     = Philox4x32-10(walk, job, hop/4, 'WALK'; seed)[hop % 4]
//   panel = largest i with T[i] <= u          (Green CDF thresholds in LDS)
//   p     = getPointFromNindex(square, panel) (FP64, reference op order)
//   p outside the strict bounds            -> walk escaped (hop not counted)
//   exact mode:   d* = nearest-rect distance (branch and bound seeded with
//                 the rect nearest to the previous point); d* <= 0 -> hit
//                 (the search's nearest rect is the point query's answer),
//                 else square = biggest square replayed from d*
//   weighted:     the exact step, from a Gaussian-surface start with a
//                 Green-gradient weight
//   bisect mode:  square = literal bisection (one Q1 per step), hit = Q2(p)
//   learned mode: the wave's 64 points go through the proposal network
//                 (bf16 MFMA); an occupancy-pyramid probe that certifies the
//                 proposed square free accepts it, otherwise the exact-mode
//                 step runs
// Tallies: int64 hit counts per original rect id + escaped + stuck bins and
// hop counts per job; optional per-walk outputs for replay checks.
#include <algorithm>

#include "frw_internal.h"
#include "frw_mlp.h"

namespace frw {
namespace {

struct WalkArgs {
    const uint32_t* thresholds;
    unsigned long long* work;
    int* err;
    unsigned long long* traffic;  // kStats: node records, rect records, start-table entries
    uint64_t seed;
    uint64_t total_walks;
    uint32_t max_hops;
    int32_t n_jobs;
    int32_t wave_ints;  // LDS per wave in 32-bit words: its lanes' traversal stacks, which the
                        // learned mode's activation tile aliases between traversals
    // harvest: (point, previous half-size, d*) of non-final exact hops, the
    // k-th record of the launch's walk g (its global index: the jobs' walks
    // one after another) at slot h_off[g] + k (offsets from a first pass of
    // the same walks, so the records come in job, walk and hop order)
    double* h_pts;
    double* h_prev;
    double* h_dstar;
    const unsigned long long* h_off;
    uint64_t h_cap;
    // learned proposal
    unsigned long long* l_stats;  // accepted, fallback
    // weighted estimator: first-hop tables gx | gy | pw | qx | qy (200 each)
    // followed by the importance thresholds (2 x 200 u32), permittivity,
    // FRW_WEIGHT_* options
    const double* wtab;
    double eps;
    uint32_t wflags;
};

constexpr int kLdsJobs = 128;                                   // job walk offsets cached in LDS
// learned-mode blocks of up to 12 waves: one block per CU holds the 12 waves
// the registers allow (3 per SIMD) with one LDS copy of the network's weights
constexpr int kLearnedMaxBlock = 768;
constexpr int kWaveChunk = 256;                                 // walk indices a wave claims per queue visit
constexpr int kTailChunk = 64;                                  // ... once the queue nears its end
constexpr uint32_t kGausDomain = 0x47415553u;                   // 'GAUS': Gaussian-surface start points

// Tallies of the walks that end on a rect other than their job's master (or
// get stuck) go to a per-block table in LDS, open addressing on (job, bin)
// keys, and reach the bins once per block.  A few rects next to a master take
// most of these walks, and one global atomic per walk on their bins
// serialised the kernel (10^4 rects, 10^7 walks: exact walks 6.3 ms, 5.1 ms
// with those atomics taken out; weighted walks, four 64-bit atomics per walk,
// 25.4 ms and 7.5 ms).  A key whose kBinProbes slots are taken by other keys
// goes straight to its bin.  Batched launches hold more slots: a block walks
// jobs one after another, and with 64 slots the first jobs' keys filled the
// table (64 layouts of 10^2-10^5 rects: exact walks 27.2 ms per launch with
// 64 slots, 20.8 with 512).  The bisection keeps 64: its walks end far less
// often than its collision queries run, and the larger table's LDS cost it
// resident waves (348 -> 375 ms on the same batch).
constexpr int kBinSlotsLog2Small = 6;
constexpr int kBinSlotsLog2Batched = 9;
__host__ __device__ constexpr int bin_slots_log2(int mode, bool single)
{
    return single || mode == kModeBisect ? kBinSlotsLog2Small : kBinSlotsLog2Batched;
}
constexpr int kBinProbes = 8;
constexpr unsigned long long kBinEmpty = ~0ull;  // no key: a job index is below 2^31

__device__ inline unsigned long long bin_key(int job, int bin)
{
    return ((unsigned long long)(uint32_t)job << 32) | (uint32_t)bin;
}

// Slot of key in the block's table of 2^kLog2 slots (claimed if the key is
// new), or -1 when every slot of its probe path holds another key.
template <int kLog2>
__device__ inline int bin_slot(unsigned long long* keys, unsigned long long key)
{
    constexpr int kBinSlots = 1 << kLog2;
    const uint32_t h = ((uint32_t)key ^ (uint32_t)(key >> 32) * 0x85EBCA6Bu) * 0x9E3779B1u;
    int s = (int)(h >> (32 - kLog2));
    for (int p = 0; p < kBinProbes; ++p) {
        const unsigned long long k = ((volatile unsigned long long*)keys)[s];
        if (k == key) return s;
        if (k == kBinEmpty) {
            const unsigned long long old = atomicCAS(&keys[s], kBinEmpty, key);
            if (old == kBinEmpty || old == key) return s;
        }
        s = (s + 1) & (kBinSlots - 1);
    }
    return -1;
}

// Exact fixed-point sums of the weighted estimator.  Each walk's weight w is
// rounded once, to q = w 2^32 as an int64 (|w| < 2^31, else an error flag),
// and its square to (q^2) >> 32; both are added as 128-bit two's-complement
// integers (lo, hi).  Integer addition is associative, so the sums, and the
// doubles they convert to at the end of the launch, do not depend on the
// order in which walks end: any launch shape gives the same bits.
struct Fx {
    unsigned long long lo, hi;
};

__device__ inline void fx_add(Fx& a, const Fx& b)
{
    const unsigned long long l = a.lo + b.lo;
    a.hi += b.hi + (l < a.lo ? 1ull : 0ull);
    a.lo = l;
}

// p[0] += lo, p[1] += hi + carry: two 64-bit atomics form one exact 128-bit add
__device__ inline void fx_atomic(unsigned long long* p, const Fx& b)
{
    if (b.lo == 0ull && b.hi == 0ull) return;
    const unsigned long long old = atomicAdd(p, b.lo);
    atomicAdd(p + 1, b.hi + (old + b.lo < old ? 1ull : 0ull));
}

__device__ inline void fx_weight(double w, Fx& sum_term, Fx& sq_term, int* err)
{
    double a = w * 4294967296.0;
    if (!(fabs(a) < 9223372036854775808.0)) {
        raise_error(err, kErrWeight);
        a = 0.0;
    }
    const long long v = __double2ll_rn(a);
    sum_term.lo = (unsigned long long)v;
    sum_term.hi = v < 0 ? ~0ull : 0ull;
    const unsigned long long u = v < 0 ? 0ull - (unsigned long long)v : (unsigned long long)v;
    const unsigned long long plo = u * u, phi = __umul64hi(u, u);
    sq_term.lo = (plo >> 32) | (phi << 32);
    sq_term.hi = phi >> 32;
}

// The fixed-point sums of a launch (4 words per bin: sum lo, hi, squares lo,
// hi) added to the caller's double bins and cleared for the next launch.  A
// sum converts as sign and magnitude, magnitude = hi 2^64 + lo, scaled by
// 2^-32.
__global__ void weighted_finish(unsigned long long* __restrict__ acc, int64_t n_bins, double* __restrict__ wsum,
                                double* __restrict__ wsq)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_bins) return;
    unsigned long long* p = acc + 4 * i;
    double v[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        unsigned long long lo = p[2 * k], hi = p[2 * k + 1];
        const bool neg = (long long)hi < 0;
        if (neg) {  // 128-bit negation
            lo = ~lo + 1ull;
            hi = ~hi + (lo == 0ull ? 1ull : 0ull);
        }
        const double mag = (double)hi * 18446744073709551616.0 + (double)lo;
        v[k] = (neg ? -mag : mag) * (1.0 / 4294967296.0);
    }
    wsum[i] += v[0];
    wsq[i] += v[1];
    p[0] = p[1] = p[2] = p[3] = 0ull;
}
constexpr int kMlpLdsPerWave = kMlpTileElems * 2;               // the bf16 activation tile

__device__ inline int find_job(const WalkJob* jobs, const uint64_t* s_begin, int n_jobs, uint64_t g)
{
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const uint64_t b = mid < kLdsJobs ? s_begin[mid] : jobs[mid].walk_begin;
        if (b <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Prologue: Gaussian surface of every job's master (one lane per job) and a
// zeroed work counter.  Part of the launch so the whole walk can be captured
// in a graph and replayed.
__global__ void __launch_bounds__(64) walk_prologue(WalkJob* jobs, int n_jobs, unsigned long long* work)
{
    extern __shared__ int32_t stack_mem[];
    const LaneStack st{stack_mem + threadIdx.x, (int)blockDim.x};
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j == 0) *work = 0ull;
    if (j >= n_jobs) return;
    WalkJob& jb = jobs[j];
    const double fx = gauss_factor(jb.tree, st, jb.master, 1.5, false);
    const double fy = gauss_factor(jb.tree, st, jb.master, 1.5, true);
    const DBox s = gauss_surface(jb.master, fx, fy);
    jb.surface[0] = s.x0;
    jb.surface[1] = s.y0;
    jb.surface[2] = s.x1;
    jb.surface[3] = s.y1;
    // a rect at the master's centre (the master, or one overlapping it): its
    // distance bounds the first hop's search, as the previous hop's nearest
    // rect bounds the later ones
    int arg = -1;
    (void)nearest_rect(jb.tree, st, ((double)jb.master.x + (double)jb.master.z) / 2.0,
                       ((double)jb.master.y + (double)jb.master.w) / 2.0, nullptr, -1, &arg);
    jb.start_hint = arg;
    jb.start_rect = arg >= 0 ? jb.tree.rects[arg] : make_int4(0, 0, 0, 0);
}

template <int kMode, bool kStats, bool kHarvest, bool kRuns, bool kSingle>
// Waves per SIMD the register budget allows: 3 for the learned mode (168
// VGPRs without spills) and for the weighted one, 4 for the others; one more
// for single-job launches (kSingle: the job's tree view, master and pyramid
// are wave-uniform and live in scalar registers), measured faster at 10^6
// rects despite a few spilled registers.  kRuns: the traversal stack holds
// runs of siblings (RunStack) instead of single nodes (PlainStack).
__global__ void __launch_bounds__(kMode == kModeLearned ? kLearnedMaxBlock : 256) __attribute__((amdgpu_waves_per_eu(
    kMode == kModeLearned ? 3 : kMode == kModeWeighted ? (kSingle ? 4 : 3) : (kSingle ? 5 : 4))))
walk_kernel(const WalkJob* __restrict__ jobs, WalkArgs a, MlpNet net)
{
    extern __shared__ int32_t stack_mem[];
    __shared__ uint32_t sT[kNSample];
    __shared__ uint64_t s_begin[kLdsJobs];
    // the block's tallies of the first kLdsJobs jobs (master, escaped, hops;
    // weighted: master and escape sums and sums of squares), added to the
    // jobs' bins once when the block ends
    __shared__ unsigned long long s_tally[3 * kLdsJobs];
    // weighted (single-job launches): the block's fixed-point master and
    // escape sums and sums of squares
    __shared__ unsigned long long s_wtally[kMode == kModeWeighted ? 8 : 1];
    // the block's tallies of other bins: keys, then per slot the count (and,
    // weighted, the fixed-point sum and sum of squares, lo and hi words)
    constexpr int kBinVals = kMode == kModeWeighted ? 5 : 1;
    constexpr int kBinLog2 = bin_slots_log2(kMode, kSingle);
    constexpr int kBinSlots = 1 << kBinLog2;
    __shared__ unsigned long long s_bkey[kBinSlots];
    __shared__ unsigned long long s_bval[kBinSlots * kBinVals];
    // learned mode: the network's biases and output weights, and its weight
    // matrices (read from LDS by every block of 16 rows the wave evaluates;
    // reloaded through the vector caches, they took 4 L2 reads per hop)
    __shared__ __align__(16) float s_small[kMode == kModeLearned ? kMlpSmallFloats : 4];
    __shared__ __align__(16) uint16_t s_w[kMode == kModeLearned ? kMlpWLdsElems : 8];
    for (int i = threadIdx.x; i < kNSample; i += blockDim.x) sT[i] = a.thresholds[i];
    for (int i = threadIdx.x; i < min(a.n_jobs, kLdsJobs); i += blockDim.x) s_begin[i] = jobs[i].walk_begin;
    for (int i = threadIdx.x; i < 3 * kLdsJobs; i += blockDim.x) s_tally[i] = 0ull;
    for (int i = threadIdx.x; i < kBinSlots; i += blockDim.x) s_bkey[i] = kBinEmpty;
    for (int i = threadIdx.x; i < kBinSlots * kBinVals; i += blockDim.x) s_bval[i] = 0ull;
    if (kMode == kModeWeighted)
        for (int i = threadIdx.x; i < 8; i += blockDim.x) s_wtally[i] = 0ull;
    if (kMode == kModeLearned) {
        mlp_stage_small(net, s_small, threadIdx.x, blockDim.x);
        mlp_stage_weights(net, s_w, threadIdx.x, blockDim.x);
    }
    __syncthreads();
    const int lane = threadIdx.x & (kWaveSize - 1);
    using Stack = typename std::conditional<kRuns, RunStack, PlainStack>::type;
    // each wave's LDS: slot i of its lanes' stacks at [64 i, 64 i + 64); the
    // learned mode's activation tile reuses it (no traversal is live
    // while the wave evaluates the network)
    int32_t* wave_lds = stack_mem + (size_t)(threadIdx.x / kWaveSize) * a.wave_ints;
    const Stack st{wave_lds + lane, kWaveSize};
    uint16_t* mlp_xs = nullptr;
    if (kMode == kModeLearned) {
        mlp_xs = (uint16_t*)wave_lds;
    }
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);

    bool active = false, done = false;
    int j = 0;
    uint64_t w = 0;  // walk index within the job
    uint32_t hop = 0;
    int hint = -1;  // sorted index of the rect nearest to the walker's previous point
    int4 hint_r = make_int4(0, 0, 0, 0);  // and its coordinates
    DevTree t = jobs[0].tree;
    int mid = jobs[0].master_id;
    MlpPyramid pyr;
    if (kMode == kModeLearned) pyr = jobs[0].pyr;
    DBox sq{0, 0, 0, 0};
    uint32_t rnd[4] = {0, 0, 0, 0};
    unsigned long long my_hops = 0;  // hops accumulated for job j
    // walks of job j that ended on its master / escaped: counted per lane and
    // added to the tallies when the lane changes job or finishes (these two
    // bins take most walks; one atomic per walk on them would serialise)
    uint32_t my_master = 0, my_escaped = 0;
    uint64_t wnext = 0, wend = 0;  // the wave's unclaimed walk indices [wnext, wend), equal in every lane
    uint64_t chunk = kWaveChunk;
    const uint64_t tail_walks = (uint64_t)gridDim.x * (blockDim.x / kWaveSize) * 4 * kWaveChunk;
    // weighted estimator: the walk's weight, its preset first point, and
    // per-lane fixed-point sums (and sums of squares) for the master and
    // escape bins
    double wgt = 0.0, fpx = 0.0, fpy = 0.0;
    bool first_pt = false, start_stuck = false;
    Fx wm_s{0, 0}, wm_q{0, 0}, we_s{0, 0}, we_q{0, 0};
    Counters cnt{0u, 0u, 0u};
    Counters* cp = kStats ? &cnt : nullptr;
    unsigned long long my_nodes = 0, my_rects = 0, my_tabs = 0, my_acc = 0, my_fb = 0;
    uint64_t my_rec = 0;  // harvest: slot of the lane's next record

    // A lane hands in its tallies (master, escape, hops; weighted sums) for
    // its current job and clears them when it moves to another job and when
    // the launch ends: into the block's LDS table for the first kLdsJobs
    // jobs, added to the jobs' bins once per block at the end (per-lane
    // atomics on a job's bins serialised: every lane of the grid leaves a
    // job at about the same time), straight to the bins beyond.
    auto hand_in = [&]() {
        if (j < kLdsJobs) {
            if (my_master) atomicAdd(&s_tally[3 * j], (unsigned long long)my_master);
            if (my_escaped) atomicAdd(&s_tally[3 * j + 1], (unsigned long long)my_escaped);
            if (my_hops) atomicAdd(&s_tally[3 * j + 2], my_hops);
        } else {
            if (my_master) atomicAdd(&jobs[j].counts[mid], (unsigned long long)my_master);
            if (my_escaped) atomicAdd(&jobs[j].counts[t.n_rects], (unsigned long long)my_escaped);
            if (my_hops) atomicAdd(jobs[j].hops, my_hops);
        }
        if (kMode == kModeWeighted) {  // single-job launches: j == 0
            fx_atomic(&s_wtally[0], wm_s);
            fx_atomic(&s_wtally[2], wm_q);
            fx_atomic(&s_wtally[4], we_s);
            fx_atomic(&s_wtally[6], we_q);
            wm_s = wm_q = we_s = we_q = Fx{0, 0};
        }
        my_master = my_escaped = 0;
        my_hops = 0;
    };

    for (;;) {
        const bool need = !active && !done;
        uint64_t g = 0;
        bool got = false;
        // lanes without a walk take the next indices of the wave's range in
        // lane order; when the range runs short one lane claims kWaveChunk
        // more from the queue (one atomic on the shared counter per 256
        // walks: a claim per few walks serialised the kernel on that
        // address), kTailChunk once fewer than 4 chunks per wave of the grid
        // are left (a late large claim keeps its wave walking after the
        // others have finished)
        const unsigned long long mask = __ballot(need);
        if (mask) {
            const uint64_t n_need = (uint64_t)__popcll(mask);
            const uint64_t avail = wend - wnext;
            const bool refill = avail < n_need;
            uint64_t base = 0;
            if (refill) {
                const int leader = __ffsll((long long)mask) - 1;
                const uint64_t left = a.total_walks > wend ? a.total_walks - wend : 0;
                chunk = left > tail_walks ? kWaveChunk : kTailChunk;
                if (lane == leader) base = atomicAdd(a.work, (unsigned long long)chunk);
                base = __shfl(base, leader);
            }
            if (need) {
                const unsigned long long below = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
                const uint64_t rank = (uint64_t)__popcll(below);
                const uint64_t g0 = rank < avail ? wnext + rank : base + (rank - avail);
                if (g0 < a.total_walks) {
                    g = g0;
                    got = true;
                } else {
                    done = true;
                }
            }
            if (refill) {
                wnext = base + (n_need - avail);
                wend = base + chunk;
            } else {
                wnext += n_need;
            }
        }
        int nj = j;
        if (!kSingle && got) nj = find_job(jobs, s_begin, a.n_jobs, g);
        // lanes moving to another job hand in the old job's tallies first
        const bool change = !kSingle && got && nj != j;
        if (change) hand_in();
        if (got) {
            active = true;
            hop = 0;
            if (kHarvest) my_rec = a.h_off[g];  // by the global walk index (batched jobs in order)
            if (change) {
                j = nj;
                t = jobs[j].tree;
                mid = jobs[j].master_id;
                if (kMode == kModeLearned) pyr = jobs[j].pyr;
            }
            w = g - jobs[j].walk_begin;
            hint = jobs[j].start_hint;
            hint_r = jobs[j].start_rect;
            sq = DBox{jobs[j].surface[0], jobs[j].surface[1], jobs[j].surface[2], jobs[j].surface[3]};
            if (kMode == kModeWeighted) {
                // start: a uniform point r on the Gaussian surface G (outward
                // normal n, perimeter L), the largest empty square at r inside
                // the bounds (half-side h), a Green-CDF panel of it (its centre
                // is the first point) and the weight -eps L (n . g_k) / (p_k 2h)
                uint32_t c[4];
                walk_counter(c, jobs[j].first_walk + w, jobs[j].job, 0u, kGausDomain);
                philox4x32_10(c, k0, k1);
                const double gw = sq.x1 - sq.x0, gh = sq.y1 - sq.y0;
                const double L = 2.0 * (gw + gh);
                const double tp = (double)c[0] * (1.0 / 4294967296.0) * L;
                double rx, ry, nx, ny;
                if (tp < gw) {
                    rx = sq.x0 + tp; ry = sq.y0; nx = 0.0; ny = -1.0;
                } else if (tp < gw + gh) {
                    rx = sq.x1; ry = sq.y0 + (tp - gw); nx = 1.0; ny = 0.0;
                } else if (tp < 2.0 * gw + gh) {
                    rx = sq.x1 - (tp - gw - gh); ry = sq.y1; nx = 0.0; ny = 1.0;
                } else {
                    rx = sq.x0; ry = sq.y1 - (tp - 2.0 * gw - gh); nx = -1.0; ny = 0.0;
                }
                int ha = -1;
                double hh = nearest_rect(t, st, rx, ry, cp, hint, &ha, &hint_r);
                hint = ha;
                hh = fmin(hh, fmin(fmin(rx - (double)t.bounds.x, (double)t.bounds.z - rx),
                                   fmin(ry - (double)t.bounds.y, (double)t.bounds.w - ry)));
                wgt = 0.0;
                start_stuck = !(hh > 0.0);
                first_pt = false;
                if (!start_stuck) {
                    walk_counter(rnd, jobs[j].first_walk + w, jobs[j].job, 0u, kWalkDomain);
                    philox4x32_10(rnd, k0, k1);
                    // the first panel: from the Green CDF (probability p_k),
                    // or, with importance sampling, proportional to |n . g_k|
                    // (probability q_k; n is axis-aligned, so the table of
                    // |gx| or of |gy|)
                    int k;
                    double prob;
                    if (a.wflags & kWeightImportance) {
                        const int axis = nx != 0.0 ? 0 : 1;
                        const uint32_t* it = reinterpret_cast<const uint32_t*>(a.wtab + 5 * kNSample);
                        k = sample_panel(it + axis * kNSample, rnd[0]);
                        prob = a.wtab[(3 + axis) * kNSample + k];
                    } else {
                        k = sample_panel(sT, rnd[0]);
                        prob = a.wtab[2 * kNSample + k];
                    }
                    const double side = 2.0 * hh;
                    const double tk = ((double)(k % (kNSample / 4)) + 0.5) / (kNSample / 4) * side;
                    const double x0 = rx - hh, y0 = ry - hh;
                    switch (k / (kNSample / 4)) {
                    case 0: fpx = x0 + tk; fpy = y0; break;
                    case 1: fpx = x0 + side; fpy = y0 + tk; break;
                    case 2: fpx = x0 + side - tk; fpy = y0 + side; break;
                    default: fpx = x0; fpy = y0 + side - tk; break;
                    }
                    wgt = -a.eps * L * (nx * a.wtab[k] + ny * a.wtab[kNSample + k]) / (prob * 2.0 * hh);
                    first_pt = true;
                }
            }
        }
        if (__ballot(active) == 0ull) break;

        double px = 0.0, py = 0.0;
        bool escaped = false;
        if (active) {
            if (kMode == kModeWeighted && first_pt) {
                px = fpx;
                py = fpy;
                first_pt = false;
            } else {
                if ((hop & 3u) == 0u) {
                    walk_counter(rnd, jobs[j].first_walk + w, jobs[j].job, hop >> 2, kWalkDomain);
                    philox4x32_10(rnd, k0, k1);
                }
                const int panel = sample_panel(sT, rnd[hop & 3u]);
                point_from_panel(sq.x0, sq.y0, sq.x1, sq.y1, panel, px, py);
            }
            escaped = !in_bounds_strict(t.bounds, px, py);
        }
        // half-size of the square the point was sampled on
        const double hprev = fmax(sq.x1 - sq.x0, sq.y1 - sq.y0) / 2;

        float y_mlp = 0.0f;
        double cell = 1.0;
        bool want = false;  // learned: this hop asks the network for a square
        if (kMode == kModeLearned) {
            // A proposal is asked for where the previous square was at least
            // the finest verification cell: below it a proposal would be
            // probed at cells as large as itself and is hardly ever
            // certified.  The lanes that ask write their feature rows to the
            // first rows of the wave's tile, and the network runs on those
            // rows only, in blocks of 16 (all 64 lanes take part).
            want = active && !escaped && hprev >= ldexp(1.0, pyr.K0 - pyr.Vmax);
            const unsigned long long ask = __ballot(want);
            if (ask) {
                const int row = __popcll(ask & (lane ? (~0ull) >> (64 - lane) : 0ull));
                if (want) {
                    float fx = 0.0f, fy = 0.0f;
                    const int level = mlp_level(pyr, hprev, &cell);
                    const uint64_t bits = mlp_window(pyr, level, cell, px, py, &fx, &fy);
                    mlp_feature_bits(mlp_xs + row * kMlpStride, bits, level, fx, fy);
                }
                y_mlp = mlp_forward_rows(s_w, mlp_xs, s_small, lane, __popcll(ask), want ? row : 64);
            }
        }
        if (!active) continue;

        int result = -3;  // walk continues
        if (kMode == kModeWeighted && start_stuck) {
            result = -2;  // no empty square at the start point
            start_stuck = false;
        } else if (escaped) {
            result = -1;
        } else {
            bool proposed = false;
            if (kMode == kModeLearned && want) {
                // outputs beyond the training targets' range [-4, 4] are not trusted;
                // 2^y in single precision (a proposal's size needs no more: the
                // probe decides whether the square is free)
                const float y = fminf(fmaxf(y_mlp, -4.0f), 4.0f);
                const double d = cell * (double)exp2f(y) * (double)net.alpha;
                if (d > 2.0 && d < 1e300) {
                    const DBox c{px - d, py - d, px + d, py + d};
                    if (mlp_certify(pyr, c.x0, c.y0, c.x1, c.y1, d)) {
                        sq = c;
                        proposed = true;
                    }
                }
            }
            // proposals (hops that evaluated the network): accepted or rejected
            if (kMode == kModeLearned && want) {
                if (proposed) my_acc++; else my_fb++;
            }
            if (!proposed) {
                if (kMode != kModeBisect) {
                    int arg = -1;
                    const double ds = nearest_rect(t, st, px, py, cp, hint, &arg, &hint_r);
                    hint = arg;
                    if (ds <= 0.0) {
                        // p lies in a rect: the search's arg is the point query's answer
                        result = gload(t.ids + arg);
                    } else {
                        if (kHarvest) {
                            const uint64_t slot = my_rec++;
                            if (slot < a.h_cap) {
                                a.h_pts[2 * slot] = px;
                                a.h_pts[2 * slot + 1] = py;
                                a.h_prev[slot] = hprev;
                                a.h_dstar[slot] = ds;
                            }
                        }
                        sq = biggest_square<false>(t, st, px, py, ds, nullptr, cp, a.err, &hint_r);
                    }
                } else {
                    sq = biggest_square<true>(t, st, px, py, 0.0, nullptr, cp, a.err);
                    const int h = q2_point(t, st, px, py, cp);
                    if (h >= 0) result = h;
                }
            }
            if (kStats) {
                my_nodes += cnt.nodes;
                my_rects += cnt.rects;
                my_tabs += cnt.tabs;
                cnt.nodes = cnt.rects = cnt.tabs = 0u;
            }
            hop++;
            if (result == -3 && hop >= a.max_hops) result = -2;
        }
        if (result != -3) {
            const int bin = result >= 0 ? result : t.n_rects + 1;  // other bins: the hit rect, or stuck
            int slot = -1;
            if (result == mid) {
                my_master++;
            } else if (result == -1) {
                my_escaped++;
            } else {
                slot = bin_slot<kBinLog2>(s_bkey, bin_key(j, bin));
                if (slot >= 0) atomicAdd(&s_bval[kBinVals * slot], 1ull);
                else atomicAdd(&jobs[j].counts[bin], 1ull);
            }
            if (kMode == kModeWeighted) {
                if (jobs[j].w_out) jobs[j].w_out[w] = wgt;
                Fx s1, q1;
                fx_weight(wgt, s1, q1, a.err);
                if (result == mid) {
                    fx_add(wm_s, s1);
                    fx_add(wm_q, q1);
                } else if (result == -1) {
                    fx_add(we_s, s1);
                    fx_add(we_q, q1);
                } else if (slot >= 0) {
                    fx_atomic(&s_bval[kBinVals * slot + 1], s1);
                    fx_atomic(&s_bval[kBinVals * slot + 3], q1);
                } else {
                    fx_atomic(jobs[j].wacc + 4 * bin, s1);
                    fx_atomic(jobs[j].wacc + 4 * bin + 2, q1);
                }
            }
            my_hops += hop;
            if (jobs[j].hit_out) jobs[j].hit_out[w] = result;
            if (jobs[j].hops_out) jobs[j].hops_out[w] = (int32_t)hop;
            active = false;
        }
    }
    hand_in();
    // the block's table to the jobs' bins, once all its waves are done
    __syncthreads();
    for (int i = threadIdx.x; i < min(a.n_jobs, kLdsJobs); i += blockDim.x) {
        const unsigned long long m = s_tally[3 * i], e = s_tally[3 * i + 1], h = s_tally[3 * i + 2];
        const int lm = jobs[i].master_id, le = jobs[i].tree.n_rects;
        if (m) atomicAdd(&jobs[i].counts[lm], m);
        if (e) atomicAdd(&jobs[i].counts[le], e);
        if (h) atomicAdd(jobs[i].hops, h);
        if (kMode == kModeWeighted && i == 0) {
            fx_atomic(jobs[0].wacc + 4 * lm, Fx{s_wtally[0], s_wtally[1]});
            fx_atomic(jobs[0].wacc + 4 * lm + 2, Fx{s_wtally[2], s_wtally[3]});
            fx_atomic(jobs[0].wacc + 4 * le, Fx{s_wtally[4], s_wtally[5]});
            fx_atomic(jobs[0].wacc + 4 * le + 2, Fx{s_wtally[6], s_wtally[7]});
        }
    }
    for (int i = threadIdx.x; i < kBinSlots; i += blockDim.x) {
        const unsigned long long k = s_bkey[i];
        if (k == kBinEmpty) continue;
        const int jj = (int)(k >> 32), bin = (int)(uint32_t)k;
        atomicAdd(&jobs[jj].counts[bin], s_bval[kBinVals * i]);
        if (kMode == kModeWeighted) {
            fx_atomic(jobs[jj].wacc + 4 * bin, Fx{s_bval[kBinVals * i + 1], s_bval[kBinVals * i + 2]});
            fx_atomic(jobs[jj].wacc + 4 * bin + 2, Fx{s_bval[kBinVals * i + 3], s_bval[kBinVals * i + 4]});
        }
    }
    if (kStats) {
        for (int off = 32; off >= 1; off >>= 1) {
            my_nodes += __shfl_xor(my_nodes, off);
            my_rects += __shfl_xor(my_rects, off);
            my_tabs += __shfl_xor(my_tabs, off);
        }
        if (lane == 0) {
            atomicAdd(&a.traffic[0], my_nodes);
            atomicAdd(&a.traffic[1], my_rects);
            atomicAdd(&a.traffic[2], my_tabs);
        }
    }
    if (kMode == kModeLearned) {
        for (int off = 32; off >= 1; off >>= 1) {
            my_acc += __shfl_xor(my_acc, off);
            my_fb += __shfl_xor(my_fb, off);
        }
        if (lane == 0 && a.l_stats) {
            atomicAdd(&a.l_stats[0], my_acc);
            atomicAdd(&a.l_stats[1], my_fb);
        }
    }
}

// Writes the single job of a plain walk into the tree's job slot.
__global__ void set_job(WalkJob* slot, WalkJob jb) { *slot = jb; }

// Launch shape when the caller gives none: the block size (256, 192, 128 or
// 64 lanes; the learned mode also 768 and 512) that keeps the most lanes
// resident per CU for this kernel's registers and its LDS, and one such full
// set of blocks per CU.
struct AutoShape {
    bool on = false;
    size_t wave_bytes = 0, fixed = 0, lds_block = 0;
    int cus = 1;
    uint64_t total = 0;
};

template <int kMode, bool kStats, bool kHarvest, bool kRuns, bool kSingle>
void launch_kernel_as(int grid, int block, size_t lds, const AutoShape& as, hipStream_t s, const WalkJob* jobs,
                      const WalkArgs& a, const MlpNet& m)
{
    auto kern = walk_kernel<kMode, kStats, kHarvest, kRuns, kSingle>;
    if (as.on) {
        int best = 0;
        for (int b : {kLearnedMaxBlock, 512, 256, 192, 128, 64}) {
            if (b > (kMode == kModeLearned ? kLearnedMaxBlock : 256)) continue;
            const size_t l = (size_t)(b / kWaveSize) * as.wave_bytes;
            if (l + as.fixed > as.lds_block) continue;
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, b, l) != hipSuccess || nb <= 0) continue;
            if (nb * b > best) {
                best = nb * b;
                block = b;
                lds = l;
                grid = nb * as.cus;
            }
        }
        const uint64_t want = (as.total + block - 1) / block;
        if (want < (uint64_t)grid) grid = (int)std::max<uint64_t>(1, want);
    }
    kern<<<grid, block, lds, s>>>(jobs, a, m);
}

template <int kMode, bool kStats, bool kHarvest>
void launch_kernel(bool runs, int grid, int block, size_t lds, const AutoShape& as, hipStream_t s,
                   const WalkJob* jobs, const WalkArgs& a, const MlpNet& m)
{
    if (a.n_jobs == 1) {
        if (runs) launch_kernel_as<kMode, kStats, kHarvest, true, true>(grid, block, lds, as, s, jobs, a, m);
        else launch_kernel_as<kMode, kStats, kHarvest, false, true>(grid, block, lds, as, s, jobs, a, m);
    } else {
        if (runs) launch_kernel_as<kMode, kStats, kHarvest, true, false>(grid, block, lds, as, s, jobs, a, m);
        else launch_kernel_as<kMode, kStats, kHarvest, false, false>(grid, block, lds, as, s, jobs, a, m);
    }
}

}  // namespace

// Static LDS of a walk block: Green thresholds, job offsets, job tallies,
// the table of other bins' tallies, the learned mode's small operands and
// weights.
size_t walk_static_lds(int mode, bool single)
{
    const size_t bin_slots = (size_t)1 << bin_slots_log2(mode, single);
    return kNSample * 4 + kLdsJobs * 8 + kLdsJobs * 24 + (mode == kModeWeighted ? 64 : 8) +
           bin_slots * 8 * (mode == kModeWeighted ? 6 : 2) +
           (mode == kModeLearned ? kMlpSmallFloats * 4 + kMlpWLdsElems * 2 : 16 + 16);
}

// LDS of one wave: its lanes' stacks, or the learned mode's activation
// tiles when larger (the two alias).
size_t walk_wave_bytes(const WalkLaunch& L, bool runs)
{
    const int entries = runs ? RunStack::entries(L.max_depth) : PlainStack::entries(L.max_depth);
    const size_t stack = (size_t)kWaveSize * entries * sizeof(int32_t);
    const size_t mlp = L.mode == kModeLearned ? (size_t)kMlpLdsPerWave : 0;
    return std::max(stack, mlp);
}

// Stack layout of a walk launch: runs of siblings (RunStack) when the
// single-node stack would hold fewer lanes per CU in LDS than the walk
// kernels' registers allow (4 SIMDs x the kernel's waves per SIMD) and runs
// take less, single nodes (PlainStack, fewer instructions per pop) otherwise.
bool walk_uses_runs(const WalkLaunch& L, size_t lds_cu, size_t fixed)
{
    if (L.max_nodes >= kMaxTreeNodes) return false;
    const bool single = L.n_jobs == 1;
    const size_t waves = L.mode == kModeLearned ? 12 : L.mode == kModeWeighted ? (single ? 16 : 12) : (single ? 20 : 16);
    const size_t plain = walk_wave_bytes(L, false);
    return walk_wave_bytes(L, true) < plain && waves * plain + (waves / 4) * fixed > lds_cu;
}

int plan_walk(const WalkLaunch& L, int* block, int* grid, size_t* lds, bool* runs)
{
    int b = L.block;
    const int bmax = L.mode == kModeLearned ? kLearnedMaxBlock : 256;
    if (b < 0 || b % kWaveSize != 0 || b > bmax)
        return set_error(FRW_E_INVALID, "block must be a multiple of 64 and <= 256 (learned mode: 768)");
    const size_t fixed = walk_static_lds(L.mode, L.n_jobs == 1);
    int lds_block = 0, lds_cu = 0, cus = 0;
    if (hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, L.device) != hipSuccess ||
        hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, L.device) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, L.device) != hipSuccess)
        return set_error(FRW_E_DEVICE, "device attributes unavailable");
    if (lds_block <= 0) lds_block = 64 * 1024;
    if (lds_cu <= 0) lds_cu = lds_block;
    const bool use_runs = walk_uses_runs(L, (size_t)lds_cu, fixed);
    if (runs) *runs = use_runs;
    const size_t wave_bytes = walk_wave_bytes(L, use_runs);
    auto need = [&](int bb) { return (size_t)(bb / kWaveSize) * wave_bytes; };
    if (b == 0) {
        b = bmax;
        while (b >= 64 && need(b) + fixed > (size_t)lds_block) b -= 64;
    }
    if (b < 64 || need(b) + fixed > (size_t)lds_block)
        return set_error(FRW_E_RANGE, "tree too deep for the LDS traversal stack at this block size");
    *block = b;
    *lds = need(b);
    int g = L.grid;
    if (g <= 0) {
        const int per_cu_lds = (int)((size_t)lds_cu / (*lds + fixed));
        const int per_cu = std::max(1, std::min(per_cu_lds, 2048 / b));
        g = std::max(1, cus * per_cu);
        const int64_t want = ((int64_t)L.total_walks + b - 1) / b;
        if (want < g) g = (int)want;
    }
    *grid = g;
    return FRW_OK;
}

int launch_walk_jobs(WalkJob* d_jobs, const WalkLaunch& L, hipStream_t s)
{
    if (L.total_walks == 0) return FRW_OK;
    int block = 0, grid = 0;
    size_t lds = 0;
    bool runs = false;
    int rc = plan_walk(L, &block, &grid, &lds, &runs);
    if (rc != FRW_OK) return rc;
    const size_t wave_bytes = walk_wave_bytes(L, runs);
    AutoShape as;
    if (L.block == 0 && L.grid == 0) {
        int lds_block = 0, cus = 0;
        if (hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, L.device) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, L.device) != hipSuccess)
            return set_error(FRW_E_DEVICE, "device attributes unavailable");
        as.on = true;
        as.wave_bytes = wave_bytes;
        as.fixed = walk_static_lds(L.mode, L.n_jobs == 1);
        as.lds_block = (size_t)lds_block;
        as.cus = std::max(cus, 1);
        as.total = L.total_walks;
    }
    const size_t plds = (size_t)64 * PlainStack::entries(L.max_depth) * sizeof(int32_t);
    walk_prologue<<<(L.n_jobs + 63) / 64, 64, plds, s>>>(d_jobs, L.n_jobs, L.work);
    if ((rc = check_hip(hipGetLastError(), "walk_prologue")) != FRW_OK) return rc;
    WalkArgs a{};
    a.thresholds = L.thresholds;
    a.work = L.work;
    a.err = L.err;
    a.traffic = L.traffic;
    a.seed = L.seed;
    a.total_walks = L.total_walks;
    a.max_hops = L.max_hops;
    a.n_jobs = L.n_jobs;
    a.wave_ints = (int32_t)(wave_bytes / sizeof(int32_t));
    a.h_pts = L.h_pts;
    a.h_prev = L.h_prev;
    a.h_dstar = L.h_dstar;
    a.h_off = L.h_off;
    a.h_cap = L.h_cap;
    a.l_stats = L.l_stats;
    a.wtab = L.wtab;
    a.eps = L.eps;
    a.wflags = L.wflags;
    MlpNet m{};
    if (L.net) m = *L.net;
    const bool st = L.traffic != nullptr;
    if (L.mode == kModeLearned) {
        if (!L.net) return set_error(FRW_E_INVALID, "learned mode needs a network");
        if (st) launch_kernel<kModeLearned, true, false>(runs, grid, block, lds, as, s, d_jobs, a, m);
        else launch_kernel<kModeLearned, false, false>(runs, grid, block, lds, as, s, d_jobs, a, m);
    } else if (L.mode == kModeBisect) {
        if (st) launch_kernel<kModeBisect, true, false>(runs, grid, block, lds, as, s, d_jobs, a, m);
        else launch_kernel<kModeBisect, false, false>(runs, grid, block, lds, as, s, d_jobs, a, m);
    } else if (L.mode == kModeWeighted) {
        if (!L.wtab) return set_error(FRW_E_INVALID, "weighted mode needs its tables");
        if (L.n_jobs != 1) return set_error(FRW_E_INVALID, "weighted walks run one job per launch");
        if (st) launch_kernel<kModeWeighted, true, false>(runs, grid, block, lds, as, s, d_jobs, a, m);
        else launch_kernel<kModeWeighted, false, false>(runs, grid, block, lds, as, s, d_jobs, a, m);
    } else if (L.h_off) {
        launch_kernel<kModeExact, false, true>(runs, grid, block, lds, as, s, d_jobs, a, m);
    } else {
        if (st) launch_kernel<kModeExact, true, false>(runs, grid, block, lds, as, s, d_jobs, a, m);
        else launch_kernel<kModeExact, false, false>(runs, grid, block, lds, as, s, d_jobs, a, m);
    }
    return check_hip(hipGetLastError(), "walk_kernel");
}

// One single-layout launch with its job record, work counter and error flag
// at the given device addresses (the tree's own, or a caller's scratch).
int launch_walk_in(const frw_tree* t, const frw_walk_params* p, const frw_walk_outputs* o, hipStream_t s,
                   const WalkScratch& ws, unsigned long long* traffic, const WalkExtras* extra)
{
    if (p->mode == kModeRefCompat) {
        if (traffic) return set_error(FRW_E_INVALID, "no traffic counters in the reference-compatible mode");
        return launch_ref_walk(t, p, o, s);
    }
    WalkJob jb{};
    jb.tree = dev_view(t);
    jb.counts = (unsigned long long*)o->counts;
    jb.hops = (unsigned long long*)o->hops;
    jb.hit_out = o->hit_out;
    jb.hops_out = o->hops_out;
    jb.walk_begin = 0;
    jb.n_walks = (uint64_t)p->n_walks;
    jb.first_walk = (uint64_t)p->first_walk;
    jb.job = p->job;
    jb.master = t->host_rects[p->master];
    jb.master_id = p->master;
    if (extra) {
        jb.wacc = extra->wacc;
        jb.w_out = extra->w_out;
    }
    if (extra && extra->pyr) jb.pyr = *extra->pyr;
    set_job<<<1, 1, 0, s>>>(ws.job, jb);
    int rc = check_hip(hipGetLastError(), "set_job");
    if (rc != FRW_OK) return rc;
    WalkLaunch L;
    L.device = t->device;
    L.n_jobs = 1;
    L.total_walks = (uint64_t)p->n_walks;
    L.max_depth = t->depth;
    L.max_nodes = (int)t->n_nodes;
    L.mode = p->mode;
    L.block = p->block;
    L.grid = p->grid;
    L.seed = p->seed;
    L.max_hops = (uint32_t)(p->max_hops > 0 ? p->max_hops : 10000000);
    L.thresholds = t->d_thresholds;
    L.work = ws.work;
    L.err = ws.err;
    L.traffic = traffic;
    if (extra) {
        L.h_pts = extra->h_pts;
        L.h_prev = extra->h_prev;
        L.h_dstar = extra->h_dstar;
        L.h_off = extra->h_off;
        L.h_cap = extra->h_cap;
        L.net = extra->net;
        L.eps = extra->eps;
        L.wflags = extra->wflags;
        L.wtab = extra->wsum ? t->d_wtab : nullptr;
        L.l_stats = extra->l_stats;
    }
    int rc2 = launch_walk_jobs(ws.job, L, s);
    if (rc2 != FRW_OK || p->mode != kModeWeighted) return rc2;
    if (!extra || !extra->wacc || !extra->wsum || !extra->wsq)
        return set_error(FRW_E_INVALID, "weighted walks need their accumulators");
    const int64_t nb = t->n_rects + 2;
    weighted_finish<<<(unsigned)((nb + 255) / 256), 256, 0, s>>>(extra->wacc, nb, extra->wsum, extra->wsq);
    return check_hip(hipGetLastError(), "weighted_finish");
}

// Walks on one tree share its job slot, work counter and error flag, so each
// launch is ordered after the tree's previous one (whatever its stream).
int launch_walk(const frw_tree* t, const frw_walk_params* p, const frw_walk_outputs* o,
                hipStream_t s, unsigned long long* traffic, const WalkExtras* extra)
{
    if (p->n_walks == 0) return FRW_OK;
    bool capturing = false;
    int rc = order_before(t->order, s, &capturing);
    if (rc != FRW_OK) return rc;
    rc = launch_walk_in(t, p, o, s, WalkScratch{t->d_job, t->d_work, t->d_err}, traffic, extra);
    const int rc2 = order_after(t->order, s, capturing);
    return rc != FRW_OK ? rc : rc2;
}

}  // namespace frw
