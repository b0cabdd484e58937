// Level-synchronous Morton quadtree build on gfx950.
//
// Replaces the reference's recursive dynamic-parallelism build
// (randomWalk/Cuda/createquadtree.cu:97-371, one block per node):
//   1. key per rect: Morton code of its centre, so each rect belongs to the
//      quadtree cell holding its centre at every level (loose quadtree).  The
//      key layout (code << 6 | level) also admits rects that stop at an inner
//      cell; with centre keys every rect sits at the deepest level.
//   2. one device radix sort of (key, rect id) (stable LSD radix sort, this
//      file).  Sorted by key, the rects of any cell form one contiguous run.
//   3. level by level: every frontier node splits its run into its four
//      quadrant runs by binary search over the sorted keys; a node becomes a
//      leaf when its run holds <= leaf_cap rects (or it is a unit cell).
//      Child groups are allocated by an exclusive scan, so the flat array is
//      in breadth-first order and deterministic.
//   4. each node's box = tight bounding box of its run (block reduction); the
//      boxes of siblings may overlap.
// Query answers depend only on the rect set, never on this structure.
#include <algorithm>
#include <climits>
#include <cstdio>

#include "frw_internal.h"

namespace frw {
namespace {

struct Frontier {
    uint64_t code;   // cell code padded with zeros (2K bits)
    int32_t node;    // index in the flat node array
    int32_t level;
    int32_t s, e;    // run of sorted rects in this subtree
};

__device__ inline uint64_t spread_bits(uint32_t v)
{
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

// Quadrant order inside a group: 0 = (low x, low y) = top-left (y down),
// 1 = top-right, 2 = bottom-left, 3 = bottom-right.
__device__ inline uint64_t morton(uint32_t ux, uint32_t uy)
{
    return spread_bits(ux) | (spread_bits(uy) << 1);
}

// Key of a rect: Morton code of its doubled centre (x0 + x1, y0 + y1)
// relative to the bounds corner (K bits per axis), tagged with the deepest
// level K.  A rect belongs to the cell that holds its centre: no rect stays
// behind in an internal node, and node boxes are the tight bounds of their
// subtrees (a loose quadtree).
__global__ void rect_keys(const int4* __restrict__ rects, int n, int4 B, int K,
                          uint64_t* __restrict__ keys, int32_t* __restrict__ vals)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int4 r = rects[i];
    const uint32_t cx = (uint32_t)((long long)r.x - B.x + (long long)r.z - B.x);
    const uint32_t cy = (uint32_t)((long long)r.y - B.y + (long long)r.w - B.y);
    keys[i] = (morton(cx, cy) << 6) | (uint64_t)K;
    vals[i] = i;
}

__global__ void gather_rects(const int4* __restrict__ rects, const int32_t* __restrict__ order,
                             int n, int4* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = rects[order[i]];
}

// first index in [lo, hi) with keys[idx] >= v
__device__ inline int lower_bound_key(const uint64_t* keys, int lo, int hi, uint64_t v)
{
    while (lo < hi) {
        const int mid = lo + ((hi - lo) >> 1);
        if (keys[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ inline int upper_bound_key(const uint64_t* keys, int lo, int hi, uint64_t v)
{
    while (lo < hi) {
        const int mid = lo + ((hi - lo) >> 1);
        if (keys[mid] <= v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__global__ void split_level(const uint64_t* __restrict__ keys, const Frontier* __restrict__ fr,
                            int nf, int K, int cap, Node* __restrict__ nodes,
                            int2* __restrict__ runs, int32_t* __restrict__ internal,
                            int32_t* __restrict__ own_end)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const Frontier f = fr[i];
    Node nd;
    nd.level = f.level;
    nd.first_child = -1;
    nd.box = make_int4(INT_MAX, INT_MAX, INT_MIN, INT_MIN);
    runs[f.node] = make_int2(f.s, f.e);
    int is_int = 0, oe = f.e;
    if (f.s == f.e) {
        nd.rbeg = nd.rend = 0;  // empty quadrant
    } else {
        const uint64_t own_key = (f.code << 6) | (uint64_t)f.level;
        oe = upper_bound_key(keys, f.s, f.e, own_key);
        if (f.e - f.s <= cap || f.level >= K || oe == f.e) {
            nd.rbeg = f.s;
            nd.rend = f.e;
        } else {
            nd.rbeg = f.s;
            nd.rend = oe;
            is_int = 1;
        }
    }
    nodes[f.node] = nd;
    internal[i] = is_int;
    own_end[i] = oe;
}

__global__ void emit_children(const uint64_t* __restrict__ keys, const Frontier* __restrict__ fr,
                              int nf, int K, const int32_t* __restrict__ internal,
                              const int32_t* __restrict__ group, const int32_t* __restrict__ own_end,
                              int base, Node* __restrict__ nodes, Frontier* __restrict__ next)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf || !internal[i]) return;
    const Frontier f = fr[i];
    const int g = group[i];
    const int first = base + 4 * g;
    nodes[f.node].first_child = first;
    const int shift = 2 * (K - f.level - 1);
    int b[5];
    b[0] = own_end[i];
    b[4] = f.e;
#pragma unroll
    for (int c = 1; c < 4; ++c)
        b[c] = lower_bound_key(keys, b[c - 1], f.e, (f.code | ((uint64_t)c << shift)) << 6);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        Frontier ch;
        ch.code = f.code | ((uint64_t)c << shift);
        ch.node = first + c;
        ch.level = f.level + 1;
        ch.s = b[c];
        ch.e = b[c + 1];
        next[4 * g + c] = ch;
    }
}

// Tight bounding boxes of one level's nodes [begin, end), deepest level
// first: a leaf's from its run of rects (an empty run gives the inverted box
// every test fails on), an internal node's as the union of its four
// children's, which the previous launch has written.  Every rect lives in a
// leaf and a node's run is the union of its children's, so this is the box
// of the node's run (one block per node over its whole run had spent 1 ms on
// the 10^6-rect root alone).
__global__ void __launch_bounds__(256) node_boxes_level(const int4* __restrict__ rects,
                                                        const int2* __restrict__ runs, int begin, int end,
                                                        Node* __restrict__ nodes)
{
    const int ni = begin + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (ni >= end) return;
    int x0 = INT_MAX, y0 = INT_MAX, x1 = INT_MIN, y1 = INT_MIN;
    const int fc = nodes[ni].first_child;
    if (fc < 0) {
        const int2 run = runs[ni];
        for (int r = run.x; r < run.y; ++r) {
            const int4 q = rects[r];
            x0 = min(x0, q.x); y0 = min(y0, q.y); x1 = max(x1, q.z); y1 = max(y1, q.w);
        }
    } else {
        for (int c = 0; c < 4; ++c) {
            const int4 b = nodes[fc + c].box;
            x0 = min(x0, b.x); y0 = min(y0, b.y); x1 = max(x1, b.z); y1 = max(y1, b.w);
        }
    }
    nodes[ni].box = make_int4(x0, y0, x1, y1);
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    int reserve(size_t n, hipStream_t s, bool keep = false, size_t keep_n = 0)
    {
        if (n <= cap) return FRW_OK;
        size_t nc = std::max<size_t>(n, cap * 2);
        T* q = nullptr;
        if (hipMalloc(&q, nc * sizeof(T)) != hipSuccess) return set_error(FRW_E_NOMEM, "hipMalloc failed in tree build");
        if (keep && p && keep_n)
            if (hipMemcpyAsync(q, p, keep_n * sizeof(T), hipMemcpyDeviceToDevice, s) != hipSuccess)
                return set_error(FRW_E_DEVICE, "copy failed in tree build");
        if (p) {
            if (hipStreamSynchronize(s) != hipSuccess) return set_error(FRW_E_DEVICE, "sync failed in tree build");
            (void)hipFree(p);
        }
        p = q;
        cap = nc;
        return FRW_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

inline int blocks_for(int64_t n, int b) { return (int)((n + b - 1) / b); }

// ---------------------------------------------------------------- scan
// Exclusive prefix sums of uint32 arrays, in three launches: per-tile sums
// (tiles of 256 threads x 16 items), one block scanning the tile sums, and
// each tile's scan offset by its tile's sum.  Integer sums: the result does
// not depend on the order of the additions.
constexpr int kScanItems = 16;
constexpr int kScanTile = 256 * kScanItems;

// Exclusive scan of one value per thread over a 256-thread block; *total
// receives the block's sum.  sh: 4 words of LDS.
__device__ inline uint32_t block_excl_scan256(uint32_t v, uint32_t* sh, uint32_t* total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(inc, off);
        if (lane >= off) inc += o;
    }
    if (lane == 63) sh[wave] = inc;
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += sh[w];
    *total = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return before + inc - v;
}

__global__ void __launch_bounds__(256) scan_tile_sums(const uint32_t* __restrict__ in, int64_t m,
                                                      uint32_t* __restrict__ sums)
{
    __shared__ uint32_t sh[4];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < m) s += in[base + k];
    uint32_t tot;
    (void)block_excl_scan256(s, sh, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Exclusive scan in place of the nb tile sums, one 256-thread block in
// chunks of 256 with a running carry.
__global__ void __launch_bounds__(256) scan_sums(uint32_t* __restrict__ sums, int nb)
{
    __shared__ uint32_t sh[4];
    uint32_t carry = 0;
    for (int c0 = 0; c0 < nb; c0 += 256) {
        const int i = c0 + (int)threadIdx.x;
        const uint32_t v = i < nb ? sums[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan256(v, sh, &tot);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
    }
}

__global__ void __launch_bounds__(256) scan_apply(const uint32_t* __restrict__ in, int64_t m,
                                                  const uint32_t* __restrict__ sums, uint32_t* __restrict__ out)
{
    __shared__ uint32_t sh[4];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < m ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = sums[blockIdx.x] + block_excl_scan256(s, sh, &tot);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < m) out[base + k] = run;
        run += v[k];
    }
}

// out[i] = sum of in[j] for j < i (out may alias in); sums: scratch of
// scan_tile_count(m) words.
inline int64_t scan_tile_count(int64_t m) { return (m + kScanTile - 1) / kScanTile; }

int exclusive_scan(const uint32_t* in, uint32_t* out, int64_t m, uint32_t* sums, hipStream_t s)
{
    if (m <= 0) return FRW_OK;
    const int nb = (int)scan_tile_count(m);
    scan_tile_sums<<<nb, 256, 0, s>>>(in, m, sums);
    scan_sums<<<1, 256, 0, s>>>(sums, nb);
    scan_apply<<<nb, 256, 0, s>>>(in, m, sums, out);
    return check_hip(hipGetLastError(), "exclusive scan");
}

// ---------------------------------------------------------------- sort
// Stable LSD radix sort of (uint64 key, int32 value) pairs on key bits
// [lo, hi), 8 bits a pass.  A pass: per-tile digit histograms (tiles of
// 256 threads x 16 items, LDS counters), the exclusive scan of the
// histograms laid out digit-major (so a tile's items of digit d land after
// every earlier tile's), and a scatter that ranks the items of each digit in
// input order: a tile goes in 16 rounds of 256 consecutive items; in a round
// each wave finds the lanes holding its digit by 8 ballots, the waves' counts
// per digit go through LDS, and the per-digit running offsets advance after
// the round.  Equal keys keep their input order.
constexpr int kSortItems = 16;
constexpr int kSortTile = 256 * kSortItems;

__device__ inline uint32_t sort_digit(uint64_t k, int shift) { return (uint32_t)(k >> shift) & 255u; }

__global__ void __launch_bounds__(256) sort_hist(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                 uint32_t* __restrict__ hist, int nb)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const int64_t tile = (int64_t)blockIdx.x * kSortTile;
#pragma unroll 4
    for (int j = 0; j < kSortItems; ++j) {
        const int64_t i = tile + (int64_t)j * 256 + threadIdx.x;
        if (i < n) atomicAdd(&h[sort_digit(keys[i], shift)], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(256) sort_scatter(const uint64_t* __restrict__ kin, const int32_t* __restrict__ vin,
                                                    uint64_t* __restrict__ kout, int32_t* __restrict__ vout,
                                                    int64_t n, int shift, const uint32_t* __restrict__ offs, int nb)
{
    __shared__ uint32_t s_base[256];
    __shared__ uint32_t s_cnt[4][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    s_base[tid] = offs[(int64_t)tid * nb + blockIdx.x];
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const int64_t tile = (int64_t)blockIdx.x * kSortTile;
    for (int j = 0; j < kSortItems; ++j) {
#pragma unroll
        for (int w = 0; w < 4; ++w) s_cnt[w][tid] = 0u;
        __syncthreads();
        const int64_t i = tile + (int64_t)j * 256 + tid;
        const bool valid = i < n;
        const uint64_t key = valid ? kin[i] : 0ull;
        const uint32_t d = sort_digit(key, shift);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t ones = __ballot(((d >> b) & 1u) != 0u);
            peers &= ((d >> b) & 1u) ? ones : ~ones;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & below);
        if (valid && rank == 0u) s_cnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = s_base[d] + rank;
            for (int w = 0; w < wave; ++w) pos += s_cnt[w][d];
            if ((int64_t)pos < n) {  // always, for a consistent histogram
                kout[pos] = key;
                vout[pos] = vin[i];
            }
        }
        __syncthreads();
        s_base[tid] += s_cnt[0][tid] + s_cnt[1][tid] + s_cnt[2][tid] + s_cnt[3][tid];
        __syncthreads();
    }
}

// Sorts (ka, va) by key bits [lo, hi) into (kb, vb); ka and va serve as the
// other half of the passes' ping-pong and are clobbered.  hist: 256 x
// (tiles) words, sums: scan scratch for them.
int sort_pairs(uint64_t* ka, int32_t* va, uint64_t* kb, int32_t* vb, int64_t n, int lo, int hi, uint32_t* hist,
               uint32_t* sums, hipStream_t s)
{
    if (n <= 0) return FRW_OK;
    const int nb = (int)((n + kSortTile - 1) / kSortTile);
    const int passes = (hi - lo + 7) / 8;
    uint64_t* kin = ka;
    int32_t* vin = va;
    uint64_t* kout = kb;
    int32_t* vout = vb;
    for (int p = 0; p < passes; ++p) {
        const int shift = lo + 8 * p;
        sort_hist<<<nb, 256, 0, s>>>(kin, n, shift, hist, nb);
        int rc = exclusive_scan(hist, hist, (int64_t)256 * nb, sums, s);
        if (rc != FRW_OK) return rc;
        sort_scatter<<<nb, 256, 0, s>>>(kin, vin, kout, vout, n, shift, hist, nb);
        rc = check_hip(hipGetLastError(), "radix sort pass");
        if (rc != FRW_OK) return rc;
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (kin != kb) {  // an even pass count ends in (ka, va)
        int rc = check_hip(hipMemcpyAsync(kb, kin, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s), "sort copy");
        if (rc == FRW_OK) rc = check_hip(hipMemcpyAsync(vb, vin, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToDevice, s), "sort copy");
        return rc;
    }
    return FRW_OK;
}

// One thread per start-table entry (all levels): the node of its cell.
__global__ void start_table(const Node* __restrict__ nodes, int levels, int64_t entries, int32_t* __restrict__ tab)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= entries) return;
    int l = 0;
    while (l + 1 < levels && start_offset(l + 1) <= i) ++l;
    const int64_t j = i - start_offset(l);
    tab[i] = start_cell_node(nodes, l, (uint32_t)(j & (((int64_t)1 << l) - 1)), (uint32_t)(j >> l));
}

}  // namespace

int build_tree(frw_tree* t, const std::vector<int4>& rects, int4 B, int leaf_cap, hipStream_t s)
{
    const int n = (int)rects.size();
    const int64_t W = (int64_t)B.z - B.x, H = (int64_t)B.w - B.y;
    int K = 0;
    while (K < 30 && ((int64_t)1 << K) <= std::max(W, H)) ++K;
    if (((int64_t)1 << K) <= std::max(W, H) || K > 28)
        return set_error(FRW_E_RANGE, "layout bounds wider than 2^28");
    K += 1;  // doubled centre coordinates take one more bit per axis
    t->morton_bits = K;

    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = check_hip(hipEventCreate(&e0), "hipEventCreate");
    if (rc == FRW_OK) rc = check_hip(hipEventCreate(&e1), "hipEventCreate");
    if (rc == FRW_OK) rc = check_hip(hipEventRecord(e0, s), "hipEventRecord");
    if (rc != FRW_OK) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        return rc;
    }
    int4* d_in = nullptr;
    uint64_t *d_keys = nullptr, *d_keys_sorted = nullptr;
    int32_t* d_vals = nullptr;
    DevBuf<uint32_t> hist, scan_sums;
    DevBuf<Frontier> fa, fb;
    DevBuf<int32_t> internal, group, own_end;
    DevBuf<int2> runs;
    DevBuf<Node> nodes;
    std::vector<Frontier> root(4);
    std::vector<int> level_start;  // nodes of depth d: [level_start[d], level_start[d + 1])
    int n_nodes = 4, nf = 4, depth = 0;
    const int nn = std::max(n, 1);

#define FRW_TRY(x)                                          \
    do {                                                    \
        if ((rc = check_hip((x), #x)) != FRW_OK) goto done; \
    } while (0)
#define FRW_TRYRC(x)                        \
    do {                                    \
        if ((rc = (x)) != FRW_OK) goto done; \
    } while (0)

    FRW_TRY(hipMalloc(&d_in, sizeof(int4) * nn));
    FRW_TRY(hipMalloc(&t->d_rects, sizeof(int4) * nn));
    FRW_TRY(hipMalloc(&t->d_ids, sizeof(int32_t) * nn));
    FRW_TRY(hipMalloc(&d_keys, sizeof(uint64_t) * nn));
    FRW_TRY(hipMalloc(&d_keys_sorted, sizeof(uint64_t) * nn));
    FRW_TRY(hipMalloc(&d_vals, sizeof(int32_t) * nn));
    if (n > 0) FRW_TRY(hipMemcpyAsync(d_in, rects.data(), sizeof(int4) * n, hipMemcpyHostToDevice, s));

    if (n > 0) {
        rect_keys<<<blocks_for(n, 256), 256, 0, s>>>(d_in, n, B, K, d_keys, d_vals);
        FRW_TRY(hipGetLastError());
        // key bits 6 .. 2K + 6: the Morton code (the low 6 bits hold the
        // same level tag in every key)
        const int64_t nb = (n + kSortTile - 1) / kSortTile;
        FRW_TRYRC(hist.reserve((size_t)(256 * nb), s));
        FRW_TRYRC(scan_sums.reserve((size_t)scan_tile_count(256 * nb), s));
        FRW_TRYRC(sort_pairs(d_keys, d_vals, d_keys_sorted, t->d_ids, n, 6, 2 * K + 6, hist.p, scan_sums.p, s));
        gather_rects<<<blocks_for(n, 256), 256, 0, s>>>(d_in, t->d_ids, n, t->d_rects);
        FRW_TRY(hipGetLastError());
    }

    // level 0: the root group (root + three empty quadrants)
    root[0] = Frontier{0ull, 0, 0, 0, n};
    for (int c = 1; c < 4; ++c) root[c] = Frontier{0ull, c, 0, 0, 0};
    FRW_TRYRC(fa.reserve(4, s));
    FRW_TRYRC(nodes.reserve(1024, s));
    FRW_TRYRC(runs.reserve(1024, s));
    FRW_TRY(hipMemcpyAsync(fa.p, root.data(), sizeof(Frontier) * 4, hipMemcpyHostToDevice, s));

    while (nf > 0) {
        level_start.push_back(n_nodes - nf);
        depth++;
        FRW_TRYRC(internal.reserve(nf, s));
        FRW_TRYRC(group.reserve(nf, s));
        FRW_TRYRC(own_end.reserve(nf, s));
        split_level<<<blocks_for(nf, 256), 256, 0, s>>>(d_keys_sorted, fa.p, nf, K, leaf_cap,
                                                        nodes.p, runs.p, internal.p, own_end.p);
        FRW_TRY(hipGetLastError());
        FRW_TRYRC(scan_sums.reserve((size_t)scan_tile_count(nf), s));
        FRW_TRYRC(exclusive_scan(reinterpret_cast<const uint32_t*>(internal.p), reinterpret_cast<uint32_t*>(group.p),
                                 nf, scan_sums.p, s));
        int last_g = 0, last_i = 0;
        FRW_TRY(hipMemcpyAsync(&last_g, group.p + nf - 1, sizeof(int), hipMemcpyDeviceToHost, s));
        FRW_TRY(hipMemcpyAsync(&last_i, internal.p + nf - 1, sizeof(int), hipMemcpyDeviceToHost, s));
        FRW_TRY(hipStreamSynchronize(s));
        const int n_int = last_g + last_i;
        if (n_int == 0) break;
        const int nf_next = 4 * n_int;
        FRW_TRYRC(nodes.reserve((size_t)n_nodes + nf_next, s, true, n_nodes));
        FRW_TRYRC(runs.reserve((size_t)n_nodes + nf_next, s, true, n_nodes));
        FRW_TRYRC(fb.reserve(nf_next, s));
        emit_children<<<blocks_for(nf, 256), 256, 0, s>>>(d_keys_sorted, fa.p, nf, K, internal.p,
                                                          group.p, own_end.p, n_nodes, nodes.p,
                                                          fb.p);
        FRW_TRY(hipGetLastError());
        n_nodes += nf_next;
        std::swap(fa, fb);
        nf = nf_next;
    }

    level_start.push_back(n_nodes);
    for (int d = (int)level_start.size() - 2; d >= 0; --d) {
        const int cnt = level_start[d + 1] - level_start[d];
        if (cnt <= 0) continue;
        node_boxes_level<<<blocks_for(cnt, 256), 256, 0, s>>>(t->d_rects, runs.p, level_start[d],
                                                               level_start[d + 1], nodes.p);
        FRW_TRY(hipGetLastError());
    }
    FRW_TRY(hipMalloc(&t->d_nodes, sizeof(Node) * n_nodes));
    FRW_TRY(hipMemcpyAsync(t->d_nodes, nodes.p, sizeof(Node) * n_nodes, hipMemcpyDeviceToDevice, s));
    {
        // start table of the nearest-rect search, levels 0 .. min(depth, kStartLevels) - 1
        const int levels = std::min(depth, kStartLevels);
        const int64_t entries = start_offset(levels);
        FRW_TRY(hipMalloc(&t->d_start, sizeof(int32_t) * entries));
        start_table<<<(unsigned)((entries + 255) / 256), 256, 0, s>>>(t->d_nodes, levels, entries, t->d_start);
        FRW_TRY(hipGetLastError());
        t->start_levels = levels;
    }
    FRW_TRY(hipEventRecord(e1, s));
    FRW_TRY(hipStreamSynchronize(s));
    {
        float ms = 0.f;
        FRW_TRY(hipEventElapsedTime(&ms, e0, e1));
        t->build_ms = ms;
    }
    t->n_nodes = n_nodes;
    t->depth = depth;

done:
#undef FRW_TRY
#undef FRW_TRYRC
    (void)hipStreamSynchronize(s);  // cleanup: the buffers below are freed either way
    if (d_in) (void)hipFree(d_in);
    if (d_keys) (void)hipFree(d_keys);
    if (d_keys_sorted) (void)hipFree(d_keys_sorted);
    if (d_vals) (void)hipFree(d_vals);
    hist.release();
    scan_sums.release();
    fa.release();
    fb.release();
    internal.release();
    group.release();
    own_end.release();
    runs.release();
    nodes.release();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

}  // namespace frw
