// Level-synchronous Morton quadtree build on gfx950.
//
// Replaces the reference's recursive dynamic-parallelism build
// (randomWalk/Cuda/createquadtree.cu:97-371, one block per node):
//   1. key per rect: Morton code of its centre, so each rect belongs to the
//      quadtree cell holding its centre at every level (loose quadtree).  The
//      key layout (code << 6 | level) also admits rects that stop at an inner
//      cell; with centre keys every rect sits at the deepest level.
//   2. one device radix sort of (key, rect id).  Sorted by key, the rects of
//      any cell form one contiguous run.
//   3. level by level: every frontier node splits its run into its four
//      quadrant runs by binary search over the sorted keys; a node becomes a
//      leaf when its run holds <= leaf_cap rects (or it is a unit cell).
//      Child groups are allocated by an exclusive scan, so the flat array is
//      in breadth-first order and deterministic.
//   4. each node's box = tight bounding box of its run (block reduction); the
//      boxes of siblings may overlap.
// Query answers depend only on the rect set, never on this structure.
#include <hipcub/hipcub.hpp>

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

// One block per node: tight bounding box of its run.
__global__ void __launch_bounds__(256) node_boxes(const int4* __restrict__ rects,
                                                  const int2* __restrict__ runs, int n_nodes,
                                                  Node* __restrict__ nodes)
{
    const int ni = blockIdx.x;
    if (ni >= n_nodes) return;
    const int2 run = runs[ni];
    int x0 = INT_MAX, y0 = INT_MAX, x1 = INT_MIN, y1 = INT_MIN;
    for (int r = run.x + (int)threadIdx.x; r < run.y; r += blockDim.x) {
        const int4 q = rects[r];
        x0 = min(x0, q.x); y0 = min(y0, q.y); x1 = max(x1, q.z); y1 = max(y1, q.w);
    }
    for (int off = 32; off >= 1; off >>= 1) {
        x0 = min(x0, __shfl_xor(x0, off));
        y0 = min(y0, __shfl_xor(y0, off));
        x1 = max(x1, __shfl_xor(x1, off));
        y1 = max(y1, __shfl_xor(y1, off));
    }
    __shared__ int4 part[4];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) part[wave] = make_int4(x0, y0, x1, y1);
    __syncthreads();
    if (threadIdx.x == 0) {
        int4 b = part[0];
        for (int w = 1; w < (int)(blockDim.x / 64); ++w) {
            b.x = min(b.x, part[w].x); b.y = min(b.y, part[w].y);
            b.z = max(b.z, part[w].z); b.w = max(b.w, part[w].w);
        }
        nodes[ni].box = b;
    }
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
    void* d_tmp = nullptr;
    size_t tmp_bytes = 0;
    DevBuf<Frontier> fa, fb;
    DevBuf<int32_t> internal, group, own_end;
    DevBuf<int2> runs;
    DevBuf<Node> nodes;
    std::vector<Frontier> root(4);
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
        const int end_bit = 2 * K + 6;
        FRW_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, d_keys, d_keys_sorted, d_vals,
                                                   t->d_ids, n, 0, end_bit, s));
        FRW_TRY(hipMalloc(&d_tmp, std::max<size_t>(tmp_bytes, 16)));
        FRW_TRY(hipcub::DeviceRadixSort::SortPairs(d_tmp, tmp_bytes, d_keys, d_keys_sorted, d_vals,
                                                   t->d_ids, n, 0, end_bit, s));
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
        depth++;
        FRW_TRYRC(internal.reserve(nf, s));
        FRW_TRYRC(group.reserve(nf, s));
        FRW_TRYRC(own_end.reserve(nf, s));
        split_level<<<blocks_for(nf, 256), 256, 0, s>>>(d_keys_sorted, fa.p, nf, K, leaf_cap,
                                                        nodes.p, runs.p, internal.p, own_end.p);
        FRW_TRY(hipGetLastError());
        size_t need = 0;
        FRW_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, need, internal.p, group.p, nf, s));
        if (need > tmp_bytes || !d_tmp) {
            if (d_tmp) {
                FRW_TRY(hipStreamSynchronize(s));
                (void)hipFree(d_tmp);
                d_tmp = nullptr;
            }
            tmp_bytes = std::max<size_t>(need, 16);
            FRW_TRY(hipMalloc(&d_tmp, tmp_bytes));
        }
        FRW_TRY(hipcub::DeviceScan::ExclusiveSum(d_tmp, need, internal.p, group.p, nf, s));
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

    node_boxes<<<n_nodes, 256, 0, s>>>(t->d_rects, runs.p, n_nodes, nodes.p);
    FRW_TRY(hipGetLastError());
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
    if (d_tmp) (void)hipFree(d_tmp);
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
