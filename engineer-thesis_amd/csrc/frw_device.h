// Device-side building blocks of the FRW hot path for gfx950 (CDNA4).
//
// Geometry conventions (reference randomWalk/CPU/rectHost.h:6-47): a rect is
// TL=(x0,y0), BR=(x1,y1), y grows downward.  Rects and tree-node boxes are
// int32 (16-byte records); walker positions and squares are FP64 so that every
// hop reproduces the reference's FP64 arithmetic operation by operation (an
// FP32 position has a 1.0 ulp at 10^7 and cannot hold the reference's p+-1 /
// 0.095 steps).  All rect tests compare an int32 (converted exactly) against
// an FP64 coordinate, or take the sign of a single rounded difference, so they
// are exact.  Every translation unit including this header is compiled with
// -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "frw_types.h"

namespace frw {

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon et al., SC'11).  Walk stream: counter = (walk id's
// low word, job, hop/4, 'WALK' ^ walk id's high word), key = seed; word
// hop%4 is the hop's uniform.  Walk ids go up to 2^48 (kWalkIdLimit); below
// 2^32 the domain word is 'WALK' itself.
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

constexpr uint32_t kWalkDomain = 0x57414C4Bu;  // 'WALK'
constexpr int64_t kWalkIdLimit = (int64_t)1 << 48;  // walk ids per (seed, job)

// Counter of walk id `id` in a stream whose domain word is `domain` ('WALK',
// or 'GAUS' for weighted starts: the two stay apart for ids below 2^48,
// since only the low 16 bits of the domain word take the id's high word).
__host__ __device__ inline void walk_counter(uint32_t c[4], uint64_t id, uint32_t job, uint32_t word2, uint32_t domain)
{
    c[0] = (uint32_t)id;
    c[1] = job;
    c[2] = word2;
    c[3] = domain ^ (uint32_t)(id >> 32);
}

// ---------------------------------------------------------------- sampling
// Panel index = largest i with T[i] <= u (T[0] = 0): an 8-step branch-free
// binary search over the 200 thresholds of the Green CDF.
__host__ __device__ inline int sample_panel(const uint32_t* T, uint32_t u)
{
    int lo = 0;
#pragma unroll
    for (int step = 128; step >= 1; step >>= 1) {
        const int probe = lo + step;
        if (probe < kNSample && T[probe] <= u) lo = probe;
    }
    return lo;
}

// Reference getPointFromNindex (randomWalk/CPU/rectHost.cpp:102-161): same
// branches, same FP64 operation order, integer-truncated perimeter and the
// int cast on the top side.
__host__ __device__ inline void point_from_panel(double sx0, double sy0, double sx1, double sy1,
                                                 int index, double& px, double& py)
{
    const double perimeter = (double)(int)(2 * (sx1 - sx0) + 2 * (sy1 - sy0));
    const double v = perimeter / (double)kNSample;
    const double h = sy1 - sy0;
    const double w = sx1 - sx0;
    const double a = index * v;
    const double b = (index - 1) * v;
    if (a < w) {
        px = (double)(int)(b + v / 2 + sx0);
        py = sy0;
    } else if (b < w && a > w) {
        px = sx1;
        py = sy0;
    } else if (a < (w + h)) {
        px = sx1;
        py = (b > w) ? (sy0 + b - w + v / 2) : (sy0 + a - w + (a - w) / 2);
    } else if (b < (w + h) && a > (w + h)) {
        px = sx1;
        py = sy1;
    } else if (a < (2 * w + h)) {
        py = sy1;
        px = (b > w + h) ? (sx1 - (b - w - h + v / 2)) : (sx1 - (a - w - h + (a - w - h) / 2));
    } else if (b < (2 * w + h) && a > (2 * w + h)) {
        px = sx0;
        py = sy1;
    } else {
        px = sx0;
        py = (b > 2 * w + h) ? (sy1 - (b - 2 * w - h + v / 2))
                             : (sy1 - (a - 2 * w - h + (a - 2 * w - h) / 2));
    }
}

// ---------------------------------------------------------------- geometry
struct DBox {
    double x0, y0, x1, y1;
};

// Tree::isInBounds(point), strict (randomWalk/CPU/quadTree.cpp:49-59).
__host__ __device__ inline bool in_bounds_strict(const int4& B, double px, double py)
{
    return px > (double)B.x && py > (double)B.y && px < (double)B.z && py < (double)B.w;
}

// Tree::isInBounds(rect), inclusive (randomWalk/CPU/quadTree.cpp:38-47).
__host__ __device__ inline bool in_bounds_box(const int4& B, double x0, double y0, double x1, double y1)
{
    return y0 >= (double)B.y && x0 >= (double)B.x && y1 <= (double)B.w && x1 <= (double)B.z;
}

// rectsCollision (randomWalk/CPU/rectHost.cpp:43-52): closed, touching collides.
__host__ __device__ inline bool box_hits_rect(const int4& r, double x0, double y0, double x1, double y1)
{
    return (double)r.z >= x0 && x1 >= (double)r.x && (double)r.w >= y0 && y1 >= (double)r.y;
}

// Signed L-infinity distance from p to a rect or node box; <= 0 exactly when
// the closed box contains p.
__host__ __device__ inline double linf(const int4& r, double px, double py)
{
    const double dx = fmax((double)r.x - px, px - (double)r.z);
    const double dy = fmax((double)r.y - py, py - (double)r.w);
    return fmax(dx, dy);
}

// Saturating int32 subtraction (v_sub_i32 with clamp on the device).
__host__ __device__ inline int sat_sub(int a, int b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_elementwise_sub_sat(a, b);
#else
    const long long r = (long long)a - (long long)b;
    return r > 2147483647LL ? 2147483647 : (r < -2147483648LL ? (int)-2147483648LL : (int)r);
#endif
}

// Integer lower bound of linf(box, p) from fx = floor(px), cx = ceil(px)
// (and fy, cy), saturated to int32: x0 - cx <= x0 - px and fx - x1 <= px - x1,
// so the bound never exceeds the exact distance; a dummy (inverted) box gets
// a huge bound.
__host__ __device__ inline int linf_lb(const int4& b, int fx, int cx, int fy, int cy)
{
    const int dx = max(sat_sub(b.x, cx), sat_sub(fx, b.z));
    const int dy = max(sat_sub(b.y, cy), sat_sub(fy, b.w));
    return max(dx, dy);
}

// Smallest int32 >= d (saturated), for comparisons of integer bounds with d.
__host__ __device__ inline int ceil_int(double d)
{
    return d >= 2147483647.0 ? 2147483647 : (d <= -2147483648.0 ? (int)-2147483648LL : (int)ceil(d));
}

__host__ __device__ inline int floor_int(double d)
{
    return d >= 2147483647.0 ? 2147483647 : (d <= -2147483648.0 ? (int)-2147483648LL : (int)floor(d));
}

// Rect equality by value (the reference compares RectHost by value when it
// ignores the master, randomWalk/CPU/quadTree.cpp:219-226).
__host__ __device__ inline bool same_rect(const int4& a, const int4& b)
{
    return a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w;
}

// Error bits: an atomic OR on the device; a plain OR in a host build of these
// functions (used by the CPU-side tests of this header).
__host__ __device__ inline void raise_error(int* err, int bit)
{
#if defined(__HIP_DEVICE_COMPILE__)
    atomicOr(err, bit);
#else
    *err |= bit;
#endif
}

// True when the predicate holds on any active lane of the wave.  A host
// build answers for the one lane it runs, or through FRW_HOST_ANY_LANE when
// the includer defines it (a test harness that lets a loop run on after the
// lane is done, as the other lanes of a wave would).
__host__ __device__ inline bool any_lane(bool v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __ballot(v) != 0ull;
#elif defined(FRW_HOST_ANY_LANE)
    return FRW_HOST_ANY_LANE(v);
#else
    return v;
#endif
}

// ---------------------------------------------------------------- tree view
// Loads from the tree's arrays as global-memory loads: the arrays live in
// device memory, and a global load (unlike the generic "flat" load the
// compiler emits for a pointer it loaded from memory) waits on the vector
// memory counter only, not on the LDS traversal stack's.  The pointer goes
// through an integer: a plain cast to the global address space is folded
// back into the generic form.
template <class T>
__host__ __device__ inline T gload(const T* p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(1))) T*)(uintptr_t)p;
#else
    return *p;
#endif
}

struct DevTree {
    const int4* rects;     // sorted by Morton key, 16 B each
    const int32_t* ids;    // original index of each sorted rect
    const Node* nodes;     // groups of 4 siblings, 128 B aligned
    int4 bounds;           // absorbing root bounds
    int32_t n_rects;
    int32_t n_nodes;
    int32_t depth;         // number of levels
    int32_t stack_cap;     // per-lane traversal stack entries (of the stack layout in use)
    double margin;         // ambiguity band of the Q3 schedule emulation
    double dist0;          // 0.05 * max(W, H), first Q3 step
    // start table of the nearest-rect search (start_levels = 0: none): for
    // Morton cell levels 0..start_levels-1, level l at entries
    // [(4^l - 1) / 3, (4^(l+1) - 1) / 3), row-major (y, x), the node whose
    // subtree holds every rect centred in the cell, -1 when none is
    const int32_t* start_tab;
    int32_t start_levels;
    int32_t morton_bits;   // K: doubled centre coordinates have K bits per axis
    int32_t ext2_max;      // largest rect width or height (doubled half-extent)
};

__host__ __device__ inline int64_t start_offset(int l) { return (((int64_t)1 << (2 * l)) - 1) / 3; }

// The node a start-table entry holds: descend from the root along the cell's
// quadrants (child c = x bit | y bit << 1) until level l or a leaf; -1 when
// that node holds no rect (an empty quadrant).
__host__ __device__ inline int32_t start_cell_node(const Node* nodes, int l, uint32_t cx, uint32_t cy)
{
    int ni = 0;
    for (int lv = 0; lv < l; ++lv) {
        const Node nd = nodes[ni];
        if (nd.first_child < 0) break;
        const int b = l - 1 - lv;
        ni = nd.first_child + (int)(((cx >> b) & 1u) | (((cy >> b) & 1u) << 1));
    }
    const Node nd = nodes[ni];
    return (nd.first_child < 0 && nd.rbeg == nd.rend) ? -1 : ni;
}

// Optional per-lane traffic counters: node records (32 B) and rect records
// (16 B) read by the traversals.
struct Counters {
    uint32_t nodes;
    uint32_t rects;
    uint32_t tabs;   // start-table entries (4 B)
};

// Per-lane traversal stacks in LDS: slot i of a lane at base[i * stride].
// Two layouts with one interface and one visiting order (a node's children:
// `lead` first when pushed, then the others highest index first):
//  * PlainStack holds one node per entry: 3 depth + 2 entries for a tree of
//    `depth` levels (three pending siblings per level, plus the start nodes);
//  * RunStack holds one run of pending siblings per entry, packed as
//    (first sibling << 6) | (lead << 4) | (mask of siblings left): depth + 6
//    entries, for trees of fewer than kMaxTreeNodes nodes, at a few more
//    instructions per pop.
// The walk kernel takes RunStack where PlainStack's LDS would limit the
// resident waves.  Built with FRW_STACK_CHECK (host builds), every slot
// access is checked against cap and an overflow aborts.
constexpr int kMaxTreeNodes = 1 << 25;

#ifdef FRW_STACK_CHECK
#define FRW_STACK_FIELDS                                             \
    int32_t* base;                                                   \
    int stride;                                                      \
    int cap;                                                         \
    __host__ __device__ int32_t& at(int slot) const                  \
    {                                                                \
        if (slot < 0 || slot >= cap) __builtin_trap();               \
        return base[slot * stride];                                  \
    }
#else
#define FRW_STACK_FIELDS                                             \
    int32_t* base;                                                   \
    int stride;                                                      \
    __host__ __device__ int32_t& at(int slot) const { return base[slot * stride]; }
#endif

struct PlainStack {
    FRW_STACK_FIELDS
    __host__ __device__ static constexpr int entries(int depth) { return 3 * depth + 2; }
    __host__ __device__ void push_node(int& sp, int node) const { at(sp++) = node; }
    __host__ __device__ void push_children(int& sp, int first_child, uint32_t mask, int lead) const
    {
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (c != lead && ((mask >> c) & 1u)) at(sp++) = first_child + c;
        if ((mask >> lead) & 1u) at(sp++) = first_child + lead;
    }
    __host__ __device__ int pop(int& sp) const { return at(--sp); }
    // children offered one at a time (first_child + c in the order 1, 2, 3,
    // 0: popped 0, 3, 2, 1), then flushed
    __host__ __device__ void offer(int& sp, uint32_t&, int first_child, int c) const { at(sp++) = first_child + c; }
    __host__ __device__ void flush(int&, uint32_t, int) const {}
};

struct RunStack {
    FRW_STACK_FIELDS
    __host__ __device__ static constexpr int entries(int depth) { return depth + 6; }
    __host__ __device__ void push_node(int& sp, int node) const { at(sp++) = (int32_t)(((uint32_t)node << 6) | 1u); }
    __host__ __device__ void push_children(int& sp, int first_child, uint32_t mask, int lead) const
    {
        if (mask) at(sp++) = (int32_t)(((uint32_t)first_child << 6) | ((uint32_t)lead << 4) | mask);
    }
    __host__ __device__ int pop(int& sp) const
    {
        const uint32_t e = (uint32_t)at(sp - 1);
        const uint32_t m = e & 15u, f = (e >> 4) & 3u;
        const uint32_t c = ((m >> f) & 1u) ? f : 31u - (uint32_t)__builtin_clz(m);
        const uint32_t rest = m ^ (1u << c);
        at(sp - 1) = (int32_t)(e ^ (1u << c));  // dead once the run is spent
        sp -= rest == 0u ? 1 : 0;
        return (int)((e >> 6) + c);
    }
    __host__ __device__ void offer(int&, uint32_t& run, int, int c) const { run |= 1u << c; }
    __host__ __device__ void flush(int& sp, uint32_t run, int first_child) const { push_children(sp, first_child, run, 0); }
};
#undef FRW_STACK_FIELDS

using LaneStack = PlainStack;

// Q1 (checkCollisions, randomWalk/CPU/quadTree.cpp:148-198 with the full
// depth-first search of randomWalk/Cuda/d_quadTree.cu:136-193): is the closed
// box outside the root bounds, or does it intersect any rect (other than
// `ignore`, by value)?
template <bool kIgnore, class Stack>
__host__ __device__ inline bool q1_box(const DevTree& t, const Stack& st, double x0, double y0,
                              double x1, double y1, int4 ignore, Counters* cnt)
{
    if (!in_bounds_box(t.bounds, x0, y0, x1, y1)) return true;
    int sp = 0;
    st.push_node(sp, 0);
    while (sp > 0) {
        const int ni = st.pop(sp);
        const Node nd = gload(t.nodes + ni);
        if (cnt) cnt->nodes += nd.first_child >= 0 ? 5u : 1u;
        for (int r = nd.rbeg; r < nd.rend; ++r) {
            const int4 rc = gload(t.rects + r);
            if (cnt) cnt->rects++;
            if (box_hits_rect(rc, x0, y0, x1, y1) && !(kIgnore && same_rect(rc, ignore)))
                return true;
        }
        if (nd.first_child >= 0) {
            uint32_t run = 0;
#pragma unroll
            for (int k = 1; k <= 4; ++k) {
                const int c = k & 3;
                const int4 cb = gload(&t.nodes[nd.first_child + c].box);
                if (box_hits_rect(cb, x0, y0, x1, y1)) st.offer(sp, run, nd.first_child, c);
            }
            st.flush(sp, run, nd.first_child);
        }
    }
    return false;
}

// Nearest rect in the L-infinity metric: min over rects of linf(rect, p),
// by branch and bound over the node boxes (children visited nearest first).
// The minimum equals a brute-force scan bit for bit: the same per-rect
// expression is evaluated, and a node is skipped only when an integer lower
// bound of its box distance cannot beat the best.
//
// `hint` (a sorted rect index, or -1) seeds the bound with that rect's
// distance, typically the rect nearest to the walker's previous point; *arg
// (optional) receives the sorted index of a nearest rect.  With hint_rect
// the hint's coordinates come from (and the nearest rect's go back to) the
// caller's registers instead of a load.
template <class Stack>
__host__ __device__ inline double nearest_rect(const DevTree& t, const Stack& st, double px, double py,
                                      Counters* cnt, int hint = -1, int* arg = nullptr, int4* hint_rect = nullptr)
{
    // node pruning and ordering use integer lower bounds (an integer lb is
    // below best exactly when it is below ceil(best)); rect distances are exact.
    // When p lies in a rect (best <= 0) the search also settles ties the way
    // the point query does (deepest containment, then lowest original id), so
    // the arg is the point query's answer: nodes with lb == ceil(best) are
    // still visited then (lim = ceil(best) + 1) and an equal distance with a
    // lower original id replaces the arg.
    const int fx = floor_int(px), cx = ceil_int(px), fy = floor_int(py), cy = ceil_int(py);
    double best = __builtin_inf();
    int lim = 2147483647;
    int besti = -1;
    int4 bestr = make_int4(0, 0, 0, 0);
    if (hint >= 0) {
        if (hint_rect) {
            bestr = *hint_rect;
        } else {
            bestr = gload(t.rects + hint);
            if (cnt) cnt->rects++;
        }
        best = linf(bestr, px, py);
        lim = ceil_int(best) + (best <= 0.0 ? 1 : 0);
        besti = hint;
    }
    int sp = 0;
    if (hint >= 0 && t.start_levels > 0) {
        // every rect within best of p has its doubled centre within
        // w = 2 best + ext2_max of 2 (p - corner) on each axis, so the
        // subtrees of the (at most 4) cells of that window at the deepest
        // level where it spans at most 2 cells per axis hold all of them
        // (doubled coordinates are below 2^morton_bits <= 2^31: int32 cells;
        // the clamped window edges are integer-valued doubles, so the
        // conversion is exact)
        const double w = 2.0 * best + (double)t.ext2_max;
        const double ux = 2.0 * (px - (double)t.bounds.x), uy = 2.0 * (py - (double)t.bounds.y);
        const double kmax = (double)(((int64_t)1 << t.morton_bits) - 1);
        auto clampk = [&](double v) -> int { return (int)fmin(fmax(v, 0.0), kmax); };
        const int x0 = clampk(floor(ux - w) - 1.0), x1 = clampk(ceil(ux + w) + 1.0);
        const int y0 = clampk(floor(uy - w) - 1.0), y1 = clampk(ceil(uy + w) + 1.0);
        // the fewest bits s with 2^s cells covering the window's span
        // max(x1 - x0, y1 - y0) + 1: the bit length of the larger difference
        const unsigned m = (unsigned)max(x1 - x0, y1 - y0);
        const int s = m ? 32 - __builtin_clz(m) : 0;
        const int l = min(max(t.morton_bits - s, 0), t.start_levels - 1);
        const int shift = t.morton_bits - l;
        // level l's entries start at (4^l - 1) / 3 = binary 0101...01 (l ones)
        const int32_t* tab = t.start_tab + (l ? (0x55555555u >> (32 - 2 * l)) : 0u);
        // the window covers at most 2 x 2 cells (cx1 - cx0 <= 1, cy1 - cy0 <= 1)
        // and p's own cell is one of its corners.  The four corner entries are
        // loaded at once (a repeated corner re-reads the same entry) and kept
        // in registers: no per-lane array, which would live in scratch memory.
        const int cx0 = x0 >> shift, cx1 = x1 >> shift;
        const int cy0 = y0 >> shift, cy1 = y1 >> shift;
        const bool ox = (clampk(floor(ux)) >> shift) != cx0, oy = (clampk(floor(uy)) >> shift) != cy0;
        const int32_t e0 = gload(tab + ((cy0 << l) + cx0)), e1 = gload(tab + ((cy0 << l) + cx1));
        const int32_t e2 = gload(tab + ((cy1 << l) + cx0)), e3 = gload(tab + ((cy1 << l) + cx1));
        if (cnt) cnt->tabs += (uint32_t)((cx1 != cx0 ? 2 : 1) * (cy1 != cy0 ? 2 : 1));
        const int32_t own = oy ? (ox ? e3 : e2) : (ox ? e1 : e0);
        // the other cells' nodes in corner order, each once (a repeated corner
        // equals an earlier entry, p's cell equals own), then p's own cell,
        // which is popped first
        if (e0 >= 0 && e0 != own) st.push_node(sp, e0);
        if (e1 >= 0 && e1 != own && e1 != e0) st.push_node(sp, e1);
        if (e2 >= 0 && e2 != own && e2 != e0 && e2 != e1) st.push_node(sp, e2);
        if (e3 >= 0 && e3 != own && e3 != e0 && e3 != e1 && e3 != e2) st.push_node(sp, e3);
        if (own >= 0) st.push_node(sp, own);  // p's cell first
    } else {
        st.push_node(sp, 0);
    }
    while (sp > 0) {
        const int ni = st.pop(sp);
        const Node nd = gload(t.nodes + ni);
        if (cnt) cnt->nodes++;
        if (!(linf_lb(nd.box, fx, cx, fy, cy) < lim)) continue;
        if (cnt) cnt->rects += (uint32_t)(nd.rend - nd.rbeg);
        if (cnt && nd.first_child >= 0) cnt->nodes += 4u;
        // the leaf's rects two at a time: two loads in flight, then the
        // distances in order (an odd tail re-reads the last rect and skips
        // it); pairs rather than fours keep 8 fewer registers per lane, which
        // measured faster at 5 waves per SIMD
        for (int r0 = nd.rbeg; r0 < nd.rend; r0 += 2) {
            int4 q[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) q[u] = gload(t.rects + (r0 + u < nd.rend ? r0 + u : nd.rend - 1));
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int r = r0 + u;
                if (r >= nd.rend) break;
                const double d = linf(q[u], px, py);
                if (d < best || (d == best && d <= 0.0 && gload(t.ids + r) < gload(t.ids + besti))) {
                    best = d;
                    lim = ceil_int(d) + (d <= 0.0 ? 1 : 0);
                    besti = r;
                    bestr = q[u];
                }
            }
        }
        if (nd.first_child >= 0) {
            int lb[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) lb[c] = linf_lb(gload(&t.nodes[nd.first_child + c].box), fx, cx, fy, cy);
            // the nearest child first, then the others
            int near = 0;
            uint32_t mask = 0;
#pragma unroll
            for (int c = 1; c < 4; ++c)
                if (lb[c] < lb[near]) near = c;
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (lb[c] < lim) mask |= 1u << c;
            st.push_children(sp, nd.first_child, mask, near);
        }
    }
    if (arg) *arg = besti;
    if (hint_rect) *hint_rect = bestr;
    return best;
}

// Q2 (checkCollisons(point), randomWalk/CPU/quadTree.cpp:240-262, answered
// exactly): original index of the rect containing p (closed), deepest
// containment first, then lowest index; -1 if none.
template <class Stack>
__host__ __device__ inline int q2_point(const DevTree& t, const Stack& st, double px, double py,
                               Counters* cnt)
{
    double best = 0.0;
    int best_id = -1;
    int sp = 0;
    st.push_node(sp, 0);
    while (sp > 0) {
        const int ni = st.pop(sp);
        const Node nd = gload(t.nodes + ni);
        if (cnt) {
            cnt->nodes += nd.first_child >= 0 ? 5u : 1u;
            cnt->rects += (uint32_t)(nd.rend - nd.rbeg);
        }
        for (int r = nd.rbeg; r < nd.rend; ++r) {
            const double d = linf(gload(t.rects + r), px, py);
            if (d <= 0.0) {
                const int id = gload(t.ids + r);
                if (best_id < 0 || d < best || (d == best && id < best_id)) {
                    best = d;
                    best_id = id;
                }
            }
        }
        if (nd.first_child >= 0) {
            uint32_t run = 0;
#pragma unroll
            for (int k = 1; k <= 4; ++k) {
                const int c = k & 3;
                if (linf(gload(&t.nodes[nd.first_child + c].box), px, py) <= 0.0)
                    st.offer(sp, run, nd.first_child, c);
            }
            st.flush(sp, run, nd.first_child);
        }
    }
    return best_id;
}

// The decided steps of the square search below in closed form.  While every
// test is decided by h alone (h >= h_hit collides, h < h_free is free), the
// search from h = 1 grows by dist0 until its first collision, at h_c = 1 +
// n1 dist0, then runs a non-restoring binary search with steps dist0 / 2,
// dist0 / 4, ... down to delta = dist0 / 2^m (the first at most MIN_DIST):
// h stays on the lattice b + j delta (b = h_c - dist0), each step keeps
// h_hit in (h - step, h + step], so after the m halvings h is the odd
// lattice point b + (2q + 1) delta within delta of h_hit (the search's
// directions are the bits of q, most significant first: 1 grows, 0
// shrinks), and one more shrink by delta follows if h still collides: the
// search ends at L, the largest lattice point below h_hit.  The loops track
// h in FP64 (its running sum may round, by at most 30 ulp of h_c over 60
// steps), so their decisions are these wherever every visited lattice point
// lies farther than eps = 64 ulp(h_c) from the band [h_free, h_hit): the
// two points next to h_hit (L and L + delta) are checked, every other one
// is delta or more away.  Otherwise a test is ambiguous, or too close to
// call, and the caller's loops run.  The decided loop stops at 60 steps:
// longer searches are left to it too.
// 1 / d to within a few ulp (the quotients below only seed searches that
// check their answer): the hardware reciprocal estimate refined by one
// Newton step on the device, a division on the host.
__host__ __device__ inline double approx_rcp(double d)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double r = __builtin_amdgcn_rcp(d);
    return fma(fma(-d, r, 1.0), r, r);
#else
    return 1.0 / d;
#endif
}

struct SquarePath {
    double L;      // final half-size
    double d0;     // first step
    double hc;     // largest half-size of the search
    int n1;        // growth steps
    int m;         // halvings
    int shrink;    // 1: a last shrink by delta
    int exact;     // 1: every h of the search is exact (delta's lowest set bit 2^eL <= 1, h_c < 2^(53 + eL))
    int eL;
    uint32_t q;    // directions of the halving steps
};

__host__ __device__ inline bool square_path(double d0, double min_dist, double h_hit, double h_free, SquarePath* sp)
{
    if (!(d0 > min_dist) || !(h_hit > 1.0) || !(h_free <= h_hit)) return false;
    int e0 = 0, em = 0;
    const double f0 = frexp(d0, &e0), fm = frexp(min_dist, &em);  // fractions in [0.5, 1)
    const int m = e0 - em + (f0 > fm ? 1 : 0);                         // halvings: d0 / 2^m <= min_dist < d0 / 2^(m-1)
    if (m < 1 || m > 31) return false;
    const double delta = ldexp(d0, -m);
    // n1: the first k >= 0 with 1 + k d0 >= h_hit
    const double inv = approx_rcp(d0);
    double k = ceil((h_hit - 1.0) * inv);
    if (!(k >= 1.0 && k < 64.0)) return false;
    if (k > 1.0 && fma(k - 1.0, d0, 1.0) >= h_hit) k -= 1.0;
    if (fma(k, d0, 1.0) < h_hit) k += 1.0;
    if (k + (double)m + 1.0 > 59.0) return false;
    const double hc = fma(k, d0, 1.0), b = fma(k - 1.0, d0, 1.0);
    int ehc = 0;
    (void)frexp(hc, &ehc);
    const double eps = ldexp(1.0, ehc - 53 + 6);  // 64 ulp(h_c)
    if (!(hc >= h_hit + eps)) return false;
    // L: the largest lattice point below h_hit, clear of the band by eps; the next one clear of h_hit
    double j = ceil((h_hit - b) * ldexp(inv, m)) - 1.0;
    if (!(j >= 0.0)) j = 0.0;
    if (fma(j, delta, b) >= h_hit) j -= 1.0;
    if (fma(j + 1.0, delta, b) < h_hit) j += 1.0;
    if (!(j >= 0.0 && j < ldexp(1.0, m))) return false;
    const double L = fma(j, delta, b);
    if (!(L < h_free - eps && fma(j + 1.0, delta, b) >= h_hit + eps)) return false;
    const uint64_t mant = __builtin_bit_cast(uint64_t, d0) & ((1ull << 52) - 1ull);
    const int tz = mant ? __builtin_ctzll(mant) : 52;
    const int eL = (e0 - 1) - 52 + tz - m;  // delta's lowest set bit
    const uint32_t ji = (uint32_t)j;
    const uint32_t odd = (ji & 1u) ? ji : ji + 1u;  // the halving steps' last point, b + odd delta
    sp->L = L;
    sp->d0 = d0;
    sp->hc = hc;
    sp->n1 = (int)k;
    sp->m = m;
    sp->shrink = (ji & 1u) ? 0 : 1;
    sp->exact = (eL <= 0 && hc < ldexp(1.0, 53 + eL)) ? 1 : 0;
    sp->eL = eL;
    sp->q = (odd - 1u) >> 1;
    return true;
}

// True when the halving steps leave a box coordinate exact: after the
// growth steps it is z (the loops' own value), and every later value is z
// plus a multiple of 2^eL (h exact on the lattice), between z and zb, its
// value at b = h_c - dist0, so a multiple of g0 = min(2^lz, 2^eL) (2^lz:
// z's lowest set bit) no larger in magnitude than max(|z|, |zb|): exact
// when that stays below 2^(53 + g0's exponent) (zb computed in FP64:
// rounding is monotone and the bound a power of two, so the test on the
// rounded value is safe).
__host__ __device__ inline bool halving_exact(double z, double zb, int eL)
{
    int g = eL;
    if (z != 0.0) {
        int ez = 0;
        (void)frexp(z, &ez);  // |z| in [2^(ez-1), 2^ez): ulp 2^(ez - 53)
        const uint64_t sig = (__builtin_bit_cast(uint64_t, z) & ((1ull << 52) - 1ull)) | (1ull << 52);
        const int lz = ez - 53 + __builtin_ctzll(sig);
        g = lz < eL ? lz : eL;
    }
    return fmax(fabs(z), fabs(zb)) < ldexp(1.0, 53 + g);
}

template <bool kLiteral, class Stack>
__host__ __device__ inline DBox biggest_square(const DevTree& t, const Stack& st, double px, double py,
                                      double dstar, uint32_t* n_q1, Counters* cnt, int* err,
                                      const int4* near = nullptr)
{
    const double MIN_DIST = (double)0.095f;
    // h tracks the nominal half-size (1, then +-dist per step).  After n < 60
    // steps every computed half-extent of the box is within (2n + 2) 2^(k-53)
    // of h (one rounding per coordinate update, per h update and per
    // difference; coordinates and h below 2^k), so a decision taken on h
    // with the band 8 m = 2^(k-45) > m + that error is the decision of the
    // exact tests below; inside the band the exact tests run.  The box meets
    // a rect when its half-extent reaches d* and leaves the bounds when it
    // exceeds the distance dB from p to the bounds, so one pair of thresholds
    // on min(d*, dB) decides both tests.  (the thresholds' own rounding,
    // below 2^(k-53), fits in the band's slack)
    const double dB = fmin(fmin(px - (double)t.bounds.x, (double)t.bounds.z - px),
                           fmin(py - (double)t.bounds.y, (double)t.bounds.w - py));
    const double dmin = fmin(dstar, dB);
    const double h_hit = dmin + 8.0 * t.margin, h_free = dmin - 8.0 * t.margin;
    double h = 1.0;
    int step = 0;
    auto collides = [&](double a, double c, double b, double e) -> bool {
        if (!kLiteral && step < 60) {
            if (h >= h_hit) return true;
            if (h < h_free) return false;
        }
        if (!in_bounds_box(t.bounds, a, c, b, e)) return true;
        if (!kLiteral) {
            if (near && box_hits_rect(*near, a, c, b, e)) return true;
            const double lo = fmin(fmin(px - a, b - px), fmin(py - c, e - py));
            const double hi = fmax(fmax(px - a, b - px), fmax(py - c, e - py));
            if (lo >= dstar + t.margin) return true;
            if (hi < dstar - t.margin) return false;
        }
        if (n_q1) (*n_q1)++;
        return q1_box<false>(t, st, a, c, b, e, make_int4(0, 0, 0, 0), cnt);
    };
    DBox o = {px - 1, py - 1, px + 1, py + 1};
    if (collides(o.x0, o.y0, o.x1, o.y1)) return o;
    const DBox init = {px - 2, py - 2, px + 2, py + 2};
    h = 2.0;
    if (collides(init.x0, init.y0, init.x1, init.y1)) return init;
    h = 1.0;
    if (!kLiteral) {
        SquarePath sp;
        if (square_path(t.dist0, MIN_DIST, h_hit, h_free, &sp)) {
#if !defined(__HIP_DEVICE_COMPILE__) && defined(FRW_HOST_CLOSED_FORM_HOOK)
            FRW_HOST_CLOSED_FORM_HOOK();
#endif
            // the growth steps as the loops add them (where a sum rounds, it
            // rounds here alike), then the halving steps: exact where every
            // value they visit is exact (halving_exact), the end z -/+ (h_c - L)
            DBox z = o;
            for (int i = 0; i < sp.n1; ++i) {
                z.x0 -= sp.d0;
                z.y0 -= sp.d0;
                z.x1 += sp.d0;
                z.y1 += sp.d0;
            }
            if (sp.exact && halving_exact(z.x0, z.x0 + sp.d0, sp.eL) && halving_exact(z.y0, z.y0 + sp.d0, sp.eL) &&
                halving_exact(z.x1, z.x1 - sp.d0, sp.eL) && halving_exact(z.y1, z.y1 - sp.d0, sp.eL)) {
#if !defined(__HIP_DEVICE_COMPILE__) && defined(FRW_HOST_CLOSED_FORM_DIRECT_HOOK)
                FRW_HOST_CLOSED_FORM_DIRECT_HOOK();
#endif
                const double D = sp.hc - sp.L;
                return DBox{z.x0 + D, z.y0 + D, z.x1 - D, z.y1 - D};
            }
            o = z;
            double d = sp.d0;
            uint32_t dirs = sp.q << (32 - sp.m);  // the directions from the top bit down (m <= 31)
            for (int i = 0; i < sp.m; ++i) {
                d *= 0.5;
                const double sd = (int32_t)dirs < 0 ? -d : d;
                dirs <<= 1;
                o.x0 += sd;
                o.y0 += sd;
                o.x1 -= sd;
                o.y1 -= sd;
            }
            if (sp.shrink) {
                o.x0 += d;
                o.y0 += d;
                o.x1 -= d;
                o.y1 -= d;
            }
            return o;
        }
    }
    double dist = t.dist0;
    bool max_reached = false;
    int it = 0;
    if (!kLiteral) {
        // The steps whose test h decides alone (step < 60, h outside the
        // band) run in a loop of their own, in step with the other lanes of
        // the wave: a lane that has stopped adds 0 (x + 0 == x: a box
        // coordinate is never -0, which only -0 + -0 gives), the loop ends
        // when no lane goes on, and the body is straight-line code with no
        // per-lane exit.  A lane stops at the schedule's end (done) or at
        // its first ambiguous step, from which the general loop below takes
        // the schedule on.  The loop carries no boolean: max_reached is
        // dist < dist0 (when dist0 > MIN_DIST, dist halves at the first
        // collision and at no step before it; a smaller dist0 leaves the
        // whole schedule to the general loop), go and done are integers.
        // (dist * 0.5 is dist / 2 exactly.)
        int go = dist > MIN_DIST ? 1 : 0, done = 0;
        const double d0 = dist;
#pragma unroll 1
        for (int s = 0; s < 60; ++s) {
            const bool g = go != 0;
            const bool mr = dist < d0;
            const bool c = h >= h_hit;
            const bool big = dist > MIN_DIST;
            const bool amb = !c & !(h < h_free);
            const bool fin = !c & mr & !big;
            const bool stop = g & (amb | fin);
            it = stop ? s : it;
            done = stop ? (int)(fin & !amb) : done;
            go = (g & !(amb | fin)) ? 1 : 0;
            if (!any_lane(go != 0)) break;
            const bool g2 = go != 0;
            const double half = dist * 0.5;
            dist = (g2 & big & (mr | c)) ? half : dist;
            const double sd = g2 ? (c ? dist : -dist) : 0.0;
            o.x0 += sd; o.y0 += sd; o.x1 -= sd; o.y1 -= sd;
            h -= sd;
        }
        if (go) it = 60;
        if (done) return o;
        max_reached = dist < d0;
    }
    for (;; ++it) {
        step = it;
        const bool c = collides(o.x0, o.y0, o.x1, o.y1);
        if (!(c || !max_reached || dist > MIN_DIST)) break;
        if (it >= kMaxBisect) {  // cannot happen for a finite layout; reported
            if (err) raise_error(err, kErrBisect);
            break;
        }
        // shrink by dist on a collision, grow otherwise; dist halves first
        // (after a collision while above MIN_DIST, after a free step once a
        // collision has been seen).  a + (-d) and a - (-d) round as a - d and
        // a + d, so the signed step gives the same coordinates branch-free.
        if (c ? dist > MIN_DIST : max_reached) dist /= 2;
        max_reached = max_reached || c;
        const double sd = c ? dist : -dist;
        o.x0 += sd; o.y0 += sd; o.x1 -= sd; o.y1 -= sd;
        h -= sd;
    }
    return o;
}

// Gaussian-surface factor search (getAdjustedGaussianFactor,
// randomWalk/CPU/quadTree.cpp:315-349) and createGaussianSurface
// (randomWalk/CPU/rectHost.cpp:84-100).
__host__ __device__ inline DBox gauss_surface(const int4& r, double fx, double fy)
{
    const double mx = ((double)r.x + (double)r.z) / 2.;
    const double my = ((double)r.y + (double)r.w) / 2.;
    double vx = (double)r.x - mx;
    double vy = (double)r.y - my;
    vx *= fx;
    vy *= fy;
    return DBox{mx + vx, my + vy, mx - vx, my - vy};
}

template <class Stack>
__host__ __device__ inline double gauss_factor(const DevTree& t, const Stack& st, const int4& r,
                                      double factor, bool axis_y)
{
    bool dividing = true, first = true;
    double f = factor, lo = 1., hi = factor;
    for (int i = 0; i < 10; i++) {
        const DBox s = axis_y ? gauss_surface(r, 1., f) : gauss_surface(r, f, 1.);
        const bool coll = !in_bounds_box(t.bounds, s.x0, s.y0, s.x1, s.y1) ||
                          q1_box<true>(t, st, s.x0, s.y0, s.x1, s.y1, r, nullptr);
        if (first && !coll) break;
        if ((coll && !dividing) || (!coll && dividing)) dividing = !dividing;
        if (dividing)
            f = hi = (lo + hi) / 2.;
        else
            f = lo = (lo + hi) / 2.;
        first = false;
    }
    return f;
}

}  // namespace frw
