// Host build of the library's device-side traversal code (frw_device.h), for
// CPU tests: a sequential restatement of the Morton build in frw_build.hip and
// of the per-walk loop in frw_walk.hip, driving the very same __host__
// __device__ functions the kernels use.  Test infrastructure only.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <vector>

// Wave-wide "any lane" tests of the traversal code answer true for
// g_pad further calls after the lane's own answer turns false, as when
// other lanes of its wave go on (0: the lane alone).
static int g_pad = 0, g_pad_left = 0;
static bool host_any_lane(bool v)
{
    if (v) return true;
    if (g_pad_left > 0) {
        --g_pad_left;
        return true;
    }
    g_pad_left = g_pad;
    return false;
}
#define FRW_HOST_ANY_LANE(v) host_any_lane(v)
// square searches answered by the closed form of their decided steps
static long long g_closed_form = 0, g_closed_direct = 0;
#define FRW_HOST_CLOSED_FORM_HOOK() (++g_closed_form)
#define FRW_HOST_CLOSED_FORM_DIRECT_HOOK() (++g_closed_direct)

#include "../../engineer-thesis_amd/csrc/frw_device.h"

using namespace frw;

namespace {

uint64_t spread(uint32_t v)
{
    uint64_t x = v, out = 0;
    for (int b = 0; b < 32; ++b) out |= ((x >> b) & 1ull) << (2 * b);
    return out;
}

struct HostTree {
    int4 bounds;
    int K = 0, depth = 0, ext2_max = 0, start_levels = 0;
    std::vector<int4> rects;
    std::vector<int32_t> ids;
    std::vector<Node> nodes;
    std::vector<int32_t> start;
    double margin = 0, dist0 = 0;
};

struct Frontier {
    uint64_t code;
    int node, level, s, e;
};

void build(HostTree& t, const int32_t* xyxy, int n, const int32_t* B, int cap)
{
    t.bounds = make_int4(B[0], B[1], B[2], B[3]);
    const int64_t W = (int64_t)B[2] - B[0], H = (int64_t)B[3] - B[1];
    int K = 0;
    while (K < 30 && ((int64_t)1 << K) <= std::max(W, H)) ++K;
    K += 1;  // doubled centre coordinates need one more bit
    t.K = K;
    std::vector<std::pair<uint64_t, int>> kv(n);
    for (int i = 0; i < n; ++i) {
        const int32_t* r = xyxy + 4 * i;
        const uint32_t cx = (uint32_t)((int64_t)r[0] - B[0] + (int64_t)r[2] - B[0]);
        const uint32_t cy = (uint32_t)((int64_t)r[1] - B[1] + (int64_t)r[3] - B[1]);
        const uint64_t code = spread(cx) | (spread(cy) << 1);
        kv[i] = {(code << 6) | (uint64_t)K, i};
    }
    std::stable_sort(kv.begin(), kv.end(),
                     [](const std::pair<uint64_t, int>& a, const std::pair<uint64_t, int>& b) { return a.first < b.first; });
    std::vector<uint64_t> keys(n);
    t.rects.resize(n);
    t.ids.resize(n);
    for (int i = 0; i < n; ++i) {
        keys[i] = kv[i].first;
        t.ids[i] = kv[i].second;
        const int32_t* r = xyxy + 4 * kv[i].second;
        t.rects[i] = make_int4(r[0], r[1], r[2], r[3]);
    }
    std::vector<Frontier> fr = {{0, 0, 0, 0, n}, {0, 1, 0, 0, 0}, {0, 2, 0, 0, 0}, {0, 3, 0, 0, 0}};
    std::vector<std::pair<int, int>> runs(4);
    t.nodes.assign(4, Node{});
    t.depth = 0;
    while (!fr.empty()) {
        t.depth++;
        std::vector<Frontier> next;
        for (const Frontier& f : fr) {
            Node nd;
            nd.level = f.level;
            nd.first_child = -1;
            nd.box = make_int4(INT_MAX, INT_MAX, INT_MIN, INT_MIN);
            runs[f.node] = {f.s, f.e};
            bool internal = false;
            int oe = f.e;
            if (f.s == f.e) {
                nd.rbeg = nd.rend = 0;
            } else {
                const uint64_t own = (f.code << 6) | (uint64_t)f.level;
                oe = (int)(std::upper_bound(keys.begin() + f.s, keys.begin() + f.e, own) - keys.begin());
                nd.rbeg = f.s;
                if (f.e - f.s <= cap || f.level >= K || oe == f.e) {
                    nd.rend = f.e;
                } else {
                    nd.rend = oe;
                    internal = true;
                }
            }
            if (internal) {
                const int first = (int)t.nodes.size() + (int)next.size();
                nd.first_child = first;
                const int shift = 2 * (K - f.level - 1);
                int b[5];
                b[0] = oe;
                b[4] = f.e;
                for (int c = 1; c < 4; ++c)
                    b[c] = (int)(std::lower_bound(keys.begin() + b[c - 1], keys.begin() + f.e,
                                                  (f.code | ((uint64_t)c << shift)) << 6) - keys.begin());
                for (int c = 0; c < 4; ++c)
                    next.push_back({f.code | ((uint64_t)c << shift), first + c, f.level + 1, b[c], b[c + 1]});
            }
            t.nodes[f.node] = nd;
        }
        if (next.empty()) break;
        t.nodes.resize(t.nodes.size() + next.size());
        runs.resize(t.nodes.size());
        fr.swap(next);
    }
    for (size_t ni = 0; ni < t.nodes.size(); ++ni) {
        int4 b = make_int4(INT_MAX, INT_MAX, INT_MIN, INT_MIN);
        for (int r = runs[ni].first; r < runs[ni].second; ++r) {
            b.x = std::min(b.x, t.rects[r].x); b.y = std::min(b.y, t.rects[r].y);
            b.z = std::max(b.z, t.rects[r].z); b.w = std::max(b.w, t.rects[r].w);
        }
        t.nodes[ni].box = b;
    }
    const double Wd = (double)B[2] - (double)B[0], Hd = (double)B[3] - (double)B[1];
    double dist = (Hd > Wd) ? Hd : Wd;
    t.dist0 = dist * 0.05;
    double cmax = 1.0;
    for (int i = 0; i < 4; ++i) cmax = std::max(cmax, std::fabs((double)B[i]));
    cmax = std::max({cmax, Wd, Hd});
    int k = 0;
    while (std::ldexp(1.0, k) <= 2.0 * cmax) ++k;
    t.margin = std::ldexp(1.0, k - 48);
    for (int i = 0; i < n; ++i)
        t.ext2_max = std::max({t.ext2_max, xyxy[4 * i + 2] - xyxy[4 * i], xyxy[4 * i + 3] - xyxy[4 * i + 1]});
    t.start_levels = std::min(t.depth, kStartLevels);
    t.start.assign((size_t)start_offset(t.start_levels), -1);
    for (int l = 0; l < t.start_levels; ++l)
        for (uint32_t cy = 0; cy < (1u << l); ++cy)
            for (uint32_t cx = 0; cx < (1u << l); ++cx)
                t.start[(size_t)(start_offset(l) + ((int64_t)cy << l) + cx)] = start_cell_node(t.nodes.data(), l, cx, cy);
}

DevTree view(const HostTree& t)
{
    DevTree v;
    v.rects = t.rects.data();
    v.ids = t.ids.data();
    v.nodes = t.nodes.data();
    v.bounds = t.bounds;
    v.n_rects = (int32_t)t.rects.size();
    v.n_nodes = (int32_t)t.nodes.size();
    v.depth = t.depth;
    v.stack_cap = LaneStack::entries(t.depth);
    v.margin = t.margin;
    v.dist0 = t.dist0;
    v.start_tab = t.start.data();
    v.start_levels = t.start_levels;
    v.morton_bits = t.K;
    v.ext2_max = t.ext2_max;
    return v;
}

// Stack layout of the traversals below: 0 single nodes (PlainStack), 1 runs
// of siblings (RunStack), as the walk kernel chooses per launch.
int g_runs = 0;

template <class S>
struct HostStack {
    std::vector<int32_t> mem;
    S st;
    explicit HostStack(int depth) : mem(S::entries(depth)), st{mem.data(), 1, S::entries(depth)} {}
};

}  // namespace

extern "C" {

void hs_set_stack(int runs) { g_runs = runs; }

void hs_set_lane_pad(int pad) { g_pad = g_pad_left = pad; }
long long hs_closed_form_count(void) { return g_closed_form; }
long long hs_closed_form_direct_count(void) { return g_closed_direct; }

void* hs_build(const int32_t* xyxy, int n, const int32_t* bounds, int cap)
{
    HostTree* t = new HostTree;
    build(*t, xyxy, n, bounds, cap);
    return t;
}

void hs_free(void* h) { delete static_cast<HostTree*>(h); }

void hs_info(void* h, int64_t* out)
{
    const HostTree* t = static_cast<HostTree*>(h);
    out[0] = (int64_t)t->nodes.size();
    out[1] = t->depth;
    out[2] = t->K;
}

void hs_export(void* h, int32_t* nodes, int32_t* rects, int32_t* ids)
{
    const HostTree* t = static_cast<HostTree*>(h);
    std::memcpy(nodes, t->nodes.data(), t->nodes.size() * sizeof(Node));
    std::memcpy(rects, t->rects.data(), t->rects.size() * sizeof(int4));
    std::memcpy(ids, t->ids.data(), t->ids.size() * sizeof(int32_t));
}

}  // extern "C"

// nearest distance, Q2 id, Q3 box (replayed and literal), Q1 fallbacks of the replay
template <class S>
static void hs_queries_impl(void* h, const double* pts, int n, double* dist, int32_t* q2, double* q3e, double* q3b,
                int32_t* n_fallback)
{
    const HostTree* t = static_cast<HostTree*>(h);
    const DevTree v = view(*t);
    HostStack<S> hs_(t->depth);
    const S& st = hs_.st;
    int err = 0;
    for (int i = 0; i < n; ++i) {
        const double px = pts[2 * i], py = pts[2 * i + 1];
        dist[i] = nearest_rect(v, st, px, py, nullptr);
        q2[i] = q2_point(v, st, px, py, nullptr);
        uint32_t nq = 0;
        DBox a = biggest_square<false>(v, st, px, py, dist[i], &nq, nullptr, &err);
        DBox b = biggest_square<true>(v, st, px, py, 0.0, nullptr, nullptr, &err);
        n_fallback[i] = (int32_t)nq;
        q3e[4 * i] = a.x0; q3e[4 * i + 1] = a.y0; q3e[4 * i + 2] = a.x1; q3e[4 * i + 3] = a.y1;
        q3b[4 * i] = b.x0; q3b[4 * i + 1] = b.y0; q3b[4 * i + 2] = b.x1; q3b[4 * i + 3] = b.y1;
    }
}

extern "C" {

void hs_queries(void* h, const double* pts, int n, double* dist, int32_t* q2, double* q3e, double* q3b,
                int32_t* n_fallback)
{
    if (g_runs) hs_queries_impl<RunStack>(h, pts, n, dist, q2, q3e, q3b, n_fallback);
    else hs_queries_impl<PlainStack>(h, pts, n, dist, q2, q3e, q3b, n_fallback);
}

}  // extern "C"

// Traffic of the nearest-rect search and of a Q1 probe of the box p +- f*d*
// (per point: nodes/rects of each), for sizing a proposal-and-verify step.
template <class S>
static void hs_probe_cost_impl(void* h, const double* pts, int n, double f, int64_t* out)
{
    const HostTree* t = static_cast<HostTree*>(h);
    const DevTree v = view(*t);
    HostStack<S> hs_(t->depth);
    const S& st = hs_.st;
    int64_t nn = 0, nr = 0, pn = 0, pr = 0, ok = 0;
    for (int i = 0; i < n; ++i) {
        const double px = pts[2 * i], py = pts[2 * i + 1];
        Counters a{0, 0}, b{0, 0};
        const double d = nearest_rect(v, st, px, py, &a);
        const double hh = f * d;
        const bool hit = q1_box<false>(v, st, px - hh, py - hh, px + hh, py + hh, make_int4(0, 0, 0, 0), &b);
        nn += a.nodes; nr += a.rects; pn += b.nodes; pr += b.rects; ok += hit ? 0 : 1;
    }
    out[0] = nn; out[1] = nr; out[2] = pn; out[3] = pr; out[4] = ok;
}

extern "C" {

void hs_probe_cost(void* h, const double* pts, int n, double f, int64_t* out)
{
    if (g_runs) hs_probe_cost_impl<RunStack>(h, pts, n, f, out);
    else hs_probe_cost_impl<PlainStack>(h, pts, n, f, out);
}

}  // extern "C"

template <class S>
static void hs_gauss_impl(void* h, int master, double* out)
{
    const HostTree* t = static_cast<HostTree*>(h);
    const DevTree v = view(*t);
    HostStack<S> hs_(t->depth);
    const S& st = hs_.st;
    int4 m = make_int4(0, 0, 0, 0);
    for (size_t i = 0; i < t->ids.size(); ++i)
        if (t->ids[i] == master) m = t->rects[i];
    const double fx = gauss_factor(v, st, m, 1.5, false);
    const double fy = gauss_factor(v, st, m, 1.5, true);
    const DBox s = gauss_surface(m, fx, fy);
    out[0] = s.x0; out[1] = s.y0; out[2] = s.x1; out[3] = s.y1;
}

extern "C" {

void hs_gauss(void* h, int master, double* out)
{
    if (g_runs) hs_gauss_impl<RunStack>(h, master, out);
    else hs_gauss_impl<PlainStack>(h, master, out);
}

int64_t hs_walk_hinted(void* h, int master, int mode, uint64_t seed, uint32_t job, uint32_t first_walk,
                       int64_t n_walks, uint32_t max_hops, const uint32_t* T, int32_t* hit_out,
                       int32_t* hops_out, int64_t* traffic, int use_hint);

// The walk_kernel loop for one walk at a time.
int64_t hs_walk(void* h, int master, int mode, uint64_t seed, uint32_t job, uint32_t first_walk,
                int64_t n_walks, uint32_t max_hops, const uint32_t* T, int32_t* hit_out,
                int32_t* hops_out, int64_t* traffic)
{
    return hs_walk_hinted(h, master, mode, seed, job, first_walk, n_walks, max_hops, T, hit_out, hops_out,
                          traffic, 0);
}

}  // extern "C"

// use_hint: seed each hop's nearest-rect search with the previous hop's nearest rect.
template <class S>
static int64_t hs_walk_hinted_impl(void* h, int master, int mode, uint64_t seed, uint32_t job, uint32_t first_walk,
                       int64_t n_walks, uint32_t max_hops, const uint32_t* T, int32_t* hit_out,
                       int32_t* hops_out, int64_t* traffic, int use_hint)
{
    const HostTree* t = static_cast<HostTree*>(h);
    const DevTree v = view(*t);
    HostStack<S> hs_(t->depth);
    const S& st = hs_.st;
    double surf[4];
    hs_gauss(h, master, surf);
    // the walk kernel's first-hop hint: a rect at the master's centre
    int start_hint = -1;
    int4 start_rect = make_int4(0, 0, 0, 0);
    if (use_hint) {
        int4 m = make_int4(0, 0, 0, 0);
        for (size_t i = 0; i < t->ids.size(); ++i)
            if (t->ids[i] == master) m = t->rects[i];
        (void)nearest_rect(v, st, ((double)m.x + (double)m.z) / 2.0, ((double)m.y + (double)m.w) / 2.0, nullptr,
                           -1, &start_hint);
        start_rect = t->rects[start_hint];
    }
    int err = 0;
    int64_t total = 0;
    Counters cnt{0, 0, 0};
    for (int64_t w = 0; w < n_walks; ++w) {
        DBox sq{surf[0], surf[1], surf[2], surf[3]};
        uint32_t rnd[4] = {0, 0, 0, 0};
        uint32_t hop = 0;
        int result = -3;
        int hint = start_hint;
        int4 hint_r = start_rect;
        while (result == -3) {
            if ((hop & 3u) == 0u) {
                walk_counter(rnd, (uint64_t)first_walk + (uint64_t)w, job, hop >> 2, kWalkDomain);
                philox4x32_10(rnd, (uint32_t)seed, (uint32_t)(seed >> 32));
            }
            const int panel = sample_panel(T, rnd[hop & 3u]);
            double px, py;
            point_from_panel(sq.x0, sq.y0, sq.x1, sq.y1, panel, px, py);
            if (!in_bounds_strict(v.bounds, px, py)) {
                result = -1;
                break;
            }
            if (mode == kModeExact) {
                int arg = -1;
                const double ds = nearest_rect(v, st, px, py, traffic ? &cnt : nullptr, use_hint ? hint : -1, &arg,
                                               use_hint ? &hint_r : nullptr);
                hint = arg;
                if (ds <= 0.0)
                    result = v.ids[arg];
                else
                    sq = biggest_square<false>(v, st, px, py, ds, nullptr, traffic ? &cnt : nullptr, &err,
                                              use_hint ? &hint_r : nullptr);
            } else {
                sq = biggest_square<true>(v, st, px, py, 0.0, nullptr, traffic ? &cnt : nullptr, &err);
                const int hh = q2_point(v, st, px, py, traffic ? &cnt : nullptr);
                if (hh >= 0) result = hh;
            }
            hop++;
            if (result == -3 && hop >= max_hops) result = -2;
        }
        hit_out[w] = result;
        hops_out[w] = (int32_t)hop;
        total += hop;
    }
    if (traffic) {
        traffic[0] = cnt.nodes;
        traffic[1] = cnt.rects;
        traffic[2] = cnt.tabs;
    }
    return err ? -1 : total;
}

extern "C" {

int64_t hs_walk_hinted(void* h, int master, int mode, uint64_t seed, uint32_t job, uint32_t first_walk,
                       int64_t n_walks, uint32_t max_hops, const uint32_t* T, int32_t* hit_out,
                       int32_t* hops_out, int64_t* traffic, int use_hint)
{
    if (g_runs) return hs_walk_hinted_impl<RunStack>(h, master, mode, seed, job, first_walk, n_walks, max_hops, T, hit_out, hops_out, traffic, use_hint);
    else return hs_walk_hinted_impl<PlainStack>(h, master, mode, seed, job, first_walk, n_walks, max_hops, T, hit_out, hops_out, traffic, use_hint);
}

}  // extern "C"
