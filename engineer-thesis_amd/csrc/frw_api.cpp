// C ABI of the FRW library (declared in include/frw.h).  Every entry point
// returns a status code; errors are recorded per thread and never thrown
// across the boundary.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "frw_internal.h"

namespace frw {
int4 padded_bounds(const std::vector<int4>& r);
int read_layout(const char* path, int layer, std::vector<int4>* rects, int* n_layers);
int validate_rects(const std::vector<int4>& r);
int generate_layout(const frw_gen_params* p, std::vector<int4>* out);
int save_layout(const std::vector<int4>& rs, const char* path);
int write_results(const std::vector<int4>& rs, const int64_t* counts, int64_t n_walks, const char* path);

namespace {
thread_local std::string g_err;
}

int set_error(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

int check_hip(hipError_t e, const char* what)
{
    if (e == hipSuccess) return FRW_OK;
    return set_error(e == hipErrorOutOfMemory ? FRW_E_NOMEM : FRW_E_DEVICE,
                     std::string(what) + ": " + hipGetErrorString(e));
}

int order_before(LaunchOrder& o, hipStream_t s, bool* capturing)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    int rc = check_hip(hipStreamIsCapturing(s, &cs), "hipStreamIsCapturing");
    if (rc != FRW_OK) return rc;
    *capturing = cs != hipStreamCaptureStatusNone;
    if (*capturing) return FRW_OK;
    if (!o.ev && (rc = check_hip(hipEventCreateWithFlags(&o.ev, hipEventDisableTiming), "hipEventCreate")) != FRW_OK)
        return rc;
    if (o.recorded) return check_hip(hipStreamWaitEvent(s, o.ev, 0), "hipStreamWaitEvent");
    return FRW_OK;
}

int order_after(LaunchOrder& o, hipStream_t s, bool capturing)
{
    if (capturing || !o.ev) return FRW_OK;
    const int rc = check_hip(hipEventRecord(o.ev, s), "hipEventRecord");
    if (rc == FRW_OK) o.recorded = true;
    return rc;
}

void order_free(LaunchOrder& o)
{
    if (o.ev) (void)hipEventDestroy(o.ev);
    o.ev = nullptr;
    o.recorded = false;
}

}  // namespace frw

using namespace frw;

#define FRW_GUARD_BEGIN try {
#define FRW_GUARD_END                                                   \
    }                                                                   \
    catch (const std::bad_alloc&) { return set_error(FRW_E_NOMEM, "host allocation failed"); } \
    catch (...) { return set_error(FRW_E_INVALID, "unexpected exception"); }

extern "C" {

const char* frw_last_error_string(void) { return g_err.c_str(); }

const char* frw_version(void) { return "frw-mi355x 0.1 (gfx950)"; }

// ---------------------------------------------------------------- layouts

static int make_layout(std::vector<int4>&& rects, frw_layout** out)
{
    int rc = validate_rects(rects);
    if (rc != FRW_OK) return rc;
    frw_layout* l = new frw_layout;
    l->rects = std::move(rects);
    l->bounds = padded_bounds(l->rects);
    *out = l;
    return FRW_OK;
}

int frw_layout_load(const char* path, int32_t layer, frw_layout** out)
{
    if (!path || !out) return set_error(FRW_E_INVALID, "null argument");
    FRW_GUARD_BEGIN
    std::vector<int4> r;
    int rc = read_layout(path, layer, &r, nullptr);
    if (rc != FRW_OK) return rc;
    return make_layout(std::move(r), out);
    FRW_GUARD_END
}

int frw_layout_layer_count(const char* path, int32_t* n_layers)
{
    if (!path || !n_layers) return set_error(FRW_E_INVALID, "null argument");
    FRW_GUARD_BEGIN
    int n = 0;
    int rc = read_layout(path, 0, nullptr, &n);
    *n_layers = n;
    return (rc == FRW_E_LAYER) ? FRW_OK : rc;
    FRW_GUARD_END
}

int frw_layout_from_rects(const int32_t* xyxy, int64_t n, frw_layout** out)
{
    if ((!xyxy && n > 0) || n < 0 || !out) return set_error(FRW_E_INVALID, "bad rect array");
    if (n > (int64_t)1 << 30) return set_error(FRW_E_RANGE, "too many rects");
    FRW_GUARD_BEGIN
    std::vector<int4> r((size_t)n);
    for (int64_t i = 0; i < n; ++i) r[i] = make_int4(xyxy[4 * i], xyxy[4 * i + 1], xyxy[4 * i + 2], xyxy[4 * i + 3]);
    return make_layout(std::move(r), out);
    FRW_GUARD_END
}

int frw_layout_generate(const frw_gen_params* p, frw_layout** out)
{
    if (!p || !out) return set_error(FRW_E_INVALID, "null argument");
    FRW_GUARD_BEGIN
    std::vector<int4> r;
    int rc = generate_layout(p, &r);
    if (rc != FRW_OK) return rc;
    return make_layout(std::move(r), out);
    FRW_GUARD_END
}

int frw_layout_save(const frw_layout* l, const char* path)
{
    if (!l || !path) return set_error(FRW_E_INVALID, "null argument");
    FRW_GUARD_BEGIN
    return save_layout(l->rects, path);
    FRW_GUARD_END
}

int64_t frw_layout_size(const frw_layout* l) { return l ? (int64_t)l->rects.size() : -1; }

int frw_layout_rects(const frw_layout* l, int32_t* xyxy_out)
{
    if (!l || !xyxy_out) return set_error(FRW_E_INVALID, "null argument");
    std::memcpy(xyxy_out, l->rects.data(), sizeof(int4) * l->rects.size());
    return FRW_OK;
}

int frw_layout_bounds(const frw_layout* l, int32_t out[4])
{
    if (!l || !out) return set_error(FRW_E_INVALID, "null argument");
    out[0] = l->bounds.x; out[1] = l->bounds.y; out[2] = l->bounds.z; out[3] = l->bounds.w;
    return FRW_OK;
}

int frw_layout_set_bounds(frw_layout* l, const int32_t b[4])
{
    if (!l || !b) return set_error(FRW_E_INVALID, "null argument");
    if (b[0] > b[2] || b[1] > b[3]) return set_error(FRW_E_RANGE, "bounds are not a box");
    for (const int4& r : l->rects)
        if (r.x < b[0] || r.y < b[1] || r.z > b[2] || r.w > b[3])
            return set_error(FRW_E_RANGE, "a rect lies outside the bounds");
    l->bounds = make_int4(b[0], b[1], b[2], b[3]);
    return FRW_OK;
}

void frw_layout_free(frw_layout* l) { delete l; }

int frw_green_thresholds(uint32_t out[FRW_NSAMPLE])
{
    if (!out) return set_error(FRW_E_INVALID, "null argument");
    green_thresholds(out);
    return FRW_OK;
}

// ---------------------------------------------------------------- tree

void frw_tree_free(frw_tree* t)
{
    if (!t) return;
    free_refcompat(t);
    if (t->d_rects) (void)hipFree(t->d_rects);
    if (t->d_ids) (void)hipFree(t->d_ids);
    if (t->d_nodes) (void)hipFree(t->d_nodes);
    if (t->d_start) (void)hipFree(t->d_start);
    if (t->d_thresholds) (void)hipFree(t->d_thresholds);
    if (t->d_wtab) (void)hipFree(t->d_wtab);
    if (t->d_work) (void)hipFree(t->d_work);
    if (t->d_err) (void)hipFree(t->d_err);
    if (t->d_surface) (void)hipFree(t->d_surface);
    if (t->d_job) (void)hipFree(t->d_job);
    if (t->d_wacc) (void)hipFree(t->d_wacc);
    order_free(t->order);
    delete t;
}

int frw_tree_build(const frw_layout* l, const frw_tree_params* p, void* stream, frw_tree** out)
{
    if (!l || !out) return set_error(FRW_E_INVALID, "null argument");
    if (l->rects.empty()) return set_error(FRW_E_RANGE, "layout has no rects");
    const int cap = (p && p->leaf_cap > 0) ? p->leaf_cap : 4;
    if (cap > 64) return set_error(FRW_E_RANGE, "leaf_cap must be <= 64");
    const int kind = p ? p->kind : FRW_TREE_MORTON;
    if (kind != FRW_TREE_MORTON && kind != FRW_TREE_REFCOMPAT) return set_error(FRW_E_INVALID, "unknown tree kind");
    FRW_GUARD_BEGIN
    hipStream_t s = (hipStream_t)stream;
    int dev = 0;
    int rc0 = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc0 != FRW_OK) return rc0;
    frw_tree* t = new frw_tree;
    t->device = dev;
    t->bounds = l->bounds;
    t->n_rects = (int64_t)l->rects.size();
    t->leaf_cap = cap;
    t->host_rects = l->rects;
    const double W = (double)l->bounds.z - (double)l->bounds.x;
    const double H = (double)l->bounds.w - (double)l->bounds.y;
    double dist = (H > W) ? H : W;
    t->dist0 = dist * 0.05;  // BIGGEST_SQUARE_INIT_FACTOR (randomWalk/defines.h:21)
    // Ambiguity band of the replayed bisection: 2^(k-48) where every
    // coordinate and distance is below 2^(k-1); FP64 rounding errors of the
    // compared quantities stay below 2^(k-53).
    const double cmax = std::max({std::fabs((double)l->bounds.x), std::fabs((double)l->bounds.y),
                                  std::fabs((double)l->bounds.z), std::fabs((double)l->bounds.w), W, H, 1.0});
    int k = 0;
    while (std::ldexp(1.0, k) <= 2.0 * cmax) ++k;
    t->margin = std::ldexp(1.0, k - 48);
    for (const int4& r : l->rects) t->ext2_max = std::max({t->ext2_max, r.z - r.x, r.w - r.y});
    int rc = build_tree(t, l->rects, l->bounds, cap, s);
    if (rc == FRW_OK) {
        uint32_t T[kNSample];
        green_thresholds(T);
        rc = check_hip(hipMalloc(&t->d_thresholds, sizeof(T)), "hipMalloc thresholds");
        if (rc == FRW_OK) rc = check_hip(hipMemcpy(t->d_thresholds, T, sizeof(T), hipMemcpyHostToDevice), "copy thresholds");
        double wt[6 * kNSample];  // 5 x 200 doubles, then 2 x 200 u32 thresholds
        green_grad_tables(wt, reinterpret_cast<uint32_t*>(wt + 5 * kNSample));
        if (rc == FRW_OK) rc = check_hip(hipMalloc(&t->d_wtab, sizeof(wt)), "hipMalloc weight tables");
        if (rc == FRW_OK) rc = check_hip(hipMemcpy(t->d_wtab, wt, sizeof(wt), hipMemcpyHostToDevice), "copy weight tables");
        if (rc == FRW_OK) rc = check_hip(hipMalloc(&t->d_work, sizeof(unsigned long long)), "hipMalloc work");
        if (rc == FRW_OK) rc = check_hip(hipMalloc(&t->d_err, sizeof(int)), "hipMalloc err");
        if (rc == FRW_OK) rc = check_hip(hipMemset(t->d_err, 0, sizeof(int)), "memset err");
        if (rc == FRW_OK) rc = check_hip(hipMalloc(&t->d_surface, 4 * sizeof(double)), "hipMalloc surface");
        if (rc == FRW_OK) rc = check_hip(hipMalloc(&t->d_job, sizeof(WalkJob)), "hipMalloc job");
        if (rc == FRW_OK && kind == FRW_TREE_REFCOMPAT) rc = build_refcompat(t);
    }
    if (rc != FRW_OK) {
        frw_tree_free(t);
        return rc;
    }
    *out = t;
    return FRW_OK;
    FRW_GUARD_END
}

int frw_tree_info_get(const frw_tree* t, frw_tree_info* info)
{
    if (!t || !info) return set_error(FRW_E_INVALID, "null argument");
    std::memset(info, 0, sizeof(*info));
    info->n_rects = t->n_rects;
    info->n_nodes = t->n_nodes;
    info->depth = t->depth;
    info->leaf_cap = t->leaf_cap;
    info->bounds[0] = t->bounds.x; info->bounds[1] = t->bounds.y;
    info->bounds[2] = t->bounds.z; info->bounds[3] = t->bounds.w;
    info->morton_bits = t->morton_bits;
    info->build_ms = t->build_ms;
    return FRW_OK;
}

int frw_tree_export(const frw_tree* t, int32_t* nodes, int32_t* rects, int32_t* ids)
{
    if (!t) return set_error(FRW_E_INVALID, "null tree");
    int rc = FRW_OK;
    if (nodes && rc == FRW_OK)
        rc = check_hip(hipMemcpy(nodes, t->d_nodes, sizeof(Node) * t->n_nodes, hipMemcpyDeviceToHost), "export nodes");
    if (rects && rc == FRW_OK)
        rc = check_hip(hipMemcpy(rects, t->d_rects, sizeof(int4) * t->n_rects, hipMemcpyDeviceToHost), "export rects");
    if (ids && rc == FRW_OK)
        rc = check_hip(hipMemcpy(ids, t->d_ids, sizeof(int32_t) * t->n_rects, hipMemcpyDeviceToHost), "export ids");
    return rc;
}

// ---------------------------------------------------------------- queries

static int check_err_flag(const frw_tree* t, hipStream_t s)
{
    int err = 0;
    int rc = check_hip(hipMemcpyAsync(&err, t->d_err, sizeof(int), hipMemcpyDeviceToHost, s), "read err flag");
    if (rc == FRW_OK) rc = check_hip(hipStreamSynchronize(s), "sync");
    if (rc != FRW_OK) return rc;
    if (err) return set_error(FRW_E_KERNEL, "kernel error flag " + std::to_string(err));
    return FRW_OK;
}


int frw_query_refcompat(const frw_tree* t, const double* pts, const double* boxes, int64_t n,
                        const int32_t* ignore, uint8_t* q1, int32_t* q2, double* q3, void* stream)
{
    if (!t || n < 0 || (n > 0 && ((q1 && !boxes) || ((q2 || q3) && !pts))))
        return set_error(FRW_E_INVALID, "bad query arguments");
    FRW_GUARD_BEGIN
    return launch_ref_queries(t, pts, boxes, n, ignore, q1, q2, q3, (hipStream_t)stream);
    FRW_GUARD_END
}

int frw_refcompat_gauss_surface(const frw_tree* t, const int32_t rect[4], double factor, double out[4])
{
    if (!t || !rect || !out || !std::isfinite(factor))
        return set_error(FRW_E_INVALID, "bad arguments");
    FRW_GUARD_BEGIN
    return refcompat_gauss(t, make_int4(rect[0], rect[1], rect[2], rect[3]), factor, out);
    FRW_GUARD_END
}

int frw_refcompat_walk_from(const frw_tree* t, const double surface[4], const frw_walk_params* p,
                            const frw_walk_outputs* o, void* stream)
{
    if (!t || !surface || !p) return set_error(FRW_E_INVALID, "null argument");
    frw_walk_params q = *p;
    q.master = 0;
    q.mode = FRW_MODE_REFCOMPAT;
    int rc = check_walk_params(t, &q);
    if (rc != FRW_OK) return rc;
    if (!o || !o->counts || !o->hops) return set_error(FRW_E_INVALID, "walk outputs need counts and hops");
    for (int k = 0; k < 4; ++k)
        if (!std::isfinite(surface[k])) return set_error(FRW_E_INVALID, "start square is not finite");
    FRW_GUARD_BEGIN
    return launch_ref_walk(t, &q, o, (hipStream_t)stream, surface);
    FRW_GUARD_END
}

int frw_refcompat_info(const frw_tree* t, int64_t out[2])
{
    if (!t || !out) return set_error(FRW_E_INVALID, "null argument");
    return refcompat_info(t, out);
}

int frw_query_collide(const frw_tree* t, const double* boxes, int64_t n, uint8_t* out, void* stream)
{
    if (!t || n < 0 || (n > 0 && (!boxes || !out))) return set_error(FRW_E_INVALID, "bad query arguments");
    return launch_q1(t, boxes, n, out, (hipStream_t)stream);
}

int frw_query_point(const frw_tree* t, const double* pts, int64_t n, int32_t* rect_id, void* stream)
{
    if (!t || n < 0 || (n > 0 && (!pts || !rect_id))) return set_error(FRW_E_INVALID, "bad query arguments");
    return launch_q2(t, pts, n, rect_id, (hipStream_t)stream);
}

int frw_query_nearest(const frw_tree* t, const double* pts, int64_t n, double* dist, void* stream)
{
    if (!t || n < 0 || (n > 0 && (!pts || !dist))) return set_error(FRW_E_INVALID, "bad query arguments");
    return launch_nearest(t, pts, n, dist, (hipStream_t)stream);
}

int frw_query_maxsquare(const frw_tree* t, const double* pts, int64_t n, int32_t mode,
                        double* boxes_out, void* stream)
{
    if (!t || n < 0 || (n > 0 && (!pts || !boxes_out))) return set_error(FRW_E_INVALID, "bad query arguments");
    if (mode != FRW_MODE_EXACT && mode != FRW_MODE_BISECT) return set_error(FRW_E_INVALID, "bad mode");
    int rc = launch_q3(t, pts, n, mode, boxes_out, (hipStream_t)stream);
    if (rc != FRW_OK) return rc;
    return check_err_flag(t, (hipStream_t)stream);
}

int frw_gauss_surface(const frw_tree* t, int32_t master, double out[4])
{
    if (!t || !out) return set_error(FRW_E_INVALID, "null argument");
    if (master < 0 || master >= t->n_rects) return set_error(FRW_E_RANGE, "master out of range");
    bool capturing = false;
    int rc = order_before(t->order, nullptr, &capturing);  // d_surface is the tree's
    if (rc == FRW_OK) rc = launch_gauss(t, t->host_rects[master], t->d_surface, nullptr);
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(out, t->d_surface, 4 * sizeof(double), hipMemcpyDeviceToHost), "copy surface");
    const int rc2 = order_after(t->order, nullptr, capturing);
    return rc != FRW_OK ? rc : rc2;
}

// ---------------------------------------------------------------- walks

}  // extern "C"

void frw::fill_summary(const frw_tree* t, const frw_walk_params* p, const int64_t* counts, int64_t hops, double ms,
                       frw_walk_summary* summary)
{
    summary->n_walks = p->n_walks;
    summary->hops = hops;
    summary->n_master = counts[p->master];
    summary->n_escaped = counts[t->n_rects];
    summary->n_stuck = counts[t->n_rects + 1];
    // getAvgPathLen (randomWalk/CPU/getAvgPathLen.cpp:96,113): hits not
    // equal (by value) to the master, escapes and stuck walks included.
    int64_t same = 0;
    const int4 m = t->host_rects[p->master];
    for (int64_t i = 0; i < t->n_rects; ++i) {
        const int4 r = t->host_rects[i];
        if (r.x == m.x && r.y == m.y && r.z == m.z && r.w == m.w) same += counts[i];
    }
    summary->fraction_other = (double)(p->n_walks - same) / (double)p->n_walks;
    summary->walk_ms = ms;
}

int frw::check_walk_params(const frw_tree* t, const frw_walk_params* p)
{
    if (!t || !p) return set_error(FRW_E_INVALID, "null argument");
    if (p->master < 0 || p->master >= t->n_rects) return set_error(FRW_E_RANGE, "master out of range");
    if (p->n_walks <= 0) return set_error(FRW_E_RANGE, "walk count must be positive");
    if (p->first_walk < 0 || p->first_walk + p->n_walks > kWalkIdLimit)
        return set_error(FRW_E_RANGE, "walk ids must lie below 2^48");
    if (p->mode != FRW_MODE_EXACT && p->mode != FRW_MODE_BISECT && p->mode != FRW_MODE_REFCOMPAT)
        return set_error(FRW_E_INVALID, "unsupported walk mode");
    if (p->mode == FRW_MODE_REFCOMPAT && !t->ref)
        return set_error(FRW_E_INVALID, "reference-compatible walks need a tree built with FRW_TREE_REFCOMPAT");
    if (p->max_hops < 0 || p->max_hops > 0xFFFFFFFFll) return set_error(FRW_E_RANGE, "bad max_hops");
    return FRW_OK;
}

extern "C" {

int frw_walk_async(const frw_tree* t, const frw_walk_params* p, const frw_walk_outputs* o, void* stream)
{
    int rc = check_walk_params(t, p);
    if (rc != FRW_OK) return rc;
    if (!o || !o->counts || !o->hops) return set_error(FRW_E_INVALID, "walk outputs need counts and hops");
    FRW_GUARD_BEGIN
    return launch_walk(t, p, o, (hipStream_t)stream);
    FRW_GUARD_END
}

void frw_weight_tables(double wt[1000], uint32_t ithr[400])
{
    if (wt && ithr) green_grad_tables(wt, ithr);
}

int frw_walk_weighted_async(const frw_tree* t, const frw_walk_params* p, const frw_walk_outputs* o, double eps,
                            uint32_t flags, double* wsum, double* wsq, double* w_out, void* stream)
{
    if (!t || !p) return set_error(FRW_E_INVALID, "null argument");
    frw_walk_params q = *p;
    q.mode = FRW_MODE_EXACT;
    int rc = check_walk_params(t, &q);
    if (rc != FRW_OK) return rc;
    if (!o || !o->counts || !o->hops || !wsum || !wsq)
        return set_error(FRW_E_INVALID, "weighted walks need counts, hops, wsum and wsq");
    if (!(eps > 0.0) || !std::isfinite(eps)) return set_error(FRW_E_INVALID, "eps must be positive");
    if (flags & ~(uint32_t)FRW_WEIGHT_IMPORTANCE) return set_error(FRW_E_INVALID, "unknown weighted-walk flags");
    FRW_GUARD_BEGIN
    if (!t->d_wacc) {  // the tree's fixed-point accumulators, zero between launches
        frw_tree* tt = const_cast<frw_tree*>(t);
        const size_t bytes = sizeof(unsigned long long) * 4 * (size_t)(t->n_rects + 2);
        rc = check_hip(hipMalloc(&tt->d_wacc, bytes), "hipMalloc weighted accumulators");
        if (rc == FRW_OK) rc = check_hip(hipMemset(tt->d_wacc, 0, bytes), "memset weighted accumulators");
        if (rc == FRW_OK) rc = check_hip(hipDeviceSynchronize(), "weighted accumulators");
        if (rc != FRW_OK) {
            (void)hipFree(tt->d_wacc);
            tt->d_wacc = nullptr;
            return rc;
        }
    }
    q.mode = FRW_MODE_WEIGHTED;
    WalkExtras ex;
    ex.wsum = wsum;
    ex.wsq = wsq;
    ex.wacc = t->d_wacc;
    ex.w_out = w_out;
    ex.eps = eps;
    ex.wflags = flags;
    return launch_walk(t, &q, o, (hipStream_t)stream, nullptr, &ex);
    FRW_GUARD_END
}

int frw_walk_weighted(const frw_tree* t, const frw_walk_params* p, double eps, uint32_t flags, int64_t* counts_host,
                      double* wsum_host, double* wsq_host, int64_t* hops)
{
    if (!counts_host || !wsum_host || !wsq_host) return set_error(FRW_E_INVALID, "null output");
    FRW_GUARD_BEGIN
    const int64_t nb = t ? t->n_rects + 2 : 0;
    int64_t *d_counts = nullptr, *d_hops = nullptr;
    double *d_ws = nullptr, *d_wq = nullptr;
    int rc = t ? FRW_OK : set_error(FRW_E_INVALID, "null tree");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&d_counts, sizeof(int64_t) * nb), "hipMalloc counts");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&d_hops, sizeof(int64_t)), "hipMalloc hops");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&d_ws, sizeof(double) * nb), "hipMalloc wsum");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&d_wq, sizeof(double) * nb), "hipMalloc wsq");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_counts, 0, sizeof(int64_t) * nb), "memset");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_hops, 0, sizeof(int64_t)), "memset");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_ws, 0, sizeof(double) * nb), "memset");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_wq, 0, sizeof(double) * nb), "memset");
    if (rc == FRW_OK) {
        frw_walk_outputs o{d_counts, d_hops, nullptr, nullptr};
        rc = frw_walk_weighted_async(t, p, &o, eps, flags, d_ws, d_wq, nullptr, nullptr);
    }
    if (rc == FRW_OK) rc = check_hip(hipDeviceSynchronize(), "weighted walk");
    if (rc == FRW_OK) rc = check_err_flag(t, nullptr);
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(counts_host, d_counts, sizeof(int64_t) * nb, hipMemcpyDeviceToHost), "copy");
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(wsum_host, d_ws, sizeof(double) * nb, hipMemcpyDeviceToHost), "copy");
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(wsq_host, d_wq, sizeof(double) * nb, hipMemcpyDeviceToHost), "copy");
    if (rc == FRW_OK && hops) rc = check_hip(hipMemcpy(hops, d_hops, sizeof(int64_t), hipMemcpyDeviceToHost), "copy");
    for (void* q : {(void*)d_counts, (void*)d_hops, (void*)d_ws, (void*)d_wq})
        if (q) (void)hipFree(q);
    return rc;
    FRW_GUARD_END
}

int frw_walk_adaptive(const frw_tree* t, const frw_walk_params* p, double rel_err, int64_t max_walks, double eps,
                      uint32_t wflags, int64_t* counts_host, double* wsum_host, double* wsq_host,
                      frw_adaptive_result* r)
{
    if (!t || !p || !counts_host || !r) return set_error(FRW_E_INVALID, "null argument");
    if (!(rel_err > 0.0) || !std::isfinite(rel_err)) return set_error(FRW_E_INVALID, "rel_err must be positive");
    const bool weighted = p->mode == FRW_MODE_WEIGHTED;
    const int64_t id_space = kWalkIdLimit - p->first_walk;
    if (max_walks <= 0 || max_walks > id_space) max_walks = id_space;
    frw_walk_params q = *p;
    if (q.n_walks == 0) q.n_walks = 65536;
    q.n_walks = std::min(q.n_walks, max_walks);
    {
        frw_walk_params c = q;
        if (weighted) c.mode = FRW_MODE_EXACT;
        const int rc = check_walk_params(t, &c);
        if (rc != FRW_OK) return rc;
    }
    if (weighted && (!(eps > 0.0) || !std::isfinite(eps))) return set_error(FRW_E_INVALID, "eps must be positive");
    FRW_GUARD_BEGIN
    const int64_t nb = t->n_rects + 2;
    // getAvgPathLen counts a hit as "other" unless the rect equals the master
    // by value (randomWalk/CPU/getAvgPathLen.cpp:96)
    std::vector<int64_t> same_ids;
    const int4 m = t->host_rects[q.master];
    for (int64_t i = 0; i < t->n_rects; ++i) {
        const int4 x = t->host_rects[i];
        if (x.x == m.x && x.y == m.y && x.z == m.z && x.w == m.w) same_ids.push_back(i);
    }
    int64_t *d_counts = nullptr, *d_hops = nullptr;
    double *d_ws = nullptr, *d_wq = nullptr;
    int rc = check_hip(hipMalloc(&d_counts, sizeof(int64_t) * nb), "hipMalloc counts");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&d_hops, sizeof(int64_t)), "hipMalloc hops");
    if (rc == FRW_OK && weighted) rc = check_hip(hipMalloc(&d_ws, sizeof(double) * nb), "hipMalloc wsum");
    if (rc == FRW_OK && weighted) rc = check_hip(hipMalloc(&d_wq, sizeof(double) * nb), "hipMalloc wsq");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_counts, 0, sizeof(int64_t) * nb), "memset counts");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_hops, 0, sizeof(int64_t)), "memset hops");
    if (rc == FRW_OK && weighted) rc = check_hip(hipMemset(d_ws, 0, sizeof(double) * nb), "memset wsum");
    if (rc == FRW_OK && weighted) rc = check_hip(hipMemset(d_wq, 0, sizeof(double) * nb), "memset wsq");
    if (rc == FRW_OK) rc = check_hip(hipDeviceSynchronize(), "adaptive set-up");
    frw_adaptive_result res{};
    res.rel_err_3sigma = INFINITY;
    const auto t0 = std::chrono::steady_clock::now();
    int64_t next = q.n_walks;
    const frw_walk_outputs o{d_counts, d_hops, nullptr, nullptr};
    while (rc == FRW_OK) {
        frw_walk_params c = q;
        c.first_walk = q.first_walk + res.n_walks;
        c.n_walks = next;
        rc = weighted ? frw_walk_weighted_async(t, &c, &o, eps, wflags, d_ws, d_wq, nullptr, nullptr)
                      : launch_walk(t, &c, &o, nullptr);
        if (rc == FRW_OK) rc = check_hip(hipDeviceSynchronize(), "adaptive round");
        if (rc == FRW_OK) rc = check_err_flag(t, nullptr);
        if (rc != FRW_OK) break;
        res.n_walks += next;
        res.rounds++;
        const double n = (double)res.n_walks;
        if (weighted) {
            double sm = 0.0, sq = 0.0;
            rc = check_hip(hipMemcpy(&sm, d_ws + q.master, sizeof(double), hipMemcpyDeviceToHost), "read wsum");
            if (rc == FRW_OK)
                rc = check_hip(hipMemcpy(&sq, d_wq + q.master, sizeof(double), hipMemcpyDeviceToHost), "read wsq");
            res.value = sm / n;
            const double var = std::max(sq / n - res.value * res.value, 0.0);
            res.rel_err_3sigma = res.value != 0.0 ? 3.0 * std::sqrt(var / n) / std::fabs(res.value) : INFINITY;
        } else {
            int64_t same = 0;
            for (int64_t id : same_ids) {
                int64_t cm = 0;
                if (rc == FRW_OK)
                    rc = check_hip(hipMemcpy(&cm, d_counts + id, sizeof(int64_t), hipMemcpyDeviceToHost), "read count");
                same += cm;
            }
            const double f = (n - (double)same) / n;
            res.value = f;
            res.rel_err_3sigma = f > 0.0 ? 3.0 * std::sqrt(f * (1.0 - f) / n) / f : INFINITY;
        }
        if (rc != FRW_OK) break;
        if (res.rel_err_3sigma <= rel_err) {
            res.converged = 1;
            break;
        }
        if (res.n_walks >= max_walks) break;
        // Yu et al. 2013 eq. 32: n_rem = n (err^2 / goal^2 - 1), with a 5 % margin
        const double e = res.rel_err_3sigma / rel_err;
        const double want = std::isfinite(e) ? n * (e * e - 1.0) * 1.05 : n;
        next = std::max<int64_t>((int64_t)std::ceil(want), std::max<int64_t>(res.n_walks / 16, 1024));
        next = std::min<int64_t>(next, max_walks - res.n_walks);
    }
    res.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc == FRW_OK)
        rc = check_hip(hipMemcpy(counts_host, d_counts, sizeof(int64_t) * nb, hipMemcpyDeviceToHost), "copy counts");
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(&res.hops, d_hops, sizeof(int64_t), hipMemcpyDeviceToHost), "copy hops");
    if (rc == FRW_OK && weighted && wsum_host)
        rc = check_hip(hipMemcpy(wsum_host, d_ws, sizeof(double) * nb, hipMemcpyDeviceToHost), "copy wsum");
    if (rc == FRW_OK && weighted && wsq_host)
        rc = check_hip(hipMemcpy(wsq_host, d_wq, sizeof(double) * nb, hipMemcpyDeviceToHost), "copy wsq");
    for (void* x : {(void*)d_counts, (void*)d_hops, (void*)d_ws, (void*)d_wq})
        if (x) (void)hipFree(x);
    if (rc == FRW_OK) *r = res;
    return rc;
    FRW_GUARD_END
}

int frw_walk(const frw_tree* t, const frw_walk_params* p, int64_t* counts_host, frw_walk_summary* summary)
{
    int rc = check_walk_params(t, p);
    if (rc != FRW_OK) return rc;
    FRW_GUARD_BEGIN
    const int64_t nb = t->n_rects + 2;
    int64_t* d_counts = nullptr;
    int64_t* d_hops = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<int64_t> counts((size_t)nb);
    int64_t hops = 0;
    float ms = 0.f;
    rc = check_hip(hipMalloc(&d_counts, sizeof(int64_t) * nb), "hipMalloc counts");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&d_hops, sizeof(int64_t)), "hipMalloc hops");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_counts, 0, sizeof(int64_t) * nb), "memset counts");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_hops, 0, sizeof(int64_t)), "memset hops");
    if (rc == FRW_OK) rc = check_hip(hipEventCreate(&e0), "hipEventCreate");
    if (rc == FRW_OK) rc = check_hip(hipEventCreate(&e1), "hipEventCreate");
    if (rc == FRW_OK) rc = check_hip(hipEventRecord(e0, nullptr), "hipEventRecord");
    if (rc == FRW_OK) {
        frw_walk_outputs o{d_counts, d_hops, nullptr, nullptr};
        rc = launch_walk(t, p, &o, nullptr);
    }
    if (rc == FRW_OK) rc = check_hip(hipEventRecord(e1, nullptr), "hipEventRecord");
    if (rc == FRW_OK) rc = check_hip(hipDeviceSynchronize(), "walk");
    if (rc == FRW_OK) rc = check_err_flag(t, nullptr);
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(counts.data(), d_counts, sizeof(int64_t) * nb, hipMemcpyDeviceToHost), "copy counts");
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(&hops, d_hops, sizeof(int64_t), hipMemcpyDeviceToHost), "copy hops");
    if (rc == FRW_OK) rc = check_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (d_counts) (void)hipFree(d_counts);
    if (d_hops) (void)hipFree(d_hops);
    if (rc != FRW_OK) return rc;
    if (counts_host) std::memcpy(counts_host, counts.data(), sizeof(int64_t) * nb);
    if (summary) fill_summary(t, p, counts.data(), hops, ms, summary);
    return FRW_OK;
    FRW_GUARD_END
}

int frw_walk_traffic(const frw_tree* t, const frw_walk_params* p, int64_t out[4])
{
    int rc = check_walk_params(t, p);
    if (rc != FRW_OK) return rc;
    if (!out) return set_error(FRW_E_INVALID, "null output");
    FRW_GUARD_BEGIN
    const int64_t nb = t->n_rects + 2;
    int64_t *d_counts = nullptr, *d_hops = nullptr, *d_tr = nullptr;
    rc = check_hip(hipMalloc(&d_counts, sizeof(int64_t) * nb), "hipMalloc counts");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&d_hops, sizeof(int64_t)), "hipMalloc hops");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&d_tr, 3 * sizeof(int64_t)), "hipMalloc traffic");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_counts, 0, sizeof(int64_t) * nb), "memset");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_hops, 0, sizeof(int64_t)), "memset");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_tr, 0, 3 * sizeof(int64_t)), "memset");
    if (rc == FRW_OK) {
        frw_walk_outputs o{d_counts, d_hops, nullptr, nullptr};
        rc = launch_walk(t, p, &o, nullptr, (unsigned long long*)d_tr);
    }
    if (rc == FRW_OK) rc = check_hip(hipDeviceSynchronize(), "walk");
    if (rc == FRW_OK) rc = check_err_flag(t, nullptr);
    int64_t tr[3] = {0, 0, 0};
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(tr, d_tr, 3 * sizeof(int64_t), hipMemcpyDeviceToHost), "copy");
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(out + 2, d_hops, sizeof(int64_t), hipMemcpyDeviceToHost), "copy");
    if (rc == FRW_OK) {
        out[0] = tr[0];
        out[1] = tr[1];
        out[3] = tr[2];
    }
    if (d_counts) (void)hipFree(d_counts);
    if (d_hops) (void)hipFree(d_hops);
    if (d_tr) (void)hipFree(d_tr);
    return rc;
    FRW_GUARD_END
}

int frw_write_results(const frw_layout* l, const int64_t* counts_host, int64_t n_walks, const char* path)
{
    if (!l || !counts_host || !path || n_walks <= 0) return set_error(FRW_E_INVALID, "bad arguments");
    FRW_GUARD_BEGIN
    return write_results(l->rects, counts_host, n_walks, path);
    FRW_GUARD_END
}

}  // extern "C"
