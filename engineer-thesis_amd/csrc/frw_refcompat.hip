// Reference-compatible mode: the reference CPU quadtree rebuilt on the host
// (Tree::insert / Tree::split, randomWalk/CPU/quadTree.cpp:73-146, with its
// redistribution behaviour), flattened to device arrays, and walks and
// queries that run the reference's own query semantics (frw_refcompat.h).
// One lane per walk: this mode exists for literal parity with the reference
// CPU FRW on small layouts, not for speed.
#include <algorithm>
#include <string>
#include <vector>

#include "frw_internal.h"
#include "frw_refcompat.h"

namespace frw {

// Host tree in the reference's shape: children UL, UR, LR, LL created
// together; each node keeps its object list with a head index (objects
// before the head were taken out during a split).
struct RefHostTree {
    struct N {
        double x0, y0, x1, y1;
        int level;
        bool split;
        int child0;
        std::vector<int32_t> obj;
        size_t head;
    };
    std::vector<N> nodes;
    const std::vector<int4>* rects = nullptr;

    static constexpr int kMaxObjects = 4;  // MAX_OBJECTS, randomWalk/defines.h:23
    static constexpr int kMaxLevels = 11;  // MAX_LEVELS, randomWalk/defines.h:20

    int add(int level, double x0, double y0, double x1, double y1)
    {
        N n;
        n.x0 = x0; n.y0 = y0; n.x1 = x1; n.y1 = y1;
        n.level = level;
        n.split = false;
        n.child0 = -1;
        n.head = 0;
        nodes.push_back(std::move(n));
        return (int)nodes.size() - 1;
    }

    size_t size(int ni) const { return nodes[ni].obj.size() - nodes[ni].head; }

    void split(int ni)
    {
        const double x0 = nodes[ni].x0, y0 = nodes[ni].y0, x1 = nodes[ni].x1, y1 = nodes[ni].y1;
        const double sw = x0 + ((x1 - x0) / 2);
        const double sh = y0 + ((y1 - y0) / 2);
        const int lv = nodes[ni].level + 1;
        const int c = add(lv, x0, y0, sw, sh);  // UL
        add(lv, sw, y0, x1, sh);                 // UR
        add(lv, sw, sh, x1, y1);                 // LR
        add(lv, x0, sh, sw, y1);                 // LL
        nodes[ni].child0 = c;
        nodes[ni].split = true;
    }

    bool fits(int ni, int rid) const
    {
        const int4 r = (*rects)[rid];
        const N& n = nodes[ni];
        return (double)r.y >= n.y0 && (double)r.x >= n.x0 && (double)r.w <= n.y1 && (double)r.z <= n.x1;
    }

    bool insert(int ni, int rid)
    {
        if (!fits(ni, rid)) return false;
        if (size(ni) < (size_t)kMaxObjects && !nodes[ni].split) {
            nodes[ni].obj.push_back(rid);
            return true;
        } else if (size(ni) >= (size_t)kMaxObjects && !nodes[ni].split) {
            split(ni);
            // each taken object is offered to the children and appended back
            // to this node; the loop bound shrinks twice when a child keeps it
            int counter = (int)size(ni);
            for (int i = 0; i < counter; ++i) {
                const int32_t tmp = nodes[ni].obj[nodes[ni].head];
                nodes[ni].head++;
                for (int c = 0; c < 4; ++c)
                    if (insert(nodes[ni].child0 + c, tmp)) {
                        counter--;
                        break;
                    }
                nodes[ni].obj.push_back(tmp);
                counter--;
            }
        }
        if (nodes[ni].level < kMaxLevels && nodes[ni].split) {
            for (int c = 0; c < 4; ++c)
                if (insert(nodes[ni].child0 + c, rid)) return true;
            nodes[ni].obj.push_back(rid);
            return true;
        }
        return false;
    }
};

struct RefCompat {
    std::vector<RefNode> h_nodes;
    std::vector<int32_t> h_objs;
    RefNode* d_nodes = nullptr;
    int32_t* d_objs = nullptr;
    int4* d_rects = nullptr;  // original order
};

namespace {

RefView device_view(const frw_tree* t)
{
    RefView v;
    v.nodes = t->ref->d_nodes;
    v.objs = t->ref->d_objs;
    v.rects = t->ref->d_rects;
    v.bounds = t->bounds;
    v.dist0 = t->dist0;
    return v;
}

__global__ void __launch_bounds__(256) ref_walk_kernel(RefView v, const uint32_t* __restrict__ thresholds,
                                                       DBox start, const double* __restrict__ surface,
                                                       uint64_t seed, uint32_t job,
                                                       uint64_t first_walk, int64_t n_walks, uint32_t max_hops,
                                                       int32_t n_rects, unsigned long long* counts,
                                                       unsigned long long* hops, int32_t* hit_out,
                                                       int32_t* hops_out)
{
    __shared__ uint32_t sT[kNSample];
    for (int i = threadIdx.x; i < kNSample; i += blockDim.x) sT[i] = thresholds[i];
    __syncthreads();
    // the start square: computed on the device (surface) or passed by value
    const DBox surf = surface ? DBox{surface[0], surface[1], surface[2], surface[3]} : start;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    unsigned long long my_hops = 0;
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < n_walks;
         w += (int64_t)gridDim.x * blockDim.x) {
        DBox sq = surf;
        uint32_t rnd[4] = {0, 0, 0, 0};
        uint32_t h = 0;
        int result;
        for (;;) {
            if (h >= max_hops) {
                result = -2;
                break;
            }
            if ((h & 3u) == 0u) {
                walk_counter(rnd, first_walk + (uint64_t)w, job, h >> 2, kWalkDomain);
                philox4x32_10(rnd, k0, k1);
            }
            double px, py;
            point_from_panel(sq.x0, sq.y0, sq.x1, sq.y1, sample_panel(sT, rnd[h & 3u]), px, py);
            if (!in_bounds_strict(v.bounds, px, py)) {
                result = -1;
                break;
            }
            sq = ref_q3(v, px, py);
            const int hit = ref_q2(v, px, py);
            h++;
            if (hit >= 0) {
                result = hit;
                break;
            }
        }
        const int bin = result >= 0 ? result : (result == -1 ? n_rects : n_rects + 1);
        atomicAdd(&counts[bin], 1ull);
        my_hops += h;
        if (hit_out) hit_out[w] = result;
        if (hops_out) hops_out[w] = (int32_t)h;
    }
    if (my_hops) atomicAdd(hops, my_hops);
}

// Tree::creatGaussianSurfFrom (quadTree.cpp:351-364), one thread.
__global__ void ref_gauss_kernel(RefView v, int4 rect, double factor, double* out)
{
    const double fx = ref_gauss_factor(v, rect, factor, false);
    const double fy = ref_gauss_factor(v, rect, factor, true);
    const DBox s = gauss_surface(rect, fx, fy);
    out[0] = s.x0;
    out[1] = s.y0;
    out[2] = s.x1;
    out[3] = s.y1;
}

__global__ void ref_query_kernel(RefView v, const double* __restrict__ pts, const double* __restrict__ boxes,
                                 int64_t n, int4 ignore, bool has_ignore, uint8_t* q1, int32_t* q2, double* q3)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (q1 && boxes)
        q1[i] = ref_q1(v, boxes[4 * i], boxes[4 * i + 1], boxes[4 * i + 2], boxes[4 * i + 3],
                       has_ignore ? &ignore : nullptr) ? 1 : 0;
    if (pts) {
        const double px = pts[2 * i], py = pts[2 * i + 1];
        if (q2) q2[i] = ref_q2(v, px, py);
        if (q3) {
            const DBox s = ref_q3(v, px, py);
            q3[4 * i] = s.x0;
            q3[4 * i + 1] = s.y0;
            q3[4 * i + 2] = s.x1;
            q3[4 * i + 3] = s.y1;
        }
    }
}

}  // namespace

int build_refcompat(frw_tree* t)
{
    RefHostTree h;
    h.rects = &t->host_rects;
    h.add(0, (double)t->bounds.x, (double)t->bounds.y, (double)t->bounds.z, (double)t->bounds.w);
    for (int i = 0; i < (int)t->host_rects.size(); ++i) h.insert(0, i);  // file order (createTree)
    RefCompat* rc = new RefCompat;
    rc->h_nodes.resize(h.nodes.size());
    for (size_t i = 0; i < h.nodes.size(); ++i) {
        const RefHostTree::N& n = h.nodes[i];
        RefNode& o = rc->h_nodes[i];
        o.x0 = n.x0; o.y0 = n.y0; o.x1 = n.x1; o.y1 = n.y1;
        o.first_child = n.split ? n.child0 : -1;
        o.level = n.level;
        o.obj_begin = (int32_t)rc->h_objs.size();
        for (size_t k = n.head; k < n.obj.size(); ++k) rc->h_objs.push_back(n.obj[k]);
        o.obj_end = (int32_t)rc->h_objs.size();
    }
    t->ref = rc;
    int e = check_hip(hipMalloc(&rc->d_nodes, sizeof(RefNode) * rc->h_nodes.size()), "hipMalloc ref nodes");
    if (e == FRW_OK)
        e = check_hip(hipMalloc(&rc->d_objs, sizeof(int32_t) * std::max<size_t>(1, rc->h_objs.size())),
                      "hipMalloc ref objects");
    if (e == FRW_OK)
        e = check_hip(hipMalloc(&rc->d_rects, sizeof(int4) * std::max<size_t>(1, t->host_rects.size())),
                      "hipMalloc ref rects");
    if (e == FRW_OK)
        e = check_hip(hipMemcpy(rc->d_nodes, rc->h_nodes.data(), sizeof(RefNode) * rc->h_nodes.size(),
                                hipMemcpyHostToDevice), "copy ref nodes");
    if (e == FRW_OK && !rc->h_objs.empty())
        e = check_hip(hipMemcpy(rc->d_objs, rc->h_objs.data(), sizeof(int32_t) * rc->h_objs.size(),
                                hipMemcpyHostToDevice), "copy ref objects");
    if (e == FRW_OK && !t->host_rects.empty())
        e = check_hip(hipMemcpy(rc->d_rects, t->host_rects.data(), sizeof(int4) * t->host_rects.size(),
                                hipMemcpyHostToDevice), "copy ref rects");
    return e;
}

void free_refcompat(frw_tree* t)
{
    if (!t->ref) return;
    if (t->ref->d_nodes) (void)hipFree(t->ref->d_nodes);
    if (t->ref->d_objs) (void)hipFree(t->ref->d_objs);
    if (t->ref->d_rects) (void)hipFree(t->ref->d_rects);
    delete t->ref;
    t->ref = nullptr;
}

int refcompat_info(const frw_tree* t, int64_t out[2])
{
    if (!t->ref) return set_error(FRW_E_INVALID, "tree was built without the reference-compatible index");
    out[0] = (int64_t)t->ref->h_nodes.size();
    out[1] = (int64_t)t->ref->h_objs.size();
    return FRW_OK;
}

// The start square into t->d_surface (stream-ordered).
static int gauss_to_device(const frw_tree* t, int4 rect, double factor, hipStream_t s)
{
    ref_gauss_kernel<<<1, 1, 0, s>>>(device_view(t), rect, factor, t->d_surface);
    return check_hip(hipGetLastError(), "ref_gauss_kernel");
}

int refcompat_gauss(const frw_tree* t, int4 rect, double factor, double out[4])
{
    if (!t->ref) return set_error(FRW_E_INVALID, "tree was built without the reference-compatible index");
    if (factor < 1.0) {  // the reference logs and returns the rect itself
        out[0] = rect.x; out[1] = rect.y; out[2] = rect.z; out[3] = rect.w;
        return FRW_OK;
    }
    int rc = gauss_to_device(t, rect, factor, nullptr);
    if (rc == FRW_OK)
        rc = check_hip(hipMemcpy(out, t->d_surface, 4 * sizeof(double), hipMemcpyDeviceToHost), "copy surface");
    return rc;
}

int launch_ref_walk(const frw_tree* t, const frw_walk_params* p, const frw_walk_outputs* o, hipStream_t s,
                    const double* surface)
{
    if (!t->ref) return set_error(FRW_E_INVALID, "tree was built without the reference-compatible index");
    if (p->n_walks == 0) return FRW_OK;
    // the start square: the master's Gaussian surface at 1.5 (getAvgPathLen
    // via RandomWalk, mainFunctions.cpp:70), or the caller's
    const DBox start = surface ? DBox{surface[0], surface[1], surface[2], surface[3]} : DBox{0, 0, 0, 0};
    if (!surface) {
        const int rc = gauss_to_device(t, t->host_rects[p->master], 1.5, s);
        if (rc != FRW_OK) return rc;
    }
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, t->device);
    const int64_t want = (p->n_walks + 255) / 256;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)std::max(cus, 1) * 8));
    const uint32_t max_hops = (uint32_t)(p->max_hops > 0 ? p->max_hops : 10000000);
    ref_walk_kernel<<<grid, 256, 0, s>>>(device_view(t), t->d_thresholds, start, surface ? nullptr : t->d_surface,
                                         p->seed, p->job,
                                         (uint64_t)p->first_walk, p->n_walks, max_hops, (int32_t)t->n_rects,
                                         (unsigned long long*)o->counts, (unsigned long long*)o->hops,
                                         o->hit_out, o->hops_out);
    return check_hip(hipGetLastError(), "ref_walk_kernel");
}

int launch_ref_queries(const frw_tree* t, const double* pts, const double* boxes, int64_t n, const int32_t* ignore,
                       uint8_t* q1, int32_t* q2, double* q3, hipStream_t s)
{
    if (!t->ref) return set_error(FRW_E_INVALID, "tree was built without the reference-compatible index");
    if (n == 0) return FRW_OK;
    const int4 ig = ignore ? make_int4(ignore[0], ignore[1], ignore[2], ignore[3]) : make_int4(0, 0, 0, 0);
    ref_query_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(device_view(t), pts, boxes, n, ig,
                                                                 ignore != nullptr, q1, q2, q3);
    int rc = check_hip(hipGetLastError(), "ref_query_kernel");
    if (rc == FRW_OK) rc = check_hip(hipStreamSynchronize(s), "ref queries");
    return rc;
}

}  // namespace frw
