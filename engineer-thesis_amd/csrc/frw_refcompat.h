// Reference-compatible tree and queries: the reference's CPU quadtree
// (randomWalk/CPU/quadTree.cpp) rebuilt on the host with its insertion
// behaviour and flattened, and its three queries as the reference runs them,
// for walks that reproduce the reference CPU FRW bit for bit, defects
// included:
//   * Q1 follows one root-to-leaf path: at each node the objects are tested,
//     then the first child (UL, UR, LR, LL) whose closed bounds meet the box
//     is taken (the pushed siblings are never popped, quadTree.cpp:153-217);
//   * Q2 descends into the first child whose closed bounds contain p and
//     returns the first object there containing p (quadTree.cpp:240-262);
//   * Q3 is the literal grow/shrink bisection on that Q1 (quadTree.cpp:
//     264-312), the Gaussian surface the 10-step factor search on it with the
//     master ignored by value (quadTree.cpp:315-364).
// Node boxes are FP64 (the reference halves them: x0 + (x1 - x0) / 2).
#pragma once

#include <stdint.h>

#include "frw_device.h"

namespace frw {

struct RefNode {
    double x0, y0, x1, y1;
    int32_t first_child;  // -1 for a leaf; children are 4 consecutive nodes
    int32_t obj_begin;    // objects (original rect ids) in objs[obj_begin, obj_end)
    int32_t obj_end;
    int32_t level;
};

struct RefView {
    const RefNode* nodes;
    const int32_t* objs;   // original rect ids, per node in the reference's list order
    const int4* rects;     // original order
    int4 bounds;
    double dist0;          // 0.05 * max(W, H)
};

// rectsCollision (rectHost.cpp:43-52), closed, FP64 box against a box.
__host__ __device__ inline bool ref_boxes_meet(double ax0, double ay0, double ax1, double ay1, double bx0,
                                               double by0, double bx1, double by1)
{
    return ax1 >= bx0 && bx1 >= ax0 && ay1 >= by0 && by1 >= ay0;
}

__host__ __device__ inline bool ref_q1(const RefView& v, double x0, double y0, double x1, double y1,
                                       const int4* ignore)
{
    if (!in_bounds_box(v.bounds, x0, y0, x1, y1)) return true;
    int ni = 0;
    while (ni >= 0) {
        const RefNode nd = v.nodes[ni];
        for (int k = nd.obj_begin; k < nd.obj_end; ++k) {
            const int4 r = v.rects[v.objs[k]];
            if (ignore && same_rect(r, *ignore)) continue;
            if (ref_boxes_meet(r.x, r.y, r.z, r.w, x0, y0, x1, y1)) return true;
        }
        int next = -1;
        if (nd.first_child >= 0) {
            for (int c = 0; c < 4; ++c) {
                const RefNode ch = v.nodes[nd.first_child + c];
                if (ref_boxes_meet(ch.x0, ch.y0, ch.x1, ch.y1, x0, y0, x1, y1)) {
                    next = nd.first_child + c;
                    break;
                }
            }
        }
        ni = next;
    }
    return false;
}

__host__ __device__ inline int ref_q2(const RefView& v, double px, double py)
{
    int cur = 0, next = -1;
    for (;;) {
        const RefNode nd = v.nodes[cur];
        if (nd.first_child >= 0) {
            for (int c = 0; c < 4; ++c) {
                const RefNode ch = v.nodes[nd.first_child + c];
                if (px >= ch.x0 && py >= ch.y0 && px <= ch.x1 && py <= ch.y1) {
                    next = nd.first_child + c;
                    break;
                }
            }
        }
        for (int k = nd.obj_begin; k < nd.obj_end; ++k) {
            const int4 r = v.rects[v.objs[k]];
            if (px >= (double)r.x && py >= (double)r.y && px <= (double)r.z && py <= (double)r.w) return v.objs[k];
        }
        if (nd.first_child < 0 || next < 0) return -1;
        cur = next;
    }
}

__host__ __device__ inline DBox ref_q3(const RefView& v, double px, double py)
{
    const double MIN_DIST = (double)0.095f;
    DBox o = {px - 1, py - 1, px + 1, py + 1};
    if (ref_q1(v, o.x0, o.y0, o.x1, o.y1, nullptr)) return o;
    const DBox init = {px - 2, py - 2, px + 2, py + 2};
    if (ref_q1(v, init.x0, init.y0, init.x1, init.y1, nullptr)) return init;
    double dist = v.dist0;
    bool max_reached = false;
    for (int it = 0; it < kMaxBisect; ++it) {
        const bool c = ref_q1(v, o.x0, o.y0, o.x1, o.y1, nullptr);
        if (!(c || !max_reached || dist > MIN_DIST)) break;
        if (c) {
            if (dist > MIN_DIST) dist /= 2;
            max_reached = true;
            o.x0 += dist; o.y0 += dist; o.x1 -= dist; o.y1 -= dist;
        } else {
            if (max_reached) dist /= 2;
            o.x0 -= dist; o.y0 -= dist; o.x1 += dist; o.y1 += dist;
        }
    }
    return o;
}

// getAdjustedGaussianFactor (quadTree.cpp:315-349) on the reference query.
__host__ __device__ inline double ref_gauss_factor(const RefView& v, const int4& m, double factor, bool axis_y)
{
    bool dividing = true, first = true;
    double f = factor, lo = 1., hi = factor;
    for (int i = 0; i < 10; i++) {
        const DBox s = axis_y ? gauss_surface(m, 1., f) : gauss_surface(m, f, 1.);
        const bool coll = !in_bounds_box(v.bounds, s.x0, s.y0, s.x1, s.y1) ||
                          ref_q1(v, s.x0, s.y0, s.x1, s.y1, &m);
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
