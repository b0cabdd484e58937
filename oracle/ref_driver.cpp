/*
 * Driver around the reference's own CPU sources, compiled into
 * oracle/_ref/libfrwref.so by oracle/Makefile (target `ref`) straight from
 * the reference tree:
 *     randomWalk/CPU/quadTree.cpp, randomWalk/CPU/rectHost.cpp,
 *     randomWalk/CPU/Parser.cpp,   randomWalk/utils/Logger.cpp
 * Test infrastructure only: it pins the C restatement in orc_*.c against the
 * original code, and serves as the reference CPU baseline.  Nothing here is
 * copied from the reference; it only calls the reference's classes.
 *
 * The hop loop of randomWalk/CPU/mainFunctions.cpp:48-90 cannot be compiled
 * here (its header chain needs <cuda.h>), so ref_walk_many() spells out the
 * same loop around the reference Tree and RectHost methods, with Philox
 * panel indices in place of the reference's per-walk /dev/urandom re-seed.
 */
#include "CPU/Parser.h"
#include "CPU/quadTree.h"
#include "CPU/rectHost.h"

#include <cstdint>
#include <cstdio>

#include "oracle.h"

static RectHost mk(const double* q)
{
    return RectHost(point(q[0], q[1]), point(q[2], q[3]));
}

static void put(const RectHost& r, double* q)
{
    q[0] = r.topLeft.x; q[1] = r.topLeft.y; q[2] = r.bottomRight.x; q[3] = r.bottomRight.y;
}

extern "C" {

/* new Tree(0, n, bounds) + createTree (insert in order). */
void* ref_tree_new(const double* xyxy, int n, const double* b)
{
    Tree* t = new Tree(0, n, mk(b));
    for (int i = 0; i < n; ++i) t->insert(mk(xyxy + 4 * i));
    return t;
}

int ref_q1(void* t, const double* box, const double* ignore)
{
    Tree* tr = static_cast<Tree*>(t);
    if (ignore) return tr->checkCollisions(mk(box), mk(ignore)) ? 1 : 0;
    return tr->checkCollisions(mk(box)) ? 1 : 0;
}

int ref_q2(void* t, double px, double py, double* out)
{
    RectHost r;
    bool hit = static_cast<Tree*>(t)->checkCollisons(point(px, py), r);
    put(r, out);
    return hit ? 1 : 0;
}

void ref_q3(void* t, double px, double py, double* out)
{
    put(static_cast<Tree*>(t)->drawBiggestSquareAtPoint(point(px, py)), out);
}

void ref_gauss_from(void* t, const double* r, double* out)
{
    put(static_cast<Tree*>(t)->creatGaussianSurfFrom(mk(r), 1.5), out);
}

int ref_in_bounds_point(void* t, double px, double py)
{
    return static_cast<Tree*>(t)->isInBounds(point(px, py)) ? 1 : 0;
}

void ref_point_from_index(const double* sq, int index, int nsample, double* out)
{
    RectHost s = mk(sq);
    point p = s.getPointFromNindex(index, nsample);
    out[0] = p.x;
    out[1] = p.y;
}

/* Parser("<<").parse(path); getLayerAt / getLayerSize / getLayerCount.
 * Returns the rect count of `layer`, or -1 if the file cannot be opened (the
 * reference would exit(0)), -2 for a missing layer. */
int ref_parse(const char* path, int layer, double* xyxy, int cap, double* bounds, int* n_layers)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return -1;
    std::fclose(f);
    char delim[] = "<<";
    Parser p(delim);
    p.parse(path);
    if (n_layers) *n_layers = p.getLayerCount();
    if (layer < 0 || layer >= p.getLayerCount()) return -2;
    Layer l = p.getLayerAt(layer);
    for (int i = 0; i < (int)l.size() && i < cap; ++i) put(l[i], xyxy + 4 * i);
    if (bounds && !l.empty()) put(p.getLayerSize(layer), bounds);
    return (int)l.size();
}

/* RandomWalk hop loop around the reference Tree; hit rect written by value
 * ((-1,-1,-1,-1) when the walk escaped).  Returns total hops. */
int64_t ref_walk_many(void* t, const double* master, uint64_t seed, uint32_t job,
                      int64_t first_walk, int64_t n_walks, int64_t max_hops,
                      double* hit_out, int32_t* hops_out, int64_t* n_other)
{
    Tree* tr = static_cast<Tree*>(t);
    uint32_t T[ORC_NSAMPLE];
    orc_green_thresholds(T);
    RectHost start = mk(master);
    int64_t total = 0, other = 0;
    for (int64_t i = 0; i < n_walks; ++i) {
        const uint64_t walk = (uint64_t)(first_walk + i);
        RectHost output;
        RectHost square = tr->creatGaussianSurfFrom(start, 1.5);
        int64_t hops = 0;
        bool hit = false;
        for (;;) {
            if (hops >= max_hops) break;
            int k = orc_sample_index(T, orc_walk_u32(seed, walk, job, (uint32_t)hops));
            point p = square.getPointFromNindex(k, ORC_NSAMPLE);
            if (!tr->isInBounds(p)) break;
            square = tr->drawBiggestSquareAtPoint(p);
            hit = tr->checkCollisons(p, output);
            hops++;
            if (hit) break;
        }
        if (!(output == start)) other++;
        if (hit_out) put(output, hit_out + 4 * i);
        if (hops_out) hops_out[i] = (int32_t)hops;
        total += hops;
    }
    if (n_other) *n_other = other;
    return total;
}

}  // extern "C"
