/*
 * CPU oracle, part 2: spatial indices, queries, biggest square, Gaussian
 * surface and the walk loop.  Test infrastructure only (see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RT_MAX_OBJECTS 4   /* randomWalk/defines.h:23 */
#define RT_MAX_LEVELS 11   /* randomWalk/defines.h:20 */

typedef struct {
    int level;
    orc_box b;
    int split;
    int child[4];
    int* obj;              /* std::list<RectHost> as array [head, len) */
    int head, len, cap;
} rt_node;

struct orc_index {
    int kind;
    int n;
    orc_box* r;
    orc_box bounds;
    rt_node* nodes;
    int nn, ncap;
};

static const orc_box ORC_NONE = { -1, -1, -1, -1 };

static int box_eq(const orc_box* a, const orc_box* b)
{
    return a->x0 == b->x0 && a->y0 == b->y0 && a->x1 == b->x1 && a->y1 == b->y1;
}

/* Tree::isInBounds(rect), quadTree.cpp:38-47: inclusive. */
static int in_bounds_rect(const orc_box* b, const orc_box* r)
{
    return r->y0 >= b->y0 && r->x0 >= b->x0 && r->y1 <= b->y1 && r->x1 <= b->x1;
}

/* Tree::isInBounds(point), quadTree.cpp:49-59: strict. */
static int in_bounds_point(const orc_box* b, double px, double py)
{
    return px > b->x0 && py > b->y0 && px < b->x1 && py < b->y1;
}

/* ---- bug-compatible Tree (randomWalk/CPU/quadTree.cpp) ------------------ */

static int rt_new_node(orc_index* ix, int level, orc_box b)
{
    if (ix->nn == ix->ncap) {
        ix->ncap = ix->ncap ? 2 * ix->ncap : 64;
        ix->nodes = (rt_node*)realloc(ix->nodes, sizeof(rt_node) * ix->ncap);
    }
    rt_node* nd = &ix->nodes[ix->nn];
    memset(nd, 0, sizeof(*nd));
    nd->level = level;
    nd->b = b;
    for (int i = 0; i < 4; ++i) nd->child[i] = -1;
    return ix->nn++;
}

static void rt_push_back(orc_index* ix, int ni, int rid)
{
    rt_node* nd = &ix->nodes[ni];
    if (nd->len == nd->cap) {
        nd->cap = nd->cap ? 2 * nd->cap : 8;
        nd->obj = (int*)realloc(nd->obj, sizeof(int) * nd->cap);
    }
    nd->obj[nd->len++] = rid;
}

static int rt_size(const rt_node* nd) { return nd->len - nd->head; }

/* Tree::split, quadTree.cpp:73-102: children UL, UR, LR, LL. */
static void rt_split(orc_index* ix, int ni)
{
    orc_box b = ix->nodes[ni].b;
    int level = ix->nodes[ni].level;
    double subWidth = b.x0 + ((b.x1 - b.x0) / 2);
    double subHeigth = b.y0 + ((b.y1 - b.y0) / 2);
    orc_box ul = { b.x0, b.y0, subWidth, subHeigth };
    orc_box ur = { subWidth, b.y0, b.x1, subHeigth };
    orc_box lr = { subWidth, subHeigth, b.x1, b.y1 };
    orc_box ll = { b.x0, subHeigth, subWidth, b.y1 };
    int c0 = rt_new_node(ix, level + 1, ul);
    int c1 = rt_new_node(ix, level + 1, ur);
    int c2 = rt_new_node(ix, level + 1, lr);
    int c3 = rt_new_node(ix, level + 1, ll);
    rt_node* nd = &ix->nodes[ni];
    nd->child[0] = c0; nd->child[1] = c1; nd->child[2] = c2; nd->child[3] = c3;
    nd->split = 1;
}

/* Tree::insert, quadTree.cpp:105-146, including the redistribution quirk:
 * every popped object is re-appended to the parent, and the loop counter is
 * decremented twice when a child accepts it. */
static int rt_insert(orc_index* ix, int ni, int rid)
{
    if (!in_bounds_rect(&ix->nodes[ni].b, &ix->r[rid])) return 0;
    if (rt_size(&ix->nodes[ni]) < RT_MAX_OBJECTS && !ix->nodes[ni].split) {
        rt_push_back(ix, ni, rid);
        return 1;
    } else if (rt_size(&ix->nodes[ni]) >= RT_MAX_OBJECTS && !ix->nodes[ni].split) {
        rt_split(ix, ni);
        int counter = rt_size(&ix->nodes[ni]);
        for (int i = 0; i < counter; ++i) {
            rt_node* nd = &ix->nodes[ni];
            int tmp = nd->obj[nd->head];
            nd->head++;
            for (int c = 0; c < 4; ++c)
                if (rt_insert(ix, ix->nodes[ni].child[c], tmp)) { counter--; break; }
            rt_push_back(ix, ni, tmp);
            counter--;
        }
    }
    if (ix->nodes[ni].level < RT_MAX_LEVELS && ix->nodes[ni].split) {
        for (int c = 0; c < 4; ++c)
            if (rt_insert(ix, ix->nodes[ni].child[c], rid)) return 1;
        rt_push_back(ix, ni, rid);
        return 1;
    }
    return 0;
}

/* checkCollisionsWithObjs, quadTree.cpp:219-226. */
static int rt_objs_collide(const orc_index* ix, const rt_node* nd, const orc_box* box,
                           const orc_box* ignore)
{
    for (int k = nd->head; k < nd->len; ++k) {
        const orc_box* o = &ix->r[nd->obj[k]];
        if (!box_eq(o, ignore) && orc_collide(o, box)) return 1;
    }
    return 0;
}

/* checkCollisions, quadTree.cpp:148-198.  The stack pointer is passed by
 * value to addNodesToStack (quadTree.cpp:210-217), so pushed siblings are
 * never seen: the query follows the first colliding child at every level and
 * stops at the first node with no colliding child. */
static int rt_q1_single_path(const orc_index* ix, const orc_box* box, const orc_box* ignore)
{
    int ni = 0;
    while (ni >= 0) {
        const rt_node* nd = &ix->nodes[ni];
        if (rt_objs_collide(ix, nd, box, ignore)) return 1;
        int next = -1;
        if (nd->split)
            for (int i = 0; i < 4; ++i)
                if (orc_collide(&ix->nodes[nd->child[i]].b, box)) { next = nd->child[i]; break; }
        ni = next;
    }
    return 0;
}

static int rt_q1_dfs(const orc_index* ix, const orc_box* box, const orc_box* ignore)
{
    int stack[4 * 64 + 8];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const rt_node* nd = &ix->nodes[stack[--sp]];
        if (rt_objs_collide(ix, nd, box, ignore)) return 1;
        if (nd->split)
            for (int i = 3; i >= 0; --i)
                if (orc_collide(&ix->nodes[nd->child[i]].b, box)) stack[sp++] = nd->child[i];
    }
    return 0;
}

/* checkCollisionObjs, quadTree.cpp:228-238: first object containing p. */
static int rt_objs_contain(const orc_index* ix, const rt_node* nd, double px, double py)
{
    for (int k = nd->head; k < nd->len; ++k)
        if (orc_contains(&ix->r[nd->obj[k]], px, py)) return nd->obj[k];
    return -1;
}

/* checkCollisons(point), quadTree.cpp:240-262: descends into the first child
 * whose closed bounds contain p. */
static int rt_q2_single_path(const orc_index* ix, double px, double py)
{
    int cur = 0, next = -1;
    for (;;) {
        const rt_node* nd = &ix->nodes[cur];
        if (nd->split)
            for (int i = 0; i < 4; ++i)
                if (orc_contains(&ix->nodes[nd->child[i]].b, px, py)) { next = nd->child[i]; break; }
        int hit = rt_objs_contain(ix, nd, px, py);
        if (hit >= 0) return hit;
        if (!nd->split || next < 0) return -1;
        cur = next;
    }
}

static int rt_q2_dfs(const orc_index* ix, double px, double py)
{
    int stack[4 * 64 + 8];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const rt_node* nd = &ix->nodes[stack[--sp]];
        int hit = rt_objs_contain(ix, nd, px, py);
        if (hit >= 0) return hit;
        if (nd->split)
            for (int i = 3; i >= 0; --i)
                if (orc_contains(&ix->nodes[nd->child[i]].b, px, py)) stack[sp++] = nd->child[i];
    }
    return -1;
}

/* ---- index construction -------------------------------------------------- */

orc_index* orc_index_new(const double* xyxy, int n, const orc_box* bounds, int kind)
{
    orc_index* ix = (orc_index*)calloc(1, sizeof(orc_index));
    ix->kind = kind;
    ix->n = n;
    ix->r = (orc_box*)malloc(sizeof(orc_box) * (n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) {
        ix->r[i].x0 = xyxy[4 * i + 0];
        ix->r[i].y0 = xyxy[4 * i + 1];
        ix->r[i].x1 = xyxy[4 * i + 2];
        ix->r[i].y1 = xyxy[4 * i + 3];
    }
    ix->bounds = *bounds;
    if (kind != ORC_EXACT) {
        /* createTree, mainFunctions.cpp:33-38: insert in file order. */
        rt_new_node(ix, 0, *bounds);
        for (int i = 0; i < n; ++i) rt_insert(ix, 0, i);
    }
    return ix;
}

void orc_index_free(orc_index* ix)
{
    if (!ix) return;
    for (int i = 0; i < ix->nn; ++i) free(ix->nodes[i].obj);
    free(ix->nodes);
    free(ix->r);
    free(ix);
}

int orc_index_kind(const orc_index* ix) { return ix->kind; }

orc_box orc_index_bounds(const orc_index* ix) { return ix->bounds; }

void orc_index_stats(const orc_index* ix, int64_t out[3])
{
    int64_t maxlvl = 0, objs = 0;
    for (int i = 0; i < ix->nn; ++i) {
        if (ix->nodes[i].level > maxlvl) maxlvl = ix->nodes[i].level;
        objs += rt_size(&ix->nodes[i]);
    }
    out[0] = ix->nn;
    out[1] = maxlvl;
    out[2] = objs;
}

/* ---- queries --------------------------------------------------------------- */

int orc_q1(const orc_index* ix, const orc_box* box, const orc_box* ignore)
{
    if (!ignore) ignore = &ORC_NONE;
    if (!in_bounds_rect(&ix->bounds, box)) return 1;
    switch (ix->kind) {
    case ORC_REFTREE: return rt_q1_single_path(ix, box, ignore);
    case ORC_REFTREE_DFS: return rt_q1_dfs(ix, box, ignore);
    default:
        for (int i = 0; i < ix->n; ++i)
            if (!box_eq(&ix->r[i], ignore) && orc_collide(&ix->r[i], box)) return 1;
        return 0;
    }
}

/* Signed L-infinity distance from p to rect r: <= 0 exactly when the closed
 * rect contains p (each difference keeps its sign under rounding). */
static double linf_dist(const orc_box* r, double px, double py)
{
    double d = r->x0 - px;
    double t = px - r->x1;
    if (t > d) d = t;
    t = r->y0 - py;
    if (t > d) d = t;
    t = py - r->y1;
    if (t > d) d = t;
    return d;
}

double orc_nearest(const orc_index* ix, double px, double py, int* argmin)
{
    double best = INFINITY;
    int arg = -1;
    for (int i = 0; i < ix->n; ++i) {
        double d = linf_dist(&ix->r[i], px, py);
        if (d < best) { best = d; arg = i; }
    }
    if (argmin) *argmin = arg;
    return best;
}

int orc_q2(const orc_index* ix, double px, double py)
{
    switch (ix->kind) {
    case ORC_REFTREE: return rt_q2_single_path(ix, px, py);
    case ORC_REFTREE_DFS: return rt_q2_dfs(ix, px, py);
    default: {
        /* Canonical answer: deepest containment, then lowest index. */
        double best = INFINITY;
        int arg = -1;
        for (int i = 0; i < ix->n; ++i) {
            double d = linf_dist(&ix->r[i], px, py);
            if (d <= 0 && d < best) { best = d; arg = i; }
        }
        return arg;
    }
    }
}

/* drawBiggestSquareAtPoint, quadTree.cpp:264-312, literally: the p+-1 and
 * p+-2 squares are returned even though they collide, otherwise a grow/shrink
 * bisection from 0.05*max(W,H) down to a step of 0.095f. */
orc_box orc_q3(const orc_index* ix, double px, double py, int* n_q1)
{
    const double MIN_DIST = (double)0.095f;
    int nq = 0;
    orc_box out = { px - 1, py - 1, px + 1, py + 1 };
    orc_box init = { px - 2, py - 2, px + 2, py + 2 };
    double W = ix->bounds.x1 - ix->bounds.x0;
    double H = ix->bounds.y1 - ix->bounds.y0;
    double dist = (H > W) ? H : W;
    dist *= 0.05;
    nq++;
    if (orc_q1(ix, &out, NULL)) { if (n_q1) *n_q1 = nq; return out; }
    nq++;
    if (orc_q1(ix, &init, NULL)) { if (n_q1) *n_q1 = nq; return init; }
    int maxReached = 0, coll;
    while ((nq++, coll = orc_q1(ix, &out, NULL)) || !maxReached || dist > MIN_DIST) {
        if (coll) {
            if (dist > MIN_DIST) dist /= 2;
            maxReached = 1;
            out.x0 += dist; out.y0 += dist; out.x1 -= dist; out.y1 -= dist;
        } else {
            if (maxReached) dist /= 2;
            out.x0 -= dist; out.y0 -= dist; out.x1 += dist; out.y1 += dist;
        }
    }
    if (n_q1) *n_q1 = nq;
    return out;
}

/* getAdjustedGaussianFactor, quadTree.cpp:315-349. */
static double gauss_factor(const orc_index* ix, const orc_box* r, double factor, int axis_y)
{
    int isDividing = 1, isFirstIt = 1;
    double f = factor, lo = 1., hi = factor;
    for (int i = 0; i < 10; i++) {
        orc_box s = axis_y ? orc_gauss_surface(r, 1., f) : orc_gauss_surface(r, f, 1.);
        int coll = (!in_bounds_rect(&ix->bounds, &s) || orc_q1(ix, &s, r));
        if (isFirstIt && !coll) break;
        if ((coll && !isDividing) || (!coll && isDividing)) isDividing = !isDividing;
        if (isDividing)
            f = hi = (lo + hi) / 2.;
        else
            f = lo = (lo + hi) / 2.;
        isFirstIt = 0;
    }
    return f;
}

/* creatGaussianSurfFrom(r, 1.5), quadTree.cpp:352-364. */
orc_box orc_gauss_from(const orc_index* ix, int master)
{
    const orc_box* r = &ix->r[master];
    double fx = gauss_factor(ix, r, 1.5, 0);
    double fy = gauss_factor(ix, r, 1.5, 1);
    return orc_gauss_surface(r, fx, fy);
}

/* ---- walks (RandomWalk, mainFunctions.cpp:48-90) ----------------------------
 * sample a panel -> point on the current square; leaving the strict root
 * bounds ends the walk as escaped (before the hop is counted); otherwise the
 * next square is the biggest square at p and the walk ends when p lies in a
 * rect. */
int orc_walk_one(const orc_index* ix, const orc_box* surface, const uint32_t T[ORC_NSAMPLE],
                 uint64_t seed, uint32_t job, uint64_t walk, int64_t max_hops, int64_t* hops)
{
    orc_box sq = *surface;
    int64_t h = 0;
    for (;;) {
        if (h >= max_hops) { *hops = h; return ORC_HIT_STUCK; }
        uint32_t u = orc_walk_u32(seed, walk, job, (uint32_t)h);
        int k = orc_sample_index(T, u);
        double px, py;
        orc_point_from_index(&sq, k, ORC_NSAMPLE, &px, &py);
        if (!in_bounds_point(&ix->bounds, px, py)) { *hops = h; return ORC_HIT_ESCAPED; }
        sq = orc_q3(ix, px, py, NULL);
        int hit = orc_q2(ix, px, py);
        h++;
        if (hit >= 0) { *hops = h; return hit; }
    }
}

int64_t orc_walk_many(const orc_index* ix, int master, uint64_t seed, uint32_t job,
                      int64_t first_walk, int64_t n_walks, int64_t max_hops,
                      int64_t* counts, int32_t* hit_out, int32_t* hops_out, int nthreads)
{
    uint32_t T[ORC_NSAMPLE];
    orc_green_thresholds(T);
    orc_box surf = orc_gauss_from(ix, master);
    const int nb = ix->n + 2;
    int64_t total = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    int64_t* part = (int64_t*)calloc((size_t)nthreads * nb, sizeof(int64_t));
#pragma omp parallel num_threads(nthreads) reduction(+ : total)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        int64_t* mine = part + (size_t)tid * nb;
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < n_walks; ++i) {
            int64_t hops = 0;
            int hit = orc_walk_one(ix, &surf, T, seed, job, (uint64_t)(first_walk + i),
                                   max_hops, &hops);
            int bin = hit >= 0 ? hit : (hit == ORC_HIT_ESCAPED ? ix->n : ix->n + 1);
            mine[bin]++;
            total += hops;
            if (hit_out) hit_out[i] = hit;
            if (hops_out) hops_out[i] = (int32_t)hops;
        }
    }
    for (int t = 0; t < nthreads; ++t)
        for (int b = 0; b < nb; ++b) counts[b] += part[(size_t)t * nb + b];
    free(part);
    return total;
}

/* Many queries at once (OpenMP over queries), for parity checks at scale:
 * q1[i] = Q1(boxes[i]), q2[i] = Q2(pts[i]) for i < n; q3[i] = Q3(pts[i]) for
 * i < n_q3 (any output may be NULL). */
void orc_queries_many(const orc_index* ix, const double* pts, const double* boxes, int64_t n, int64_t n_q3,
                      uint8_t* q1, int32_t* q2, double* q3, int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        if (q1 && boxes) {
            const orc_box b = {boxes[4 * i], boxes[4 * i + 1], boxes[4 * i + 2], boxes[4 * i + 3]};
            q1[i] = (uint8_t)orc_q1(ix, &b, NULL);
        }
        if (q2) q2[i] = orc_q2(ix, pts[2 * i], pts[2 * i + 1]);
        if (q3 && i < n_q3) {
            const orc_box s = orc_q3(ix, pts[2 * i], pts[2 * i + 1], NULL);
            q3[4 * i] = s.x0;
            q3[4 * i + 1] = s.y0;
            q3[4 * i + 2] = s.x1;
            q3[4 * i + 3] = s.y1;
        }
    }
}
