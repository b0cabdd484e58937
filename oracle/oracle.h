/*
 * CPU oracle for the floating-random-walk (FRW) hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker.
 *
 * It restates, in plain C with FP64 geometry, the reference CPU algorithm of
 * piogrzej/engineer-thesis (paths relative to the reference repository):
 *
 *   randomWalk/CPU/rectHost.cpp      geometry: collision, containment,
 *                                    Gaussian surface, panel -> point
 *   randomWalk/CPU/quadTree.cpp      Tree insert/split, checkCollisions,
 *                                    checkCollisons(point),
 *                                    drawBiggestSquareAtPoint,
 *                                    creatGaussianSurfFrom
 *   randomWalk/CPU/mainFunctions.cpp getIndex, RandomWalk hop loop
 *   randomWalk/CPU/getAvgPathLen.cpp tallies and the "other rect" fraction
 *   randomWalk/CPU/Parser.cpp        layout file reader and padded bounds
 *
 * Two deliberate replacements (the reference parts are licence-barred or
 * non-reproducible):
 *   - RNG: Philox4x32-10 (Salmon et al., SC'11) instead of JKISS seeded from
 *     /dev/urandom (randomWalk/CPU/mainFunctions.cpp:58-62).
 *   - Green CDF: re-derived from the Poisson kernel of the square
 *     (series below) instead of randomWalk/green/green.cpp.
 *
 * Three spatial indices answer the same three queries:
 *   ORC_EXACT        brute force over all rectangles (ground truth)
 *   ORC_REFTREE      bug-compatible restatement of the reference Tree:
 *                    single-path rectangle query, insert split quirk,
 *                    first-containing-child point query
 *   ORC_REFTREE_DFS  the same tree with a full depth-first rectangle query
 *                    and an all-containing-children point query
 *
 * Pinning (see tests/): Philox against Random123 known-answer vectors and
 * rocRAND; the Green CDF against closed-form invariants and the reference's
 * randomWalk/rand.data panel sample; geometry, tree queries, the biggest-square
 * search and whole walks against the reference's own CPU sources compiled into
 * oracle/_ref (bit-for-bit).
 */
#ifndef FRW_ORACLE_H
#define FRW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NSAMPLE 200

/* FP64 rectangle, reference RectHost: TL=(x0,y0), BR=(x1,y1), y grows down. */
typedef struct { double x0, y0, x1, y1; } orc_box;

/* ---- RNG ---------------------------------------------------------------- */
void     orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* One uniform u32 for hop `hop` of walk `walk` in job `job`. */
uint32_t orc_walk_u32(uint64_t seed, uint64_t walk, uint32_t job, uint32_t hop);

/* ---- Green CDF -------------------------------------------------------- */
void orc_green_cdf(double intg[ORC_NSAMPLE + 1]);
void orc_green_thresholds(uint32_t T[ORC_NSAMPLE]);
int  orc_sample_index(const uint32_t T[ORC_NSAMPLE], uint32_t u);
int  orc_get_index_fp(const double intg[ORC_NSAMPLE + 1], double r);

/* ---- geometry ----------------------------------------------------------- */
int     orc_collide(const orc_box* a, const orc_box* b);
int     orc_contains(const orc_box* r, double px, double py);
void    orc_point_from_index(const orc_box* sq, int index, int nsample,
                             double* px, double* py);
orc_box orc_gauss_surface(const orc_box* r, double fx, double fy);

/* ---- layout file ---------------------------------------------------------- */
/* Returns the number of rects of `layer` (>= 0) or a negative error; writes at
 * most `cap` rects (4 doubles each) into xyxy and the layer count. */
int  orc_parse(const char* path, int layer, double* xyxy, int cap, int* n_layers);
void orc_bounds(const double* xyxy, int n, orc_box* out);

/* ---- spatial index ------------------------------------------------------ */
enum { ORC_EXACT = 0, ORC_REFTREE = 1, ORC_REFTREE_DFS = 2 };
typedef struct orc_index orc_index;

orc_index* orc_index_new(const double* xyxy, int n, const orc_box* bounds, int kind);
void       orc_index_free(orc_index* ix);
int        orc_index_kind(const orc_index* ix);
orc_box    orc_index_bounds(const orc_index* ix);
/* Tree shape summary: nodes, max level, stored object count (with duplicates). */
void       orc_index_stats(const orc_index* ix, int64_t out[3]);

/* Q1: closed box outside the root bounds, or intersecting a rect whose value
 * differs from *ignore (NULL = the reference's (-1,-1,-1,-1) sentinel). */
int     orc_q1(const orc_index* ix, const orc_box* box, const orc_box* ignore);
/* Q2: index of a rect containing (px,py) (closed), or -1. */
int     orc_q2(const orc_index* ix, double px, double py);
/* Signed L-infinity distance from p to the nearest rect (<= 0: inside); exact
 * index only.  argmin = lowest index among rects at that distance. */
double  orc_nearest(const orc_index* ix, double px, double py, int* argmin);
/* Q3: drawBiggestSquareAtPoint, literal restatement; *n_q1 = Q1 calls. */
orc_box orc_q3(const orc_index* ix, double px, double py, int* n_q1);
void    orc_queries_many(const orc_index* ix, const double* pts, const double* boxes, int64_t n,
                         int64_t n_q3, uint8_t* q1, int32_t* q2, double* q3, int nthreads);
/* Starting surface of a walk from `master` (creatGaussianSurfFrom(r, 1.5)). */
orc_box orc_gauss_from(const orc_index* ix, int master);

/* ---- walks ---------------------------------------------------------------- */
#define ORC_HIT_ESCAPED (-1)
#define ORC_HIT_STUCK   (-2)

int orc_walk_one(const orc_index* ix, const orc_box* surface, const uint32_t T[ORC_NSAMPLE],
                 uint64_t seed, uint32_t job, uint64_t walk, int64_t max_hops,
                 int64_t* hops);

/* Runs walks first_walk .. first_walk+n_walks-1 from `master`.
 * counts[n_rects] = escaped, counts[n_rects+1] = stuck.  hit_out/hops_out may
 * be NULL.  nthreads <= 0: all OpenMP threads.  Returns total hops. */
int64_t orc_walk_many(const orc_index* ix, int master, uint64_t seed, uint32_t job,
                      int64_t first_walk, int64_t n_walks, int64_t max_hops,
                      int64_t* counts, int32_t* hit_out, int32_t* hops_out,
                      int nthreads);

/* ---- learned square-size proposal (orc_mlp.c) -------------------------- */
float orc_bf16(float x);
void  orc_mlp_set_rounding(int on);
int   orc_mlp_level(const orc_box* bounds, double h, double* cell);
void  orc_mlp_features(const double* xyxy, int n, const orc_box* bounds, double px, double py,
                       double h, float out[67]);
int   orc_mlp_vmax(const orc_box* bounds, int n_rects);
int   orc_mlp_vmax_capped(const orc_box* bounds, int n_rects, int cap);
int   orc_mlp_certify(const double* xyxy, int n, const orc_box* bounds, double px, double py, double d);
int   orc_mlp_param_count(void);
void  orc_mlp_forward(const float* params, const float* X, int B, float* y);
float orc_mlp_backward(const float* params, const float* X, const float* t, int B, float* grad);

/* ---- weighted (physical) estimator (orc_weight.c) ------------------------ */
double  orc_square_density(double s, double u, double v);
void    orc_green_grad(double gx[ORC_NSAMPLE], double gy[ORC_NSAMPLE]);
void    orc_panel_centre(double x0, double y0, double a, int k, double* px, double* py);
void    orc_gauss_point(const orc_box* G, uint64_t seed, uint64_t walk, uint32_t job, double* rx, double* ry,
                        double* nx, double* ny, double* L);
void    orc_importance_table(const double g[ORC_NSAMPLE], uint32_t T[ORC_NSAMPLE], double q[ORC_NSAMPLE]);
int     orc_walk_weighted_one(const orc_index* ix, const orc_box* G, const uint32_t T[ORC_NSAMPLE],
                              const double gx[ORC_NSAMPLE], const double gy[ORC_NSAMPLE],
                              const double pw[ORC_NSAMPLE], const uint32_t* tx, const uint32_t* ty,
                              const double* qx, const double* qy, uint64_t seed, uint32_t job, uint64_t walk,
                              int64_t max_hops, double eps, int64_t* hops, double* weight);
int64_t orc_walk_weighted_many(const orc_index* ix, int master, uint64_t seed, uint32_t job,
                               int64_t first_walk, int64_t n_walks, int64_t max_hops, double eps,
                               int importance, int32_t* hit_out, int32_t* hops_out, double* w_out);

#ifdef __cplusplus
}
#endif
#endif
