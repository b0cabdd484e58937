/*
 * frw.h — C ABI of the MI355X (gfx950) floating-random-walk library.
 *
 * Drop-in boundary for the hot path of piogrzej/engineer-thesis: the layout
 * reader, the quadtree build, the three tree queries and the random-walk
 * driver that the reference exposes through randomWalk/randomwalk.h.  Each
 * entry point names the reference interface it replaces (file:line, relative
 * to the reference repository).
 *
 * Conventions
 *   - Every function returns an int status (FRW_OK = 0, negative on error);
 *     frw_last_error_string() describes the last error of the calling thread.
 *     Nothing calls exit() or throws across this boundary.
 *   - Rects are int32 {x0, y0, x1, y1}: top-left (x0,y0), bottom-right
 *     (x1,y1), y grows downward, closed intervals (touching = collision).
 *   - Pointers documented as "device" are HIP device pointers; functions
 *     taking a `stream` (a hipStream_t, NULL = default stream) are
 *     asynchronous on it.  Host pointers are read or written before return.
 *   - Handles are owned by the caller and may be used from one host thread
 *     at a time.  Launches on one tree (or one batch) share that handle's
 *     device scratch, so a launch on any stream first waits for the handle's
 *     previous launch (an event recorded after each one); launches under
 *     graph capture are not ordered this way.
 */
#ifndef FRW_H
#define FRW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FRW_OK = 0,
    FRW_E_INVALID = -1, /* bad argument (null handle, negative size, ...) */
    FRW_E_IO = -2,      /* layout file cannot be opened */
    FRW_E_PARSE = -3,   /* malformed "rect" line, or non-integer coordinate */
    FRW_E_LAYER = -4,   /* no such layer */
    FRW_E_RANGE = -5,   /* master id / walk count / coordinates out of range */
    FRW_E_DEVICE = -6,  /* HIP runtime error */
    FRW_E_NOMEM = -7,   /* allocation failed */
    FRW_E_KERNEL = -8   /* a kernel raised an internal error flag */
};

enum { FRW_MODE_EXACT = 0, FRW_MODE_BISECT = 1, FRW_MODE_LEARNED = 2, FRW_MODE_WEIGHTED = 3, FRW_MODE_REFCOMPAT = 4 };
enum { FRW_TREE_MORTON = 0, FRW_TREE_REFCOMPAT = 1 };
enum { FRW_GEN_UNIFORM = 0, FRW_GEN_CLUSTERED = 1 };
enum { FRW_NSAMPLE = 200 };

typedef struct frw_layout frw_layout;
typedef struct frw_tree frw_tree;

const char* frw_last_error_string(void);
const char* frw_version(void);

/* ---------------------------------------------------------------- layouts */

/* Parser("<<").parse(path) + getLayerAt(layer) + getLayerSize(layer)
 * (randomWalk/CPU/Parser.h:13-30, randomWalk/CPU/Parser.cpp:15-125).  Same
 * file format and line rules; errors are status codes instead of exit(0). */
int frw_layout_load(const char* path, int32_t layer, frw_layout** out);
/* Parser::getLayerCount (randomWalk/CPU/Parser.h:22). */
int frw_layout_layer_count(const char* path, int32_t* n_layers);
/* A layout from n rects (4 int32 each, host memory). */
int frw_layout_from_rects(const int32_t* xyxy, int64_t n, frw_layout** out);

typedef struct {
    int32_t kind;         /* FRW_GEN_UNIFORM or FRW_GEN_CLUSTERED */
    int32_t box;          /* coordinates in [0, box] */
    int64_t n;            /* rect count */
    int32_t side_min;     /* sides ~ U{side_min..side_max} */
    int32_t side_max;
    double cluster_frac;  /* fraction of rects drawn around Gaussian blobs */
    double sigma;         /* blob standard deviation */
    int32_t n_clusters;   /* blob centres, uniform in the box */
    int32_t reserved;
    uint64_t seed;        /* Philox key */
} frw_gen_params;
/* Non-overlapping, non-touching random layout (rejection sampling, the
 * semantics of randomWalk/utils/TestGenerator.cpp:47-110 with Philox draws). */
int frw_layout_generate(const frw_gen_params* p, frw_layout** out);
/* Writes the reference's layout format (TestGenerator.cpp:65,103-106). */
int frw_layout_save(const frw_layout* l, const char* path);
int64_t frw_layout_size(const frw_layout* l);
int frw_layout_rects(const frw_layout* l, int32_t* xyxy_out);
/* Padded absorbing bounds = Parser::getLayerSize (Parser.cpp:51-87). */
int frw_layout_bounds(const frw_layout* l, int32_t out[4]);
/* Replaces the bounds with the caller's, as the reference's
 * Tree(level, count, bounds) takes them (randomWalk/CPU/quadTree.h:51,
 * getAvgPathLen.cpp:63); every rect must lie inside them (inclusive). */
int frw_layout_set_bounds(frw_layout* l, const int32_t b[4]);
void frw_layout_free(frw_layout* l);

/* Panel thresholds of the Green CDF: T[i] = round(intg[i] * 2^32), the table
 * precompute_unit_square_green feeds getIndex (getAvgPathLen.cpp:83-84,
 * mainFunctions.cpp:40-46). */
int frw_green_thresholds(uint32_t out[FRW_NSAMPLE]);

/* ---------------------------------------------------------------- tree */

typedef struct {
    int32_t leaf_cap;     /* max rects per leaf (0 = default 4, <= 64) */
    int32_t kind;         /* FRW_TREE_MORTON, or FRW_TREE_REFCOMPAT: also build the reference's own
                             CPU tree (randomWalk/CPU/quadTree.cpp:73-146) for FRW_MODE_REFCOMPAT */
    int32_t reserved[6];
} frw_tree_params;

typedef struct {
    int64_t n_rects;
    int64_t n_nodes;      /* including dummy quadrants */
    int32_t depth;        /* levels */
    int32_t leaf_cap;
    int32_t bounds[4];
    int32_t morton_bits;  /* K: the root cell is 2^K wide */
    int32_t reserved;
    double build_ms;      /* device time of the last build */
} frw_tree_info;

/* createQuadTree(layer, spaceSize, RECT_ID, doCheck)
 * (randomWalk/Cuda/mainkernels.h:10, randomWalk/Cuda/createquadtree.cu:16-87):
 * a level-synchronous Morton sort-and-split build on the device. */
int frw_tree_build(const frw_layout* l, const frw_tree_params* p, void* stream,
                   frw_tree** out);
int frw_tree_info_get(const frw_tree* t, frw_tree_info* info);
/* Copies the flat tree to host memory: nodes (8 int32 each: x0 y0 x1 y1
 * first_child rbeg rend level), rects in tree order (4 int32 each) and their
 * original ids.  Any pointer may be NULL.  (The structural check of
 * checkQuadTree, randomWalk/Cuda/createquadtree.cu:373-425, runs on this.) */
int frw_tree_export(const frw_tree* t, int32_t* nodes, int32_t* rects, int32_t* ids);
/* freeQuadTreeManager (randomWalk/Cuda/freequadtreemanager.cu:3-11). */
void frw_tree_free(frw_tree* t);

/* ---------------------------------------------------------------- queries
 * Batched parity hooks; inputs/outputs are device arrays of n items. */

/* Q1 Tree::checkCollisions(rect) (randomWalk/CPU/quadTree.cpp:148-198,
 * randomWalk/Cuda/d_quadTree.cu:136-193): boxes = n x {x0,y0,x1,y1} FP64;
 * out[i] = 1 if box i leaves the bounds or touches any rect. */
int frw_query_collide(const frw_tree* t, const double* boxes, int64_t n, uint8_t* out,
                      void* stream);
/* Q2 Tree::checkCollisons(point, rect) (quadTree.cpp:240-262,
 * d_quadTree.cu:28-84): original id of the rect containing each point, or -1. */
int frw_query_point(const frw_tree* t, const double* pts, int64_t n, int32_t* rect_id,
                    void* stream);
/* Signed L-infinity distance from each point to the nearest rect (<= 0 when
 * the point is inside one): the quantity the biggest-square search bisects. */
int frw_query_nearest(const frw_tree* t, const double* pts, int64_t n, double* dist,
                      void* stream);
/* Q3 Tree::drawBiggestSquareAtPoint (quadTree.cpp:264-312,
 * d_quadTree.cu:86-134): boxes_out = n x {x0,y0,x1,y1}.  mode
 * FRW_MODE_EXACT replays the bisection from one nearest-rect search,
 * FRW_MODE_BISECT runs it literally (one Q1 per step). */
int frw_query_maxsquare(const frw_tree* t, const double* pts, int64_t n, int32_t mode,
                        double* boxes_out, void* stream);
/* Tree::creatGaussianSurfFrom(rect(master), 1.5) (quadTree.cpp:352-364);
 * out is host memory. */
int frw_gauss_surface(const frw_tree* t, int32_t master, double out[4]);

/* ---------------------------------------------------------------- walks */

typedef struct {
    int32_t master;      /* RECT_ID */
    int32_t mode;        /* FRW_MODE_* */
    int64_t n_walks;     /* ITER_NUM */
    int64_t first_walk;  /* id of the first walk (Philox counter; ids
                          * first_walk .. first_walk + n_walks - 1 lie below
                          * 2^48: the low word is counter word 0, the high
                          * word is XORed into the stream's domain word) */
    uint64_t seed;       /* Philox key */
    uint32_t job;        /* Philox counter word 1 (layout/master id in a batch) */
    int32_t block;       /* threads per block: 0 = tuned default, else a multiple of 64,
                          * at most 256 (learned walks: 768) */
    int32_t grid;        /* blocks, 0 = fill the device */
    int32_t reserved;
    int64_t max_hops;    /* per-walk hop cap (0 = 10^7); capped walks count as stuck */
} frw_walk_params;

typedef struct {
    int64_t* counts;    /* device, n_rects + 2: hits per original rect id, then
                           escaped, then stuck; accumulated (caller zeroes) */
    int64_t* hops;      /* device, 1 element, accumulated */
    int32_t* hit_out;   /* device, n_walks or NULL: rect id, -1 escaped, -2 stuck */
    int32_t* hops_out;  /* device, n_walks or NULL */
} frw_walk_outputs;

/* randomWalkCudaWrapper(n, tree, output, rectOutput, seed)
 * (randomWalk/Cuda/mainkernels.h:9, randomWalk/Cuda/randomwalk.cu:64-134):
 * n_walks walks from the master's Gaussian surface, one persistent kernel. */
int frw_walk_async(const frw_tree* t, const frw_walk_params* p, const frw_walk_outputs* o,
                   void* stream);

typedef struct {
    int64_t n_walks;
    int64_t hops;
    int64_t n_master;       /* walks that ended on the master */
    int64_t n_escaped;
    int64_t n_stuck;
    double fraction_other;  /* getAvgPathLen's return value: #(hit != master)/N */
    double walk_ms;         /* device time of the walk kernel */
} frw_walk_summary;

/* getAvgPathLenCUDA's walk phase (randomWalk/Cuda/getAvgPathLenCUDA.cu:50-80)
 * as one synchronous call: counts_host (n_rects + 2, may be NULL) and the
 * summary are written on return. */
int frw_walk(const frw_tree* t, const frw_walk_params* p, int64_t* counts_host,
             frw_walk_summary* summary);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the walks of frw_walk (FRW_MODE_EXACT or
 * FRW_MODE_BISECT), synchronous on `stream`, with every scratch buffer (job
 * record, work counter, error flag, hop counter and the n_rects + 2 count
 * bins) in one fixed device array of the library: 8 MiB + 2 KiB per GPU,
 * held until the process ends, bins for at most FRW_NOALLOC_MAX_RECTS rects
 * (FRW_E_RANGE beyond).  No hipMalloc/hipFree and no allocating library call.
 * counts_host (n_rects + 2, required when summary is given) and the summary
 * as for frw_walk; summary->walk_ms is the host wall time of the call.  The
 * tree's device must be current.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other threads, other streams) race. */
enum { FRW_NOALLOC_MAX_RECTS = 1 << 20 };
int frw_walk_noalloc(const frw_tree* t, const frw_walk_params* p, void* stream, int64_t* counts_host,
                     frw_walk_summary* summary);

/* ---------------------------------------------------------------- batches
 * Many layouts and masters, each with its own walk count, in one persistent
 * kernel launch.  No reference counterpart: the reference runs one layout and
 * one master per process (randomWalk/Cuda/getAvgPathLenCUDA.cu:50-80). */

typedef struct frw_batch frw_batch;

typedef struct {
    int32_t tree;        /* index into the trees array */
    int32_t master;      /* RECT_ID in that layout */
    int64_t n_walks;
    int64_t first_walk;  /* Philox walk id of the job's first walk (the
                          * job's ids below 2^48) */
    uint32_t job;        /* Philox counter word 1 */
    int32_t reserved;
} frw_batch_job;

typedef struct {
    int32_t mode;        /* FRW_MODE_EXACT or FRW_MODE_BISECT (learned: frw_batch_walk_learned) */
    int32_t block;       /* 0 = default; a multiple of 64, at most 256 (learned: 768) */
    int32_t grid;        /* 0 = fill the device */
    int32_t use_graph;   /* capture the launch as a HIP graph once, then replay */
    uint64_t seed;
    int64_t max_hops;    /* 0 = 10^7 */
} frw_batch_params;

/* The trees must outlive the batch. */
int frw_batch_create(const frw_tree* const* trees, int32_t n_trees, const frw_batch_job* jobs,
                     int32_t n_jobs, frw_batch** out);
/* Size of the counts array: sum over jobs of (n_rects + 2). */
int64_t frw_batch_counts_size(const frw_batch* b);
/* Offset of job j's bins (n_rects hits, escaped, stuck) in the counts array. */
int frw_batch_counts_offset(const frw_batch* b, int32_t job, int64_t* offset);
/* counts: device int64[frw_batch_counts_size], hops: device int64[n_jobs];
 * both accumulated (caller zeroes).  Asynchronous on `stream`. */
int frw_batch_walk(frw_batch* b, const frw_batch_params* p, int64_t* counts, int64_t* hops,
                   void* stream);
/* Times each launch shape (block x waves per CU) on the whole batch and
 * returns the fastest in *best (walk steps/s in *best_hops_per_s). */
int frw_batch_autotune(frw_batch* b, const frw_batch_params* base, frw_batch_params* best,
                       double* best_hops_per_s);
/* FRW_E_KERNEL if a kernel of this batch raised an error flag. */
int frw_batch_check(const frw_batch* b);
void frw_batch_free(frw_batch* b);

/* ---------------------------------------------------------------- learned proposal
 * A small network that proposes the half-size of each hop's square; one probe
 * of the layout's occupancy pyramid (a complete quadtree of cell occupancy,
 * stored per level as a summed-area table) verifies the proposal, and a
 * failed probe falls back to the exact search.  No reference counterpart (the
 * reference always bisects, randomWalk/CPU/quadTree.cpp:264-312).  Features:
 * the 8x8 occupancy window around the point at a scale set by the previous
 * square, the level and the point's position in its cell; network
 * 67-64-64-64-1 (input padded to 72), GELU, bf16 weights, FP32 accumulation (MFMA). */

typedef struct frw_mlp frw_mlp;

/* Occupancy pyramid of the tree's layout and Philox-initialised weights.
 * max_level (0 = 15) caps the finest pyramid level: a level-l summed-area
 * table takes (2^l + 1)^2 x 4 bytes (4.3 GB at 15, 268 MB at 13).  The tree
 * must outlive the network. */
int frw_mlp_create(const frw_tree* t, uint64_t seed, int32_t max_level, frw_mlp** out);
void frw_mlp_free(frw_mlp* m);
/* Number of FP32 parameters: W1[64][72] b1 W2[64][64] b2 W3[64][64] b3 w4[64] b4. */
int frw_mlp_param_count(void);
int frw_mlp_get_params(const frw_mlp* m, float* out, int32_t n);
int frw_mlp_set_params(frw_mlp* m, const float* in, int32_t n);
/* Copies one occupancy level (2^l x 2^l bits, row-major, 32-bit words). */
int frw_mlp_occupancy(const frw_mlp* m, int32_t level, uint32_t* out_words, int64_t n_words);
/* out = {K0 (2^K0 > bounds extent), Lmax (finest window level), Vmax (finest
 * verification level), finest bitmap level}. */
int frw_mlp_levels(const frw_mlp* m, int32_t out[4]);
/* Verification probe on device arrays: out[i] = 1 when the closed square of
 * half-size d[i] centred at pts[i] is certified free (inside the bounds and
 * no occupied cell in its covering block at the verification level). */
int frw_mlp_certify(const frw_mlp* m, const double* pts, const double* d, int64_t n, uint8_t* out,
                    void* stream);
/* Device arrays: pts n x 2, h n (previous half-size) -> feats n x 72 (FP32). */
int frw_mlp_features(const frw_mlp* m, const double* pts, const double* h, int64_t n, float* feats,
                     void* stream);
/* Device arrays: X n x 72 -> y n, the MFMA inference path of the walk. */
int frw_mlp_forward(const frw_mlp* m, const float* X, int64_t n, float* y, void* stream);
/* One forward + backward pass over device X (n x 72) and targets t (n, a
 * multiple of 32); grad (host, param_count) = gradient of mean((y-t)^2)/2. */
int frw_mlp_gradient(frw_mlp* m, const float* X, const float* t, int64_t n, float* grad, float* loss);
/* Training set from device arrays (pts n x 2, h n, d* n): features and
 * targets log2(d* over cell) are computed on the device. */
int frw_mlp_set_samples(frw_mlp* m, const double* pts, const double* h, const double* dstar, int64_t n);
/* Appends samples whose features come from another network's pyramid
 * (features_of; NULL = m's own): one network trained on several layouts. */
int frw_mlp_add_samples(frw_mlp* m, const frw_mlp* features_of, const double* pts, const double* h,
                        const double* dstar, int64_t n);
/* Adam over the training set (batch a multiple of 32, at most 2^31
 * samples).  Each epoch visits the samples in an order fixed by (seed,
 * epoch): a keyed Feistel bijection of the sample indices.  One optimiser
 * step is two launches (forward and backward of 16-row tiles with per-tile
 * partial gradients, then their sum in a fixed order and the Adam update),
 * so the same samples and seed give the same parameters bit for bit.
 * last_loss = mean batch loss of the last epoch. */
int frw_mlp_train(frw_mlp* m, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* last_loss);
/* Runs exact walks and records, for up to `capacity` hops that did not end
 * the walk, the point, the previous half-size and d* (device arrays pts
 * capacity x 2, h capacity, dstar capacity); *n_out = records written.  The
 * records are in walk-id and hop order (two passes over the same walks), so
 * the same arguments give the same records in every run. */
int frw_walk_harvest(const frw_tree* t, const frw_walk_params* p, double* pts, double* h, double* dstar,
                     int64_t capacity, int64_t* n_out);
/* Walks with the learned proposal (FRW_MODE_LEARNED semantics): a hop whose
 * previous square reaches the finest verification cell asks the network,
 * which proposes alpha * s * 2^y; a probe of the occupancy pyramid that
 * certifies the square free accepts it, otherwise (and on every hop that
 * does not ask) the exact search decides the square.  stats (host, may be
 * NULL) = {accepted proposals, rejected proposals}: their sum is the number
 * of network evaluations, and hops - accepted the number of exact searches.
 * Synchronous. */
int frw_walk_learned(const frw_tree* t, const frw_mlp* m, const frw_walk_params* p,
                     const frw_walk_outputs* o, float alpha, int64_t stats[2]);
/* Asynchronous form on `stream`: no allocation, no synchronisation; stats_dev
 * (device, 2 x uint64, may be NULL) accumulates {accepted, rejected}. */
int frw_walk_learned_async(const frw_tree* t, const frw_mlp* m, const frw_walk_params* p,
                           const frw_walk_outputs* o, float alpha, uint64_t* stats_dev, void* stream);
/* Measurement hook: learned walks with traffic counters compiled in.
 * out = {node records read, rect records read, hops, accepted proposals,
 * rejected proposals, start-table entries read}. */
int frw_walk_learned_traffic(const frw_tree* t, const frw_mlp* m, const frw_walk_params* p, float alpha,
                             int64_t out[6]);

/* Learned walks of a batch.  frw_batch_set_proposal gives each tree of the
 * batch its occupancy pyramid (per_tree[i] must have been created for tree
 * i; only its pyramid is used).  One network `net` proposes for every job of
 * a launch (its features are scale-free; a wave evaluates one weight set).
 * stats_dev as in frw_walk_learned_async; graph capture and the autotuner
 * work as for frw_batch_walk. */
int frw_batch_set_proposal(frw_batch* b, const frw_mlp* const* per_tree, int32_t n_trees);
int frw_batch_walk_learned(frw_batch* b, const frw_batch_params* p, const frw_mlp* net, float alpha,
                           int64_t* counts, int64_t* hops, uint64_t* stats_dev, void* stream);
int frw_batch_autotune_learned(frw_batch* b, const frw_batch_params* base, const frw_mlp* net, float alpha,
                               frw_batch_params* best, double* best_hops_per_s);
/* frw_walk_harvest for every job of the batch in one launch per pass (the
 * jobs' exact walks, p->mode ignored): the records of job 0's walks first, in
 * walk and hop order, then job 1's, and so on; up to `capacity` records in
 * all (the first in that order).  n_per_job (host int64[n_jobs]) receives each
 * job's number of records written.  Synchronous; the next walk launch of the
 * batch uploads its job table again.  No reference counterpart (training
 * data for the learned proposal). */
int frw_batch_harvest(frw_batch* b, const frw_batch_params* p, double* pts, double* h, double* dstar,
                      int64_t capacity, int64_t* n_per_job);

/* Weighted (physical) estimator of the capacitance row of the master (no
 * reference counterpart: the reference computes Green's-function gradients
 * but never uses them, randomWalk/CPU/getAvgPathLen.cpp:83-84).  Each walk
 * starts at a uniform point r on the master's Gaussian surface G (perimeter L,
 * outward normal n), takes the largest empty square at r (half-side h), a
 * Green-CDF panel k of it (the panel centre is the first point) and the
 * weight w = -eps L (n . g_k) / (p_k 2h) (g_k: gradient of the panel
 * probability for the unit square, p_k: its draw probability), then walks as
 * FRW_MODE_EXACT.  wsum[j] / n_walks estimates C_mj (per unit length; the
 * escape bin is the grounded boundary), wsq[j] the second moment.
 * counts/hops as for frw_walk_async; wsum, wsq: device double[n_rects + 2],
 * accumulated (caller zeroes); w_out: optional device double[n_walks].
 * The sums are exact: each weight is rounded once to a multiple of 2^-32
 * (q = rint(w 2^32), |w| < 2^31, else FRW_E_KERNEL), the q and q^2 >> 32 of
 * the walks are added as 128-bit integers in the tree's accumulators, and
 * each launch adds their conversion to wsum and wsq, so the results are the
 * same bits for any launch shape or run.
 * flags: FRW_WEIGHT_IMPORTANCE draws the first panel with probability q_k
 * proportional to |n . g_k| instead of p_k (Yu et al. 2013, importance
 * sampling of the first hop); the weight divides by q_k, so every weight of
 * one side of G has the same magnitude and the estimate's variance drops.
 * The expectation is unchanged. */
/* The weighted estimator's first-hop tables, host side (no device needed):
 * wt = gx | gy | pw | qx | qy (5 x 200 doubles: Green-gradient panel weights
 * of the unit square, draw probabilities of the Green CDF and of the two
 * importance tables), ithr = thresholds drawing panels proportional to |gx|
 * then |gy| (2 x 200 u32). */
void frw_weight_tables(double wt[1000], uint32_t ithr[400]);
enum { FRW_WEIGHT_IMPORTANCE = 1 };
int frw_walk_weighted_async(const frw_tree* t, const frw_walk_params* p, const frw_walk_outputs* o, double eps,
                            uint32_t flags, double* wsum, double* wsq, double* w_out, void* stream);
/* Synchronous form into host arrays (counts int64[n+2], wsum/wsq double[n+2]). */
int frw_walk_weighted(const frw_tree* t, const frw_walk_params* p, double eps, uint32_t flags,
                      int64_t* counts_host, double* wsum_host, double* wsq_host, int64_t* hops);

/* Adaptive stopping (Yu et al. 2013 eq. 32, no reference counterpart: the
 * reference runs a fixed ITER_NUM, randomWalk/CPU/getAvgPathLen.cpp:86-111):
 * walks in rounds until the controlled estimate reaches a 3-sigma relative
 * error of rel_err, each round sized n_rem = n (err^2 / rel_err^2 - 1) from
 * the walks so far (at least 1/16 of them, at most what max_walks leaves).
 * The controlled estimate is f = #(hit != master) / n with
 * sigma^2 = f (1 - f) / n for FRW_MODE_EXACT, FRW_MODE_BISECT and
 * FRW_MODE_REFCOMPAT, and the master's self term C_mm = wsum[m] / n with the
 * sample variance of the weights for FRW_MODE_WEIGHTED (eps, wflags as for
 * frw_walk_weighted).  p->n_walks is the first round (0: 65 536); rounds take
 * consecutive walk ids from p->first_walk, so the tallies equal one launch of
 * the total.  Outputs: counts_host int64[n+2] (required), wsum_host and
 * wsq_host double[n+2] (weighted mode, may be NULL), *r.  Synchronous. */
typedef struct {
    int64_t n_walks;        /* walks run */
    int64_t hops;
    int32_t rounds;         /* launches */
    int32_t converged;      /* 1: the goal was met within max_walks */
    double value;           /* f or C_mm */
    double rel_err_3sigma;  /* 3-sigma relative error of value */
    double wall_ms;         /* host wall time of the rounds (launches, waits, reads) */
} frw_adaptive_result;
int frw_walk_adaptive(const frw_tree* t, const frw_walk_params* p, double rel_err, int64_t max_walks, double eps,
                      uint32_t wflags, int64_t* counts_host, double* wsum_host, double* wsq_host,
                      frw_adaptive_result* r);

/* Reference-compatible mode (FRW_MODE_REFCOMPAT walks on a tree built with
 * FRW_TREE_REFCOMPAT): the reference CPU tree with its insertion behaviour
 * and its queries as written — Q1 along one root-to-leaf path
 * (randomWalk/CPU/quadTree.cpp:148-217), Q2 in the first child containing p
 * (quadTree.cpp:240-262), Q3 the literal bisection (quadTree.cpp:264-312) —
 * so walks reproduce the reference CPU FRW (randomWalk/CPU/mainFunctions.cpp:
 * 48-90) walk for walk for the same random stream.  One lane per walk.
 * Query hook: q1[i] for boxes[i] (x0,y0,x1,y1), ignoring a rect equal to
 * *ignore (host int32[4], may be NULL) as Tree::checkCollisions(r, ignore)
 * does (quadTree.cpp:148); q2[i] / q3[i] (4 doubles) for pts[i]; device
 * arrays; any output may be NULL.  Synchronous. */
int frw_query_refcompat(const frw_tree* t, const double* pts, const double* boxes, int64_t n,
                        const int32_t* ignore, uint8_t* q1, int32_t* q2, double* q3, void* stream);
/* Tree::creatGaussianSurfFrom(r, factor) (quadTree.cpp:351-364) on the
 * reference tree, computed on the device: rect (host int32[4]) scaled per axis
 * by the 10-step factor search (quadTree.cpp:315-349) from `factor`, the rect
 * itself ignored by value; a factor below 1 gives the rect itself, as in the
 * reference.  out: host double[4].  Synchronous. */
int frw_refcompat_gauss_surface(const frw_tree* t, const int32_t rect[4], double factor, double out[4]);
/* RandomWalk(R, tree, pointCount, intg, iterId) (randomWalk/CPU/
 * mainFunctions.cpp:48-90) from an explicit start square (host double[4],
 * normally frw_refcompat_gauss_surface of R at 1.5): p->n_walks walks with
 * ids from p->first_walk (p->seed, p->job, p->max_hops as for frw_walk; master
 * and mode are not used) into the outputs of frw_walk_async. */
int frw_refcompat_walk_from(const frw_tree* t, const double surface[4], const frw_walk_params* p,
                            const frw_walk_outputs* o, void* stream);
/* {nodes, object entries} of the reference-compatible tree. */
int frw_refcompat_info(const frw_tree* t, int64_t out[2]);

/* Measurement hook (no reference counterpart): the same walks as frw_walk
 * with per-hop traffic counters compiled in.  out = {node records read
 * (32 B each), rect records read (16 B each), hops, start-table entries read
 * (4 B each)}.  Slower than frw_walk; used to state the algorithmic bytes per
 * hop. */
int frw_walk_traffic(const frw_tree* t, const frw_walk_params* p, int64_t out[4]);

/* Measurement hook (no reference counterpart): read bandwidth in GB/s of a
 * buffer (the largest power of two of 1 KiB chunks within `bytes`) read
 * `repeats` times over by every CU in scattered, coalesced 1 KiB chunks,
 * after one warm-up pass.  A buffer within one XCD's L2 (4 MiB) gives the
 * L2 read ceiling, one well beyond the 256 MiB Infinity Cache the HBM one. */
int frw_probe_read_bandwidth(int64_t bytes, int32_t repeats, double* gbps);

/* Writes <path> in the reference result format (getAvgPathLen.cpp:13-33):
 * "x0 y0 x1 y1: P%" per hit rect and the escaped percentage. */
int frw_write_results(const frw_layout* l, const int64_t* counts_host, int64_t n_walks,
                      const char* path);

#ifdef __cplusplus
}
#endif
#endif /* FRW_H */
