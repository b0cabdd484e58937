// Internal host-side declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/frw.h"
#include "frw_device.h"
#include "frw_mlp.h"
#include "frw_types.h"

namespace frw {

// One row of a walk launch: a layout tree, its master and a range of walks.
struct WalkJob {
    DevTree tree;
    int4 master;
    double surface[4];             // Gaussian surface, written by the prologue
    unsigned long long* counts;    // n_rects + 2 bins
    unsigned long long* hops;      // 1 element
    int32_t* hit_out;              // optional, indexed by walk within the job
    int32_t* hops_out;
    uint64_t walk_begin;           // global index of the job's first walk
    uint64_t n_walks;
    uint64_t first_walk;           // Philox walk id of the job's first walk
    uint32_t job;                  // Philox counter word 1
    int32_t master_id;             // original index of the master (its hits are counted per lane)
    int32_t start_hint;            // sorted index of a rect at the master's centre (prologue): the
                                   // first hop's nearest-rect search starts from its distance
    int4 start_rect;               // that rect
    unsigned long long* wacc;      // weighted mode: per-bin fixed-point sums of weights and of squared
                                   // weights (4 words per bin), converted at the end of the launch
    double* w_out;                 // weighted mode: optional per-walk weights
    MlpPyramid pyr;                // learned mode: the layout's occupancy pyramid
};

struct WalkLaunch {
    int device = 0;
    int n_jobs = 1;
    uint64_t total_walks = 0;
    int max_depth = 1;
    int max_nodes = 0;  // largest tree's node count (the packed stack needs < kMaxTreeNodes)
    int mode = 0;
    int block = 0;
    int grid = 0;
    uint64_t seed = 1;
    uint32_t max_hops = 10000000;
    const uint32_t* thresholds = nullptr;
    unsigned long long* work = nullptr;
    int* err = nullptr;
    unsigned long long* traffic = nullptr;
    // exact-mode training-set harvest
    double* h_pts = nullptr;
    double* h_prev = nullptr;
    double* h_dstar = nullptr;
    const unsigned long long* h_off = nullptr;  // per-walk offsets of the records (walk order)
    uint64_t h_cap = 0;
    // learned mode
    const MlpNet* net = nullptr;
    unsigned long long* l_stats = nullptr;
    // weighted mode
    const double* wtab = nullptr;
    double eps = 1.0;
    uint32_t wflags = 0;
};

// Optional extras of a single-layout walk launch.
struct WalkExtras {
    double* h_pts = nullptr;
    double* h_prev = nullptr;
    double* h_dstar = nullptr;
    const unsigned long long* h_off = nullptr;  // per-walk offsets of the records (walk order)
    uint64_t h_cap = 0;
    const MlpNet* net = nullptr;
    const MlpPyramid* pyr = nullptr;
    unsigned long long* l_stats = nullptr;
    double* wsum = nullptr;  // weighted mode outputs
    double* wsq = nullptr;
    unsigned long long* wacc = nullptr;  // and the launch's fixed-point accumulators (4 words per bin, zero)
    double* w_out = nullptr;
    double eps = 1.0;
    uint32_t wflags = 0;  // FRW_WEIGHT_* options
};

// Orders the launches that share one handle's device scratch (job slot, work
// counter, error flag, start square): each launch first waits for the
// handle's previous launch, on whatever stream that ran, and then becomes the
// one the next launch waits for.  While the launch stream is being captured
// into a graph neither step is taken (the graph orders its own nodes).
struct LaunchOrder {
    hipEvent_t ev = nullptr;
    bool recorded = false;
};
int order_before(LaunchOrder& o, hipStream_t s, bool* capturing);
int order_after(LaunchOrder& o, hipStream_t s, bool capturing);
void order_free(LaunchOrder& o);

}  // namespace frw

struct frw_layout {
    std::vector<int4> rects;  // file order = original ids
    int4 bounds;              // padded absorbing bounds
};

namespace frw {
struct RefCompat;
}

struct frw_tree {
    int device = 0;
    frw::RefCompat* ref = nullptr;  // reference-compatible index (FRW_TREE_REFCOMPAT)
    int4 bounds{};
    int64_t n_rects = 0;
    int64_t n_nodes = 0;
    int32_t depth = 0;
    int32_t leaf_cap = 16;
    int32_t morton_bits = 0;
    double margin = 0.0;
    double dist0 = 0.0;
    double build_ms = 0.0;
    std::vector<int4> host_rects;  // original order, for master lookup
    int4* d_rects = nullptr;       // tree order
    int32_t* d_ids = nullptr;
    frw::Node* d_nodes = nullptr;
    int32_t* d_start = nullptr;        // start table of the nearest-rect search
    int32_t start_levels = 0;
    int32_t ext2_max = 0;              // largest rect width or height
    uint32_t* d_thresholds = nullptr;  // Green CDF panel thresholds
    double* d_wtab = nullptr;          // weighted estimator: gx | gy | pw | qx | qy (5 x 200), then
                                       // the importance thresholds (2 x 200 u32)
    // walk scratch owned by the tree handle
    unsigned long long* d_work = nullptr;  // walk work-queue counter
    int* d_err = nullptr;                  // kernel error bits
    double* d_surface = nullptr;           // Gaussian surface of the master
    frw::WalkJob* d_job = nullptr;         // job slot of a plain walk
    unsigned long long* d_wacc = nullptr;  // weighted walks' fixed-point accumulators (4 words per bin, kept zero
                                           // between launches), allocated by the first weighted walk
    mutable frw::LaunchOrder order;        // launches on this tree run one after another
};

namespace frw {

// thread-local error reporting
int set_error(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);

// host tables
void green_thresholds(uint32_t T[kNSample]);
// Weighted-estimator tables.  wt (5 x 200 doubles): gx | gy (Green-gradient
// panel weights, unit square, source at the centre, y down) | pw (Green-CDF
// draw probabilities) | qx | qy (draw probabilities of the importance tables);
// ithr (2 x 200 u32): thresholds drawing panels proportional to |gx| and |gy|.
void green_grad_tables(double* wt, uint32_t* ithr);

// device build (frw_build.hip)
int build_tree(frw_tree* t, const std::vector<int4>& rects, int4 bounds, int leaf_cap,
               hipStream_t s);

DevTree dev_view(const frw_tree* t);
int stack_bytes_per_lane(const frw_tree* t);
int lds_per_block(const frw_tree* t);
int block_for_stack(const frw_tree* t, int want, size_t extra_lds);

// query launchers (frw_query.hip)
int launch_q1(const frw_tree* t, const double* boxes, int64_t n, uint8_t* out, hipStream_t s);
int launch_q2(const frw_tree* t, const double* pts, int64_t n, int32_t* ids, hipStream_t s);
int launch_nearest(const frw_tree* t, const double* pts, int64_t n, double* d, hipStream_t s);
int launch_q3(const frw_tree* t, const double* pts, int64_t n, int mode, double* out,
              hipStream_t s);
int launch_gauss(const frw_tree* t, int4 master, double* d_out, hipStream_t s);

// walk launchers (frw_walk.hip)
int plan_walk(const WalkLaunch& L, int* block, int* grid, size_t* lds, bool* runs = nullptr);
int launch_walk_jobs(WalkJob* d_jobs, const WalkLaunch& L, hipStream_t s);
int launch_walk(const frw_tree* t, const frw_walk_params* p, const frw_walk_outputs* o,
                hipStream_t s, unsigned long long* traffic = nullptr, const WalkExtras* extra = nullptr);
// Device scratch of one single-layout launch.
struct WalkScratch {
    WalkJob* job;
    unsigned long long* work;
    int* err;
};
// launch_walk without the tree's launch ordering, on the given scratch
// (the reference-compatible mode always uses the tree's own).
int launch_walk_in(const frw_tree* t, const frw_walk_params* p, const frw_walk_outputs* o, hipStream_t s,
                   const WalkScratch& ws, unsigned long long* traffic = nullptr, const WalkExtras* extra = nullptr);
// Argument checks of the single-layout walk entry points.
int check_walk_params(const frw_tree* t, const frw_walk_params* p);
// The summary of frw_walk from host counts (n_rects + 2) and the hop total.
void fill_summary(const frw_tree* t, const frw_walk_params* p, const int64_t* counts, int64_t hops, double ms,
                  frw_walk_summary* summary);

// reference-compatible index (frw_refcompat.hip)
int build_refcompat(frw_tree* t);
void free_refcompat(frw_tree* t);
int refcompat_info(const frw_tree* t, int64_t out[2]);
int refcompat_gauss(const frw_tree* t, int4 rect, double factor, double out[4]);
// surface: start square (host double[4]); NULL = the master's Gaussian surface
int launch_ref_walk(const frw_tree* t, const frw_walk_params* p, const frw_walk_outputs* o, hipStream_t s,
                    const double* surface = nullptr);
int launch_ref_queries(const frw_tree* t, const double* pts, const double* boxes, int64_t n, const int32_t* ignore,
                       uint8_t* q1, int32_t* q2, double* q3, hipStream_t s);

// learned proposal views (frw_mlp.hip)
MlpPyramid mlp_pyramid(const frw_mlp* m);
MlpNet mlp_net(const frw_mlp* m, float alpha);
const frw_tree* mlp_tree(const frw_mlp* m);

}  // namespace frw
