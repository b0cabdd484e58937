// Batched walks: many layouts and masters, each with its own walk count, in
// one persistent kernel launch (one job table, one work queue), optionally
// captured once as a HIP graph and replayed; plus the launch-shape autotuner.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "frw_internal.h"

using namespace frw;

struct frw_batch {
    int device = 0;
    std::vector<const frw_tree*> trees;
    std::vector<frw_batch_job> jobs;
    std::vector<int64_t> count_off;  // offset of each job's bins in the counts array
    int64_t counts_size = 0;
    uint64_t total_walks = 0;
    int max_depth = 1;
    int max_nodes = 0;
    WalkJob* d_jobs = nullptr;
    unsigned long long* d_work = nullptr;
    int* d_err = nullptr;
    uint32_t* d_thresholds = nullptr;
    // table currently uploaded / graph currently captured
    int64_t* up_counts = nullptr;
    int64_t* up_hops = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    frw_batch_params graph_params{};
    const frw_mlp* graph_net = nullptr;  // learned-mode arguments the graph was captured with
    float graph_alpha = 0.0f;
    uint64_t* graph_stats = nullptr;
    hipStream_t capture = nullptr;  // private stream the graph is captured on
    std::vector<MlpPyramid> pyrs;   // per tree; occ == nullptr until a proposal is set
    LaunchOrder order;              // launches of this batch run one after another
};

namespace {

void drop_graph(frw_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int upload_jobs(frw_batch* b, int64_t* counts, int64_t* hops, hipStream_t s)
{
    if (b->up_counts == counts && b->up_hops == hops) return FRW_OK;
    std::vector<WalkJob> tab(b->jobs.size());
    uint64_t begin = 0;
    for (size_t j = 0; j < b->jobs.size(); ++j) {
        const frw_batch_job& q = b->jobs[j];
        const frw_tree* t = b->trees[q.tree];
        WalkJob& w = tab[j];
        std::memset(&w, 0, sizeof(w));
        w.tree = dev_view(t);
        w.master = t->host_rects[q.master];
        w.master_id = q.master;
        w.counts = (unsigned long long*)(counts + b->count_off[j]);
        w.hops = (unsigned long long*)(hops + j);
        w.walk_begin = begin;
        w.n_walks = (uint64_t)q.n_walks;
        w.first_walk = (uint64_t)q.first_walk;
        w.job = q.job;
        w.pyr = b->pyrs[q.tree];
        begin += (uint64_t)q.n_walks;
    }
    int rc = check_hip(hipMemcpyAsync(b->d_jobs, tab.data(), sizeof(WalkJob) * tab.size(),
                                      hipMemcpyHostToDevice, s), "upload job table");
    if (rc == FRW_OK) rc = check_hip(hipStreamSynchronize(s), "upload job table");
    if (rc != FRW_OK) return rc;
    b->up_counts = counts;
    b->up_hops = hops;
    drop_graph(b);
    return FRW_OK;
}

WalkLaunch launch_of(const frw_batch* b, const frw_batch_params* p)
{
    WalkLaunch L;
    L.device = b->device;
    L.n_jobs = (int)b->jobs.size();
    L.total_walks = b->total_walks;
    L.max_depth = b->max_depth;
    L.max_nodes = b->max_nodes;
    L.mode = p->mode;
    L.block = p->block;
    L.grid = p->grid;
    L.seed = p->seed;
    L.max_hops = (uint32_t)(p->max_hops > 0 ? p->max_hops : 10000000);
    L.thresholds = b->d_thresholds;
    L.work = b->d_work;
    L.err = b->d_err;
    return L;
}

bool same_params(const frw_batch_params& a, const frw_batch_params& b)
{
    return a.mode == b.mode && a.block == b.block && a.grid == b.grid && a.seed == b.seed &&
           a.max_hops == b.max_hops;
}

}  // namespace

extern "C" {

int frw_batch_create(const frw_tree* const* trees, int32_t n_trees, const frw_batch_job* jobs,
                     int32_t n_jobs, frw_batch** out)
{
    if (!trees || n_trees <= 0 || !jobs || n_jobs <= 0 || !out)
        return set_error(FRW_E_INVALID, "bad batch arguments");
    try {
        int dev = 0;
        const int rc0 = check_hip(hipGetDevice(&dev), "hipGetDevice");
        if (rc0 != FRW_OK) return rc0;
        frw_batch* b = new frw_batch;
        b->device = dev;
        b->trees.assign(trees, trees + n_trees);
        b->pyrs.assign(n_trees, MlpPyramid{});
        b->jobs.assign(jobs, jobs + n_jobs);
        for (int32_t j = 0; j < n_jobs; ++j) {
            const frw_batch_job& q = jobs[j];
            if (q.tree < 0 || q.tree >= n_trees || !trees[q.tree]) {
                delete b;
                return set_error(FRW_E_RANGE, "job " + std::to_string(j) + ": bad tree index");
            }
            const frw_tree* t = trees[q.tree];
            if (t->device != b->device) {
                delete b;
                return set_error(FRW_E_INVALID, "all trees of a batch must live on the current device");
            }
            if (q.master < 0 || q.master >= t->n_rects || q.n_walks <= 0 || q.first_walk < 0 ||
                q.first_walk + q.n_walks > kWalkIdLimit) {
                delete b;
                return set_error(FRW_E_RANGE, "job " + std::to_string(j) + ": bad master or walk range");
            }
            b->count_off.push_back(b->counts_size);
            b->counts_size += t->n_rects + 2;
            b->total_walks += (uint64_t)q.n_walks;
            b->max_depth = std::max(b->max_depth, (int)t->depth);
            b->max_nodes = std::max(b->max_nodes, (int)t->n_nodes);
        }
        uint32_t T[kNSample];
        green_thresholds(T);
        int rc = check_hip(hipMalloc(&b->d_jobs, sizeof(WalkJob) * n_jobs), "hipMalloc jobs");
        if (rc == FRW_OK) rc = check_hip(hipMalloc(&b->d_work, sizeof(unsigned long long)), "hipMalloc work");
        if (rc == FRW_OK) rc = check_hip(hipMalloc(&b->d_err, sizeof(int)), "hipMalloc err");
        if (rc == FRW_OK) rc = check_hip(hipMemset(b->d_err, 0, sizeof(int)), "memset err");
        if (rc == FRW_OK) rc = check_hip(hipMalloc(&b->d_thresholds, sizeof(T)), "hipMalloc thresholds");
        if (rc == FRW_OK) rc = check_hip(hipMemcpy(b->d_thresholds, T, sizeof(T), hipMemcpyHostToDevice), "copy thresholds");
        if (rc != FRW_OK) {
            frw_batch_free(b);
            return rc;
        }
        *out = b;
        return FRW_OK;
    } catch (...) {
        return set_error(FRW_E_NOMEM, "host allocation failed");
    }
}

int64_t frw_batch_counts_size(const frw_batch* b) { return b ? b->counts_size : -1; }

int frw_batch_counts_offset(const frw_batch* b, int32_t job, int64_t* offset)
{
    if (!b || !offset || job < 0 || job >= (int32_t)b->jobs.size())
        return set_error(FRW_E_INVALID, "bad job index");
    *offset = b->count_off[job];
    return FRW_OK;
}

}  // extern "C"

namespace {

// One batched launch (or graph replay) in any mode; net/alpha/stats only for
// the learned mode.
int batch_walk_on(frw_batch* b, const frw_batch_params* p, const frw_mlp* net, float alpha, uint64_t* stats,
                  int64_t* counts, int64_t* hops, hipStream_t s)
{
    int rc = upload_jobs(b, counts, hops, s);
    if (rc != FRW_OK) return rc;
    WalkLaunch L = launch_of(b, p);
    MlpNet mnet{};
    if (p->mode == FRW_MODE_LEARNED) {
        mnet = mlp_net(net, alpha);
        L.net = &mnet;
        L.l_stats = (unsigned long long*)stats;
    }
    if (!p->use_graph) return launch_walk_jobs(b->d_jobs, L, s);
    if (!b->exec || !same_params(b->graph_params, *p) || b->graph_net != net || b->graph_alpha != alpha ||
        b->graph_stats != stats) {
        drop_graph(b);
        int block = 0, grid = 0;
        size_t lds = 0;
        if ((rc = plan_walk(L, &block, &grid, &lds)) != FRW_OK) return rc;
        // capture on a private stream (the legacy default stream cannot be
        // captured); the graph is then launched on the caller's stream
        if (!b->capture &&
            (rc = check_hip(hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking), "capture stream")) != FRW_OK)
            return rc;
        if ((rc = check_hip(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal), "begin capture")) != FRW_OK)
            return rc;
        const int lrc = launch_walk_jobs(b->d_jobs, L, b->capture);
        hipGraph_t g = nullptr;
        rc = check_hip(hipStreamEndCapture(b->capture, &g), "end capture");
        if (lrc != FRW_OK) {
            if (g) (void)hipGraphDestroy(g);
            return lrc;
        }
        if (rc != FRW_OK) return rc;
        b->graph = g;
        if ((rc = check_hip(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "instantiate")) != FRW_OK) {
            drop_graph(b);
            return rc;
        }
        b->graph_params = *p;
        b->graph_net = net;
        b->graph_alpha = alpha;
        b->graph_stats = stats;
    }
    return check_hip(hipGraphLaunch(b->exec, s), "graph launch");
}

// Launches of one batch share its job table, work counter and error flag:
// each is ordered after the batch's previous launch, whatever its stream.
int batch_walk(frw_batch* b, const frw_batch_params* p, const frw_mlp* net, float alpha, uint64_t* stats,
               int64_t* counts, int64_t* hops, hipStream_t s)
{
    if (p->max_hops < 0 || p->max_hops > 0xFFFFFFFFll) return set_error(FRW_E_RANGE, "bad max_hops");
    bool capturing = false;
    int rc = order_before(b->order, s, &capturing);
    if (rc != FRW_OK) return rc;
    rc = batch_walk_on(b, p, net, alpha, stats, counts, hops, s);
    const int rc2 = order_after(b->order, s, capturing);
    return rc != FRW_OK ? rc : rc2;
}

int check_learned_batch(const frw_batch* b, const frw_mlp* net, float alpha)
{
    if (!net) return set_error(FRW_E_INVALID, "learned mode needs a network");
    if (!(alpha > 0.0f)) return set_error(FRW_E_INVALID, "alpha must be positive");
    for (const frw_batch_job& q : b->jobs)
        if (!b->pyrs[q.tree].occ) return set_error(FRW_E_INVALID, "a job's layout has no occupancy pyramid");
    return FRW_OK;
}

// Times each launch shape (block x waves per CU) with `run` and keeps the fastest.
template <class Run>
int autotune(frw_batch* b, const frw_batch_params* base, frw_batch_params* best, double* best_hops_per_s,
             Run run)
{
    int64_t *d_counts = nullptr, *d_hops = nullptr;
    int rc = check_hip(hipMalloc(&d_counts, sizeof(int64_t) * b->counts_size), "hipMalloc counts");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&d_hops, sizeof(int64_t) * b->jobs.size()), "hipMalloc hops");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == FRW_OK) rc = check_hip(hipEventCreate(&e0), "hipEventCreate");
    if (rc == FRW_OK) rc = check_hip(hipEventCreate(&e1), "hipEventCreate");
    double best_rate = -1.0;
    if (base->max_hops < 0 || base->max_hops > 0xFFFFFFFFll) rc = set_error(FRW_E_RANGE, "bad max_hops");
    frw_batch_params cand = *base;
    cand.use_graph = 0;
    int cus = 0;
    if (rc == FRW_OK) rc = check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device), "CUs");
    // (block, waves per CU) shapes: the learned mode's blocks hold the
    // network's weights in LDS, so it also tries blocks of 8 and 12 waves
    std::vector<std::pair<int, int>> shapes;
    for (int blk : {64, 128, 192, 256})
        for (int wv : {4, 8, 16, 32}) shapes.emplace_back(blk, wv);
    if (base->mode == FRW_MODE_LEARNED)
        for (const std::pair<int, int>& sh : {std::make_pair(512, 8), std::make_pair(512, 16), std::make_pair(768, 12),
                                              std::make_pair(768, 24)})
            shapes.push_back(sh);
    for (size_t si = 0; si < shapes.size() && rc == FRW_OK; ++si) {
        cand.block = shapes[si].first;
        cand.grid = std::max(1, cus * shapes[si].second / (shapes[si].first / kWaveSize));
        WalkLaunch L = launch_of(b, &cand);
        int blk = 0, grd = 0;
        size_t lds = 0;
        if (plan_walk(L, &blk, &grd, &lds) != FRW_OK) continue;  // does not fit: skip
        rc = check_hip(hipMemsetAsync(d_counts, 0, sizeof(int64_t) * b->counts_size, nullptr), "memset counts");
        if (rc == FRW_OK) rc = check_hip(hipMemsetAsync(d_hops, 0, sizeof(int64_t) * b->jobs.size(), nullptr), "memset hops");
        if (rc == FRW_OK) rc = check_hip(hipEventRecord(e0, nullptr), "hipEventRecord");
        if (rc == FRW_OK) rc = run(&cand, d_counts, d_hops);
        if (rc == FRW_OK) rc = check_hip(hipEventRecord(e1, nullptr), "hipEventRecord");
        if (rc != FRW_OK) break;
        if ((rc = check_hip(hipDeviceSynchronize(), "autotune run")) != FRW_OK) break;
        float ms = 0.f;
        if ((rc = check_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) != FRW_OK) break;
        std::vector<int64_t> h(b->jobs.size());
        if ((rc = check_hip(hipMemcpy(h.data(), d_hops, sizeof(int64_t) * h.size(), hipMemcpyDeviceToHost),
                            "copy hops")) != FRW_OK)
            break;
        int64_t tot = 0;
        for (int64_t v : h) tot += v;
        const double rate = tot / (ms * 1e-3);
        if (rate > best_rate) {
            best_rate = rate;
            *best = cand;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (d_counts) (void)hipFree(d_counts);
    if (d_hops) (void)hipFree(d_hops);
    b->up_counts = b->up_hops = nullptr;  // the tuning buffers are gone
    if (rc != FRW_OK) return rc;
    if (best_rate < 0) return set_error(FRW_E_RANGE, "no launch shape fits");
    best->use_graph = base->use_graph;
    if (best_hops_per_s) *best_hops_per_s = best_rate;
    return FRW_OK;
}

}  // namespace

extern "C" {

int frw_batch_walk(frw_batch* b, const frw_batch_params* p, int64_t* counts, int64_t* hops,
                   void* stream)
{
    if (!b || !p || !counts || !hops) return set_error(FRW_E_INVALID, "null argument");
    if (p->mode != FRW_MODE_EXACT && p->mode != FRW_MODE_BISECT)
        return set_error(FRW_E_INVALID, "unsupported walk mode");
    return batch_walk(b, p, nullptr, 0.0f, nullptr, counts, hops, (hipStream_t)stream);
}

int frw_batch_harvest(frw_batch* b, const frw_batch_params* p, double* pts, double* h, double* dstar,
                      int64_t capacity, int64_t* n_per_job)
{
    if (!b || !p || !pts || !h || !dstar || capacity <= 0 || !n_per_job)
        return set_error(FRW_E_INVALID, "bad arguments");
    if (p->max_hops < 0 || p->max_hops > 0xFFFFFFFFll) return set_error(FRW_E_RANGE, "bad max_hops");
    frw_batch_params q = *p;
    q.mode = FRW_MODE_EXACT;
    q.use_graph = 0;
    hipStream_t s = nullptr;
    bool capturing = false;
    int rc = order_before(b->order, s, &capturing);
    if (rc != FRW_OK) return rc;
    if (capturing) {
        (void)order_after(b->order, s, capturing);
        return set_error(FRW_E_INVALID, "harvest is synchronous: not under stream capture");
    }
    // Two passes over the same walks of every job, as frw_walk_harvest does
    // for one: the first gives each walk's hop count and end, hence its
    // number of records (hops - 1 when it ended on a rect, hops otherwise);
    // the second writes the launch's walk g (the jobs' walks one after
    // another) at the exclusive prefix sum of those counts.  Records beyond
    // the capacity are dropped, so a small capacity keeps the first jobs'.
    const uint64_t nw = b->total_walks;
    const size_t nj = b->jobs.size();
    char* scratch = nullptr;
    unsigned long long* d_off = nullptr;
    int64_t *d_counts = nullptr, *d_hops = nullptr;
    int32_t *d_hit = nullptr, *d_hp = nullptr;
    std::vector<WalkJob> tab(nj);
    std::vector<int32_t> hit, hp;
    std::vector<unsigned long long> off;
    uint64_t total = 0;
    try {
        rc = check_hip(hipMalloc(&scratch, (size_t)(nw * 8 + (b->counts_size + (int64_t)nj) * 8 + nw * 8)),
                       "hipMalloc harvest scratch");
        if (rc == FRW_OK) {
            d_off = reinterpret_cast<unsigned long long*>(scratch);
            d_counts = reinterpret_cast<int64_t*>(scratch + nw * 8);
            d_hops = d_counts + b->counts_size;
            d_hit = reinterpret_cast<int32_t*>(d_hops + nj);
            d_hp = d_hit + nw;
            rc = check_hip(hipMemset(d_counts, 0, (size_t)(b->counts_size + (int64_t)nj) * 8), "memset");
        }
        if (rc == FRW_OK) {
            // the job table with each job's per-walk outputs
            uint64_t begin = 0;
            for (size_t j = 0; j < nj; ++j) {
                const frw_batch_job& bj = b->jobs[j];
                const frw_tree* t = b->trees[bj.tree];
                WalkJob& w = tab[j];
                std::memset(&w, 0, sizeof(w));
                w.tree = dev_view(t);
                w.master = t->host_rects[bj.master];
                w.master_id = bj.master;
                w.counts = (unsigned long long*)(d_counts + b->count_off[j]);
                w.hops = (unsigned long long*)(d_hops + j);
                w.hit_out = d_hit + begin;
                w.hops_out = d_hp + begin;
                w.walk_begin = begin;
                w.n_walks = (uint64_t)bj.n_walks;
                w.first_walk = (uint64_t)bj.first_walk;
                w.job = bj.job;
                begin += (uint64_t)bj.n_walks;
            }
            // the table of later walk launches is uploaded again
            b->up_counts = b->up_hops = nullptr;
            drop_graph(b);
            rc = check_hip(hipMemcpy(b->d_jobs, tab.data(), sizeof(WalkJob) * nj, hipMemcpyHostToDevice),
                           "upload job table");
        }
        if (rc == FRW_OK) rc = launch_walk_jobs(b->d_jobs, launch_of(b, &q), s);
        if (rc == FRW_OK) rc = check_hip(hipDeviceSynchronize(), "harvest pass 1");
        if (rc == FRW_OK) {
            hit.resize((size_t)nw);
            hp.resize((size_t)nw);
            off.resize((size_t)nw);
            rc = check_hip(hipMemcpy(hit.data(), d_hit, nw * 4, hipMemcpyDeviceToHost), "copy hits");
            if (rc == FRW_OK) rc = check_hip(hipMemcpy(hp.data(), d_hp, nw * 4, hipMemcpyDeviceToHost), "copy hops");
        }
        if (rc == FRW_OK) {
            uint64_t g = 0;
            for (size_t j = 0; j < nj; ++j) {
                const uint64_t first = total;
                for (int64_t w = 0; w < b->jobs[j].n_walks; ++w, ++g) {
                    off[(size_t)g] = total;
                    total += (uint64_t)hp[(size_t)g] - (hit[(size_t)g] >= 0 ? 1u : 0u);
                }
                const uint64_t lo = std::min<uint64_t>(first, (uint64_t)capacity);
                const uint64_t hi = std::min<uint64_t>(total, (uint64_t)capacity);
                n_per_job[j] = (int64_t)(hi - lo);
            }
            rc = check_hip(hipMemcpy(d_off, off.data(), (size_t)nw * 8, hipMemcpyHostToDevice), "copy offsets");
        }
        if (rc == FRW_OK) {
            WalkLaunch L = launch_of(b, &q);
            L.h_pts = pts;
            L.h_prev = h;
            L.h_dstar = dstar;
            L.h_off = d_off;
            L.h_cap = (uint64_t)capacity;
            rc = launch_walk_jobs(b->d_jobs, L, s);
        }
        if (rc == FRW_OK) rc = check_hip(hipDeviceSynchronize(), "harvest pass 2");
    } catch (...) {
        rc = set_error(FRW_E_NOMEM, "host allocation failed");
    }
    if (scratch) (void)hipFree(scratch);
    const int rc2 = order_after(b->order, s, capturing);
    return rc != FRW_OK ? rc : rc2;
}

int frw_batch_set_proposal(frw_batch* b, const frw_mlp* const* per_tree, int32_t n_trees)
{
    if (!b || !per_tree || n_trees != (int32_t)b->trees.size())
        return set_error(FRW_E_INVALID, "one network (pyramid source) per tree is needed");
    for (int32_t i = 0; i < n_trees; ++i)
        if (!per_tree[i] || mlp_tree(per_tree[i]) != b->trees[i])
            return set_error(FRW_E_INVALID, "network " + std::to_string(i) + " was built for another tree");
    for (int32_t i = 0; i < n_trees; ++i) b->pyrs[i] = mlp_pyramid(per_tree[i]);
    b->up_counts = b->up_hops = nullptr;  // re-upload the job table
    drop_graph(b);
    return FRW_OK;
}

int frw_batch_walk_learned(frw_batch* b, const frw_batch_params* p, const frw_mlp* net, float alpha,
                           int64_t* counts, int64_t* hops, uint64_t* stats_dev, void* stream)
{
    if (!b || !p || !counts || !hops) return set_error(FRW_E_INVALID, "null argument");
    int rc = check_learned_batch(b, net, alpha);
    if (rc != FRW_OK) return rc;
    frw_batch_params q = *p;
    q.mode = FRW_MODE_LEARNED;
    return batch_walk(b, &q, net, alpha, stats_dev, counts, hops, (hipStream_t)stream);
}

int frw_batch_autotune_learned(frw_batch* b, const frw_batch_params* base, const frw_mlp* net, float alpha,
                               frw_batch_params* best, double* best_hops_per_s)
{
    if (!b || !base || !best) return set_error(FRW_E_INVALID, "null argument");
    int rc = check_learned_batch(b, net, alpha);
    if (rc != FRW_OK) return rc;
    frw_batch_params q = *base;
    q.mode = FRW_MODE_LEARNED;
    return autotune(b, &q, best, best_hops_per_s, [&](const frw_batch_params* c, int64_t* cn, int64_t* hp) {
        return batch_walk(b, c, net, alpha, nullptr, cn, hp, nullptr);
    });
}

int frw_batch_autotune(frw_batch* b, const frw_batch_params* base, frw_batch_params* best,
                       double* best_hops_per_s)
{
    if (!b || !base || !best) return set_error(FRW_E_INVALID, "null argument");
    if (base->mode != FRW_MODE_EXACT && base->mode != FRW_MODE_BISECT)
        return set_error(FRW_E_INVALID, "unsupported walk mode");
    return autotune(b, base, best, best_hops_per_s, [&](const frw_batch_params* c, int64_t* cn, int64_t* hp) {
        return batch_walk(b, c, nullptr, 0.0f, nullptr, cn, hp, nullptr);
    });
}

int frw_batch_check(const frw_batch* b)
{
    if (!b) return set_error(FRW_E_INVALID, "null batch");
    int err = 0;
    int rc = check_hip(hipMemcpy(&err, b->d_err, sizeof(int), hipMemcpyDeviceToHost), "read err flag");
    if (rc != FRW_OK) return rc;
    if (err) return set_error(FRW_E_KERNEL, "kernel error flag " + std::to_string(err));
    return FRW_OK;
}

void frw_batch_free(frw_batch* b)
{
    if (!b) return;
    drop_graph(b);
    order_free(b->order);
    if (b->capture) (void)hipStreamDestroy(b->capture);
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->d_work) (void)hipFree(b->d_work);
    if (b->d_err) (void)hipFree(b->d_err);
    if (b->d_thresholds) (void)hipFree(b->d_thresholds);
    delete b;
}

}  // extern "C"
