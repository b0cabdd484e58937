// Sanitizer driver of the library's host code (test infrastructure): the
// layout reader, bounds, validation, generators, writers and Green tables of
// engineer-thesis_amd/csrc/frw_layout.cpp, and the host build of the device
// traversal code (frw_hostsim.cpp: Morton build, queries, walks with both
// stack layouts, FRW_STACK_CHECK on).  Built with AddressSanitizer and
// UndefinedBehaviorSanitizer by `make -C tests/hostsim san` and run by the
// CPU test-suite; exit status 0 and no sanitizer report is the pass condition.
//   san_main <layout> <master> <walks> <tmp dir>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../engineer-thesis_amd/csrc/frw_internal.h"

namespace frw {
int4 padded_bounds(const std::vector<int4>& r);
int read_layout(const char* path, int layer, std::vector<int4>* rects, int* n_layers);
int validate_rects(const std::vector<int4>& r);
int generate_layout(const frw_gen_params* p, std::vector<int4>* out);
int save_layout(const std::vector<int4>& rs, const char* path);
int write_results(const std::vector<int4>& rs, const int64_t* counts, int64_t n_walks, const char* path);

// frw_layout.cpp reports errors through the library's error slot; this
// driver only needs the status codes
int set_error(int code, const std::string& msg)
{
    std::fprintf(stderr, "(status %d: %s)\n", code, msg.c_str());
    return code;
}
}  // namespace frw

extern "C" {
void hs_set_stack(int runs);
void* hs_build(const int32_t* xyxy, int n, const int32_t* bounds, int cap);
void hs_free(void* h);
void hs_info(void* h, int64_t* out);
void hs_export(void* h, int32_t* nodes, int32_t* rects, int32_t* ids);
void hs_queries(void* h, const double* pts, int n, double* dist, int32_t* q2, double* q3e, double* q3b,
                int32_t* n_fallback);
void hs_probe_cost(void* h, const double* pts, int n, double f, int64_t* out);
void hs_gauss(void* h, int master, double* out);
int64_t hs_walk_hinted(void* h, int master, int mode, uint64_t seed, uint32_t job, uint32_t first_walk,
                       int64_t n_walks, uint32_t max_hops, const uint32_t* T, int32_t* hit_out,
                       int32_t* hops_out, int64_t* traffic, int use_hint);
}

using namespace frw;

int main(int argc, char** argv)
{
    if (argc < 5) return 2;
    const int master = std::atoi(argv[2]), n_walks = std::atoi(argv[3]);
    const std::string tmp = argv[4];
    std::vector<int4> r;
    int n_layers = 0;
    if (read_layout(argv[1], 0, &r, &n_layers) != FRW_OK || r.empty() || master >= (int)r.size()) return 3;
    if (read_layout(argv[1], n_layers + 3, &r, nullptr) != FRW_E_LAYER) return 4;  // no such layer
    if (read_layout(argv[1], 0, &r, nullptr) != FRW_OK || validate_rects(r) != FRW_OK) return 5;
    const int4 B = padded_bounds(r);
    // writers round-trip through the reader
    if (save_layout(r, (tmp + "/layout").c_str()) != FRW_OK) return 6;
    std::vector<int4> back;
    if (read_layout((tmp + "/layout").c_str(), 0, &back, nullptr) != FRW_OK || back.size() != r.size()) return 7;
    std::vector<int64_t> counts(r.size() + 2, 1);
    if (write_results(r, counts.data(), (int64_t)r.size() + 2, (tmp + "/wyniki.txt").c_str()) != FRW_OK) return 8;
    // generators, uniform and clustered
    for (int kind : {FRW_GEN_UNIFORM, FRW_GEN_CLUSTERED}) {
        frw_gen_params g{};
        g.kind = kind;
        g.box = 60000;
        g.n = 2000;
        g.side_min = 10;
        g.side_max = 400;
        g.cluster_frac = kind == FRW_GEN_CLUSTERED ? 0.3 : 0.0;
        g.sigma = 3000;
        g.n_clusters = 7;
        g.seed = 5;
        std::vector<int4> gen;
        if (generate_layout(&g, &gen) != FRW_OK || validate_rects(gen) != FRW_OK) return 9;
    }
    uint32_t T[kNSample];
    green_thresholds(T);
    double wt[6 * kNSample];
    green_grad_tables(wt, reinterpret_cast<uint32_t*>(wt + 5 * kNSample));
    // host build of the device traversal code
    const int32_t b[4] = {B.x, B.y, B.z, B.w};
    long long check = 0;
    for (int cap : {1, 8, 64}) {
        void* h = hs_build(reinterpret_cast<const int32_t*>(r.data()), (int)r.size(), b, cap);
        int64_t info[3];
        hs_info(h, info);
        std::vector<int32_t> nodes((size_t)info[0] * 8), rects(r.size() * 4), ids(r.size());
        hs_export(h, nodes.data(), rects.data(), ids.data());
        const int m = 300;
        std::vector<double> pts(2 * m), dist(m), q3e(4 * m), q3b(4 * m);
        std::vector<int32_t> q2(m), nf(m);
        for (int i = 0; i < m; ++i) {
            pts[2 * i] = B.x + (B.z - B.x) * ((i * 37) % m + 0.5) / m;
            pts[2 * i + 1] = B.y + (B.w - B.y) * ((i * 91) % m + 0.5) / m;
        }
        for (int runs : {0, 1}) {
            hs_set_stack(runs);
            hs_queries(h, pts.data(), m, dist.data(), q2.data(), q3e.data(), q3b.data(), nf.data());
            int64_t pc[5] = {0, 0, 0, 0, 0};  // nodes, rects of the search and of the probe, free probes
            hs_probe_cost(h, pts.data(), m, 0.9, pc);
            double g[4];
            hs_gauss(h, master, g);
            std::vector<int32_t> hit(n_walks), hp(n_walks);
            int64_t tr[3] = {0, 0, 0};
            for (int mode : {0, 1})
                for (int hint : {0, 1})
                    check += hs_walk_hinted(h, master, mode, 1, 0, 0, n_walks, 1000000, T, hit.data(), hp.data(),
                                            tr, hint);
            for (int i = 0; i < m; ++i) check += q2[i] + nf[i];
        }
        hs_free(h);
    }
    std::printf("ok %lld\n", check);
    return 0;
}
