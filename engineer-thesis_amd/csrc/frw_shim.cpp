// Reference-signature C++ shim (include/randomwalk.h) over the C ABI.
#include "../../include/randomwalk.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/frw.h"

struct QuadTreeManager {
    frw_tree* tree = nullptr;
    frw_layout* layout = nullptr;
    int master = 0;
};

namespace {

uint64_t shim_seed()
{
    const char* s = std::getenv("FRW_SEED");
    return s ? std::strtoull(s, nullptr, 10) : 1ull;
}

RectHost to_rect(const int32_t* q)
{
    return RectHost(point(q[0], q[1]), point(q[2], q[3]));
}

bool to_ints(const std::vector<RectHost>& layer, std::vector<int32_t>* out)
{
    out->resize(layer.size() * 4);
    for (size_t i = 0; i < layer.size(); ++i) {
        const double c[4] = {layer[i].topLeft.x, layer[i].topLeft.y, layer[i].bottomRight.x,
                             layer[i].bottomRight.y};
        for (int k = 0; k < 4; ++k) {
            if (c[k] != (double)(int32_t)c[k]) return false;
            (*out)[4 * i + k] = (int32_t)c[k];
        }
    }
    return true;
}

// Parse + build + walk; returns the fraction of walks not ending on the master
// (getAvgPathLen.cpp:35-114 / getAvgPathLenCUDA.cu:50-80), -1 on error.
floatingPoint run_engine(char* path, int iters, int rect_id, bool measure, int layer, int mode,
                         const char* tag)
{
    frw_layout* l = nullptr;
    if (frw_layout_load(path, layer, &l) != FRW_OK) {
        std::fprintf(stderr, "%s %s\n", tag, frw_last_error_string());
        return -1.0;
    }
    if (rect_id < 0 || rect_id >= frw_layout_size(l) || iters <= 0) {
        std::fprintf(stderr, "%s Incorrect args!\n", tag);
        frw_layout_free(l);
        return -1.0;
    }
    frw_tree* t = nullptr;
    frw_tree_params tp{};
    tp.kind = mode == FRW_MODE_REFCOMPAT ? FRW_TREE_REFCOMPAT : FRW_TREE_MORTON;
    if (frw_tree_build(l, &tp, nullptr, &t) != FRW_OK) {
        std::fprintf(stderr, "%s %s\n", tag, frw_last_error_string());
        frw_layout_free(l);
        return -1.0;
    }
    frw_walk_params p{};
    p.master = rect_id;
    p.mode = mode;
    p.n_walks = iters;
    p.seed = shim_seed();
    std::vector<int64_t> counts((size_t)frw_layout_size(l) + 2);
    frw_walk_summary s{};
    floatingPoint result = -1.0;
    if (frw_walk(t, &p, counts.data(), &s) == FRW_OK) {
        const std::string fn = std::to_string((unsigned long)std::time(nullptr)) + "wyniki.txt";
        if (frw_write_results(l, counts.data(), iters, fn.c_str()) == FRW_OK)
            std::printf("Wynik zapisany do %s\n", fn.c_str());
        if (measure) {
            frw_tree_info info{};
            frw_tree_info_get(t, &info);
            std::printf("%s tree build %.3f ms, walks %.3f ms, %lld hops\n", tag, info.build_ms,
                        s.walk_ms, (long long)s.hops);
        }
        result = s.fraction_other;
    } else {
        std::fprintf(stderr, "%s %s\n", tag, frw_last_error_string());
    }
    frw_tree_free(t);
    frw_layout_free(l);
    return result;
}

}  // namespace

Parser::Parser(const char* delimiter) : delimiter(delimiter ? delimiter : "<<") {}

void Parser::parse(std::string const& fileName)
{
    layers.clear();
    bounds.clear();
    int32_t n_layers = 0;
    if (frw_layout_layer_count(fileName.c_str(), &n_layers) != FRW_OK) {
        std::fprintf(stderr, "Parser: %s\n", frw_last_error_string());
        return;
    }
    for (int k = 0; k < n_layers; ++k) {
        frw_layout* l = nullptr;
        if (frw_layout_load(fileName.c_str(), k, &l) != FRW_OK) {
            std::fprintf(stderr, "Parser: %s\n", frw_last_error_string());
            return;
        }
        std::vector<int32_t> r((size_t)frw_layout_size(l) * 4);
        frw_layout_rects(l, r.data());
        Layer layer;
        for (size_t i = 0; i < r.size() / 4; ++i) layer.push_back(to_rect(&r[4 * i]));
        int32_t b[4];
        frw_layout_bounds(l, b);
        layers.push_back(layer);
        bounds.push_back(to_rect(b));
        frw_layout_free(l);
    }
}

RectHost Parser::getLayerSize(int layerIt)
{
    if (layerIt < 0 || layerIt >= (int)bounds.size()) return RectHost();
    return bounds[layerIt];
}

void runRandomWalk(char* path, int ITER_NUM, int RECT_ID, bool GPU_FLAG, bool measure, int layer)
{
    const auto t1 = std::chrono::high_resolution_clock::now();
    const char* name = GPU_FLAG ? "[GPU]" : "[CPU]";
    const floatingPoint result = GPU_FLAG ? getAvgPathLenCUDA(path, ITER_NUM, RECT_ID, measure, layer)
                                          : getAvgPathLen(path, ITER_NUM, RECT_ID, measure, layer);
    const auto t2 = std::chrono::high_resolution_clock::now();
    if (result < 0) return;
    std::printf("%s Ile sciezek trafiło do innego elementu: %f%%\n", name, result * 100.);
    std::printf("Czas wykonania: %lld ms\n",
                (long long)std::chrono::duration_cast<std::chrono::milliseconds>(t2 - t1).count());
}

floatingPoint getAvgPathLen(char* path, int ITER_NUM, int RECT_ID, bool measure, int layerID)
{
    return run_engine(path, ITER_NUM, RECT_ID, measure, layerID, FRW_MODE_REFCOMPAT, "[CPU]");
}

floatingPoint getAvgPathLenCUDA(char* path, int ITER_NUM, int RECT_ID, bool measure, int layerID)
{
    return run_engine(path, ITER_NUM, RECT_ID, measure, layerID, FRW_MODE_EXACT, "[GPU]");
}

QuadTreeManager* createQuadTree(const std::vector<RectHost>& layer, RectHost const& spaceSize,
                                int RECT_ID, bool doCheck)
{
    (void)spaceSize;  // the bounds are recomputed from the rects (Parser::getLayerSize)
    (void)doCheck;    // the build is checked by its own invariants in the test-suite
    std::vector<int32_t> r;
    if (!to_ints(layer, &r)) {
        std::fprintf(stderr, "createQuadTree: coordinates must be integers\n");
        return nullptr;
    }
    QuadTreeManager* q = new QuadTreeManager;
    q->master = RECT_ID;
    if (frw_layout_from_rects(r.data(), (int64_t)layer.size(), &q->layout) != FRW_OK ||
        frw_tree_build(q->layout, nullptr, nullptr, &q->tree) != FRW_OK) {
        std::fprintf(stderr, "createQuadTree: %s\n", frw_last_error_string());
        freeQuadTreeManager(q);
        return nullptr;
    }
    return q;
}

QuadTreeManager* randomWalkCudaInit(char* path, bool measure, int RECT_ID, int layerID)
{
    (void)measure;
    Parser p("<<");
    p.parse(path);
    if (layerID < 0 || layerID >= p.getLayerCount()) return nullptr;
    return createQuadTree(p.getLayerAt(layerID), p.getLayerSize(layerID), RECT_ID, true);
}

void randomWalkCudaWrapper(int n, QuadTreeManager* q, unsigned int* output, RectHost* rectOutput,
                           unsigned long long randomSeed)
{
    if (!q || n <= 0) return;
    const int64_t nr = frw_layout_size(q->layout);
    int64_t *d_counts = nullptr, *d_hops = nullptr;
    int32_t* d_hit = nullptr;
    hipMalloc(&d_counts, sizeof(int64_t) * (nr + 2));
    hipMalloc(&d_hops, sizeof(int64_t));
    hipMalloc(&d_hit, sizeof(int32_t) * n);
    hipMemset(d_counts, 0, sizeof(int64_t) * (nr + 2));
    hipMemset(d_hops, 0, sizeof(int64_t));
    frw_walk_params p{};
    p.master = q->master;
    p.mode = FRW_MODE_EXACT;
    p.n_walks = n;
    p.seed = randomSeed;
    frw_walk_outputs o{d_counts, d_hops, d_hit, nullptr};
    if (frw_walk_async(q->tree, &p, &o, nullptr) == FRW_OK && hipDeviceSynchronize() == hipSuccess) {
        std::vector<int32_t> hit((size_t)n), rects((size_t)nr * 4);
        hipMemcpy(hit.data(), d_hit, sizeof(int32_t) * n, hipMemcpyDeviceToHost);
        frw_layout_rects(q->layout, rects.data());
        std::vector<unsigned int> flag((size_t)n);
        std::vector<RectHost> out((size_t)n);
        const RectHost master = to_rect(&rects[4 * (size_t)q->master]);
        for (int i = 0; i < n; ++i) {
            out[i] = hit[i] >= 0 ? to_rect(&rects[4 * (size_t)hit[i]]) : RectHost();
            flag[i] = (out[i] == master) ? 0u : 1u;
        }
        if (output) hipMemcpy(output, flag.data(), sizeof(unsigned int) * n, hipMemcpyHostToDevice);
        if (rectOutput) hipMemcpy(rectOutput, out.data(), sizeof(RectHost) * n, hipMemcpyHostToDevice);
    } else {
        std::fprintf(stderr, "randomWalkCudaWrapper: %s\n", frw_last_error_string());
    }
    hipFree(d_counts);
    hipFree(d_hops);
    hipFree(d_hit);
}

bool initCuda(int argc, char** argv)
{
    (void)argc;
    (void)argv;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        std::fprintf(stderr, "initCuda: no HIP device\n");
        return false;
    }
    return hipSetDevice(0) == hipSuccess;
}

void freeQuadTreeManager(QuadTreeManager* qtm)
{
    if (!qtm) return;
    frw_tree_free(qtm->tree);
    frw_layout_free(qtm->layout);
    delete qtm;
}

floatingPoint countAvg(unsigned int output[], int ITER_NUM)
{
    floatingPoint out = 0;
    for (int i = 0; i < ITER_NUM; ++i) out += output[i];
    return out / ITER_NUM;
}
