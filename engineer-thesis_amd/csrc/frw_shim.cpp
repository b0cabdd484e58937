// Reference-signature C++ shim (include/randomwalk.h) over the C ABI.
#include "../../include/randomwalk.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
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

namespace {

// Structural check of a built tree (the role of checkQuadTree,
// randomWalk/Cuda/createquadtree.cu:373-425, on this library's flat layout,
// where rects live only in leaves): the sorted rects are a permutation of the
// layout's; an inner node owns no rects and its subtree's rects are its
// children's, in quadrant order; the leaves' rect ranges partition [0, n) in
// order, each within the leaf capacity unless the leaf is at the last level;
// every node box is the tight bounding box of its subtree's rects, and a
// subtree without rects has an inverted box.
struct TreeCheck {
    const std::vector<int32_t>& nodes;
    const std::vector<int32_t>& rects;
    int64_t n_nodes, next_leaf = 0;
    int leaf_cap, depth;
    const char* error = nullptr;
    int64_t error_node = -1;

    bool fail(const char* what, int64_t i)
    {
        error = what;
        error_node = i;
        return false;
    }

    // Checks the subtree of node i; *empty / box = its rects' bounding box.
    bool subtree(int64_t i, int level, bool* empty, int64_t box[4])
    {
        if (i < 0 || i >= n_nodes) return fail("node index out of range", i);
        if (level > depth) return fail("deeper than the tree's depth", i);
        const int32_t* q = &nodes[(size_t)i * 8];
        const int32_t fc = q[4], rb = q[5], re = q[6];
        box[0] = box[1] = INT64_MAX;
        box[2] = box[3] = INT64_MIN;
        *empty = true;
        if (fc < 0) {
            if (rb < 0 || re < rb || (size_t)re * 4 > rects.size()) return fail("rect range out of bounds", i);
            if (re - rb > leaf_cap && level + 1 < depth) return fail("leaf over capacity", i);
            if (rb != re && rb != next_leaf) return fail("leaf ranges do not partition the rects in order", i);
            if (rb != re) next_leaf = re;
            for (int32_t r = rb; r < re; ++r) {
                box[0] = std::min<int64_t>(box[0], rects[4 * (size_t)r]);
                box[1] = std::min<int64_t>(box[1], rects[4 * (size_t)r + 1]);
                box[2] = std::max<int64_t>(box[2], rects[4 * (size_t)r + 2]);
                box[3] = std::max<int64_t>(box[3], rects[4 * (size_t)r + 3]);
                *empty = false;
            }
        } else {
            if (rb != re) return fail("inner node owns rects", i);
            if (fc + 4 > n_nodes) return fail("child group out of range", i);
            for (int c = 0; c < 4; ++c) {
                bool ce = true;
                int64_t cb[4];
                if (!subtree(fc + c, level + 1, &ce, cb)) return false;
                if (ce) continue;
                *empty = false;
                box[0] = std::min(box[0], cb[0]);
                box[1] = std::min(box[1], cb[1]);
                box[2] = std::max(box[2], cb[2]);
                box[3] = std::max(box[3], cb[3]);
            }
        }
        if (*empty) return (q[0] > q[2] || q[1] > q[3]) ? true : fail("empty subtree with a box", i);
        if (q[0] != box[0] || q[1] != box[1] || q[2] != box[2] || q[3] != box[3])
            return fail("node box is not the bounding box of its subtree", i);
        return true;
    }
};

bool check_tree(frw_tree* t, const std::vector<int32_t>& layout_rects)
{
    frw_tree_info info{};
    if (frw_tree_info_get(t, &info) != FRW_OK) return false;
    const int64_t n = info.n_rects;
    std::vector<int32_t> nodes((size_t)info.n_nodes * 8), rects((size_t)n * 4), ids((size_t)n);
    if (frw_tree_export(t, nodes.data(), rects.data(), ids.data()) != FRW_OK) return false;
    std::vector<char> seen((size_t)n, 0);
    for (int64_t i = 0; i < n; ++i) {
        const int32_t id = ids[i];
        bool ok = id >= 0 && id < n && !seen[id];
        if (ok) seen[id] = 1;
        for (int k = 0; ok && k < 4; ++k) ok = rects[4 * i + k] == layout_rects[4 * (size_t)id + k];
        if (!ok) {
            std::fprintf(stderr, "CreateTree: sorted rect %lld is not a distinct layout rect\n", (long long)i);
            return false;
        }
    }
    TreeCheck c{nodes, rects, info.n_nodes, 0, info.leaf_cap, info.depth};
    bool empty = true;
    int64_t box[4];
    bool ok = c.subtree(0, 0, &empty, box);
    if (ok && c.next_leaf != n) ok = c.fail("leaves do not cover every rect", -1);
    if (!ok) std::fprintf(stderr, "CreateTree: error in node %lld: %s\n", (long long)c.error_node, c.error);
    return ok;
}

}  // namespace

QuadTreeManager* createQuadTree(const std::vector<RectHost>& layer, RectHost const& spaceSize,
                                int RECT_ID, bool doCheck)
{
    std::vector<int32_t> r;
    int32_t b[4];
    const std::vector<RectHost> space{spaceSize};
    std::vector<int32_t> bi;
    if (!to_ints(layer, &r) || !to_ints(space, &bi)) {
        std::fprintf(stderr, "createQuadTree: coordinates must be integers\n");
        return nullptr;
    }
    for (int k = 0; k < 4; ++k) b[k] = bi[k];
    QuadTreeManager* q = new QuadTreeManager;
    q->master = RECT_ID;
    // the root bounds are the caller's spaceSize (createquadtree.cu:32)
    if (frw_layout_from_rects(r.data(), (int64_t)layer.size(), &q->layout) != FRW_OK ||
        frw_layout_set_bounds(q->layout, b) != FRW_OK ||
        frw_tree_build(q->layout, nullptr, nullptr, &q->tree) != FRW_OK) {
        std::fprintf(stderr, "createQuadTree: %s\n", frw_last_error_string());
        freeQuadTreeManager(q);
        return nullptr;
    }
    if (doCheck) {
        // createquadtree.cu:61-74: the tree is checked and the outcome logged
        std::fprintf(stderr, "Sprawdzanie drzewa.\n");
        std::fprintf(stderr, check_tree(q->tree, r) ? "Stworzono pomyslnie\n" : "Blad tworzenia drzewa\n");
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
    if (hipMalloc(&d_counts, sizeof(int64_t) * (nr + 2)) != hipSuccess ||
        hipMalloc(&d_hops, sizeof(int64_t)) != hipSuccess || hipMalloc(&d_hit, sizeof(int32_t) * n) != hipSuccess ||
        hipMemset(d_counts, 0, sizeof(int64_t) * (nr + 2)) != hipSuccess ||
        hipMemset(d_hops, 0, sizeof(int64_t)) != hipSuccess) {
        std::fprintf(stderr, "randomWalkCudaWrapper: device allocation failed\n");
        (void)hipFree(d_counts);
        (void)hipFree(d_hops);
        (void)hipFree(d_hit);
        return;
    }
    frw_walk_params p{};
    p.master = q->master;
    p.mode = FRW_MODE_EXACT;
    p.n_walks = n;
    p.seed = randomSeed;
    frw_walk_outputs o{d_counts, d_hops, d_hit, nullptr};
    std::vector<int32_t> hit((size_t)n), rects((size_t)nr * 4);
    if (frw_walk_async(q->tree, &p, &o, nullptr) == FRW_OK && hipDeviceSynchronize() == hipSuccess &&
        hipMemcpy(hit.data(), d_hit, sizeof(int32_t) * n, hipMemcpyDeviceToHost) == hipSuccess) {
        frw_layout_rects(q->layout, rects.data());
        std::vector<unsigned int> flag((size_t)n);
        std::vector<RectHost> out((size_t)n);
        const RectHost master = to_rect(&rects[4 * (size_t)q->master]);
        for (int i = 0; i < n; ++i) {
            out[i] = hit[i] >= 0 ? to_rect(&rects[4 * (size_t)hit[i]]) : RectHost();
            flag[i] = (out[i] == master) ? 0u : 1u;
        }
        if ((output && hipMemcpy(output, flag.data(), sizeof(unsigned int) * n, hipMemcpyHostToDevice) != hipSuccess) ||
            (rectOutput && hipMemcpy(rectOutput, out.data(), sizeof(RectHost) * n, hipMemcpyHostToDevice) != hipSuccess))
            std::fprintf(stderr, "randomWalkCudaWrapper: copy to the output arrays failed\n");
    } else {
        std::fprintf(stderr, "randomWalkCudaWrapper: %s\n", frw_last_error_string());
    }
    (void)hipFree(d_counts);
    (void)hipFree(d_hops);
    (void)hipFree(d_hit);
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

// ---------------------------------------------------------------- Tree
namespace {

// Device staging of one Tree query or walk.
struct TreeScratch {
    double* in = nullptr;      // box or point
    double* q3 = nullptr;      // 4 doubles
    int32_t* q2 = nullptr;
    uint8_t* q1 = nullptr;
    int64_t* counts = nullptr; // n_rects + 2
    int64_t* hops = nullptr;
    int32_t* hit = nullptr;
    int32_t* hop_n = nullptr;
};

void free_scratch(TreeScratch* s)
{
    if (!s) return;
    for (void* p : {(void*)s->in, (void*)s->q3, (void*)s->q2, (void*)s->q1, (void*)s->counts, (void*)s->hops,
                    (void*)s->hit, (void*)s->hop_n})
        if (p) (void)hipFree(p);
    delete s;
}

bool is_int(double v) { return v == (double)(int32_t)v; }

bool rect_ints(const RectHost& r, int32_t out[4])
{
    const double c[4] = {r.topLeft.x, r.topLeft.y, r.bottomRight.x, r.bottomRight.y};
    for (int k = 0; k < 4; ++k) {
        if (!is_int(c[k])) return false;
        out[k] = (int32_t)c[k];
    }
    return true;
}

bool dev_ok(hipError_t e, const char* what)
{
    if (e == hipSuccess) return true;
    std::fprintf(stderr, "Tree: %s: %s\n", what, hipGetErrorString(e));
    return false;
}

}  // namespace

Tree::Tree(int pLevel, int nodeSCount, RectHost const& b) : bounds(b)
{
    (void)pLevel;      // the root is level 0 (quadTree.cpp:5-17); sub-trees are internal here
    (void)nodeSCount;  // a capacity hint in the reference
}

Tree::~Tree() { clear(); }

bool Tree::isInBounds(point const& p)
{
    return p.x > bounds.topLeft.x && p.y > bounds.topLeft.y && p.x < bounds.bottomRight.x &&
           p.y < bounds.bottomRight.y;
}

bool Tree::isInBounds(RectHost const& r)
{
    return r.topLeft.y >= bounds.topLeft.y && r.topLeft.x >= bounds.topLeft.x &&
           r.bottomRight.y <= bounds.bottomRight.y && r.bottomRight.x <= bounds.bottomRight.x;
}

void Tree::clear()
{
    rects.clear();
    if (tree) frw_tree_free(tree);
    if (layout) frw_layout_free(layout);
    free_scratch(static_cast<TreeScratch*>(scratch));
    tree = nullptr;
    layout = nullptr;
    scratch = nullptr;
    dirty = true;
}

bool Tree::insert(RectHost const& r)
{
    if (!isInBounds(r)) return false;
    rects.push_back(r);
    dirty = true;
    return true;
}

frw_tree* Tree::deviceTree()
{
    if (!dirty) return tree;
    const Layer keep = rects;
    clear();
    rects = keep;
    std::vector<int32_t> r;
    int32_t b[4];
    if (rects.empty() || !to_ints(rects, &r) || !rect_ints(bounds, b)) {
        std::fprintf(stderr, "Tree: needs at least one rect and integer coordinates\n");
        return nullptr;
    }
    frw_tree_params tp{};
    tp.kind = FRW_TREE_REFCOMPAT;
    if (frw_layout_from_rects(r.data(), (int64_t)rects.size(), &layout) != FRW_OK ||
        frw_layout_set_bounds(layout, b) != FRW_OK || frw_tree_build(layout, &tp, nullptr, &tree) != FRW_OK) {
        std::fprintf(stderr, "Tree: %s\n", frw_last_error_string());
        clear();
        rects = keep;
        return nullptr;
    }
    TreeScratch* s = new TreeScratch;
    scratch = s;
    const size_t nb = rects.size() + 2;
    if (!dev_ok(hipMalloc(&s->in, 4 * sizeof(double)), "hipMalloc") ||
        !dev_ok(hipMalloc(&s->q3, 4 * sizeof(double)), "hipMalloc") ||
        !dev_ok(hipMalloc(&s->q2, sizeof(int32_t)), "hipMalloc") ||
        !dev_ok(hipMalloc(&s->q1, sizeof(uint8_t)), "hipMalloc") ||
        !dev_ok(hipMalloc(&s->counts, nb * sizeof(int64_t)), "hipMalloc") ||
        !dev_ok(hipMalloc(&s->hops, sizeof(int64_t)), "hipMalloc") ||
        !dev_ok(hipMalloc(&s->hit, sizeof(int32_t)), "hipMalloc") ||
        !dev_ok(hipMalloc(&s->hop_n, sizeof(int32_t)), "hipMalloc") ||
        !dev_ok(hipMemset(s->counts, 0, nb * sizeof(int64_t)), "hipMemset") ||
        !dev_ok(hipMemset(s->hops, 0, sizeof(int64_t)), "hipMemset")) {
        clear();
        rects = keep;
        return nullptr;
    }
    dirty = false;
    return tree;
}

bool Tree::checkCollisions(RectHost const& r, const RectHost& ignore)
{
    frw_tree* t = deviceTree();
    if (!t) return !isInBounds(r);
    TreeScratch* s = static_cast<TreeScratch*>(scratch);
    const double box[4] = {r.topLeft.x, r.topLeft.y, r.bottomRight.x, r.bottomRight.y};
    int32_t ig[4];
    const bool has_ig = rect_ints(ignore, ig);  // a non-integer rect equals no stored rect
    uint8_t out = 0;
    if (!dev_ok(hipMemcpy(s->in, box, sizeof(box), hipMemcpyHostToDevice), "copy box")) return true;
    if (frw_query_refcompat(t, nullptr, s->in, 1, has_ig ? ig : nullptr, s->q1, nullptr, nullptr, nullptr) !=
        FRW_OK) {
        std::fprintf(stderr, "Tree: %s\n", frw_last_error_string());
        return true;
    }
    if (!dev_ok(hipMemcpy(&out, s->q1, 1, hipMemcpyDeviceToHost), "read Q1")) return true;
    return out != 0;
}

bool Tree::checkCollisons(point p, RectHost& r)
{
    frw_tree* t = deviceTree();
    if (!t) return false;
    TreeScratch* s = static_cast<TreeScratch*>(scratch);
    const double pt[2] = {p.x, p.y};
    int32_t id = -1;
    if (!dev_ok(hipMemcpy(s->in, pt, sizeof(pt), hipMemcpyHostToDevice), "copy point")) return false;
    if (frw_query_refcompat(t, s->in, nullptr, 1, nullptr, nullptr, s->q2, nullptr, nullptr) != FRW_OK) {
        std::fprintf(stderr, "Tree: %s\n", frw_last_error_string());
        return false;
    }
    if (!dev_ok(hipMemcpy(&id, s->q2, sizeof(id), hipMemcpyDeviceToHost), "read Q2")) return false;
    if (id < 0) return false;
    r = rects[(size_t)id];
    return true;
}

RectHost Tree::drawBiggestSquareAtPoint(point p)
{
    frw_tree* t = deviceTree();
    if (!t) return RectHost();
    TreeScratch* s = static_cast<TreeScratch*>(scratch);
    const double pt[2] = {p.x, p.y};
    double sq[4];
    if (!dev_ok(hipMemcpy(s->in, pt, sizeof(pt), hipMemcpyHostToDevice), "copy point")) return RectHost();
    if (frw_query_refcompat(t, s->in, nullptr, 1, nullptr, nullptr, nullptr, s->q3, nullptr) != FRW_OK) {
        std::fprintf(stderr, "Tree: %s\n", frw_last_error_string());
        return RectHost();
    }
    if (!dev_ok(hipMemcpy(sq, s->q3, sizeof(sq), hipMemcpyDeviceToHost), "read Q3")) return RectHost();
    return RectHost(point(sq[0], sq[1]), point(sq[2], sq[3]));
}

RectHost Tree::creatGaussianSurfFrom(RectHost const& r, floatingPoint const factor)
{
    if (factor < 1) return r;  // the reference logs and returns r (quadTree.cpp:353-357)
    frw_tree* t = deviceTree();
    int32_t ri[4];
    if (!t || !rect_ints(r, ri)) return r;
    double g[4];
    if (frw_refcompat_gauss_surface(t, ri, factor, g) != FRW_OK) {
        std::fprintf(stderr, "Tree: %s\n", frw_last_error_string());
        return r;
    }
    return RectHost(point(g[0], g[1]), point(g[2], g[3]));
}

void Tree::walkBuffers(int64_t** counts, int64_t** hops, int32_t** hit, int32_t** hop_n)
{
    TreeScratch* s = static_cast<TreeScratch*>(scratch);
    *counts = s ? s->counts : nullptr;
    *hops = s ? s->hops : nullptr;
    *hit = s ? s->hit : nullptr;
    *hop_n = s ? s->hop_n : nullptr;
}

void createTree(Tree* mainTree, Layer const& layer)
{
    if (!mainTree) return;
    for (const RectHost& r : layer) mainTree->insert(r);
}

int getIndex(REAL64_t intg[NSAMPLE + 1], floatingPoint rand)
{
    for (int i = 0; i < NSAMPLE; ++i)
        if (intg[i] <= rand && intg[i + 1] > rand) return i;
    return -1;
}

RectHost RandomWalk(RectHost const& R, Tree* mainTree, int& pointCount, REAL64_t intg[NSAMPLE + 1], int iterId)
{
    (void)intg;
    pointCount = 0;
    if (!mainTree || iterId < 0) return RectHost();
    frw_tree* t = mainTree->deviceTree();
    if (!t) return RectHost();
    const RectHost g = mainTree->creatGaussianSurfFrom(R, 1.5);
    const double surface[4] = {g.topLeft.x, g.topLeft.y, g.bottomRight.x, g.bottomRight.y};
    // the device staging of this tree (counts accumulate there unread)
    int32_t hit = -1, hops = 0;
    int64_t* counts = nullptr;
    int64_t* hop_sum = nullptr;
    int32_t *d_hit = nullptr, *d_hops = nullptr;
    mainTree->walkBuffers(&counts, &hop_sum, &d_hit, &d_hops);
    frw_walk_params p{};
    p.n_walks = 1;
    p.first_walk = iterId;
    p.seed = shim_seed();
    frw_walk_outputs o{counts, hop_sum, d_hit, d_hops};
    if (frw_refcompat_walk_from(t, surface, &p, &o, nullptr) != FRW_OK) {
        std::fprintf(stderr, "RandomWalk: %s\n", frw_last_error_string());
        return RectHost();
    }
    if (!dev_ok(hipMemcpy(&hit, d_hit, sizeof(hit), hipMemcpyDeviceToHost), "read hit") ||
        !dev_ok(hipMemcpy(&hops, d_hops, sizeof(hops), hipMemcpyDeviceToHost), "read hops"))
        return RectHost();
    pointCount = hops;
    return hit >= 0 ? mainTree->objects()[(size_t)hit] : RectHost();
}
