// Test driver for the reference-signature shim (include/randomwalk.h): runs
// the reference's own call sequence (Parser -> Tree -> createTree ->
// RandomWalk, randomWalk/CPU/getAvgPathLen.cpp:35-111) and the Tree queries,
// printing one result per line for the test-suite to compare.
//   shim_driver walk <layout> <master> <n>        i hops x0 y0 x1 y1 per walk
//   shim_driver query <layout> <points>           per point line "x y h":
//                                                 q1(box p+-h) q2 x0 y0 x1 y1 q3...
//   shim_driver gauss <layout> <rect id> <factor> x0 y0 x1 y1
//   shim_driver getindex <cdf file> <values file> one index per value
// and the GPU half (randomWalk/Cuda/mainkernels.h:9-15):
//   shim_driver cudawalk <layout> <master> <n> <seed>
//       randomWalkCudaInit -> randomWalkCudaWrapper -> countAvg ->
//       freeQuadTreeManager: "flag x0 y0 x1 y1" per walk, then "avg <countAvg>"
//   shim_driver cudatree <layout> <master> <n> <seed> <pad> <check>
//       createQuadTree(layer, layer size grown by pad on every side, master,
//       check) and the same walks
//   shim_driver avg <layout> <master> <n> <gpu>   getAvgPathLen(CUDA)'s value
//   shim_driver run <layout> <master> <n> <gpu>   runRandomWalk
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "randomwalk.h"

static Tree* load(const char* path, Layer* layer)
{
    Parser p("<<");
    p.parse(path);
    if (p.getLayerCount() < 1) return nullptr;
    *layer = p.getLayerAt(0);
    Tree* t = new Tree(0, (int)layer->size(), p.getLayerSize(0));
    createTree(t, *layer);
    return t;
}

static void print_rect(const RectHost& r)
{
    std::printf(" %.17g %.17g %.17g %.17g", r.topLeft.x, r.topLeft.y, r.bottomRight.x, r.bottomRight.y);
}

// randomWalkCudaWrapper into device arrays, copied back and printed
static int cuda_walks(QuadTreeManager* q, int n, unsigned long long seed)
{
    if (!q) return 4;
    unsigned int* d_out = nullptr;
    RectHost* d_rects = nullptr;
    if (hipMalloc(&d_out, sizeof(unsigned int) * n) != hipSuccess ||
        hipMalloc(&d_rects, sizeof(RectHost) * n) != hipSuccess)
        return 5;
    randomWalkCudaWrapper(n, q, d_out, d_rects, seed);
    std::vector<unsigned int> out((size_t)n);
    std::vector<RectHost> rects((size_t)n);
    if (hipMemcpy(out.data(), d_out, sizeof(unsigned int) * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(rects.data(), d_rects, sizeof(RectHost) * n, hipMemcpyDeviceToHost) != hipSuccess)
        return 6;
    for (int i = 0; i < n; ++i) {
        std::printf("%u", out[i]);
        print_rect(rects[i]);
        std::printf("\n");
    }
    std::printf("avg %.17g\n", countAvg(out.data(), n));
    (void)hipFree(d_out);
    (void)hipFree(d_rects);
    freeQuadTreeManager(q);
    return 0;
}

int main(int argc, char** argv)
{
    if (argc < 4) return 2;
    const std::string cmd = argv[1];
    if (cmd == "cudawalk" && argc >= 6) {
        if (!initCuda(argc, argv)) return 3;
        return cuda_walks(randomWalkCudaInit(argv[2], false, std::atoi(argv[3]), 0), std::atoi(argv[4]),
                          std::strtoull(argv[5], nullptr, 10));
    }
    if (cmd == "cudatree" && argc >= 8) {
        if (!initCuda(argc, argv)) return 3;
        Parser p("<<");
        p.parse(argv[2]);
        if (p.getLayerCount() < 1) return 3;
        const RectHost b = p.getLayerSize(0);
        const double pad = std::atof(argv[6]);
        const RectHost grown(point(b.topLeft.x - pad, b.topLeft.y - pad),
                             point(b.bottomRight.x + pad, b.bottomRight.y + pad));
        return cuda_walks(createQuadTree(p.getLayerAt(0), grown, std::atoi(argv[3]), std::atoi(argv[7]) != 0),
                          std::atoi(argv[4]), std::strtoull(argv[5], nullptr, 10));
    }
    if ((cmd == "avg" || cmd == "run") && argc >= 6) {
        const bool gpu = std::atoi(argv[5]) != 0;
        if (gpu && !initCuda(argc, argv)) return 3;
        if (cmd == "run") {
            runRandomWalk(argv[2], std::atoi(argv[4]), std::atoi(argv[3]), gpu, false, 0);
            return 0;
        }
        const floatingPoint v = gpu ? getAvgPathLenCUDA(argv[2], std::atoi(argv[4]), std::atoi(argv[3]), false, 0)
                                    : getAvgPathLen(argv[2], std::atoi(argv[4]), std::atoi(argv[3]), false, 0);
        std::printf("%.17g\n", v);
        return 0;
    }
    if (cmd == "getindex") {
        REAL64_t intg[NSAMPLE + 1];
        std::ifstream c(argv[2]);
        for (int i = 0; i <= NSAMPLE; ++i) c >> intg[i];
        std::ifstream v(argv[3]);
        double r;
        while (v >> r) std::printf("%d\n", getIndex(intg, r));
        return 0;
    }
    Layer layer;
    Tree* t = load(argv[2], &layer);
    if (!t) return 3;
    int rc = 0;
    if (cmd == "walk" && argc >= 5) {
        const int master = std::atoi(argv[3]), n = std::atoi(argv[4]);
        REAL64_t intg[NSAMPLE + 1] = {0};
        for (int i = 0; i < n; ++i) {
            int hops = -1;
            const RectHost h = RandomWalk(layer.at(master), t, hops, intg, i);
            std::printf("%d %d", i, hops);
            print_rect(h);
            std::printf("\n");
        }
    } else if (cmd == "query") {
        std::ifstream f(argv[3]);
        double x, y, h;
        while (f >> x >> y >> h) {
            const RectHost box(point(x - h, y - h), point(x + h, y + h));
            RectHost hit;
            const bool q1 = t->checkCollisions(box);
            const bool q2 = t->checkCollisons(point(x, y), hit);
            std::printf("%d %d", q1 ? 1 : 0, q2 ? 1 : 0);
            print_rect(q2 ? hit : RectHost());
            print_rect(t->drawBiggestSquareAtPoint(point(x, y)));
            std::printf("\n");
        }
    } else if (cmd == "gauss" && argc >= 5) {
        const RectHost g = t->creatGaussianSurfFrom(layer.at(std::atoi(argv[3])), std::atof(argv[4]));
        print_rect(g);
        std::printf("\n");
    } else {
        rc = 2;
    }
    delete t;
    return rc;
}
