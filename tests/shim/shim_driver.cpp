// Test driver for the reference-signature shim (include/randomwalk.h): runs
// the reference's own call sequence (Parser -> Tree -> createTree ->
// RandomWalk, randomWalk/CPU/getAvgPathLen.cpp:35-111) and the Tree queries,
// printing one result per line for the test-suite to compare.
//   shim_driver walk <layout> <master> <n>        i hops x0 y0 x1 y1 per walk
//   shim_driver query <layout> <points>           per point line "x y h":
//                                                 q1(box p+-h) q2 x0 y0 x1 y1 q3...
//   shim_driver gauss <layout> <rect id> <factor> x0 y0 x1 y1
//   shim_driver getindex <cdf file> <values file> one index per value
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

int main(int argc, char** argv)
{
    if (argc < 4) return 2;
    const std::string cmd = argv[1];
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
