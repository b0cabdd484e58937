// randomwalk.h — C++ drop-in for the entry points that the reference exports
// through randomWalk/randomwalk.h (randomWalk/CPU/mainFunctions.h:15-18,
// randomWalk/CPU/getAvgPathLen.h:5, randomWalk/CPU/Parser.h:13-30,
// randomWalk/Cuda/mainkernels.h:9-15).  Same names, argument meanings and
// result scalar; implemented on the gfx950 library declared in frw.h.
//
// Differences a caller can observe (kept on purpose):
//   - errors that the reference answers with exit(0) print a message to
//     stderr and make the function return (runRandomWalk) or return -1.0;
//   - both engines run on the GPU: GPU_FLAG=true runs the exact-search walk,
//     GPU_FLAG=false runs the reference-compatible mode (the reference CPU
//     tree rebuilt with its insertion behaviour, its single-path queries and
//     literal bisection), which reproduces the reference CPU FRW walk for
//     walk under the same random stream;
//   - walk streams are Philox4x32-10 keyed by FRW_SEED (default 1), so runs
//     are reproducible; the reference re-seeds from /dev/urandom.
#ifndef FRW_RANDOMWALK_H
#define FRW_RANDOMWALK_H

#include <stdint.h>

#include <string>
#include <vector>

typedef double floatingPoint;

struct point {
    floatingPoint x, y;
    point() : x(0), y(0) {}
    point(floatingPoint px, floatingPoint py) : x(px), y(py) {}
};

struct RectHost {
    point topLeft, bottomRight;
    RectHost() : topLeft(-1, -1), bottomRight(-1, -1) {}  // "no rect" sentinel
    RectHost(point tl, point br) : topLeft(tl), bottomRight(br) {}
    bool operator==(const RectHost& o) const
    {
        return topLeft.x == o.topLeft.x && topLeft.y == o.topLeft.y &&
               bottomRight.x == o.bottomRight.x && bottomRight.y == o.bottomRight.y;
    }
    bool operator!=(const RectHost& o) const { return !(*this == o); }
};

typedef std::vector<RectHost> Layer;
typedef std::vector<Layer> Layers;

// randomWalk/CPU/Parser.h:13-30
class Parser {
public:
    explicit Parser(const char* delimiter);
    void parse(std::string const& fileName);
    RectHost getLayerSize(int layerIt);
    Layer getLayerAt(int i) { return layers.at(i); }
    Layers getAllLayers() { return layers; }
    int getLayerCount() { return (int)layers.size(); }

private:
    Layers layers;
    std::vector<RectHost> bounds;
    std::string delimiter;
};

typedef double REAL64_t;
#ifndef NSAMPLE
#define NSAMPLE 200  // panels per square (randomWalk/defines.h:25)
#endif

struct frw_layout;
struct frw_tree;

// randomWalk/CPU/quadTree.h:31-77: the reference's CPU quadtree, served by the
// device reference-compatible index (the reference tree rebuilt with its
// insertion behaviour, its single-path queries and literal bisection).
// insert() collects rects in call order; the device index is built from them
// at the first query after an insert.  Every query is one device call.
// Differences: coordinates must be integers (the library's rect type);
// insert() returns whether the rect lies inside the bounds (the reference
// also returns false for a rect its level cap drops, quadTree.cpp:139-145).
class Tree {
public:
    Tree(int pLevel, int nodeSCount, RectHost const& bounds);
    ~Tree();
    Tree(const Tree&) = delete;
    Tree& operator=(const Tree&) = delete;
    RectHost getBounds() { return bounds; }
    bool isInBounds(point const& p);       // strict (quadTree.cpp:49-59)
    bool isInBounds(RectHost const& r);    // inclusive (quadTree.cpp:38-47)
    void clear();
    bool insert(RectHost const& r);
    bool checkCollisions(RectHost const& r, const RectHost& ignore = RectHost());
    bool checkCollisons(point p, RectHost& r);
    RectHost drawBiggestSquareAtPoint(point p);
    RectHost creatGaussianSurfFrom(RectHost const& r, floatingPoint const factor);
    // not in the reference class: the device index (built on demand, NULL on
    // error), the stored rects, and the device buffers RandomWalk writes
    frw_tree* deviceTree();
    const Layer& objects() const { return rects; }
    void walkBuffers(int64_t** counts, int64_t** hops, int32_t** hit, int32_t** hop_n);

private:
    RectHost bounds;
    Layer rects;
    frw_layout* layout = nullptr;
    frw_tree* tree = nullptr;
    void* scratch = nullptr;  // device staging for one query
    bool dirty = true;
};

// randomWalk/CPU/mainFunctions.h:15-17 (mainFunctions.cpp:33-90).  RandomWalk
// runs walk iterId (Philox stream keyed by FRW_SEED) from the Gaussian surface
// of R at 1.5 on the device and returns the rect it ended on (RectHost() when
// it escaped); pointCount = its hops.  intg is accepted for the signature:
// the walk samples the library's Green table (the same CDF re-derived from its
// series).  getIndex is the reference's linear CDF scan (-1 when rand lies
// outside the table).
void createTree(Tree* mainTree, Layer const& layer);
RectHost RandomWalk(RectHost const& R, Tree* mainTree, int& pointCount, REAL64_t intg[NSAMPLE + 1], int iterId);
int getIndex(REAL64_t intg[NSAMPLE + 1], floatingPoint rand);

// Opaque device tree + master (randomWalk/Cuda/d_quadtree.h:36-47).
struct QuadTreeManager;

// randomWalk/CPU/mainFunctions.h:18
void runRandomWalk(char* path, int ITER_NUM, int RECT_ID, bool GPU_FLAG, bool measure, int layer);
// randomWalk/CPU/getAvgPathLen.h:5
floatingPoint getAvgPathLen(char* path, int ITER_NUM, int RECT_ID, bool measure, int layerID);
// randomWalk/Cuda/mainkernels.h:9-15
floatingPoint getAvgPathLenCUDA(char* path, int ITER_NUM, int RECT_ID, bool measure, int layerID);
QuadTreeManager* createQuadTree(const std::vector<RectHost>& layer, RectHost const& spaceSize,
                                int RECT_ID, bool doCheck);
QuadTreeManager* randomWalkCudaInit(char* path, bool measure, int RECT_ID, int layerID);
// output[i] = 1 when walk i did not end on the master; rectOutput[i] = the
// rect it ended on ((-1,-1,-1,-1) when it escaped).  Both are device arrays.
void randomWalkCudaWrapper(int n, QuadTreeManager* quadTree, unsigned int* output,
                           RectHost* rectOutput, unsigned long long randomSeed);
bool initCuda(int argc, char** argv);
void freeQuadTreeManager(QuadTreeManager* qtm);
floatingPoint countAvg(unsigned int output[], int ITER_NUM);

#endif  // FRW_RANDOMWALK_H
