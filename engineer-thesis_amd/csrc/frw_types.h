// Plain data types and constants shared by host code and gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace frw {

constexpr int kNSample = 200;        // panels per square (randomWalk/defines.h:25)
constexpr int kMaxBisect = 4096;     // safety cap on one biggest-square search
constexpr int kWaveSize = 64;
constexpr int kStartLevels = 11;     // Morton levels 0..10 in the nearest-search start table (5.6 MB)

// Error bits reported by kernels.
constexpr int kErrBisect = 1;        // biggest-square search exceeded kMaxBisect
constexpr int kErrStack = 2;         // traversal stack would overflow
constexpr int kErrWeight = 4;        // a weighted walk's weight beyond the fixed-point range (|w| >= 2^31)

// Flattened quadtree node, 32 bytes.  Children are stored as a group of four
// consecutive nodes (quadrants in Morton order; an empty quadrant is a dummy
// node with an inverted box), so one 128-byte line holds all four child
// boxes.  `box` is the tight bounding box of every rect in the subtree.
struct alignas(16) Node {
    int4 box;             // x0, y0, x1, y1 (dummy: INT_MAX, INT_MAX, INT_MIN, INT_MIN)
    int32_t first_child;  // index of the child group, -1 for a leaf
    int32_t rbeg;         // own rects [rbeg, rend) in the sorted rect array
    int32_t rend;
    int32_t level;
};
static_assert(sizeof(Node) == 32, "node must stay 32 bytes");

// Walk modes.
enum WalkMode : int32_t {
    kModeExact = 0,    // nearest-rect distance + replayed bisection schedule
    kModeBisect = 1,   // reference bisection, one Q1 per step
    kModeLearned = 2,  // network proposal verified by a pyramid probe, exact fallback
    kModeWeighted = 3, // exact walks from a Gaussian-surface start with first-hop weights
    kModeRefCompat = 4,  // the reference CPU tree and its queries, defects included
};

// Weighted-walk options (FRW_WEIGHT_* in the C ABI).
constexpr uint32_t kWeightImportance = 1u;  // first panel drawn proportional to |n . grad G|

}  // namespace frw
