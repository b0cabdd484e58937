// Internal host-side declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/frw.h"
#include "frw_device.h"
#include "frw_types.h"

struct frw_layout {
    std::vector<int4> rects;  // file order = original ids
    int4 bounds;              // padded absorbing bounds
};

struct frw_tree {
    int device = 0;
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
    uint32_t* d_thresholds = nullptr;  // Green CDF panel thresholds
    // walk scratch owned by the tree handle
    unsigned long long* d_work = nullptr;  // walk work-queue counter
    int* d_err = nullptr;                  // kernel error bits
    double* d_surface = nullptr;           // Gaussian surface of the master
};

namespace frw {

// thread-local error reporting
int set_error(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);

// host tables
void green_thresholds(uint32_t T[kNSample]);

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

// walk launcher (frw_walk.hip)
int launch_walk(const frw_tree* t, const frw_walk_params* p, const frw_walk_outputs* o,
                hipStream_t s, unsigned long long* traffic = nullptr);

}  // namespace frw
