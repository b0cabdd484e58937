// Batched tree queries (the parity hooks) and the Gaussian-surface kernel.
// One lane per query; each lane's traversal stack lives in LDS.
#include "frw_internal.h"

namespace frw {
namespace {

constexpr int kQueryBlock = 256;

__device__ inline LaneStack lane_stack()
{
    extern __shared__ int32_t stack_mem[];
    return LaneStack{stack_mem + threadIdx.x, (int)blockDim.x};
}

__global__ void __launch_bounds__(kQueryBlock) q1_kernel(DevTree t, const double4* __restrict__ boxes,
                                                        int64_t n, uint8_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const LaneStack st = lane_stack();
    if (i >= n) return;
    const double4 b = boxes[i];
    out[i] = q1_box<false>(t, st, b.x, b.y, b.z, b.w, make_int4(0, 0, 0, 0), nullptr) ? 1 : 0;
}

__global__ void __launch_bounds__(kQueryBlock) q2_kernel(DevTree t, const double2* __restrict__ pts,
                                                        int64_t n, int32_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const LaneStack st = lane_stack();
    if (i >= n) return;
    const double2 p = pts[i];
    out[i] = q2_point(t, st, p.x, p.y, nullptr);
}

__global__ void __launch_bounds__(kQueryBlock) nearest_kernel(DevTree t, const double2* __restrict__ pts,
                                                             int64_t n, double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const LaneStack st = lane_stack();
    if (i >= n) return;
    const double2 p = pts[i];
    out[i] = nearest_rect(t, st, p.x, p.y, nullptr);
}

template <bool kLiteral>
__global__ void __launch_bounds__(kQueryBlock) q3_kernel(DevTree t, const double2* __restrict__ pts,
                                                        int64_t n, double4* __restrict__ out, int* err)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const LaneStack st = lane_stack();
    if (i >= n) return;
    const double2 p = pts[i];
    const double ds = kLiteral ? 0.0 : nearest_rect(t, st, p.x, p.y, nullptr);
    const DBox b = biggest_square<kLiteral>(t, st, p.x, p.y, ds, nullptr, nullptr, err);
    out[i] = make_double4(b.x0, b.y0, b.x1, b.y1);
}

__global__ void __launch_bounds__(64) gauss_kernel(DevTree t, int4 master, double* __restrict__ out)
{
    const LaneStack st = lane_stack();
    if (threadIdx.x != 0) return;
    const double fx = gauss_factor(t, st, master, 1.5, false);
    const double fy = gauss_factor(t, st, master, 1.5, true);
    const DBox s = gauss_surface(master, fx, fy);
    out[0] = s.x0;
    out[1] = s.y0;
    out[2] = s.x1;
    out[3] = s.y1;
}

}  // namespace

// LDS budget of one block (bytes); the traversal stacks must fit in it.
int lds_per_block(const frw_tree* t)
{
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, t->device) != hipSuccess || v <= 0)
        v = 64 * 1024;
    return v;
}

// Largest block (<= want, multiple of 64) whose traversal stacks fit in LDS;
// 0 if not even one wave fits.
int block_for_stack(const frw_tree* t, int want, size_t extra_lds)
{
    const size_t budget = (size_t)lds_per_block(t);
    int b = want;
    while (b >= 64 && (size_t)b * stack_bytes_per_lane(t) + extra_lds > budget) b -= 64;
    return b >= 64 ? b : 0;
}

namespace {

struct QueryLaunch {
    int block;
    int grid;
    size_t lds;
};

int plan_query(const frw_tree* t, int64_t n, QueryLaunch* q)
{
    q->block = block_for_stack(t, kQueryBlock, 0);
    if (q->block == 0) return set_error(FRW_E_RANGE, "tree too deep for the LDS traversal stack");
    q->grid = (int)((n + q->block - 1) / q->block);
    q->lds = (size_t)q->block * stack_bytes_per_lane(t);
    return FRW_OK;
}

}  // namespace

int stack_bytes_per_lane(const frw_tree* t) { return LaneStack::entries(t->depth) * (int)sizeof(int32_t); }

DevTree dev_view(const frw_tree* t)
{
    DevTree v;
    v.rects = t->d_rects;
    v.ids = t->d_ids;
    v.nodes = t->d_nodes;
    v.bounds = t->bounds;
    v.n_rects = (int32_t)t->n_rects;
    v.n_nodes = (int32_t)t->n_nodes;
    v.depth = t->depth;
    v.stack_cap = LaneStack::entries(t->depth);
    v.margin = t->margin;
    v.dist0 = t->dist0;
    v.start_tab = t->d_start;
    v.start_levels = t->d_start ? t->start_levels : 0;
    v.morton_bits = t->morton_bits;
    v.ext2_max = t->ext2_max;
    return v;
}

int launch_q1(const frw_tree* t, const double* boxes, int64_t n, uint8_t* out, hipStream_t s)
{
    if (n == 0) return FRW_OK;
    QueryLaunch ql;
    if (plan_query(t, n, &ql) != FRW_OK) return FRW_E_RANGE;
    q1_kernel<<<ql.grid, ql.block, ql.lds, s>>>(dev_view(t), (const double4*)boxes, n, out);
    return check_hip(hipGetLastError(), "q1_kernel");
}

int launch_q2(const frw_tree* t, const double* pts, int64_t n, int32_t* ids, hipStream_t s)
{
    if (n == 0) return FRW_OK;
    QueryLaunch ql;
    if (plan_query(t, n, &ql) != FRW_OK) return FRW_E_RANGE;
    q2_kernel<<<ql.grid, ql.block, ql.lds, s>>>(dev_view(t), (const double2*)pts, n, ids);
    return check_hip(hipGetLastError(), "q2_kernel");
}

int launch_nearest(const frw_tree* t, const double* pts, int64_t n, double* d, hipStream_t s)
{
    if (n == 0) return FRW_OK;
    QueryLaunch ql;
    if (plan_query(t, n, &ql) != FRW_OK) return FRW_E_RANGE;
    nearest_kernel<<<ql.grid, ql.block, ql.lds, s>>>(dev_view(t), (const double2*)pts, n, d);
    return check_hip(hipGetLastError(), "nearest_kernel");
}

int launch_q3(const frw_tree* t, const double* pts, int64_t n, int mode, double* out, hipStream_t s)
{
    if (n == 0) return FRW_OK;
    QueryLaunch ql;
    if (plan_query(t, n, &ql) != FRW_OK) return FRW_E_RANGE;
    if (mode == kModeBisect)
        q3_kernel<true><<<ql.grid, ql.block, ql.lds, s>>>(dev_view(t), (const double2*)pts, n,
                                                               (double4*)out, t->d_err);
    else
        q3_kernel<false><<<ql.grid, ql.block, ql.lds, s>>>(dev_view(t), (const double2*)pts, n,
                                                                (double4*)out, t->d_err);
    return check_hip(hipGetLastError(), "q3_kernel");
}

int launch_gauss(const frw_tree* t, int4 master, double* d_out, hipStream_t s)
{
    const size_t lds = (size_t)64 * stack_bytes_per_lane(t);
    gauss_kernel<<<1, 64, lds, s>>>(dev_view(t), master, d_out);
    return check_hip(hipGetLastError(), "gauss_kernel");
}

}  // namespace frw
