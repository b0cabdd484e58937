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

inline int grid_for(int64_t n) { return (int)((n + kQueryBlock - 1) / kQueryBlock); }

}  // namespace

int stack_bytes_per_lane(const frw_tree* t) { return (3 * t->depth + 2) * (int)sizeof(int32_t); }

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
    v.stack_cap = 3 * t->depth + 2;
    v.margin = t->margin;
    v.dist0 = t->dist0;
    return v;
}

int launch_q1(const frw_tree* t, const double* boxes, int64_t n, uint8_t* out, hipStream_t s)
{
    if (n == 0) return FRW_OK;
    const size_t lds = (size_t)kQueryBlock * stack_bytes_per_lane(t);
    q1_kernel<<<grid_for(n), kQueryBlock, lds, s>>>(dev_view(t), (const double4*)boxes, n, out);
    return check_hip(hipGetLastError(), "q1_kernel");
}

int launch_q2(const frw_tree* t, const double* pts, int64_t n, int32_t* ids, hipStream_t s)
{
    if (n == 0) return FRW_OK;
    const size_t lds = (size_t)kQueryBlock * stack_bytes_per_lane(t);
    q2_kernel<<<grid_for(n), kQueryBlock, lds, s>>>(dev_view(t), (const double2*)pts, n, ids);
    return check_hip(hipGetLastError(), "q2_kernel");
}

int launch_nearest(const frw_tree* t, const double* pts, int64_t n, double* d, hipStream_t s)
{
    if (n == 0) return FRW_OK;
    const size_t lds = (size_t)kQueryBlock * stack_bytes_per_lane(t);
    nearest_kernel<<<grid_for(n), kQueryBlock, lds, s>>>(dev_view(t), (const double2*)pts, n, d);
    return check_hip(hipGetLastError(), "nearest_kernel");
}

int launch_q3(const frw_tree* t, const double* pts, int64_t n, int mode, double* out, hipStream_t s)
{
    if (n == 0) return FRW_OK;
    const size_t lds = (size_t)kQueryBlock * stack_bytes_per_lane(t);
    if (mode == kModeBisect)
        q3_kernel<true><<<grid_for(n), kQueryBlock, lds, s>>>(dev_view(t), (const double2*)pts, n,
                                                               (double4*)out, t->d_err);
    else
        q3_kernel<false><<<grid_for(n), kQueryBlock, lds, s>>>(dev_view(t), (const double2*)pts, n,
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
