// Learned square-size proposal: occupancy-window features and a 67-64-64-64-1
// GELU network evaluated with bf16 MFMA (v_mfma_f32_32x32x8_bf16) and FP32
// accumulation, and the occupancy-pyramid probe that verifies a proposal.
// No counterpart in the reference: the reference sizes every square by
// bisection (randomWalk/CPU/quadTree.cpp:264-312).
//
// Occupancy pyramid: level l is a 2^l x 2^l grid of cells of side
// s_l = 2^(K0-l) from the bounds corner (K0 = smallest k with 2^k > max(W, H));
// a cell is occupied when some closed rect meets the half-open cell.  Levels
// are stored as row-major bitmaps (features) and, up to level Vmax, as
// summed-area tables of (2^l + 1)^2 uint32 counts with a zero first row and
// column (verification).
//
// Features of a point p whose previous square had half-size h:
//   Lmax = min(K0, 15), level l = clamp(K0 - floor(log2 max(h, 1)) + 1, 0, Lmax),
//   cell s = s_l, (ci, cj) = floor((p - bounds corner) / s);
//   f[v*8+u] = occupancy of cell (ci-4+u, cj-4+v) at level l (1 outside the
//   grid), f[64] = l / 16, f[65], f[66] = fractional position of p in its
//   cell (x, y), f[67..71] = 0; every feature is rounded to bf16.
// Output y estimates log2(d* / s); the proposed half-size is d = s 2^y alpha.
//
// Verification of the square [p - d, p + d]^2 (closed, as the walk builds it):
// level v = clamp(K0 - floor(log2 d) + 4, 0, Vmax), so cells are at most d/16;
// the square is certified empty when it lies inside the bounds and the cell
// block covering it, floor((corner - bounds corner) / s_v) per axis, holds no
// occupied cell (four loads from the summed-area table).  Any rect point in
// the square lies in a cell of that block, so a certified square is free.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "frw_device.h"

namespace frw {

constexpr int kMlpIn = 67;   // features (window 64, level, position in cell x and y)
constexpr int kMlpInPad = 72;   // K of the first layer, a multiple of 8
constexpr int kMlpHid = 64;
constexpr int kMlpStride = 72;  // LDS row stride (bf16 elements) of activations
// a wave's activation tile: 64 rows of kMlpStride, then 16 bf16 per lane of
// staging for the layers' first halves
constexpr int kMlpTileElems = 64 * kMlpStride + 64 * 16;
constexpr int kMlpMaxLevel = 15;  // feature-window levels
constexpr int kPyrMaxLevel = 15;  // bitmap / summed-area-table levels
constexpr int kCertShift = 4;     // verification cells are at most d / 2^4

// FP32 parameter block layout (row-major [out][in]).
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpHid * kMlpInPad;
constexpr int kOffW2 = kOffB1 + kMlpHid;
constexpr int kOffB2 = kOffW2 + kMlpHid * kMlpHid;
constexpr int kOffW3 = kOffB2 + kMlpHid;
constexpr int kOffB3 = kOffW3 + kMlpHid * kMlpHid;
constexpr int kOffW4 = kOffB3 + kMlpHid;
constexpr int kOffB4 = kOffW4 + kMlpHid;
constexpr int kMlpParams = kOffB4 + 1;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Occupancy pyramid of one layout, passed by value (walk jobs carry one each).
struct MlpPyramid {
    const uint32_t* occ;      // bitmaps, level l at word occ_off[l]
    const uint32_t* sat;      // summed-area tables, level l at entry sat_off[l]
    const int64_t* occ_off;   // device arrays of Pmax + 1 / Vmax + 1 offsets
    const int64_t* sat_off;
    int4 bounds;
    int K0;
    int Lmax;
    int Vmax;
    int pad;
};

// Network weights (bf16 copies of the FP32 parameters) and the safety factor.
struct MlpNet {
    const uint16_t* W1;   // bf16 [64][72]
    const uint16_t* W2;   // bf16 [64][64], input columns in mlp_col order (the walk copy)
    const uint16_t* W3;   // bf16 [64][64], input columns in mlp_col order (the walk copy)
    const uint16_t* w4;   // bf16 [64]
    const float* b1;
    const float* b2;
    const float* b3;
    const float* b4;      // device scalar
    float alpha;
};

__host__ __device__ inline uint16_t f2bf(float x)
{
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bfx2 __attribute__((ext_vector_type(2)));
// Four floats to bf16 with the hardware's packed round-to-nearest-even
// conversion (v_cvt_pk_bf16_f32, two values per instruction); equal to f2bf
// for every finite value (a host build of this header uses f2bf itself).
__host__ __device__ inline bf16x4 to_bf16x4(float a, float b, float c, float d)
{
    bf16x4 v;
#if defined(__HIP_DEVICE_COMPILE__)
    const bfx2 lo = __builtin_convertvector((f32x2){a, b}, bfx2);
    const bfx2 hi = __builtin_convertvector((f32x2){c, d}, bfx2);
    const uint32_t l = __builtin_bit_cast(uint32_t, lo), h = __builtin_bit_cast(uint32_t, hi);
    v[0] = (short)(l & 0xFFFFu);
    v[1] = (short)(l >> 16);
    v[2] = (short)(h & 0xFFFFu);
    v[3] = (short)(h >> 16);
#else
    v[0] = (short)f2bf(a);
    v[1] = (short)f2bf(b);
    v[2] = (short)f2bf(c);
    v[3] = (short)f2bf(d);
#endif
    return v;
}

// The network's small operands staged in LDS by every block that evaluates
// it: b1 | b2 | b3 | w4 (as FP32) | b4.
constexpr int kMlpSmallB1 = 0, kMlpSmallB2 = 64, kMlpSmallB3 = 128, kMlpSmallW4 = 192, kMlpSmallB4 = 256;
constexpr int kMlpSmallFloats = 260;

// erf by Abramowitz & Stegun 7.1.26 (Hastings' rational approximation,
// |error| <= 1.5e-7 in exact arithmetic, well inside a bf16 activation's
// rounding): erf(x) = 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-x^2)
// with t = 1 / (1 + p x) for x >= 0, odd in x.  One reciprocal and one
// exponential instead of a full-accuracy erf.
__device__ inline float erf_as(float x)
{
    const float ax = fabsf(x);
    const float t = __builtin_amdgcn_rcpf(__builtin_fmaf(0.3275911f, ax, 1.0f));
    float poly = __builtin_fmaf(t, 1.061405429f, -1.453152027f);
    poly = __builtin_fmaf(t, poly, 1.421413741f);
    poly = __builtin_fmaf(t, poly, -0.284496736f);
    poly = __builtin_fmaf(t, poly, 0.254829592f);
    const float e = __builtin_fmaf(-t * poly, __expf(-ax * ax), 1.0f);
    return copysignf(e, x);
}

// GELU(z) = z Phi(z) = max(z, 0) - t Phi(-t), t = |z|, with Phi(-t) replaced by
// a degree-10 polynomial in x = 1 - t / 4.5 that vanishes at t = 4.5 (where
// t Phi(-t) < 1.6e-5), t clamped there: FMA and min/max only, no
// transcendental; within 1.7e-5 of the exact GELU, far inside a bf16
// activation's rounding.  Coefficients from fit_gelu.py (a minimax fit).
__device__ inline float gelu_f(float z)
{
    const float t = fminf(fabsf(z), 4.5f);
    const float x = __builtin_fmaf(t, -1.0f / 4.5f, 1.0f);
    float p = -7.503951923e-03f;
    p = __builtin_fmaf(p, x, 1.011371091e+01f);
    p = __builtin_fmaf(p, x, -3.778662722e+01f);
    p = __builtin_fmaf(p, x, 5.380414907e+01f);
    p = __builtin_fmaf(p, x, -3.819883533e+01f);
    p = __builtin_fmaf(p, x, 1.585222026e+01f);
    p = __builtin_fmaf(p, x, -3.753199773e+00f);
    p = __builtin_fmaf(p, x, 5.077477714e-01f);
    p = __builtin_fmaf(p, x, -3.255581238e-02f);
    p = __builtin_fmaf(p, x, 9.095615919e-04f);
    return fmaxf(z, 0.0f) - t * (p * x);
}

__device__ inline float gelu_grad_f(float z)
{
    const float cdf = 0.5f * (1.0f + erf_as(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * __expf(-0.5f * z * z);
    return cdf + z * pdf;
}

// The feature level of a point whose previous square had half-size h, and
// its cell side s = 2^(K0 - l).
__device__ inline int mlp_level(const MlpPyramid& m, double h, double* cell)
{
    const int e = h >= 1.0 ? ilogb(h) : 0;
    int l = m.K0 - e + 1;
    l = l < 0 ? 0 : (l > m.Lmax ? m.Lmax : l);
    *cell = ldexp(1.0, m.K0 - l);
    return l;
}

// 64 occupancy bits of the 8x8 window (bit v*8+u) at level l around p.  Each
// window row is one unaligned 8-bit field of the level's row-major bitmap,
// read as two words (every level carries one padding word at its end).
__device__ inline uint64_t mlp_window(const MlpPyramid& m, int l, double cell, double px, double py, float* frac_x,
                                      float* frac_y)
{
    const int64_t n = (int64_t)1 << l;
    // cell is 2^(K0 - l): multiplying by its exact inverse equals the division
    (void)cell;
    const double inv = ldexp(1.0, l - m.K0);
    const double qx = (px - (double)m.bounds.x) * inv;
    const double qy = (py - (double)m.bounds.y) * inv;
    const double fx = floor(qx);
    const double fy = floor(qy);
    *frac_x = 0.0f;
    *frac_y = 0.0f;
    // far outside the grid (or not a number): every window cell reads as occupied
    if (!(fx > -16.0 && fy > -16.0 && fx < (double)(n + 16) && fy < (double)(n + 16))) return ~0ull;
    *frac_x = (float)(qx - fx);
    *frac_y = (float)(qy - fy);
    const int64_t ci = (int64_t)fx;
    const int64_t cj = (int64_t)fy;
    const uint32_t* base = m.occ + m.occ_off[l];
    const int64_t i0 = ci - 4;
    const int64_t lo = i0 > 0 ? i0 : 0;
    const int64_t hi = i0 + 8 < n ? i0 + 8 : n;
    uint32_t inside = 0;  // window columns inside the grid
    if (lo < hi) inside = ((1u << (hi - lo)) - 1u) << (lo - i0);
    uint64_t bits = 0;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const int64_t j = cj - 4 + v;
        uint32_t row = 0xFFu;
        if (inside && j >= 0 && j < n) {
            const int64_t idx = j * n + lo;
            const uint64_t w = ((uint64_t)base[(idx >> 5) + 1] << 32) | base[idx >> 5];
            const uint32_t field = (uint32_t)(w >> (idx & 31)) & ((1u << (hi - lo)) - 1u);
            row = (field << (lo - i0)) | (~inside & 0xFFu);
        }
        bits |= (uint64_t)row << (v * 8);
    }
    return bits;
}

// Writes one feature row (72 bf16) of this lane's point into LDS.
__device__ inline void mlp_feature_row(uint16_t* row, uint64_t bits, int level, float frac_x, float frac_y)
{
    const short one = 0x3F80;  // bf16(1.0)
    // four elements per 8-byte store (rows are 8-byte aligned)
#pragma unroll
    for (int k = 0; k < 64; k += 4) {
        bf16x4 v;
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = ((bits >> (k + u)) & 1u) ? one : (short)0;
        *(bf16x4*)(row + k) = v;
    }
    bf16x4 a;
    a[0] = (short)f2bf((float)level / 16.0f);
    a[1] = (short)f2bf(frac_x);
    a[2] = (short)f2bf(frac_y);
    a[3] = 0;
    *(bf16x4*)(row + 64) = a;
    static_assert(kMlpInPad == 72, "feature row layout");
    *(bf16x4*)(row + 68) = bf16x4{0, 0, 0, 0};
}

// Verification probe: true when the closed square (x0, y0, x1, y1) of
// half-size d (> 0) lies in the bounds and its covering cell block at level
// clamp(K0 - floor(log2 d) + 4, 0, Vmax) holds no occupied cell.
__device__ inline bool mlp_certify(const MlpPyramid& m, double x0, double y0, double x1, double y1, double d)
{
    if (!in_bounds_box(m.bounds, x0, y0, x1, y1)) return false;
    int v = m.K0 - ilogb(d) + kCertShift;
    v = v < 0 ? 0 : (v > m.Vmax ? m.Vmax : v);
    const double inv = ldexp(1.0, v - m.K0);  // the exact inverse of the cell side 2^(K0 - v)
    const int64_t n = (int64_t)1 << v;
    const double a0 = floor((x0 - (double)m.bounds.x) * inv), a1 = floor((x1 - (double)m.bounds.x) * inv);
    const double b0 = floor((y0 - (double)m.bounds.y) * inv), b1 = floor((y1 - (double)m.bounds.y) * inv);
    if (!(a0 >= 0.0 && b0 >= 0.0 && a1 < (double)n && b1 < (double)n)) return false;
    const int64_t i0 = (int64_t)a0, i1 = (int64_t)a1 + 1, j0 = (int64_t)b0, j1 = (int64_t)b1 + 1;
    const uint32_t* S = m.sat + m.sat_off[v];
    const int64_t w = n + 1;
    const uint32_t occ = S[j1 * w + i1] - S[j0 * w + i1] - S[j1 * w + i0] + S[j0 * w + i0];
    return occ == 0u;
}

// Orders this wave's LDS accesses (the other lanes' rows) without a block barrier.
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16v8 __attribute__((ext_vector_type(8)));

// Column of the activation tile that holds hidden neuron n (0..63) of a
// layer's output: n = 32 nj + 8 g + 4 q + u is stored at 16 (2 nj + g/2) +
// 8 q + 4 (g mod 2) + u (bits 2 and 3 of n swapped; an involution).  Lane
// (r, q) of a 32-row block then writes exactly the columns it reads as the
// next v_mfma_f32_32x32x16_bf16's B operand, 16 k + 8 q .. 16 k + 8 q + 7.
// The next layer's weights are stored with their input columns in the same
// order (the walk copy of W2 and W3).
__host__ __device__ constexpr int mlp_col(int n) { return (n & ~12) | ((n & 4) << 1) | ((n & 8) >> 1); }

// 32x32 outputs (neurons nj*32 + 8 g + 4 q + u of tile rows r) of W * buf^T
// for one 32-row block: four v_mfma_f32_32x32x16_bf16 steps over columns
// 0..63 (lane (r, q) holds columns 16 k + 8 q .. + 7) and, for K = 72, one
// v_mfma_f32_32x32x8_bf16 step over columns 64..71 (lane holds 64 + 4 q .. + 3).
template <int K>
__device__ inline f32x16 mlp_block(const uint16_t* row, const uint16_t* wrow, int q)
{
    f32x16 acc = {0};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        const bf16x8 x = *(const bf16x8*)(row + 16 * kb + 8 * q);
        const bf16x8 w = *(const bf16x8*)(wrow + 16 * kb + 8 * q);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16v8, w), __builtin_bit_cast(bf16v8, x),
                                                      acc, 0, 0, 0);
    }
    if (K == 72) {
        const bf16x4 x = *(const bf16x4*)(row + 64 + 4 * q);
        const bf16x4 w = *(const bf16x4*)(wrow + 64 + 4 * q);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(w, x, acc, 0, 0, 0);
    }
    return acc;
}

// One hidden layer for the wave's 64 rows, in place: buf = bf16(gelu(buf *
// W^T + b)) with the outputs in mlp_col order.  Computed transposed (W *
// buf^T, weights as the MFMA's A operand): each lane then holds four
// consecutive output neurons of one row per group, stored as one packed
// 8-byte write.  Lane (r, q) of a 32-row block overwrites only columns that
// it alone read (mlp_col); both 32-neuron halves read the whole block, so
// the first half's outputs wait in the lane's staging slots until the second
// half has been computed.
template <int K>
__device__ inline void mlp_layer(uint16_t* buf, const uint16_t* W, const float* __restrict__ bias, int lane)
{
    const int r32 = lane & 31, q = lane >> 5;
    bf16x4* held = (bf16x4*)(buf + 64 * kMlpStride) + 4 * lane;
#pragma unroll 1
    for (int mi = 0; mi < 2; ++mi) {
        uint16_t* row = buf + (mi * 32 + r32) * kMlpStride;
#pragma unroll 1
        for (int nj = 0; nj < 2; ++nj) {
            const f32x16 acc = mlp_block<K>(row, W + (nj * 32 + r32) * K, q);
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int n0 = nj * 32 + 8 * g + 4 * q;
                const float4 b = *(const float4*)(bias + n0);  // one LDS read for four biases
                const bf16x4 v = to_bf16x4(gelu_f(acc[4 * g] + b.x), gelu_f(acc[4 * g + 1] + b.y),
                                           gelu_f(acc[4 * g + 2] + b.z), gelu_f(acc[4 * g + 3] + b.w));
                if (nj == 0)
                    held[g] = v;
                else
                    *(bf16x4*)(row + mlp_col(n0)) = v;
            }
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) *(bf16x4*)(row + mlp_col(8 * g + 4 * q)) = held[g];
    }
}

// Third hidden layer and output neuron for the wave's 64 rows: y = w4 .
// bf16(gelu(buf * W3^T + b3)) + b4, returned in lane `row`.  Each lane sums
// the 32 neurons it holds of rows r and 32 + r (lane = r + 32 q); the other
// 32 of each row are in lane r + 32 (1 - q), which sends its sum for the
// row this lane returns.  The activations are not stored.
__device__ inline float mlp_last_layer(const uint16_t* buf, const MlpNet& m, const float* __restrict__ small, int lane)
{
    const int r32 = lane & 31, q = lane >> 5;
    float p0 = 0.0f, p1 = 0.0f;
#pragma unroll 1
    for (int mi = 0; mi < 2; ++mi) {
        const uint16_t* row = buf + (mi * 32 + r32) * kMlpStride;
        float part = 0.0f;
#pragma unroll 1
        for (int nj = 0; nj < 2; ++nj) {
            const f32x16 acc = mlp_block<kMlpHid>(row, m.W3 + (nj * 32 + r32) * kMlpHid, q);
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int n0 = nj * 32 + 8 * g + 4 * q;
                const float4 b = *(const float4*)(small + kMlpSmallB3 + n0);
                const float4 w = *(const float4*)(small + kMlpSmallW4 + n0);
                const bf16x4 a = to_bf16x4(gelu_f(acc[4 * g] + b.x), gelu_f(acc[4 * g + 1] + b.y),
                                           gelu_f(acc[4 * g + 2] + b.z), gelu_f(acc[4 * g + 3] + b.w));
                part += w.x * bf2f((uint16_t)a[0]);
                part += w.y * bf2f((uint16_t)a[1]);
                part += w.z * bf2f((uint16_t)a[2]);
                part += w.w * bf2f((uint16_t)a[3]);
            }
        }
        if (mi == 0)
            p0 = part;
        else
            p1 = part;
    }
    const float other = __shfl_xor(q ? p0 : p1, 32);
    return (q ? p1 : p0) + other + small[kMlpSmallB4];
}

// Stages the network's biases, output weights and output bias in LDS
// (kMlpSmallFloats floats); the caller synchronises the block afterwards.
__device__ inline void mlp_stage_small(const MlpNet& m, float* small, int tid, int n_threads)
{
    for (int i = tid; i < kMlpSmallFloats; i += n_threads) {
        float v = 0.0f;
        if (i < kMlpSmallB2) v = m.b1[i];
        else if (i < kMlpSmallB3) v = m.b2[i - kMlpSmallB2];
        else if (i < kMlpSmallW4) v = m.b3[i - kMlpSmallB3];
        else if (i < kMlpSmallB4) v = bf2f(m.w4[i - kMlpSmallW4]);
        else if (i == kMlpSmallB4) v = *m.b4;
        small[i] = v;
    }
}

// Forward pass of the wave's 64 feature rows in xs (LDS, kMlpTileElems:
// 64 x kMlpStride, overwritten by the hidden activations, and the staging
// slots), with the small operands staged by mlp_stage_small.  Returns row
// `lane`'s output.  Must be called by all 64 lanes together.
__device__ inline float mlp_forward_wave(const MlpNet& m, uint16_t* xs, const float* small, int lane)
{
    wave_sync();
    mlp_layer<kMlpInPad>(xs, m.W1, small + kMlpSmallB1, lane);
    wave_sync();
    mlp_layer<kMlpHid>(xs, m.W2, small + kMlpSmallB2, lane);
    wave_sync();
    const float y = mlp_last_layer(xs, m, small, lane);
    wave_sync();
    return y;
}

}  // namespace frw
