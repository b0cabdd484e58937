// Learned square-size proposal: occupancy-window features and a 67-64-64-64-1
// GELU network evaluated with bf16 MFMA (v_mfma_f32_16x16x32_bf16 on 16-row
// blocks) and FP32 accumulation, and the occupancy-pyramid probe that
// verifies a proposal.
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
constexpr int kMlpStride = 72;  // LDS row stride (bf16 elements) of activations and weights
// a wave's activation tile: 64 rows of kMlpStride, then the 64 FP32 outputs
constexpr int kMlpTileElems = 64 * kMlpStride + 64 * 2;
// a block's copy of the weights in LDS: W1 [64][72], W2 and W3 [64][64] (the
// walk copies), each row at a stride of kMlpStride (144 bytes, so the 16 rows
// a wave reads at once fall on distinct banks, as in the activation tile)
constexpr int kMlpWElems = 64 * kMlpStride;
constexpr int kMlpWLdsElems = 3 * kMlpWElems;  // 27 KB
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
    const uint64_t* occ;      // bitmaps in 8x8-cell tiles (occ_tiles), level l at word occ_off[l]
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
// x P(x), P a degree-8 polynomial in x = 1 - t / 4.5 (x P vanishes at t = 4.5,
// where t Phi(-t) < 1.6e-5), t clamped there: GELU = fma(-(t x), P(x),
// max(z, 0)), 13 FMA / min / max / mul instructions, no transcendental;
// within 1.7e-5 of the exact GELU, far inside a bf16 activation's rounding.
// Coefficients from fit_gelu.py (a minimax fit).
__device__ inline float gelu_f(float z)
{
    const float t = fminf(fabsf(z), 4.5f);
    const float x = __builtin_fmaf(t, -1.0f / 4.5f, 1.0f);
    const float q = t * x;
    float p = 1.007601249e+01f;
    p = __builtin_fmaf(p, x, -3.770581612e+01f);
    p = __builtin_fmaf(p, x, 5.370768662e+01f);
    p = __builtin_fmaf(p, x, -3.812867467e+01f);
    p = __builtin_fmaf(p, x, 1.582026759e+01f);
    p = __builtin_fmaf(p, x, -3.744195627e+00f);
    p = __builtin_fmaf(p, x, 5.062563132e-01f);
    p = __builtin_fmaf(p, x, -3.242632146e-02f);
    p = __builtin_fmaf(p, x, 9.052214061e-04f);
    return __builtin_fmaf(-q, p, fmaxf(z, 0.0f));
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

// Occupancy bitmaps are stored in tiles of 8x8 cells, one 64-bit word each:
// bit 8 y + x of tile (tx, ty) is cell (8 tx + x, 8 ty + y), tiles row-major,
// occ_tiles(l) per side (levels below 3 use the low cells of one tile).
__host__ __device__ constexpr int64_t occ_tiles(int l) { return l < 3 ? 1 : (int64_t)1 << (l - 3); }

// Bytes 0..k-1 of a 64-bit word set (k = 0..8).
__host__ __device__ inline uint64_t occ_low_bytes(int k) { return k >= 8 ? ~0ull : (1ull << (8 * k)) - 1ull; }

// 64 occupancy bits of the 8x8 window (bit v*8+u) at level l around p: the
// cells (c - 4 + u, c' - 4 + v) of p's cell (c, c').  The window meets at
// most 2x2 tiles: four word loads, merged by byte-wise shifts.  Cells
// outside the grid read as occupied.
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
    const int64_t i0 = (int64_t)fx - 4, j0 = (int64_t)fy - 4;  // the window's first cell
    const int64_t T = occ_tiles(l);
    const int64_t tx = i0 >> 3, ty = j0 >> 3;  // floor(i0 / 8), also below 0
    const int ox = (int)(i0 & 7), oy = (int)(j0 & 7);
    const uint64_t* base = m.occ + m.occ_off[l];
    auto tile = [&](int64_t x, int64_t y) -> uint64_t {
        return (x >= 0 && x < T && y >= 0 && y < T) ? base[y * T + x] : 0ull;
    };
    const uint64_t t00 = tile(tx, ty), t10 = tile(tx + 1, ty), t01 = tile(tx, ty + 1), t11 = tile(tx + 1, ty + 1);
    // window column u: column ox + u of the left tile (< 8) or ox + u - 8 of
    // the right one, in every byte (row) at once
    const uint64_t left = 0x0101010101010101ull * (0xFFu >> ox);
    auto merge = [&](uint64_t a, uint64_t b) -> uint64_t { return ((a >> ox) & left) | ((b << (8 - ox)) & ~left); };
    const uint64_t top = merge(t00, t10), bottom = merge(t01, t11);
    // window row v: row oy + v of the top pair (< 8) or oy + v - 8 of the bottom one
    uint64_t bits = (top >> (8 * oy)) | (oy ? bottom << (64 - 8 * oy) : 0ull);
    // window columns [c0, c1) and rows [r0, r1) lie inside the grid
    const int c0 = (int)(i0 < 0 ? -i0 : 0), c1 = (int)(n - i0 < 8 ? (n - i0 > c0 ? n - i0 : c0) : 8);
    const int r0 = (int)(j0 < 0 ? -j0 : 0), r1 = (int)(n - j0 < 8 ? (n - j0 > r0 ? n - j0 : r0) : 8);
    const uint32_t cols = ((1u << c1) - 1u) & ~((1u << c0) - 1u);
    const uint64_t inside = (0x0101010101010101ull * cols) & occ_low_bytes(r1) & ~occ_low_bytes(r0);
    bits = (bits & inside) | ~inside;
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

// The walk kernel's feature row in its activation tile: columns 64..67 as in
// the 72-wide row (level / 16, position in the cell, 0), the 64 window bits
// packed in columns 68..71 (8 bytes) instead of written as 64 bf16 values;
// the first layer expands them as it loads its operand (mlp_expand_bits).
constexpr int kMlpBitsCol = 68;
__device__ inline void mlp_feature_bits(uint16_t* row, uint64_t bits, int level, float frac_x, float frac_y)
{
    bf16x4 a;
    a[0] = (short)f2bf((float)level / 16.0f);
    a[1] = (short)f2bf(frac_x);
    a[2] = (short)f2bf(frac_y);
    a[3] = 0;
    *(bf16x4*)(row + 64) = a;
    *(uint64_t*)(row + kMlpBitsCol) = bits;
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
typedef float f32x4 __attribute__((ext_vector_type(4)));

// The network runs on blocks of 16 rows of the wave's activation tile with
// v_mfma_f32_16x16x32_bf16, computed transposed (W * rows^T, weights as the A
// operand): lane l = j + 16 h (j = row in the block, h = 0..3) supplies
// columns 32 k + 8 h .. + 7 of row j as the B operand and receives neurons
// 16 t + 4 h + r (t = 0..3 tiles of 16 neurons, r = 0..3) of row j.
//
// Column of the tile that holds hidden neuron n (0..63) of a layer's output:
// n = 16 t + 4 h + r goes to 32 (t / 2) + 8 h + 4 (t mod 2) + r, so that
// lane (j, h) holds exactly the columns it supplies as the next layer's B
// operand: a hidden layer's activations stay in the lane's registers (no
// other lane needs them, hence no LDS round trip and no barrier between
// layers).  The next layer's weights are stored with their input columns in
// the same order (the walk copy of W2 and W3).
__host__ __device__ constexpr int mlp_col(int n) { return (n & 0x23) | ((n & 0x0C) << 1) | ((n & 0x10) >> 2); }

// Eight window bits (bit k of b = the k-th column) as a bf16x8 operand, 1.0
// (0x3F80) where the bit is set, two columns per 32-bit word: for the bit
// pair x = 2 b1 + b0, x * 0x8001 holds b0 at bits 0 and 15 and b1 at bits 1
// and 16, the mask keeps bits 0 and 16, and * 0x3F80 puts bf16(1) in the
// halves whose bit is set (every product below 2^24: 24-bit multiplies).
__device__ __forceinline__ bf16x8 mlp_expand_bits(uint32_t b)
{
    uint32_t w[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const uint32_t x = (b >> (2 * p)) & 3u;
        w[p] = __umul24(__umul24(x, 0x8001u) & 0x10001u, 0x3F80u);
    }
    return __builtin_bit_cast(bf16x8, w);
}

// The four 16x16 output tiles (neurons 16 t + 4 h + r of row j) of W * rows^T
// over 64 input columns, W in LDS at a row stride of kMlpStride, the lane's
// B operand in registers: x[k] = columns 32 k + 8 h .. + 7 of row j.
__device__ __forceinline__ void mlp_tiles64(const bf16x8 x[2], const uint16_t* W, int j, int h, f32x4 acc[4])
{
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const bf16x8 w = *(const bf16x8*)(W + (16 * t + j) * kMlpStride + 32 * kb + 8 * h);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16v8, w),
                                                             __builtin_bit_cast(bf16v8, x[kb]), acc[t], 0, 0, 0);
        }
    }
}

// The first layer's tiles from the walk kernel's feature row in LDS: the
// window bits (packed in columns 68..71) expanded to the operand of columns
// 0..63 as it is loaded, then one v_mfma_f32_16x16x16_bf16 step over
// columns 64..79 of which only columns 64..67 (lanes with h = 0) are
// non-zero (column 67 and W1's columns 67..71 are zero).
__device__ __forceinline__ void mlp_tiles_first(const uint16_t* row, const uint16_t* W, int j, int h, f32x4 acc[4])
{
    const uint64_t bits = *(const uint64_t*)(row + kMlpBitsCol);
    bf16x8 x[2];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) x[kb] = mlp_expand_bits((uint32_t)(bits >> (32 * kb + 8 * h)) & 0xFFu);
    mlp_tiles64(x, W, j, h, acc);
    const bool in = h == 0;
    const bf16x4 zero = {0, 0, 0, 0};
    const bf16x4 xt = in ? *(const bf16x4*)(row + 64) : zero;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const bf16x4 w = in ? *(const bf16x4*)(W + (16 * t + j) * kMlpStride + 64) : zero;
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w, xt, acc[t], 0, 0, 0);
    }
}

// A hidden layer's activations bf16(gelu(acc + b)) as the next layer's B
// operand: x[k] holds neurons 16 (2 k) + 4 h + r, then 16 (2 k + 1) + 4 h + r
// (r = 0..3), i.e. the columns 32 k + 8 h .. + 7 in mlp_col order.
__device__ __forceinline__ void mlp_activate(const f32x4 acc[4], const float* __restrict__ bias, int h, bf16x8 x[2])
{
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        bf16x4 half[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int t = 2 * kb + u;
            const float4 b = *(const float4*)(bias + 16 * t + 4 * h);  // one LDS read for four biases
            half[u] = to_bf16x4(gelu_f(acc[t][0] + b.x), gelu_f(acc[t][1] + b.y), gelu_f(acc[t][2] + b.z),
                                gelu_f(acc[t][3] + b.w));
        }
        x[kb] = __builtin_shufflevector(half[0], half[1], 0, 1, 2, 3, 4, 5, 6, 7);
    }
}

// Third hidden layer and output neuron for a 16-row block: y = w4 .
// bf16(gelu(x * W3^T + b3)) + b4 of row j, in every lane j + 16 h, x the
// second layer's activations in the lane's registers.  Each lane sums its 16
// neurons; the four lanes of a row add their sums.
__device__ __forceinline__ float mlp_last_layer(const bf16x8 x[2], const uint16_t* W3, const float* __restrict__ small, int j,
                                                int h)
{
    f32x4 acc[4];
    mlp_tiles64(x, W3, j, h, acc);
    float part = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int n0 = 16 * t + 4 * h;
        const float4 b = *(const float4*)(small + kMlpSmallB3 + n0);
        const float4 w = *(const float4*)(small + kMlpSmallW4 + n0);
        const bf16x4 a = to_bf16x4(gelu_f(acc[t][0] + b.x), gelu_f(acc[t][1] + b.y), gelu_f(acc[t][2] + b.z),
                                   gelu_f(acc[t][3] + b.w));
        part += w.x * bf2f((uint16_t)a[0]);
        part += w.y * bf2f((uint16_t)a[1]);
        part += w.z * bf2f((uint16_t)a[2]);
        part += w.w * bf2f((uint16_t)a[3]);
    }
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    return part + small[kMlpSmallB4];
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

// Copies the three weight matrices into a block's LDS (kMlpWLdsElems bf16,
// rows at a stride of kMlpStride), 16 bytes per copy; the caller
// synchronises the block afterwards.
__device__ inline void mlp_stage_weights(const MlpNet& m, uint16_t* w, int tid, int n_threads)
{
    constexpr int kW1Chunks = 64 * kMlpInPad / 8;  // 9 chunks of 8 per W1 row
    constexpr int kWChunks = 64 * kMlpHid / 8;     // 8 per W2 or W3 row
    for (int i = tid; i < kW1Chunks + 2 * kWChunks; i += n_threads) {
        const uint16_t* src;
        uint16_t* dst;
        if (i < kW1Chunks) {
            const int r = i / (kMlpInPad / 8), c = i % (kMlpInPad / 8);
            src = m.W1 + r * kMlpInPad + 8 * c;
            dst = w + r * kMlpStride + 8 * c;
        } else {
            const int k = i - kW1Chunks, mat = k / kWChunks, q = k % kWChunks;
            const int r = q / (kMlpHid / 8), c = q % (kMlpHid / 8);
            src = (mat ? m.W3 : m.W2) + r * kMlpHid + 8 * c;
            dst = w + (1 + mat) * kMlpWElems + r * kMlpStride + 8 * c;
        }
        *(uint4*)dst = *(const uint4*)src;
    }
}

// Forward pass of the first n_rows (0..64) feature rows of the wave's tile xs
// (LDS, kMlpTileElems: 64 x kMlpStride rows as mlp_feature_bits writes them,
// then 64 floats of outputs), in ceil(n_rows / 16) blocks of 16 rows (the
// hidden activations in registers), with the weights staged in LDS by
// mlp_stage_weights (w) and the small operands by mlp_stage_small.  Returns
// the output of row `row` (any lane; rows at or beyond n_rows give 0).  Must
// be called by all 64 lanes together with the same n_rows.
__device__ __forceinline__ float mlp_forward_rows(const uint16_t* w, uint16_t* xs, const float* small, int lane, int n_rows,
                                         int row)
{
    float* ys = (float*)(xs + 64 * kMlpStride);
    const int j = lane & 15, h = lane >> 4;
    wave_sync();  // the feature rows written by other lanes
    const int nb = (n_rows + 15) >> 4;
#pragma unroll 1
    for (int b = 0; b < nb; ++b) {
        uint16_t* block = xs + 16 * b * kMlpStride;
        // the weight fragments are read again for every block rather than
        // held across the loop (all three layers' would take about 100
        // registers per lane): an offset the compiler cannot see through
        // keeps the reads in the loop
        int off = 0;
        asm volatile("" : "+s"(off));
        const uint16_t* wb = w + off;
        f32x4 acc[4];
        bf16x8 x[2];
        mlp_tiles_first(block + j * kMlpStride, wb, j, h, acc);
        mlp_activate(acc, small + kMlpSmallB1, h, x);
        mlp_tiles64(x, wb + kMlpWElems, j, h, acc);
        mlp_activate(acc, small + kMlpSmallB2, h, x);
        const float y = mlp_last_layer(x, wb + 2 * kMlpWElems, small, j, h);
        if (lane < 16) ys[16 * b + lane] = y;
    }
    wave_sync();
    return row < n_rows ? ys[row] : 0.0f;
}

}  // namespace frw
