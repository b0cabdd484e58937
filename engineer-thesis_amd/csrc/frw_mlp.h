// Learned square-size proposal: occupancy-window features and a 65-64-64-64-1
// GELU network evaluated with bf16 MFMA (v_mfma_f32_32x32x8_bf16) and FP32
// accumulation.  No counterpart in the reference: the reference sizes every
// square by bisection (randomWalk/CPU/quadTree.cpp:264-312).
//
// Features of a point p whose previous square had half-size h:
//   K0 = smallest k with 2^k > max(W, H) of the bounds, Lmax = min(K0, 12),
//   level l = clamp(K0 - floor(log2 max(h, 1)) + 1, 0, Lmax), cell s = 2^(K0-l),
//   (ci, cj) = floor((p - bounds corner) / s);
//   f[v*8+u] = occupancy of cell (ci-4+u, cj-4+v) at level l (1 outside the
//   grid), f[64] = l / 16, f[65..71] = 0.
// Occupancy bit of a cell: some closed rect meets the half-open cell.
// Output y estimates log2(d* / s); the proposed half-size is s * 2^y * alpha.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "frw_device.h"

namespace frw {

constexpr int kMlpIn = 65;
constexpr int kMlpInPad = 72;   // K of the first layer, a multiple of 8
constexpr int kMlpHid = 64;
constexpr int kMlpStride = 72;  // LDS row stride (bf16 elements) of activations
constexpr int kMlpMaxLevel = 12;

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

struct MlpView {
    const uint16_t* W1;   // bf16 [64][72]
    const uint16_t* W2;   // bf16 [64][64]
    const uint16_t* W3;   // bf16 [64][64]
    const uint16_t* w4;   // bf16 [64]
    const float* b1;
    const float* b2;
    const float* b3;
    float b4;
    float alpha;          // safety factor on the proposal
    const uint32_t* occ;  // occupancy pyramid, level l at word occ_off[l]
    int64_t occ_off[kMlpMaxLevel + 1];
    int4 bounds;
    int K0;
    int Lmax;
};

__host__ __device__ inline uint16_t f2bf(float x)
{
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }

__device__ inline float gelu_f(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ inline float gelu_grad_f(float z)
{
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

__device__ inline int mlp_level(const MlpView& m, double h, double* cell)
{
    const int e = h >= 1.0 ? ilogb(h) : 0;
    int l = m.K0 - e + 1;
    l = l < 0 ? 0 : (l > m.Lmax ? m.Lmax : l);
    *cell = ldexp(1.0, m.K0 - l);
    return l;
}

// 64 occupancy bits of the 8x8 window (bit v*8+u) at level l around p.
__device__ inline uint64_t mlp_window(const MlpView& m, int l, double cell, double px, double py)
{
    const int64_t n = (int64_t)1 << l;
    const int64_t ci = (int64_t)floor((px - (double)m.bounds.x) / cell);
    const int64_t cj = (int64_t)floor((py - (double)m.bounds.y) / cell);
    const uint32_t* base = m.occ + m.occ_off[l];
    uint64_t bits = 0;
    for (int v = 0; v < 8; ++v) {
        const int64_t j = cj - 4 + v;
        for (int u = 0; u < 8; ++u) {
            const int64_t i = ci - 4 + u;
            uint64_t b = 1;
            if (i >= 0 && j >= 0 && i < n && j < n) {
                const int64_t idx = j * n + i;
                b = (base[idx >> 5] >> (idx & 31)) & 1u;
            }
            bits |= b << (v * 8 + u);
        }
    }
    return bits;
}

// Writes one feature row (72 bf16) of this lane's point into LDS.
__device__ inline void mlp_feature_row(uint16_t* row, uint64_t bits, int level)
{
    const uint16_t one = 0x3F80u;  // bf16(1.0)
    for (int k = 0; k < 64; ++k) row[k] = ((bits >> k) & 1u) ? one : (uint16_t)0;
    row[64] = f2bf((float)level / 16.0f);
    for (int k = 65; k < kMlpInPad; ++k) row[k] = 0;
}

// Orders this wave's LDS accesses (the other lanes' rows) without a block barrier.
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One hidden layer for the wave's 64 rows: out = bf16(gelu(in * W^T + b)).
// 2x2 tiles of 32x32, K/8 MFMA steps each.  Lane l feeds row/column l%32 and
// k-quad l/32; its 16 results sit at rows 8*(r/4) + 4*(l/32) + r%4.
template <int K>
__device__ inline void mlp_layer(const uint16_t* in, uint16_t* out, const uint16_t* W, const float* bias,
                                 int lane)
{
    const int r32 = lane & 31, q = lane >> 5;
#pragma unroll
    for (int mi = 0; mi < 2; ++mi) {
#pragma unroll
        for (int nj = 0; nj < 2; ++nj) {
            f32x16 acc = {0};
#pragma unroll
            for (int kb = 0; kb < K / 8; ++kb) {
                const int kk = kb * 8 + 4 * q;
                const bf16x4 a = *(const bf16x4*)(in + (mi * 32 + r32) * kMlpStride + kk);
                const bf16x4 b = *(const bf16x4*)(W + (nj * 32 + r32) * K + kk);
                acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
            }
            const int col = nj * 32 + r32;
            const float bb = bias[col];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = mi * 32 + 8 * (r >> 2) + 4 * q + (r & 3);
                out[row * kMlpStride + col] = f2bf(gelu_f(acc[r] + bb));
            }
        }
    }
}

// Forward pass of the wave's 64 feature rows in xs (LDS, 64 x kMlpStride);
// hs is a second LDS buffer of the same size.  Returns row `lane`'s output.
// Must be called by all 64 lanes together.
__device__ inline float mlp_forward_wave(const MlpView& m, uint16_t* xs, uint16_t* hs, int lane)
{
    wave_sync();
    mlp_layer<kMlpInPad>(xs, hs, m.W1, m.b1, lane);
    wave_sync();
    mlp_layer<kMlpHid>(hs, xs, m.W2, m.b2, lane);
    wave_sync();
    mlp_layer<kMlpHid>(xs, hs, m.W3, m.b3, lane);
    wave_sync();
    float y = 0.0f;
    for (int j = 0; j < kMlpHid; ++j) y += bf2f(m.w4[j]) * bf2f(hs[lane * kMlpStride + j]);
    wave_sync();
    return y + m.b4;
}

}  // namespace frw
