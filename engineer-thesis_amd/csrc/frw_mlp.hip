// Learned square-size proposal: occupancy pyramid, training-set harvest,
// bf16-MFMA training (forward, backward, Adam) and the batched inference hook.
// The network definition and feature layout are in frw_mlp.h.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "frw_internal.h"
#include "frw_mlp.h"

struct frw_mlp {
    const frw_tree* tree = nullptr;
    int device = 0;
    int K0 = 0, Lmax = 0, Vmax = 0, Pmax = 0;  // Pmax = max(Lmax, Vmax)
    std::vector<int64_t> occ_off;  // bitmap word offsets, levels 0..Pmax
    std::vector<int64_t> sat_off;  // summed-area-table entry offsets, levels 0..Vmax
    uint64_t* d_occ = nullptr;     // tiles of 8x8 cells
    uint32_t* d_sat = nullptr;
    int64_t* d_occ_off = nullptr;  // device copies of the offsets
    int64_t* d_sat_off = nullptr;
    float* d_params = nullptr;     // FP32 master parameters
    uint16_t* d_bf = nullptr;      // bf16 copy of W1 | W2 | W3 | w4 (training)
    uint16_t* d_bfw = nullptr;     // bf16 W2 | W3 with input columns in mlp_col order (the walk's network)
    float* d_adam_m = nullptr;
    float* d_adam_v = nullptr;
    float* d_grad = nullptr;       // summed gradient (frw_mlp_gradient)
    uint16_t* d_bfT = nullptr;     // bf16 W2^T | W3^T, [in][out] (the backward pass)
    float* d_part = nullptr;       // partial sums, one row of kPartStride floats per 32 batch rows
    int64_t part_rows = 0;
    float* d_loss = nullptr;       // loss sum, batches summed
    int64_t adam_step = 0;
    // training set
    int64_t n_samples = 0;
    int64_t cap_samples = 0;       // rows allocated (grown geometrically)
    uint16_t* d_X = nullptr;       // [cap][72] bf16
    float* d_t = nullptr;          // [cap] targets
};

namespace frw {
namespace {

constexpr int kBfW1 = 0;
constexpr int kBfW2 = kBfW1 + kMlpHid * kMlpInPad;
constexpr int kBfW3 = kBfW2 + kMlpHid * kMlpHid;
constexpr int kBfW4 = kBfW3 + kMlpHid * kMlpHid;
constexpr int kBfCount = kBfW4 + kMlpHid;

// ---------------------------------------------------------------- occupancy
// Levels are stored in tiles of 8x8 cells (occ_tiles, one 64-bit word each).
// One lane per rect: sets the bits of every cell of level L the closed rect
// meets, one 64-bit mask per tile.  Only the finest level is marked this
// way; occ_reduce derives the coarser ones (marking every level here had put
// every rect's atomics on the few words of the coarse levels).
__global__ void occ_mark(const int4* __restrict__ rects, int n, int4 B, int K0, int L,
                         const int64_t* __restrict__ off, uint64_t* __restrict__ occ)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int4 q = rects[r];
    const int64_t x0 = (int64_t)q.x - B.x, y0 = (int64_t)q.y - B.y;
    const int64_t x1 = (int64_t)q.z - B.x, y1 = (int64_t)q.w - B.y;
    const int sh = K0 - L;
    const int64_t nl = (int64_t)1 << L;
    const int64_t i0 = x0 >> sh, i1 = min(x1 >> sh, nl - 1);
    const int64_t j0 = y0 >> sh, j1 = min(y1 >> sh, nl - 1);
    const int64_t T = occ_tiles(L);
    uint64_t* base = occ + off[L];
    for (int64_t ty = j0 >> 3; ty <= j1 >> 3; ++ty) {
        const int r0 = (int)max<int64_t>(j0 - 8 * ty, 0), r1 = (int)min<int64_t>(j1 - 8 * ty, 7);  // rows r0..r1
        const uint64_t rows = occ_low_bytes(r1 + 1) & ~occ_low_bytes(r0);
        for (int64_t tx = i0 >> 3; tx <= i1 >> 3; ++tx) {
            const int c0 = (int)max<int64_t>(i0 - 8 * tx, 0), c1 = (int)min<int64_t>(i1 - 8 * tx, 7);
            const uint32_t cols = ((2u << c1) - 1u) & ~((1u << c0) - 1u);  // columns c0..c1
            atomicOr((unsigned long long*)(base + ty * T + tx),
                     (unsigned long long)(rows & (0x0101010101010101ull * cols)));
        }
    }
}

// One lane per tile of level l from the (up to) 2x2 tiles of level l + 1
// under it: cell (i, j) is occupied when one of its four children (2i + a,
// 2j + b) is.  The children's half-open cells tile the parent's, and the
// finest level's cell ranges, shifted right by one bit, are the parent
// level's, so this equals marking level l directly.  A child tile's 8x8
// cells become 4x4: rows ORed in pairs, then columns, the even bits of the
// even rows compacted into one nibble per row, placed in the parent tile's
// quadrant.  Child tiles beyond level l + 1's grid (levels below 3) are empty.
__global__ void occ_reduce(const uint64_t* __restrict__ child, int l, uint64_t* __restrict__ parent)
{
    const int64_t Tp = occ_tiles(l), Tc = occ_tiles(l + 1);
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= Tp * Tp) return;
    const int64_t TX = w % Tp, TY = w / Tp;
    uint64_t out = 0ull;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int qx = q & 1, qy = q >> 1;
        const int64_t cx = 2 * TX + qx, cy = 2 * TY + qy;
        if (cx >= Tc || cy >= Tc) continue;
        uint64_t x = child[cy * Tc + cx];
        x |= x >> 8;  // rows 2v and 2v + 1 in row 2v
        x |= x >> 1;  // columns 2u and 2u + 1 in column 2u
        x &= 0x0055005500550055ull;
        x = (x | (x >> 1)) & 0x0033003300330033ull;
        x = (x | (x >> 2)) & 0x000F000F000F000Full;  // parent row v: the nibble of byte 2v
#pragma unroll
        for (int v = 0; v < 4; ++v) out |= ((x >> (16 * v)) & 0xFull) << ((4 * qy + v) * 8 + 4 * qx);
    }
    parent[w] = out;
}

// ---------------------------------------------------------------- summed-area tables
// Row pass, one block of 256 lanes per grid row j of a level's tiles: S[(j+1)(n+1) + i + 1] =
// occupied cells (a, j) with a <= i; S[(j+1)(n+1)] = 0.  Row 0 is zeroed by
// the caller.
__global__ void __launch_bounds__(256) sat_rows(const uint64_t* __restrict__ tiles, int l,
                                                uint32_t* __restrict__ S)
{
    __shared__ uint32_t part[256];
    const int64_t n = (int64_t)1 << l, T = occ_tiles(l);
    const int64_t j = blockIdx.x;
    const int t = threadIdx.x;
    const int64_t c = (n + 255) / 256;
    const int64_t a = min(n, (int64_t)t * c), b = min(n, a + c);
    // row j of the level is byte j mod 8 of its tile row's words: the lane
    // reads one word per 8 cells of its range [a, b)
    const uint64_t* trow = tiles + (j >> 3) * T;
    const int sh = 8 * (int)(j & 7);
    const int64_t k0 = a >> 3, k1 = b > a ? ((b - 1) >> 3) + 1 : k0;
    auto cells = [&](int64_t k, int* lo, int* hi) -> uint32_t {  // the tile's byte and the lane's cells in it
        *lo = (int)max<int64_t>(a - 8 * k, 0);
        *hi = (int)min<int64_t>(b - 8 * k, 8);
        return (uint32_t)(trow[k] >> sh) & 0xFFu;
    };
    uint32_t cnt = 0;
    for (int64_t k = k0; k < k1; ++k) {
        int lo, hi;
        const uint32_t by = cells(k, &lo, &hi);
        cnt += (uint32_t)__popc(by & ((1u << hi) - 1u) & ~((1u << lo) - 1u));
    }
    part[t] = cnt;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - cnt;
    uint32_t* row = S + (j + 1) * (n + 1);
    if (t == 0) row[0] = 0u;
    for (int64_t k = k0; k < k1; ++k) {
        int lo, hi;
        const uint32_t by = cells(k, &lo, &hi);
        for (int u = lo; u < hi; ++u) {
            run += (by >> u) & 1u;
            row[8 * k + u + 1] = run;
        }
    }
}

// Column pass: one lane per column, running sums down the rows.
__global__ void sat_cols(int64_t n, uint32_t* __restrict__ S)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    const int64_t w = n + 1;
    uint32_t run = 0;
    int64_t j = 1;
    for (; j + 8 <= n + 1; j += 8) {  // eight loads in flight per lane
        uint32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = S[(j + k) * w + i];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            run += v[k];
            S[(j + k) * w + i] = run;
        }
    }
    for (; j <= n; ++j) {
        run += S[j * w + i];
        S[j * w + i] = run;
    }
}

// Verification probe for host-supplied squares (test hook).
__global__ void certify_kernel(MlpPyramid m, const double* __restrict__ pts, const double* __restrict__ d,
                               int64_t n, uint8_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = pts[2 * i], py = pts[2 * i + 1], h = d[i];
    bool ok = false;
    if (h > 0.0 && h < 1e300) ok = mlp_certify(m, px - h, py - h, px + h, py + h, h);
    out[i] = ok ? 1 : 0;
}

// ---------------------------------------------------------------- features
// One lane per point: level, cell and the 72-wide bf16 feature row.
__global__ void features_kernel(MlpPyramid m, const double* __restrict__ pts, const double* __restrict__ hs,
                                int64_t n, uint16_t* __restrict__ X, const double* __restrict__ dstar,
                                float* __restrict__ target, double* __restrict__ cell_out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = pts[2 * i], py = pts[2 * i + 1];
    double cell;
    const int l = mlp_level(m, hs[i], &cell);
    float fx, fy;
    const uint64_t bits = mlp_window(m, l, cell, px, py, &fx, &fy);
    mlp_feature_row(X + i * kMlpInPad, bits, l, fx, fy);
    if (target) {
        // log2(d* / s), with d* below s/16 (no proposal could be used) lumped together
        const float t = (float)log2(fmax(dstar[i], cell / 16.0) / cell);
        target[i] = fminf(fmaxf(t, -4.0f), 4.0f);
    }
    if (cell_out) cell_out[i] = cell;
}

// ---------------------------------------------------------------- training
// One optimiser step is two launches.  train_tile: one wave per 16 rows of
// the batch runs the forward pass, the loss and the backward pass of its rows
// on bf16 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation) entirely in LDS
// and registers, and writes its share of the gradient (and of the loss) to a
// row of partial sums.  train_apply: eight lanes per parameter add the rows
// (each a fixed, interleaved eighth of them, in row order; then the eight
// sums in lane order), take the Adam step and write the bf16 weight copies.
// The result does not depend on scheduling: each partial is a fixed sequence
// of operations, and the rows are added in a fixed order.
//
// Rounding points (those of the FP32/bf16 restatement the tests compare
// against): inputs, activations and hidden-layer gradients are bf16, every
// product accumulates in FP32; Z = A W^T + b, A' = bf16(gelu(Z)); y = w4 . A3
// + b4; dy = (y - t) / B; D3 = bf16(dy w4 gelu'(Z3)); F = D W,
// D' = bf16(F gelu'(Z)); dW = D^T A; db = column sums of D; dw4 = dy^T A3.
//
// MFMA 16x16xK operand layout (lane l, i = l % 16, h = l / 16): the A
// operand holds row i, k = 8h .. 8h + 7 (K = 32; 4h .. 4h + 3 for K = 16)
// of a matrix stored [row][k]; the B operand holds column i, the same k, of
// a matrix stored [column][k]; the result holds rows 4h .. 4h + 3 of column
// i.  Every product is set up so that both operands are read as rows:
// forward Z^T = W A^T (weights [out][in], activations [row][in]); backward
// data F^T = W^T D^T (transposed weights [in][out], gradients [row][out]);
// weight gradients dW = D^T A over the 16 rows (v_mfma_f32_16x16x16_bf16;
// gradients and activations transposed, [neuron][row]).  16 rows a wave
// (64 waves for a 1024-row batch): the step is a chain of dependent layers
// in one wave per CU, so fewer rows per wave shorten it.
constexpr int kTrRows = 16;                 // rows per wave
constexpr int kTrS = 72;                    // row-major LDS stride (bf16)
constexpr int kTrT = 24;                    // transposed LDS stride (16 rows + 8)
constexpr int kPartStride = 13120;          // floats per partial row: parameters, loss, padding

static_assert(kMlpParams + 1 <= kPartStride, "partial row too short");

// A 32-bit mix of Wellons' "lowbias32" integer hash (xor-shift-multiply).
__device__ inline uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

// Sample order of an epoch: a 4-round Feistel network on 2 * half bits
// keyed by the epoch's keys, walked until it lands below n (cycle walking,
// Black & Rogaway 2002): a bijection of [0, n).  The walk ends, since the
// orbit of i < n under a bijection returns to i.  half = 0: the identity.
__device__ inline uint32_t epoch_order(uint32_t i, uint32_t n, int half, uint4 key)
{
    if (half == 0) return i;
    const uint32_t mask = (1u << half) - 1u;
    uint32_t v = i;
    do {
        uint32_t L = v >> half, R = v & mask;
        const uint32_t k[4] = {key.x, key.y, key.z, key.w};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t nr = L ^ (mix32(R ^ k[r]) & mask);
            L = R;
            R = nr;
        }
        v = (L << half) | R;
    } while (v >= n);
    return v;
}

__device__ inline f32x4 mfma16(const bf16x8& a, const bf16x8& b, const f32x4& c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16v8, a), __builtin_bit_cast(bf16v8, b), c,
                                                  0, 0, 0);
}

// acc[t] = W[16t .. 16t + 15][:] . In[0 .. 15][:]^T over K inputs: W global
// bf16 [64][K] (K = 72: a 16x16x16 step takes inputs 64 .. 71), In LDS bf16
// [16][kTrS].
template <int K>
__device__ inline void tr_product(const uint16_t* __restrict__ W, const uint16_t* In, f32x4 acc[4], int i, int h)
{
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        const bf16x8 b = *(const bf16x8*)(In + i * kTrS + 32 * kb + 8 * h);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const bf16x8 a = *(const bf16x8*)(W + (16 * t + i) * K + 32 * kb + 8 * h);
            acc[t] = mfma16(a, b, acc[t]);
        }
    }
    if (K == 72) {
        const bool in = h < 2;
        const bf16x4 zero = {0, 0, 0, 0};
        const bf16x4 b = in ? *(const bf16x4*)(In + i * kTrS + 64 + 4 * h) : zero;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const bf16x4 a = in ? *(const bf16x4*)(W + (16 * t + i) * K + 64 + 4 * h) : zero;
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
    }
}

// Weight gradient dW[n][k] = sum over the 16 rows of DT[n][r] AT[k][r] for
// k < kmax (KT = k tiles of 16), into the partial row at out (row-major,
// ld columns).
template <int KT>
__device__ inline void tr_wgrad(const uint16_t* DT, const uint16_t* AT, float* __restrict__ out, int ld, int kmax,
                                int i, int h)
{
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const bf16x4 a = *(const bf16x4*)(DT + (16 * t + i) * kTrT + 4 * h);
#pragma unroll
        for (int u = 0; u < KT; ++u) {
            const bf16x4 b = *(const bf16x4*)(AT + (16 * u + i) * kTrT + 4 * h);
            const f32x4 c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
            const int k = 16 * u + i;
            if (k < kmax) {
#pragma unroll
                for (int q = 0; q < 4; ++q) out[(16 * t + 4 * h + q) * ld + k] = c[q];
            }
        }
    }
}

// Stores bf16 values v of neurons n0 .. n0 + 3, row r: row-major at M and
// transposed at MT (either may be null).
__device__ inline void tr_store(uint16_t* M, uint16_t* MT, int r, int n0, const bf16x4& v)
{
    if (M) *(bf16x4*)(M + r * kTrS + n0) = v;
    if (MT) {
#pragma unroll
        for (int q = 0; q < 4; ++q) MT[(n0 + q) * kTrT + r] = (uint16_t)v[q];
    }
}

// Column sums over the 16 rows of a row-major bf16 tile, lane n < 64 sums
// column n: returns its sum (FP32, rows in order).
__device__ inline float tr_colsum(const uint16_t* M, int n)
{
    float acc = 0.0f;
    for (int r = 0; r < kTrRows; ++r) acc += bf2f(M[r * kTrS + n]);
    return acc;
}

__global__ void __launch_bounds__(64) train_tile(const uint16_t* __restrict__ X, const float* __restrict__ tgt,
                                                 uint32_t n, uint32_t b0, int half, uint4 key, int B,
                                                 const float* __restrict__ P, const uint16_t* __restrict__ bf,
                                                 const uint16_t* __restrict__ bfT, float* __restrict__ part)
{
    __shared__ __align__(16) uint16_t sX[kTrRows * kTrS];
    __shared__ __align__(16) uint16_t sXT[80 * kTrT];
    __shared__ __align__(16) uint16_t sA[3][kTrRows * kTrS];
    __shared__ __align__(16) uint16_t sAT[2][kMlpHid * kTrT];
    __shared__ __align__(16) uint16_t sD[kTrRows * kTrS];
    __shared__ __align__(16) uint16_t sDT[kMlpHid * kTrT];
    __shared__ float s_dy[kTrRows], s_loss[kTrRows];
    __shared__ uint32_t s_idx[kTrRows];
    const int lane = threadIdx.x, i = lane & 15, h = lane >> 4;
    float* out = part + (int64_t)blockIdx.x * kPartStride;
    const uint32_t row0 = b0 + (uint32_t)blockIdx.x * kTrRows;
    if (lane < kTrRows) s_idx[lane] = epoch_order(row0 + (uint32_t)lane, n, half, key);
    // inputs 72 .. 79 of the transposed copy are zero (the fifth k tile of dW1)
    for (int e = lane; e < 8 * kTrT; e += 64) sXT[72 * kTrT + e] = 0;
    __syncthreads();
    // gather: 16 rows of 72 bf16 as 9 16-byte pieces each, and the targets
    float tr = 0.0f;
    if (lane < kTrRows) tr = tgt[s_idx[lane]];
    for (int c = lane; c < kTrRows * 9; c += 64) {
        const int r = c / 9, p = c % 9;
        const bf16x8 v = *(const bf16x8*)(X + (int64_t)s_idx[r] * kMlpInPad + 8 * p);
        *(bf16x8*)(sX + r * kTrS + 8 * p) = v;
#pragma unroll
        for (int e = 0; e < 8; ++e) sXT[(8 * p + e) * kTrT + r] = (uint16_t)v[e];
    }
    tr = __shfl(tr, i);  // row i's target in each of its four lanes
    __syncthreads();
    // forward; Z of each layer stays in registers for the backward pass
    f32x4 Z[3][4], acc[4];
    const float* bias[3] = {P + kOffB1, P + kOffB2, P + kOffB3};
    const uint16_t* W[3] = {bf + kBfW1, bf + kBfW2, bf + kBfW3};
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        if (l == 0) tr_product<kMlpInPad>(W[0], sX, acc, i, h);
        else tr_product<kMlpHid>(W[l], sA[l - 1], acc, i, h);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int n0 = 16 * t + 4 * h;
            const float4 b = *(const float4*)(bias[l] + n0);
            f32x4 z = acc[t];
            z[0] += b.x;
            z[1] += b.y;
            z[2] += b.z;
            z[3] += b.w;
            Z[l][t] = z;
            tr_store(sA[l], l < 2 ? sAT[l] : nullptr, i, n0,
                     to_bf16x4(gelu_f(z[0]), gelu_f(z[1]), gelu_f(z[2]), gelu_f(z[3])));
        }
        __syncthreads();
    }
    // output neuron, loss, D3 = bf16(dy w4 gelu'(Z3)); row i in lanes i + 16 h
    float w4v[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) w4v[t][q] = bf2f(bf[kBfW4 + 16 * t + 4 * h + q]);
    {
        float y = 0.0f;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const bf16x4 a = *(const bf16x4*)(sA[2] + i * kTrS + 16 * t + 4 * h);
#pragma unroll
            for (int q = 0; q < 4; ++q) y += w4v[t][q] * bf2f((uint16_t)a[q]);
        }
        y += __shfl_xor(y, 16);
        y += __shfl_xor(y, 32);
        y += P[kOffB4];
        const float res = y - tr;
        const float dy = res / (float)B;
        if (h == 0) {
            s_dy[i] = dy;
            s_loss[i] = 0.5f * res * res / (float)B;
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const f32x4 z = Z[2][t];
            tr_store(sD, sDT, i, 16 * t + 4 * h,
                     to_bf16x4(dy * w4v[t][0] * gelu_grad_f(z[0]), dy * w4v[t][1] * gelu_grad_f(z[1]),
                               dy * w4v[t][2] * gelu_grad_f(z[2]), dy * w4v[t][3] * gelu_grad_f(z[3])));
        }
    }
    __syncthreads();
    // output layer gradients: dw4 = dy^T A3, db4 = sum dy, loss
    {
        float acc4 = 0.0f;
        for (int r = 0; r < kTrRows; ++r) acc4 += s_dy[r] * bf2f(sA[2][r * kTrS + lane]);
        out[kOffW4 + lane] = acc4;
        if (lane == 0) {
            float sdy = 0.0f, sl = 0.0f;
            for (int r = 0; r < kTrRows; ++r) {
                sdy += s_dy[r];
                sl += s_loss[r];
            }
            out[kOffB4] = sdy;
            out[kMlpParams] = sl;
        }
    }
    // layers 3, 2, 1: dW = D^T A_prev, db = colsum D, then D_prev =
    // bf16((D W) gelu'(Z_prev)) for the layer below
    const int offW[3] = {kOffW1, kOffW2, kOffW3}, offB[3] = {kOffB1, kOffB2, kOffB3};
#pragma unroll
    for (int l = 2; l >= 0; --l) {
        if (l == 0) tr_wgrad<5>(sDT, sXT, out + offW[0], kMlpInPad, kMlpInPad, i, h);
        else tr_wgrad<4>(sDT, sAT[l - 1], out + offW[l], kMlpHid, kMlpHid, i, h);
        out[offB[l] + lane] = tr_colsum(sD, lane);
        if (l == 0) break;
        tr_product<kMlpHid>(bfT + (l - 1) * kMlpHid * kMlpHid, sD, acc, i, h);
        __syncthreads();  // every lane has read D and DT
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const f32x4 z = Z[l - 1][t], f = acc[t];
            tr_store(sD, sDT, i, 16 * t + 4 * h,
                     to_bf16x4(f[0] * gelu_grad_f(z[0]), f[1] * gelu_grad_f(z[1]), f[2] * gelu_grad_f(z[2]),
                               f[3] * gelu_grad_f(z[3])));
        }
        __syncthreads();
    }
}

// bf16 copies of parameter i (FP32 value p): W1 (padding columns zero), W2
// and W3 row-major, in the walk's column order (mlp_col) and transposed (the
// backward pass's copies), w4.
__device__ inline void write_copies(int i, float p, uint16_t* __restrict__ bf, uint16_t* __restrict__ bfw,
                                    uint16_t* __restrict__ bfT)
{
    const uint16_t v = f2bf(p);
    if (i < kOffB1) {
        bf[kBfW1 + i] = (i % kMlpInPad) >= kMlpIn ? (uint16_t)0 : v;
    } else if (i >= kOffW2 && i < kOffB2) {
        const int j = i - kOffW2, o = j / kMlpHid, c = j % kMlpHid;
        bf[kBfW2 + j] = v;
        bfw[o * kMlpHid + mlp_col(c)] = v;
        bfT[c * kMlpHid + o] = v;
    } else if (i >= kOffW3 && i < kOffB3) {
        const int j = i - kOffW3, o = j / kMlpHid, c = j % kMlpHid;
        bf[kBfW3 + j] = v;
        bfw[kMlpHid * kMlpHid + o * kMlpHid + mlp_col(c)] = v;
        bfT[kMlpHid * kMlpHid + c * kMlpHid + o] = v;
    } else if (i >= kOffW4 && i < kOffB4) {
        bf[kBfW4 + i - kOffW4] = v;
    }
}

// Sums the G partial rows of each parameter in row order.  adam: the Adam
// step on the FP32 parameters and their bf16 copies, the loss added to
// loss[0] (and 1 to loss[1]) when record is set; otherwise the gradient goes
// to grad and the loss to loss[0].
// Eight lanes per parameter: lane c of a parameter's eight sums rows c, c +
// 8, c + 16, ... in that order, then lane 0 adds the eight sums in lane
// order (a fixed order, so the step does not depend on scheduling).  32
// parameters per block of 256.  With one lane per parameter the 52 blocks'
// loads of a 256-row batch had made this launch latency-bound.
constexpr int kApplyLanes = 8;
constexpr int kApplyParamsPerBlock = 256 / kApplyLanes;

__global__ void __launch_bounds__(256) train_apply(const float* __restrict__ part, int G, float* __restrict__ P,
                                                   float* __restrict__ mom, float* __restrict__ vel, float lr,
                                                   float b1c, float b2c, int adam, int record,
                                                   float* __restrict__ grad, float* __restrict__ loss,
                                                   uint16_t* __restrict__ bf, uint16_t* __restrict__ bfw,
                                                   uint16_t* __restrict__ bfT)
{
    __shared__ float s_sum[kApplyLanes][kApplyParamsPerBlock];
    const int pi = threadIdx.x % kApplyParamsPerBlock, c = threadIdx.x / kApplyParamsPerBlock;
    const int i = blockIdx.x * kApplyParamsPerBlock + pi;
    // the chunk's loads issued 8 at a time, then added in row order
    float g = 0.0f;
    if (i <= kMlpParams) {
        for (int w0 = c; w0 < G; w0 += 8 * kApplyLanes) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int w = w0 + u * kApplyLanes;
                v[u] = w < G ? part[(int64_t)w * kPartStride + i] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (w0 + u * kApplyLanes < G) g += v[u];
        }
    }
    s_sum[c][pi] = g;
    __syncthreads();
    if (c != 0 || i > kMlpParams) return;
    g = s_sum[0][pi];
#pragma unroll
    for (int k = 1; k < kApplyLanes; ++k) g += s_sum[k][pi];
    if (i == kMlpParams) {
        if (!adam) {
            loss[0] = g;
        } else if (record) {
            loss[0] += g;
            loss[1] += 1.0f;
        }
        return;
    }
    if (!adam) {
        grad[i] = g;
        return;
    }
    const float m = 0.9f * mom[i] + 0.1f * g;
    const float v = 0.999f * vel[i] + 0.001f * g * g;
    mom[i] = m;
    vel[i] = v;
    const float p = P[i] - lr * (m / b1c) / (sqrtf(v / b2c) + 1e-8f);
    P[i] = p;
    write_copies(i, p, bf, bfw, bfT);
}

__global__ void copies_kernel(const float* __restrict__ P, uint16_t* __restrict__ bf, uint16_t* __restrict__ bfw,
                              uint16_t* __restrict__ bfT)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kMlpParams) write_copies(i, P[i], bf, bfw, bfT);
}

__global__ void widen_bf16(const uint16_t* __restrict__ X, float* __restrict__ F, int64_t total)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) F[i] = bf2f(X[i]);
}

__global__ void narrow_bf16(const float* __restrict__ F, uint16_t* __restrict__ X, int64_t total)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) X[i] = f2bf(F[i]);
}

// Wave-level MFMA forward over rows of X (the walk kernel's inference path):
// one wave per 64 rows, each 72-wide row repacked into the walk kernel's tile
// row (columns 0..63, which features hold as 0 or 1, as packed bits; columns
// 64..67 as they are).
__global__ void __launch_bounds__(64) forward_waves(MlpNet m, const uint16_t* __restrict__ X, int64_t n,
                                                    float* __restrict__ y)
{
    __shared__ __align__(16) uint16_t xs[kMlpTileElems];
    __shared__ __align__(16) float small[kMlpSmallFloats];
    __shared__ __align__(16) uint16_t w[kMlpWLdsElems];
    const int lane = threadIdx.x;
    const int64_t row = (int64_t)blockIdx.x * 64 + lane;
    mlp_stage_small(m, small, lane, 64);
    mlp_stage_weights(m, w, lane, 64);
    __syncthreads();
    uint16_t* r = xs + lane * kMlpStride;
    uint64_t bits = 0;
    for (int k = 0; k < 64; ++k)
        if (row < n && (X[row * kMlpInPad + k] & 0x7FFFu) != 0u) bits |= 1ull << k;
    for (int k = 64; k < kMlpBitsCol; ++k) r[k] = row < n ? X[row * kMlpInPad + k] : (uint16_t)0;
    *(uint64_t*)(r + kMlpBitsCol) = bits;
    const float out = mlp_forward_rows(w, xs, small, lane, 64, lane);
    if (row < n) y[row] = out;
}

inline int blocks(int64_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

MlpPyramid mlp_pyramid(const frw_mlp* m)
{
    MlpPyramid v;
    v.occ = m->d_occ;
    v.sat = m->d_sat;
    v.occ_off = m->d_occ_off;
    v.sat_off = m->d_sat_off;
    v.bounds = m->tree->bounds;
    v.K0 = m->K0;
    v.Lmax = m->Lmax;
    v.Vmax = m->Vmax;
    v.pad = 0;
    return v;
}

MlpNet mlp_net(const frw_mlp* m, float alpha)
{
    MlpNet v;
    v.W1 = m->d_bf + kBfW1;
    v.W2 = m->d_bfw;
    v.W3 = m->d_bfw + kMlpHid * kMlpHid;
    v.w4 = m->d_bf + kBfW4;
    v.b1 = m->d_params + kOffB1;
    v.b2 = m->d_params + kOffB2;
    v.b3 = m->d_params + kOffB3;
    v.b4 = m->d_params + kOffB4;
    v.alpha = alpha;
    return v;
}

const frw_tree* mlp_tree(const frw_mlp* m) { return m->tree; }

namespace {

// Partial-sum rows for batches of up to B rows (B / 32 rows of kPartStride
// floats).
int ensure_part(frw_mlp* m, int64_t B)
{
    const int64_t G = B / kTrRows;
    if (m->part_rows >= G) return FRW_OK;
    (void)hipFree(m->d_part);
    m->d_part = nullptr;
    m->part_rows = 0;
    int rc = check_hip(hipMalloc(&m->d_part, (size_t)G * kPartStride * 4), "hipMalloc partial sums");
    if (rc == FRW_OK) m->part_rows = G;
    return rc;
}

}  // namespace
}  // namespace frw

using namespace frw;

extern "C" {

int frw_mlp_create(const frw_tree* t, uint64_t seed, int32_t max_level, frw_mlp** out)
{
    if (!t || !out) return set_error(FRW_E_INVALID, "null argument");
    if (max_level < 0 || max_level > kPyrMaxLevel) return set_error(FRW_E_RANGE, "max_level must be in 0..15");
    const int cap = max_level ? max_level : kPyrMaxLevel;
    frw_mlp* m = new frw_mlp;
    m->tree = t;
    m->device = t->device;
    const int64_t W = (int64_t)t->bounds.z - t->bounds.x, H = (int64_t)t->bounds.w - t->bounds.y;
    while (((int64_t)1 << m->K0) <= std::max(W, H)) ++m->K0;
    m->Lmax = std::min({m->K0, kMlpMaxLevel, cap});
    // verification levels: finest cell about 1/128 of the mean rect spacing
    // sqrt(W H / n), at most level `cap` (15 unless the caller limits memory:
    // a level-l table takes (2^l + 1)^2 x 4 bytes)
    {
        const double spacing = std::sqrt((double)std::max<int64_t>(W, 1) * (double)std::max<int64_t>(H, 1) /
                                         (double)std::max<int64_t>(t->n_rects, 1));
        const double target = std::max(1.0, spacing / 128.0);
        const int v = m->K0 - (int)std::floor(std::log2(target));
        m->Vmax = std::max(0, std::min({v, m->K0, cap}));
    }
    m->Pmax = std::max(m->Lmax, m->Vmax);
    int64_t words = 0, entries = 0;
    for (int l = 0; l <= m->Pmax; ++l) {
        m->occ_off.push_back(words);
        words += occ_tiles(l) * occ_tiles(l);
    }
    for (int l = 0; l <= m->Vmax; ++l) {
        m->sat_off.push_back(entries);
        entries += (((int64_t)1 << l) + 1) * (((int64_t)1 << l) + 1);
    }
    int rc = check_hip(hipMalloc(&m->d_occ, words * 8), "hipMalloc occupancy");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_sat, entries * 4), "hipMalloc summed-area tables");
    if (rc == FRW_OK) rc = check_hip(hipMemset(m->d_occ, 0, words * 8), "memset occupancy");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_occ_off, sizeof(int64_t) * m->occ_off.size()), "hipMalloc off");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_sat_off, sizeof(int64_t) * m->sat_off.size()), "hipMalloc off");
    if (rc == FRW_OK)
        rc = check_hip(hipMemcpy(m->d_occ_off, m->occ_off.data(), sizeof(int64_t) * m->occ_off.size(),
                                 hipMemcpyHostToDevice), "copy off");
    if (rc == FRW_OK)
        rc = check_hip(hipMemcpy(m->d_sat_off, m->sat_off.data(), sizeof(int64_t) * m->sat_off.size(),
                                 hipMemcpyHostToDevice), "copy off");
    if (rc == FRW_OK) {
        occ_mark<<<blocks(t->n_rects, 256), 256>>>(t->d_rects, (int)t->n_rects, t->bounds, m->K0, m->Pmax,
                                                   m->d_occ_off, m->d_occ);
        rc = check_hip(hipGetLastError(), "occ_mark");
    }
    for (int l = m->Pmax - 1; l >= 0 && rc == FRW_OK; --l) {
        const int64_t tiles = occ_tiles(l) * occ_tiles(l);
        occ_reduce<<<blocks(tiles, 256), 256>>>(m->d_occ + m->occ_off[l + 1], l, m->d_occ + m->occ_off[l]);
        rc = check_hip(hipGetLastError(), "occ_reduce");
    }
    for (int l = 0; l <= m->Vmax && rc == FRW_OK; ++l) {
        const int64_t n = (int64_t)1 << l;
        uint32_t* S = m->d_sat + m->sat_off[l];
        rc = check_hip(hipMemset(S, 0, (size_t)(n + 1) * 4), "memset table row 0");
        if (rc != FRW_OK) break;
        sat_rows<<<(unsigned)n, 256>>>(m->d_occ + m->occ_off[l], l, S);
        sat_cols<<<blocks(n + 1, 256), 256>>>(n, S);
        rc = check_hip(hipGetLastError(), "summed-area tables");
    }
    if (rc == FRW_OK) rc = check_hip(hipDeviceSynchronize(), "occupancy pyramid");
    // Xavier-uniform initial weights from Philox, zero biases
    std::vector<float> P(kMlpParams, 0.0f);
    auto fill = [&](int off, int rows, int cols, int fan_in, int fan_out, uint32_t tag) {
        const float a = std::sqrt(6.0f / (float)(fan_in + fan_out));
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c) {
                uint32_t w[4] = {(uint32_t)(r * cols + c), tag, 0u, 0x4D4C5030u};
                philox4x32_10(w, (uint32_t)seed, (uint32_t)(seed >> 32));
                const float u = (float)(w[0] >> 8) / 16777216.0f;
                P[off + r * cols + c] = (c < (tag == 1 ? kMlpIn : cols)) ? a * (2.0f * u - 1.0f) : 0.0f;
            }
    };
    fill(kOffW1, kMlpHid, kMlpInPad, kMlpIn, kMlpHid, 1);
    fill(kOffW2, kMlpHid, kMlpHid, kMlpHid, kMlpHid, 2);
    fill(kOffW3, kMlpHid, kMlpHid, kMlpHid, kMlpHid, 3);
    fill(kOffW4, 1, kMlpHid, kMlpHid, 1, 4);
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_params, kMlpParams * 4), "hipMalloc params");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_grad, kMlpParams * 4), "hipMalloc grad");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_bfT, 2 * kMlpHid * kMlpHid * 2), "hipMalloc bf16 transposed copy");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_adam_m, kMlpParams * 4), "hipMalloc adam");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_adam_v, kMlpParams * 4), "hipMalloc adam");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_bf, kBfCount * 2), "hipMalloc bf16");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_bfw, 2 * kMlpHid * kMlpHid * 2), "hipMalloc bf16 walk copy");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_loss, 2 * 4), "hipMalloc loss");
    if (rc == FRW_OK) rc = check_hip(hipMemset(m->d_adam_m, 0, kMlpParams * 4), "memset");
    if (rc == FRW_OK) rc = check_hip(hipMemset(m->d_adam_v, 0, kMlpParams * 4), "memset");
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(m->d_params, P.data(), kMlpParams * 4, hipMemcpyHostToDevice), "copy params");
    if (rc == FRW_OK) rc = check_hip(hipMemset(m->d_bf, 0, kBfCount * 2), "memset bf16");
    if (rc == FRW_OK) {
        copies_kernel<<<blocks(kMlpParams, 256), 256>>>(m->d_params, m->d_bf, m->d_bfw, m->d_bfT);
        rc = check_hip(hipDeviceSynchronize(), "bf16 copies");
    }
    if (rc != FRW_OK) {
        frw_mlp_free(m);
        return rc;
    }
    *out = m;
    return FRW_OK;
}

void frw_mlp_free(frw_mlp* m)
{
    if (!m) return;
    for (void* p : {(void*)m->d_occ, (void*)m->d_sat, (void*)m->d_occ_off, (void*)m->d_sat_off, (void*)m->d_params,
                    (void*)m->d_bf, (void*)m->d_bfw, (void*)m->d_bfT, (void*)m->d_adam_m, (void*)m->d_adam_v,
                    (void*)m->d_grad, (void*)m->d_part, (void*)m->d_X, (void*)m->d_t, (void*)m->d_loss})
        if (p) (void)hipFree(p);
    delete m;
}

int frw_mlp_param_count(void) { return kMlpParams; }

int frw_mlp_get_params(const frw_mlp* m, float* out, int32_t n)
{
    if (!m || !out || n != kMlpParams) return set_error(FRW_E_INVALID, "bad parameter buffer");
    return check_hip(hipMemcpy(out, m->d_params, kMlpParams * 4, hipMemcpyDeviceToHost), "copy params");
}

int frw_mlp_set_params(frw_mlp* m, const float* in, int32_t n)
{
    if (!m || !in || n != kMlpParams) return set_error(FRW_E_INVALID, "bad parameter buffer");
    int rc = check_hip(hipMemcpy(m->d_params, in, kMlpParams * 4, hipMemcpyHostToDevice), "copy params");
    if (rc != FRW_OK) return rc;
    copies_kernel<<<blocks(kMlpParams, 256), 256>>>(m->d_params, m->d_bf, m->d_bfw, m->d_bfT);
    return check_hip(hipDeviceSynchronize(), "bf16 copies");
}

int frw_mlp_occupancy(const frw_mlp* m, int32_t level, uint32_t* out_words, int64_t n_words)
{
    if (!m || !out_words || level < 0 || level > m->Pmax) return set_error(FRW_E_INVALID, "bad level");
    const int64_t n = (int64_t)1 << level, T = occ_tiles(level);
    const int64_t need = (n * n + 31) / 32;
    if (n_words < need) return set_error(FRW_E_INVALID, "output too small");
    std::vector<uint64_t> tiles((size_t)(T * T));
    const int rc = check_hip(hipMemcpy(tiles.data(), m->d_occ + m->occ_off[level], tiles.size() * 8,
                                       hipMemcpyDeviceToHost), "copy occupancy");
    if (rc != FRW_OK) return rc;
    // the tiles as the row-major bitmap (bit j n + i = cell (i, j))
    std::memset(out_words, 0, (size_t)need * 4);
    if (n >= 8) {
        // a tile's row is one byte of the row-major bitmap (n is a multiple of 8)
        uint8_t* bytes = reinterpret_cast<uint8_t*>(out_words);
        for (int64_t ty = 0; ty < T; ++ty)
            for (int64_t tx = 0; tx < T; ++tx)
                for (int y = 0; y < 8; ++y)
                    bytes[((8 * ty + y) * n + 8 * tx) / 8] = (uint8_t)(tiles[(size_t)(ty * T + tx)] >> (8 * y));
    } else {
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < n; ++i)
                if ((tiles[0] >> (8 * j + i)) & 1u) out_words[(j * n + i) >> 5] |= 1u << ((j * n + i) & 31);
    }
    return FRW_OK;
}

int frw_mlp_levels(const frw_mlp* m, int32_t out[4])
{
    if (!m || !out) return set_error(FRW_E_INVALID, "null argument");
    out[0] = m->K0;
    out[1] = m->Lmax;
    out[2] = m->Vmax;
    out[3] = m->Pmax;
    return FRW_OK;
}

int frw_mlp_certify(const frw_mlp* m, const double* pts, const double* d, int64_t n, uint8_t* out, void* stream)
{
    if (!m || n < 0 || (n > 0 && (!pts || !d || !out))) return set_error(FRW_E_INVALID, "bad arguments");
    if (n == 0) return FRW_OK;
    hipStream_t s = (hipStream_t)stream;
    certify_kernel<<<blocks(n, 256), 256, 0, s>>>(mlp_pyramid(m), pts, d, n, out);
    return check_hip(hipStreamSynchronize(s), "certify");
}

int frw_mlp_features(const frw_mlp* m, const double* pts, const double* h, int64_t n, float* feats,
                     void* stream)
{
    if (!m || n < 0 || (n > 0 && (!pts || !h || !feats))) return set_error(FRW_E_INVALID, "bad arguments");
    if (n == 0) return FRW_OK;
    hipStream_t s = (hipStream_t)stream;
    uint16_t* X = nullptr;
    int rc = check_hip(hipMalloc(&X, (size_t)n * kMlpInPad * 2), "hipMalloc features");
    if (rc != FRW_OK) return rc;
    features_kernel<<<blocks(n, 256), 256, 0, s>>>(mlp_pyramid(m), pts, h, n, X, nullptr, nullptr, nullptr);
    widen_bf16<<<blocks(n * kMlpInPad, 256), 256, 0, s>>>(X, feats, n * kMlpInPad);
    rc = check_hip(hipStreamSynchronize(s), "features");
    (void)hipFree(X);
    return rc;
}

int frw_mlp_forward(const frw_mlp* m, const float* X, int64_t n, float* y, void* stream)
{
    if (!m || n < 0 || (n > 0 && (!X || !y))) return set_error(FRW_E_INVALID, "bad arguments");
    if (n == 0) return FRW_OK;
    hipStream_t s = (hipStream_t)stream;
    uint16_t* Xb = nullptr;
    int rc = check_hip(hipMalloc(&Xb, (size_t)n * kMlpInPad * 2), "hipMalloc forward");
    if (rc != FRW_OK) return rc;
    narrow_bf16<<<blocks(n * kMlpInPad, 256), 256, 0, s>>>(X, Xb, n * kMlpInPad);
    forward_waves<<<blocks(n, 64), 64, 0, s>>>(mlp_net(m, 1.0f), Xb, n, y);
    rc = check_hip(hipStreamSynchronize(s), "mlp forward");
    (void)hipFree(Xb);
    return rc;
}

int frw_mlp_gradient(frw_mlp* m, const float* X, const float* t, int64_t n, float* grad, float* loss)
{
    if (!m || !X || !t || !grad || n <= 0 || n % 32 || n > ((int64_t)1 << 31))
        return set_error(FRW_E_INVALID, "n must be a positive multiple of 32, at most 2^31");
    int rc = ensure_part(m, n);
    if (rc != FRW_OK) return rc;
    uint16_t* Xb = nullptr;
    rc = check_hip(hipMalloc(&Xb, (size_t)n * kMlpInPad * 2), "hipMalloc gradient");
    if (rc != FRW_OK) return rc;
    narrow_bf16<<<blocks(n * kMlpInPad, 256), 256>>>(X, Xb, n * kMlpInPad);
    train_tile<<<(unsigned)(n / kTrRows), 64>>>(Xb, t, (uint32_t)n, 0u, 0, make_uint4(0, 0, 0, 0), (int)n,
                                                m->d_params, m->d_bf, m->d_bfT, m->d_part);
    train_apply<<<blocks(kMlpParams + 1, kApplyParamsPerBlock), 256>>>(
        m->d_part, (int)(n / kTrRows), m->d_params, m->d_adam_m, m->d_adam_v, 0.0f, 1.0f, 1.0f, 0, 0, m->d_grad,
        m->d_loss, m->d_bf, m->d_bfw, m->d_bfT);
    rc = check_hip(hipDeviceSynchronize(), "mlp gradient");
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(grad, m->d_grad, kMlpParams * 4, hipMemcpyDeviceToHost), "copy grad");
    if (rc == FRW_OK && loss) rc = check_hip(hipMemcpy(loss, m->d_loss, 4, hipMemcpyDeviceToHost), "copy loss");
    (void)hipFree(Xb);
    return rc;
}

int frw_mlp_add_samples(frw_mlp* m, const frw_mlp* features_of, const double* pts, const double* h,
                        const double* dstar, int64_t n)
{
    if (!m || !pts || !h || !dstar || n <= 0) return set_error(FRW_E_INVALID, "bad arguments");
    const frw_mlp* src = features_of ? features_of : m;
    const int64_t old_n = m->n_samples, tot = old_n + n;
    int rc = FRW_OK;
    if (tot > m->cap_samples) {
        // at least double the capacity, so that appending in many calls
        // copies the rows already held a bounded number of times
        const int64_t cap = std::max(tot, 2 * m->cap_samples);
        uint16_t* X = nullptr;
        float* T = nullptr;
        rc = check_hip(hipMalloc(&X, (size_t)cap * kMlpInPad * 2), "hipMalloc samples");
        if (rc == FRW_OK) rc = check_hip(hipMalloc(&T, (size_t)cap * 4), "hipMalloc targets");
        if (rc == FRW_OK && old_n) {
            rc = check_hip(hipMemcpy(X, m->d_X, (size_t)old_n * kMlpInPad * 2, hipMemcpyDeviceToDevice), "copy samples");
            if (rc == FRW_OK) rc = check_hip(hipMemcpy(T, m->d_t, (size_t)old_n * 4, hipMemcpyDeviceToDevice), "copy targets");
        }
        if (rc != FRW_OK) {
            if (X) (void)hipFree(X);
            if (T) (void)hipFree(T);
            return rc;
        }
        (void)hipFree(m->d_X);
        (void)hipFree(m->d_t);
        m->d_X = X;
        m->d_t = T;
        m->cap_samples = cap;
    }
    features_kernel<<<blocks(n, 256), 256>>>(mlp_pyramid(src), pts, h, n, m->d_X + old_n * kMlpInPad, dstar,
                                             m->d_t + old_n, nullptr);
    rc = check_hip(hipDeviceSynchronize(), "features");
    if (rc != FRW_OK) return rc;
    m->n_samples = tot;
    return FRW_OK;
}

int frw_mlp_set_samples(frw_mlp* m, const double* pts, const double* h, const double* dstar, int64_t n)
{
    if (!m || !pts || !h || !dstar || n <= 0) return set_error(FRW_E_INVALID, "bad arguments");
    m->n_samples = 0;  // the rows are rewritten in place (the capacity is kept)
    return frw_mlp_add_samples(m, nullptr, pts, h, dstar, n);
}

int frw_mlp_train(frw_mlp* m, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* last_loss)
{
    if (!m || epochs <= 0 || batch <= 0 || batch % 32 || lr <= 0) return set_error(FRW_E_INVALID, "bad training arguments");
    if (m->n_samples < batch) return set_error(FRW_E_RANGE, "fewer samples than one batch");
    if (m->n_samples > ((int64_t)1 << 31)) return set_error(FRW_E_RANGE, "more than 2^31 samples");
    int rc = ensure_part(m, batch);
    if (rc != FRW_OK) return rc;
    const int64_t n = m->n_samples;
    // the epoch order's Feistel network works on 2 * half >= log2(n) bits
    int k = 0;
    while (((int64_t)1 << k) < n) ++k;
    const int half = std::max(1, (k + 1) / 2);
    const int G = batch / kTrRows;
    rc = check_hip(hipMemsetAsync(m->d_loss, 0, 2 * 4, nullptr), "memset loss");
    for (int ep = 0; ep < epochs && rc == FRW_OK; ++ep) {
        uint32_t w[4] = {(uint32_t)ep, 0u, 0u, 0x5348464Cu};  // 'SHFL'
        philox4x32_10(w, (uint32_t)seed, (uint32_t)(seed >> 32));
        const uint4 key = make_uint4(w[0], w[1], w[2], w[3]);
        for (int64_t b0 = 0; b0 + batch <= n; b0 += batch) {
            train_tile<<<G, 64>>>(m->d_X, m->d_t, (uint32_t)n, (uint32_t)b0, half, key, batch, m->d_params,
                                  m->d_bf, m->d_bfT, m->d_part);
            m->adam_step++;
            const float b1c = 1.0f - std::pow(0.9f, (float)m->adam_step);
            const float b2c = 1.0f - std::pow(0.999f, (float)m->adam_step);
            train_apply<<<blocks(kMlpParams + 1, kApplyParamsPerBlock), 256>>>(
                m->d_part, G, m->d_params, m->d_adam_m, m->d_adam_v, lr, b1c, b2c, 1, ep == epochs - 1 ? 1 : 0,
                nullptr, m->d_loss, m->d_bf, m->d_bfw, m->d_bfT);
        }
        rc = check_hip(hipGetLastError(), "training step");
    }
    float l[2] = {0.0f, 0.0f};
    if (rc == FRW_OK) rc = check_hip(hipDeviceSynchronize(), "train");
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(l, m->d_loss, 2 * 4, hipMemcpyDeviceToHost), "read loss");
    if (rc == FRW_OK && last_loss) *last_loss = l[1] > 0.0f ? l[0] / l[1] : 0.0f;
    return rc;
}

}  // extern "C"

extern "C" {

int frw_walk_harvest(const frw_tree* t, const frw_walk_params* p, double* pts, double* h, double* dstar,
                     int64_t capacity, int64_t* n_out)
{
    if (!t || !p || !pts || !h || !dstar || capacity <= 0) return set_error(FRW_E_INVALID, "bad arguments");
    frw_walk_params q = *p;
    q.mode = FRW_MODE_EXACT;
    int rc = check_walk_params(t, &q);
    if (rc != FRW_OK) return rc;
    // Two passes over the same walks: the first gives each walk's hop count
    // and end, hence its number of records (every hop that did not end the
    // walk: hops - 1 when it ended on a rect, hops otherwise); the second
    // writes walk w's records at the exclusive prefix sum of those counts.
    // The records thus come in walk and hop order whatever the scheduling,
    // and a capacity below the total keeps the first walks' records.
    const int64_t nb = t->n_rects + 2, nw = q.n_walks;
    // one allocation: walk offsets (8 B each), bins and the hop total (8 B
    // each), hits and hop counts (4 B each)
    char* scratch = nullptr;
    rc = check_hip(hipMalloc(&scratch, (size_t)(nw * 8 + (nb + 1) * 8 + nw * 8)), "hipMalloc harvest scratch");
    unsigned long long* d_off = nullptr;
    int64_t *d_counts = nullptr, *d_hops = nullptr;
    int32_t *d_hit = nullptr, *d_hp = nullptr;
    if (rc == FRW_OK) {
        d_off = reinterpret_cast<unsigned long long*>(scratch);
        d_counts = reinterpret_cast<int64_t*>(scratch + nw * 8);
        d_hops = d_counts + nb;
        d_hit = reinterpret_cast<int32_t*>(d_hops + 1);
        d_hp = d_hit + nw;
        rc = check_hip(hipMemset(d_counts, 0, (nb + 1) * 8), "memset");
    }
    if (rc == FRW_OK) {
        const frw_walk_outputs o{d_counts, d_hops, d_hit, d_hp};
        rc = launch_walk(t, &q, &o, nullptr);
    }
    if (rc == FRW_OK) rc = check_hip(hipDeviceSynchronize(), "harvest pass 1");
    std::vector<int32_t> hit, hp;
    std::vector<unsigned long long> off;
    int64_t used = 0;  // walks of the second pass
    uint64_t total = 0;
    if (rc == FRW_OK) {
        hit.resize((size_t)nw);
        hp.resize((size_t)nw);
        off.resize((size_t)nw);
        rc = check_hip(hipMemcpy(hit.data(), d_hit, nw * 4, hipMemcpyDeviceToHost), "copy hits");
        if (rc == FRW_OK) rc = check_hip(hipMemcpy(hp.data(), d_hp, nw * 4, hipMemcpyDeviceToHost), "copy hops");
    }
    if (rc == FRW_OK) {
        for (int64_t w = 0; w < nw && total < (uint64_t)capacity; ++w) {
            off[(size_t)w] = total;
            total += (uint64_t)hp[(size_t)w] - (hit[(size_t)w] >= 0 ? 1u : 0u);
            used = w + 1;
        }
        rc = check_hip(hipMemcpy(d_off, off.data(), (size_t)used * 8, hipMemcpyHostToDevice), "copy offsets");
    }
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_counts, 0, nb * 8), "memset");
    if (rc == FRW_OK && used > 0) {
        frw_walk_params q2 = q;
        q2.n_walks = used;
        const frw_walk_outputs o{d_counts, d_hops, nullptr, nullptr};
        WalkExtras ex;
        ex.h_pts = pts;
        ex.h_prev = h;
        ex.h_dstar = dstar;
        ex.h_off = d_off;
        ex.h_cap = (uint64_t)capacity;
        rc = launch_walk(t, &q2, &o, nullptr, nullptr, &ex);
    }
    if (rc == FRW_OK) rc = check_hip(hipDeviceSynchronize(), "harvest pass 2");
    if (rc == FRW_OK && n_out) *n_out = (int64_t)std::min<uint64_t>(total, (uint64_t)capacity);
    if (scratch) (void)hipFree(scratch);
    return rc;
}

namespace {

int check_learned(const frw_tree* t, const frw_mlp* m, const frw_walk_params* p, float alpha)
{
    if (!t || !m || !p) return set_error(FRW_E_INVALID, "bad arguments");
    if (m->tree != t) return set_error(FRW_E_INVALID, "network built for another tree");
    if (p->master < 0 || p->master >= t->n_rects || p->n_walks <= 0 || p->first_walk < 0 ||
        p->first_walk + p->n_walks > kWalkIdLimit)
        return set_error(FRW_E_RANGE, "bad master or walk range");
    if (!(alpha > 0.0f)) return set_error(FRW_E_INVALID, "alpha must be positive");
    if (p->max_hops < 0 || p->max_hops > 0xFFFFFFFFll) return set_error(FRW_E_RANGE, "bad max_hops");
    return FRW_OK;
}

int learned_sync(const frw_tree* t, const frw_mlp* m, const frw_walk_params* p, const frw_walk_outputs* o,
                 float alpha, unsigned long long* d_traffic, int64_t stats[2])
{
    frw_walk_params q = *p;
    q.mode = FRW_MODE_LEARNED;
    unsigned long long* d_st = nullptr;
    int rc = check_hip(hipMalloc(&d_st, 16), "hipMalloc");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_st, 0, 16), "memset");
    const MlpNet net = mlp_net(m, alpha);
    const MlpPyramid pyr = mlp_pyramid(m);
    if (rc == FRW_OK) {
        WalkExtras ex;
        ex.net = &net;
        ex.pyr = &pyr;
        ex.l_stats = d_st;
        rc = launch_walk(t, &q, o, nullptr, d_traffic, &ex);
    }
    if (rc == FRW_OK) rc = check_hip(hipDeviceSynchronize(), "learned walk");
    unsigned long long st[2] = {0, 0};
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(st, d_st, 16, hipMemcpyDeviceToHost), "copy stats");
    if (rc == FRW_OK && stats) {
        stats[0] = (int64_t)st[0];
        stats[1] = (int64_t)st[1];
    }
    if (d_st) (void)hipFree(d_st);
    int err = 0;
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(&err, t->d_err, sizeof(int), hipMemcpyDeviceToHost), "err flag");
    if (rc == FRW_OK && err) rc = set_error(FRW_E_KERNEL, "kernel error flag " + std::to_string(err));
    return rc;
}

}  // namespace

int frw_walk_learned_async(const frw_tree* t, const frw_mlp* m, const frw_walk_params* p,
                           const frw_walk_outputs* o, float alpha, uint64_t* stats_dev, void* stream)
{
    int rc = check_learned(t, m, p, alpha);
    if (rc != FRW_OK) return rc;
    if (!o || !o->counts || !o->hops) return set_error(FRW_E_INVALID, "null outputs");
    frw_walk_params q = *p;
    q.mode = FRW_MODE_LEARNED;
    const MlpNet net = mlp_net(m, alpha);
    const MlpPyramid pyr = mlp_pyramid(m);
    WalkExtras ex;
    ex.net = &net;
    ex.pyr = &pyr;
    ex.l_stats = (unsigned long long*)stats_dev;
    return launch_walk(t, &q, o, (hipStream_t)stream, nullptr, &ex);
}

int frw_walk_learned(const frw_tree* t, const frw_mlp* m, const frw_walk_params* p, const frw_walk_outputs* o,
                     float alpha, int64_t stats[2])
{
    int rc = check_learned(t, m, p, alpha);
    if (rc != FRW_OK) return rc;
    if (!o || !o->counts || !o->hops) return set_error(FRW_E_INVALID, "null outputs");
    return learned_sync(t, m, p, o, alpha, nullptr, stats);
}

int frw_walk_learned_traffic(const frw_tree* t, const frw_mlp* m, const frw_walk_params* p, float alpha,
                             int64_t out[6])
{
    int rc = check_learned(t, m, p, alpha);
    if (rc != FRW_OK) return rc;
    if (!out) return set_error(FRW_E_INVALID, "null output");
    const int64_t nb = t->n_rects + 2;
    int64_t *d_counts = nullptr, *d_hops = nullptr, *d_tr = nullptr;
    rc = check_hip(hipMalloc(&d_counts, sizeof(int64_t) * nb), "hipMalloc counts");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&d_hops, sizeof(int64_t)), "hipMalloc hops");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&d_tr, 3 * sizeof(int64_t)), "hipMalloc traffic");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_counts, 0, sizeof(int64_t) * nb), "memset");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_hops, 0, sizeof(int64_t)), "memset");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_tr, 0, 3 * sizeof(int64_t)), "memset");
    if (rc == FRW_OK) {
        frw_walk_outputs o{d_counts, d_hops, nullptr, nullptr};
        rc = learned_sync(t, m, p, &o, alpha, (unsigned long long*)d_tr, out + 3);
    }
    int64_t tr[3] = {0, 0, 0};
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(tr, d_tr, 3 * sizeof(int64_t), hipMemcpyDeviceToHost), "copy");
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(out + 2, d_hops, sizeof(int64_t), hipMemcpyDeviceToHost), "copy");
    if (rc == FRW_OK) {
        out[0] = tr[0];
        out[1] = tr[1];
        out[5] = tr[2];
    }
    if (d_counts) (void)hipFree(d_counts);
    if (d_hops) (void)hipFree(d_hops);
    if (d_tr) (void)hipFree(d_tr);
    return rc;
}

}  // extern "C"
