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
    uint32_t* d_occ = nullptr;
    uint32_t* d_sat = nullptr;
    int64_t* d_occ_off = nullptr;  // device copies of the offsets
    int64_t* d_sat_off = nullptr;
    float* d_params = nullptr;     // FP32 master parameters
    uint16_t* d_bf = nullptr;      // bf16 copy of W1 | W2 | W3 | w4 (training)
    uint16_t* d_bfw = nullptr;     // bf16 W2 | W3 with input columns in mlp_col order (the walk's network)
    float* d_adam_m = nullptr;
    float* d_adam_v = nullptr;
    float* d_grad = nullptr;
    float* d_gs = nullptr;         // weight-gradient split partials, kMaxSplits x 64 x 72
    int64_t adam_step = 0;
    // training set
    int64_t n_samples = 0;
    uint16_t* d_X = nullptr;       // [n][72] bf16
    float* d_t = nullptr;          // [n] targets
    // training scratch (batch-sized)
    int64_t ws_batch = 0;
    float *d_Z1 = nullptr, *d_Z2 = nullptr, *d_Z3 = nullptr, *d_F = nullptr;
    float *d_dy = nullptr, *d_lossv = nullptr;
    uint16_t *d_A1 = nullptr, *d_A2 = nullptr, *d_A3 = nullptr, *d_D = nullptr, *d_Xb = nullptr;
    float* d_tb = nullptr;
    float* d_loss = nullptr;
    int32_t* d_idx = nullptr;
};

namespace frw {
namespace {

constexpr int kBfW1 = 0;
constexpr int kBfW2 = kBfW1 + kMlpHid * kMlpInPad;
constexpr int kBfW3 = kBfW2 + kMlpHid * kMlpHid;
constexpr int kBfW4 = kBfW3 + kMlpHid * kMlpHid;
constexpr int kBfCount = kBfW4 + kMlpHid;
constexpr int kMaxSplits = 16;  // batch slices of the weight-gradient products

// ---------------------------------------------------------------- occupancy
// One lane per rect: sets the bits of every cell the closed rect meets, one
// word mask per bitmap row segment.
__global__ void occ_mark(const int4* __restrict__ rects, int n, int4 B, int K0, int Lmax,
                         const int64_t* __restrict__ off, uint32_t* __restrict__ occ)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int4 q = rects[r];
    const int64_t x0 = (int64_t)q.x - B.x, y0 = (int64_t)q.y - B.y;
    const int64_t x1 = (int64_t)q.z - B.x, y1 = (int64_t)q.w - B.y;
    for (int l = 0; l <= Lmax; ++l) {
        const int sh = K0 - l;
        const int64_t nl = (int64_t)1 << l;
        const int64_t i0 = x0 >> sh, i1 = min(x1 >> sh, nl - 1);
        const int64_t j0 = y0 >> sh, j1 = min(y1 >> sh, nl - 1);
        uint32_t* base = occ + off[l];
        for (int64_t j = j0; j <= j1; ++j) {
            int64_t a = j * nl + i0;
            const int64_t e = j * nl + i1;  // inclusive
            while (a <= e) {
                const int64_t wd = a >> 5;
                const int b0 = (int)(a & 31);
                const int b1 = (int)min<int64_t>(31, b0 + (e - a));
                const uint32_t mask = (b1 == 31 ? 0xFFFFFFFFu : ((1u << (b1 + 1)) - 1u)) & ~((1u << b0) - 1u);
                atomicOr(base + wd, mask);
                a += b1 - b0 + 1;
            }
        }
    }
}

// ---------------------------------------------------------------- summed-area tables
// Row pass, one block of 256 lanes per grid row j: S[(j+1)(n+1) + i + 1] =
// occupied cells (a, j) with a <= i; S[(j+1)(n+1)] = 0.  Row 0 is zeroed by
// the caller.
__global__ void __launch_bounds__(256) sat_rows(const uint32_t* __restrict__ bits, int64_t n,
                                                uint32_t* __restrict__ S)
{
    __shared__ uint32_t part[256];
    const int64_t j = blockIdx.x;
    const int t = threadIdx.x;
    const int64_t c = (n + 255) / 256;
    const int64_t a = min(n, (int64_t)t * c), b = min(n, a + c);
    auto bit = [&](int64_t i) -> uint32_t {
        const int64_t idx = j * n + i;
        return (bits[idx >> 5] >> (idx & 31)) & 1u;
    };
    uint32_t cnt = 0;
    for (int64_t i = a; i < b; ++i) cnt += bit(i);
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
    for (int64_t i = a; i < b; ++i) {
        run += bit(i);
        row[i + 1] = run;
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

// ---------------------------------------------------------------- GEMM
// C[m][n] = sum_k A(m,k) * B(k,n) with bf16 operands and FP32 accumulation;
// A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn].  One wave per 32x32
// tile of C; rows/columns beyond M/N are masked.  Split K: blockIdx.y takes
// k in [y*kc, (y+1)*kc) and writes its partial product to C + y*cstride
// (kc a multiple of 8).
__global__ void __launch_bounds__(64) gemm_bf16(const uint16_t* __restrict__ A, int64_t sam, int64_t sak,
                                                const uint16_t* __restrict__ B, int64_t sbk, int64_t sbn,
                                                float* __restrict__ C, int64_t ldc, int M, int N, int K,
                                                int kc, int64_t cstride)
{
    const int tiles_n = (N + 31) / 32;
    const int tm = blockIdx.x / tiles_n, tn = blockIdx.x % tiles_n;
    const int lane = threadIdx.x, r32 = lane & 31, q = lane >> 5;
    const int m = tm * 32 + r32, nn = tn * 32 + r32;
    const int kb = blockIdx.y * kc;
    const int ke = min(K, kb + kc);
    float* Cs = C + blockIdx.y * cstride;
    f32x16 acc = {0};
    for (int k0 = kb; k0 < ke; k0 += 8) {
        bf16x4 a, b;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int k = k0 + 4 * q + e;
            a[e] = (short)(m < M ? A[m * sam + k * sak] : 0);
            b[e] = (short)(nn < N ? B[k * sbk + nn * sbn] : 0);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = tm * 32 + 8 * (r >> 2) + 4 * q + (r & 3);
        if (row < M && nn < N) Cs[row * ldc + nn] = acc[r];
    }
}

// dst[i] += sum over the split partials src[y*n + i], y = 0..splits-1, in order.
__global__ void add_splits(float* __restrict__ dst, const float* __restrict__ src, int n, int splits)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = 0.0f;
    for (int y = 0; y < splits; ++y) acc += src[(int64_t)y * n + i];
    dst[i] += acc;
}

// Z += bias; A = bf16(gelu(Z)).  Z and A are [B][64] / [B][64].
__global__ void bias_gelu(float* __restrict__ Z, const float* __restrict__ bias,
                          uint16_t* __restrict__ A, int64_t total)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const float z = Z[i] + bias[i & (kMlpHid - 1)];
    Z[i] = z;
    A[i] = f2bf(gelu_f(z));
}

// Output layer: y = w4.a3 + b4, r = y - t; per sample dy = r/B and the loss
// term r^2/(2B); D = bf16(dy * w4 * gelu'(z3)).  The sums over samples are
// taken by out_grads.
__global__ void output_layer(const uint16_t* __restrict__ A3, const float* __restrict__ Z3,
                             const uint16_t* __restrict__ w4, const float* __restrict__ b4p,
                             const float* __restrict__ t, int B, float* __restrict__ y_out,
                             uint16_t* __restrict__ D, float* __restrict__ dy_out,
                             float* __restrict__ loss_out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    float y = 0.0f;
    for (int j = 0; j < kMlpHid; ++j) y += bf2f(w4[j]) * bf2f(A3[s * kMlpHid + j]);
    y += *b4p;
    if (y_out) y_out[s] = y;
    if (!t) return;
    const float r = y - t[s];
    const float dy = r / (float)B;
    dy_out[s] = dy;
    loss_out[s] = 0.5f * r * r / (float)B;
    for (int j = 0; j < kMlpHid; ++j)
        D[s * kMlpHid + j] = f2bf(dy * bf2f(w4[j]) * gelu_grad_f(Z3[s * kMlpHid + j]));
}

// Block sum of one value per lane (256 lanes).
__device__ inline float block_sum256(float v, float* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    return sh[0];
}

// Sums over the batch, one block per output: blocks 0..63 dw4[j] = sum dy a3[j],
// block 64 db4 = sum dy, block 65 the loss.
__global__ void __launch_bounds__(256) out_grads(const uint16_t* __restrict__ A3, const float* __restrict__ dy,
                                                 const float* __restrict__ lossv, int B,
                                                 float* __restrict__ grad, float* __restrict__ loss)
{
    __shared__ float sh[256];
    const int j = blockIdx.x;
    float acc = 0.0f;
    for (int s = threadIdx.x; s < B; s += 256) {
        if (j < kMlpHid) acc += dy[s] * bf2f(A3[s * kMlpHid + j]);
        else if (j == kMlpHid) acc += dy[s];
        else acc += lossv[s];
    }
    const float tot = block_sum256(acc, sh);
    if (threadIdx.x == 0) {
        if (j < kMlpHid) grad[kOffW4 + j] += tot;
        else if (j == kMlpHid) grad[kOffB4] += tot;
        else *loss += tot;
    }
}

// D = bf16(F * gelu'(Z)) for the hidden layer below.
__global__ void back_gelu(const float* __restrict__ F, const float* __restrict__ Z, uint16_t* __restrict__ D,
                          int64_t total)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    D[i] = f2bf(F[i] * gelu_grad_f(Z[i]));
}

// Bias gradient: column sums of D [B][64], one block per column.
__global__ void __launch_bounds__(256) bias_grad(const uint16_t* __restrict__ D, int B, float* __restrict__ g)
{
    __shared__ float sh[256];
    const int c = blockIdx.x;
    float acc = 0.0f;
    for (int s = threadIdx.x; s < B; s += 256) acc += bf2f(D[s * kMlpHid + c]);
    const float tot = block_sum256(acc, sh);
    if (threadIdx.x == 0) g[c] += tot;
}

__global__ void adam_step(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ mom,
                          float* __restrict__ vel, int n, float lr, float b1c, float b2c)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float m = 0.9f * mom[i] + 0.1f * gi;
    const float v = 0.999f * vel[i] + 0.001f * gi * gi;
    mom[i] = m;
    vel[i] = v;
    p[i] -= lr * (m / b1c) / (sqrtf(v / b2c) + 1e-8f);
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

// bf16 copies of the weights; the padding columns of W1 stay zero.  The walk
// copy bfw holds W2 and W3 with input column n at mlp_col(n), the column of
// the activation tile where the walk kernel's previous layer stores neuron n.
__global__ void to_bf16(const float* __restrict__ p, uint16_t* __restrict__ bf, uint16_t* __restrict__ bfw)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kMlpHid * kMlpInPad) {
        const bool pad = (i % kMlpInPad) >= kMlpIn;
        bf[kBfW1 + i] = pad ? (uint16_t)0 : f2bf(p[kOffW1 + i]);
    }
    if (i < kMlpHid * kMlpHid) {
        const uint16_t w2 = f2bf(p[kOffW2 + i]), w3 = f2bf(p[kOffW3 + i]);
        bf[kBfW2 + i] = w2;
        bf[kBfW3 + i] = w3;
        const int o = i / kMlpHid, n = i % kMlpHid;
        bfw[o * kMlpHid + mlp_col(n)] = w2;
        bfw[kMlpHid * kMlpHid + o * kMlpHid + mlp_col(n)] = w3;
    }
    if (i < kMlpHid) bf[kBfW4 + i] = f2bf(p[kOffW4 + i]);
}

__global__ void gather_rows(const uint16_t* __restrict__ X, const float* __restrict__ t,
                            const int32_t* __restrict__ idx, int B, uint16_t* __restrict__ Xb,
                            float* __restrict__ tb)
{
    const int s = blockIdx.x;
    const int k = threadIdx.x;
    const int64_t src = idx[s];
    if (k < kMlpInPad) Xb[(int64_t)s * kMlpInPad + k] = X[src * kMlpInPad + k];
    if (k == 0) tb[s] = t[src];
}

// Wave-level MFMA forward over rows of X (the walk kernel's inference path):
// one wave per 64 rows.
__global__ void __launch_bounds__(64) forward_waves(MlpNet m, const uint16_t* __restrict__ X, int64_t n,
                                                    float* __restrict__ y)
{
    __shared__ __align__(16) uint16_t xs[kMlpTileElems];
    __shared__ __align__(16) float small[kMlpSmallFloats];
    const int lane = threadIdx.x;
    const int64_t row = (int64_t)blockIdx.x * 64 + lane;
    mlp_stage_small(m, small, lane, 64);
    __syncthreads();
    for (int k = 0; k < kMlpInPad; ++k)
        xs[lane * kMlpStride + k] = row < n ? X[row * kMlpInPad + k] : (uint16_t)0;
    const float out = mlp_forward_rows(m, xs, small, lane, 64, lane);
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

int ensure_ws(frw_mlp* m, int64_t B)
{
    if (m->ws_batch >= B) return FRW_OK;
    (void)hipFree(m->d_Z1); (void)hipFree(m->d_Z2); (void)hipFree(m->d_Z3); (void)hipFree(m->d_F);
    (void)hipFree(m->d_A1); (void)hipFree(m->d_A2); (void)hipFree(m->d_A3); (void)hipFree(m->d_D); (void)hipFree(m->d_Xb);
    (void)hipFree(m->d_tb); (void)hipFree(m->d_idx); (void)hipFree(m->d_dy); (void)hipFree(m->d_lossv);
    const size_t h = (size_t)B * kMlpHid;
    int rc = FRW_OK;
    rc = rc ? rc : check_hip(hipMalloc(&m->d_Z1, h * 4), "mlp ws");
    rc = rc ? rc : check_hip(hipMalloc(&m->d_Z2, h * 4), "mlp ws");
    rc = rc ? rc : check_hip(hipMalloc(&m->d_Z3, h * 4), "mlp ws");
    rc = rc ? rc : check_hip(hipMalloc(&m->d_F, (size_t)B * kMlpInPad * 4), "mlp ws");
    rc = rc ? rc : check_hip(hipMalloc(&m->d_A1, h * 2), "mlp ws");
    rc = rc ? rc : check_hip(hipMalloc(&m->d_A2, h * 2), "mlp ws");
    rc = rc ? rc : check_hip(hipMalloc(&m->d_A3, h * 2), "mlp ws");
    rc = rc ? rc : check_hip(hipMalloc(&m->d_D, h * 2), "mlp ws");
    rc = rc ? rc : check_hip(hipMalloc(&m->d_Xb, (size_t)B * kMlpInPad * 2), "mlp ws");
    rc = rc ? rc : check_hip(hipMalloc(&m->d_tb, (size_t)B * 4), "mlp ws");
    rc = rc ? rc : check_hip(hipMalloc(&m->d_idx, (size_t)B * 4), "mlp ws");
    rc = rc ? rc : check_hip(hipMalloc(&m->d_dy, (size_t)B * 4), "mlp ws");
    rc = rc ? rc : check_hip(hipMalloc(&m->d_lossv, (size_t)B * 4), "mlp ws");
    if (rc == FRW_OK) m->ws_batch = B;
    return rc;
}

void gemm(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk, int64_t sbn, float* C,
          int64_t ldc, int M, int N, int K, hipStream_t s)
{
    const int tiles = ((M + 31) / 32) * ((N + 31) / 32);
    gemm_bf16<<<dim3(tiles, 1), 64, 0, s>>>(A, sam, sak, B, sbk, sbn, C, ldc, M, N, K, K, 0);
}

// Weight gradient dst (M x N, row-major) += A^T-shaped product over the batch
// (K = B rows), split over kMaxSplits slices of the batch into the scratch and
// summed in a fixed order.
void gemm_wgrad(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk, int64_t sbn,
                float* dst, float* scratch, int M, int N, int K, hipStream_t s)
{
    int splits = std::max(1, std::min(kMaxSplits, K / 64));
    int kc = (K + splits - 1) / splits;
    kc = (kc + 7) / 8 * 8;
    splits = (K + kc - 1) / kc;
    const int tiles = ((M + 31) / 32) * ((N + 31) / 32);
    gemm_bf16<<<dim3(tiles, splits), 64, 0, s>>>(A, sam, sak, B, sbk, sbn, scratch, N, M, N, K, kc,
                                                 (int64_t)M * N);
    add_splits<<<blocks((int64_t)M * N, 256), 256, 0, s>>>(dst, scratch, M * N, splits);
}

// Forward + backward of batch (Xb, tb) of size B (a multiple of 32): adds the
// gradient into d_grad and the mean loss into d_loss.  Hidden-layer gradients
// are rounded to bf16 before they enter the MFMA weight-gradient products.
int gradient(frw_mlp* m, const uint16_t* Xb, const float* tb, int B, hipStream_t s)
{
    const uint16_t* W1 = m->d_bf + kBfW1;
    const uint16_t* W2 = m->d_bf + kBfW2;
    const uint16_t* W3 = m->d_bf + kBfW3;
    const uint16_t* w4 = m->d_bf + kBfW4;
    float* P = m->d_params;
    float* G = m->d_grad;
    float* GS = m->d_gs;  // split-K partials of the weight gradients
    const int64_t h = (int64_t)B * kMlpHid;
    const int H = kMlpHid, IN = kMlpInPad;
    // forward: Z = A_prev * W^T + b, A = bf16(gelu(Z))
    gemm(Xb, IN, 1, W1, 1, IN, m->d_Z1, H, B, H, IN, s);
    bias_gelu<<<blocks(h, 256), 256, 0, s>>>(m->d_Z1, P + kOffB1, m->d_A1, h);
    gemm(m->d_A1, H, 1, W2, 1, H, m->d_Z2, H, B, H, H, s);
    bias_gelu<<<blocks(h, 256), 256, 0, s>>>(m->d_Z2, P + kOffB2, m->d_A2, h);
    gemm(m->d_A2, H, 1, W3, 1, H, m->d_Z3, H, B, H, H, s);
    bias_gelu<<<blocks(h, 256), 256, 0, s>>>(m->d_Z3, P + kOffB3, m->d_A3, h);
    // output layer: loss, dw4, db4 and D3
    output_layer<<<blocks(B, 256), 256, 0, s>>>(m->d_A3, m->d_Z3, w4, P + kOffB4, tb, B, nullptr, m->d_D,
                                                m->d_dy, m->d_lossv);
    out_grads<<<kMlpHid + 2, 256, 0, s>>>(m->d_A3, m->d_dy, m->d_lossv, B, G, m->d_loss);
    // layer 3: dW3 = D3^T A2, db3 = colsum D3, F = D3 W3, D2 = bf16(F gelu'(Z2))
    gemm_wgrad(m->d_D, 1, H, m->d_A2, H, 1, G + kOffW3, GS, H, H, B, s);
    bias_grad<<<H, 256, 0, s>>>(m->d_D, B, G + kOffB3);
    gemm(m->d_D, H, 1, W3, H, 1, m->d_F, H, B, H, H, s);
    back_gelu<<<blocks(h, 256), 256, 0, s>>>(m->d_F, m->d_Z2, m->d_D, h);
    // layer 2
    gemm_wgrad(m->d_D, 1, H, m->d_A1, H, 1, G + kOffW2, GS, H, H, B, s);
    bias_grad<<<H, 256, 0, s>>>(m->d_D, B, G + kOffB2);
    gemm(m->d_D, H, 1, W2, H, 1, m->d_F, H, B, H, H, s);
    back_gelu<<<blocks(h, 256), 256, 0, s>>>(m->d_F, m->d_Z1, m->d_D, h);
    // layer 1: dW1 = D1^T X ([64][72]), db1
    gemm_wgrad(m->d_D, 1, H, Xb, IN, 1, G + kOffW1, GS, H, IN, B, s);
    bias_grad<<<H, 256, 0, s>>>(m->d_D, B, G + kOffB1);
    return check_hip(hipGetLastError(), "mlp gradient");
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
        const int64_t bits = ((int64_t)1 << l) * ((int64_t)1 << l);
        words += (bits + 31) / 32 + 1;  // +1: the window reader loads word pairs
    }
    for (int l = 0; l <= m->Vmax; ++l) {
        m->sat_off.push_back(entries);
        entries += (((int64_t)1 << l) + 1) * (((int64_t)1 << l) + 1);
    }
    int rc = check_hip(hipMalloc(&m->d_occ, words * 4), "hipMalloc occupancy");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_sat, entries * 4), "hipMalloc summed-area tables");
    if (rc == FRW_OK) rc = check_hip(hipMemset(m->d_occ, 0, words * 4), "memset occupancy");
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
    for (int l = 0; l <= m->Vmax && rc == FRW_OK; ++l) {
        const int64_t n = (int64_t)1 << l;
        uint32_t* S = m->d_sat + m->sat_off[l];
        rc = check_hip(hipMemset(S, 0, (size_t)(n + 1) * 4), "memset table row 0");
        if (rc != FRW_OK) break;
        sat_rows<<<(unsigned)n, 256>>>(m->d_occ + m->occ_off[l], n, S);
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
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_gs, (size_t)kMaxSplits * kMlpHid * kMlpInPad * 4), "hipMalloc grad scratch");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_adam_m, kMlpParams * 4), "hipMalloc adam");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_adam_v, kMlpParams * 4), "hipMalloc adam");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_bf, kBfCount * 2), "hipMalloc bf16");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_bfw, 2 * kMlpHid * kMlpHid * 2), "hipMalloc bf16 walk copy");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&m->d_loss, 4), "hipMalloc loss");
    if (rc == FRW_OK) rc = check_hip(hipMemset(m->d_adam_m, 0, kMlpParams * 4), "memset");
    if (rc == FRW_OK) rc = check_hip(hipMemset(m->d_adam_v, 0, kMlpParams * 4), "memset");
    if (rc == FRW_OK) rc = check_hip(hipMemcpy(m->d_params, P.data(), kMlpParams * 4, hipMemcpyHostToDevice), "copy params");
    if (rc == FRW_OK) {
        to_bf16<<<blocks(kMlpHid * kMlpInPad, 256), 256>>>(m->d_params, m->d_bf, m->d_bfw);
        rc = check_hip(hipDeviceSynchronize(), "to_bf16");
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
    for (void* p : {(void*)m->d_occ, (void*)m->d_sat, (void*)m->d_occ_off, (void*)m->d_sat_off, (void*)m->d_params, (void*)m->d_bf, (void*)m->d_bfw, (void*)m->d_adam_m,
                    (void*)m->d_adam_v, (void*)m->d_grad, (void*)m->d_gs, (void*)m->d_X, (void*)m->d_t, (void*)m->d_Z1,
                    (void*)m->d_Z2, (void*)m->d_Z3, (void*)m->d_F, (void*)m->d_A1, (void*)m->d_A2,
                    (void*)m->d_A3, (void*)m->d_D, (void*)m->d_Xb, (void*)m->d_tb, (void*)m->d_loss,
                    (void*)m->d_idx, (void*)m->d_dy, (void*)m->d_lossv})
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
    to_bf16<<<blocks(kMlpHid * kMlpInPad, 256), 256>>>(m->d_params, m->d_bf, m->d_bfw);
    return check_hip(hipDeviceSynchronize(), "to_bf16");
}

int frw_mlp_occupancy(const frw_mlp* m, int32_t level, uint32_t* out_words, int64_t n_words)
{
    if (!m || !out_words || level < 0 || level > m->Pmax) return set_error(FRW_E_INVALID, "bad level");
    const int64_t bits = ((int64_t)1 << level) * ((int64_t)1 << level);
    const int64_t need = (bits + 31) / 32;
    if (n_words < need) return set_error(FRW_E_INVALID, "output too small");
    return check_hip(hipMemcpy(out_words, m->d_occ + m->occ_off[level], need * 4, hipMemcpyDeviceToHost),
                     "copy occupancy");
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
    if (!m || !X || !t || !grad || n <= 0 || n % 32) return set_error(FRW_E_INVALID, "n must be a positive multiple of 32");
    int rc = ensure_ws(m, n);
    if (rc != FRW_OK) return rc;
    uint16_t* Xb = nullptr;
    rc = check_hip(hipMalloc(&Xb, (size_t)n * kMlpInPad * 2), "hipMalloc gradient");
    if (rc != FRW_OK) return rc;
    narrow_bf16<<<blocks(n * kMlpInPad, 256), 256>>>(X, Xb, n * kMlpInPad);
    rc = check_hip(hipMemset(m->d_grad, 0, kMlpParams * 4), "memset grad");
    if (rc == FRW_OK) rc = check_hip(hipMemset(m->d_loss, 0, 4), "memset loss");
    if (rc == FRW_OK) rc = gradient(m, Xb, t, (int)n, nullptr);
    if (rc == FRW_OK) rc = check_hip(hipDeviceSynchronize(), "mlp gradient");
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
    uint16_t* X = nullptr;
    float* T = nullptr;
    int rc = check_hip(hipMalloc(&X, (size_t)tot * kMlpInPad * 2), "hipMalloc samples");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&T, (size_t)tot * 4), "hipMalloc targets");
    if (rc == FRW_OK && old_n) {
        rc = check_hip(hipMemcpy(X, m->d_X, (size_t)old_n * kMlpInPad * 2, hipMemcpyDeviceToDevice), "copy samples");
        if (rc == FRW_OK) rc = check_hip(hipMemcpy(T, m->d_t, (size_t)old_n * 4, hipMemcpyDeviceToDevice), "copy targets");
    }
    if (rc == FRW_OK) {
        features_kernel<<<blocks(n, 256), 256>>>(mlp_pyramid(src), pts, h, n, X + old_n * kMlpInPad, dstar,
                                                 T + old_n, nullptr);
        rc = check_hip(hipDeviceSynchronize(), "features");
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
    m->n_samples = tot;
    return FRW_OK;
}

int frw_mlp_set_samples(frw_mlp* m, const double* pts, const double* h, const double* dstar, int64_t n)
{
    if (!m || !pts || !h || !dstar || n <= 0) return set_error(FRW_E_INVALID, "bad arguments");
    (void)hipFree(m->d_X);
    (void)hipFree(m->d_t);
    m->d_X = nullptr;
    m->d_t = nullptr;
    m->n_samples = 0;
    return frw_mlp_add_samples(m, nullptr, pts, h, dstar, n);
}

int frw_mlp_train(frw_mlp* m, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* last_loss)
{
    if (!m || epochs <= 0 || batch <= 0 || batch % 32 || lr <= 0) return set_error(FRW_E_INVALID, "bad training arguments");
    if (m->n_samples < batch) return set_error(FRW_E_RANGE, "fewer samples than one batch");
    int rc = ensure_ws(m, batch);
    if (rc != FRW_OK) return rc;
    const int64_t n = m->n_samples;
    std::vector<int32_t> perm((size_t)n);
    std::iota(perm.begin(), perm.end(), 0);
    int32_t* d_perm = nullptr;
    rc = check_hip(hipMalloc(&d_perm, (size_t)n * 4), "hipMalloc perm");
    float loss_sum = 0.f;
    int64_t loss_n = 0;
    for (int ep = 0; ep < epochs && rc == FRW_OK; ++ep) {
        // Fisher-Yates with Philox draws: a fresh order every epoch
        for (int64_t i = n - 1; i > 0; --i) {
            uint32_t w[4] = {(uint32_t)i, (uint32_t)ep, 0u, 0x5348464Cu};
            philox4x32_10(w, (uint32_t)seed, (uint32_t)(seed >> 32));
            const int64_t j = (int64_t)(((uint64_t)w[0] * (uint64_t)(i + 1)) >> 32);
            std::swap(perm[i], perm[j]);
        }
        rc = check_hip(hipMemcpy(d_perm, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice), "copy perm");
        loss_sum = 0.f;
        loss_n = 0;
        for (int64_t b0 = 0; b0 + batch <= n && rc == FRW_OK; b0 += batch) {
            gather_rows<<<batch, 128>>>(m->d_X, m->d_t, d_perm + b0, batch, m->d_Xb, m->d_tb);
            rc = check_hip(hipMemsetAsync(m->d_grad, 0, kMlpParams * 4, nullptr), "memset grad");
            if (rc == FRW_OK) rc = check_hip(hipMemsetAsync(m->d_loss, 0, 4, nullptr), "memset loss");
            if (rc == FRW_OK) rc = gradient(m, m->d_Xb, m->d_tb, batch, nullptr);
            if (rc != FRW_OK) break;
            m->adam_step++;
            const float b1c = 1.0f - std::pow(0.9f, (float)m->adam_step);
            const float b2c = 1.0f - std::pow(0.999f, (float)m->adam_step);
            adam_step<<<blocks(kMlpParams, 256), 256>>>(m->d_params, m->d_grad, m->d_adam_m, m->d_adam_v,
                                                        kMlpParams, lr, b1c, b2c);
            to_bf16<<<blocks(kMlpHid * kMlpInPad, 256), 256>>>(m->d_params, m->d_bf, m->d_bfw);
            if (((b0 / batch) & 63) == 63 || b0 + 2 * batch > n) {
                float l = 0.f;
                rc = check_hip(hipMemcpy(&l, m->d_loss, 4, hipMemcpyDeviceToHost), "read loss");
                loss_sum += l;
                loss_n++;
            }
        }
    }
    if (d_perm) (void)hipFree(d_perm);
    if (rc == FRW_OK) rc = check_hip(hipDeviceSynchronize(), "train");
    if (rc == FRW_OK && last_loss) *last_loss = loss_n ? loss_sum / loss_n : 0.f;
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
    int64_t *d_counts = nullptr, *d_hops = nullptr;
    int32_t *d_hit = nullptr, *d_hp = nullptr;
    unsigned long long* d_off = nullptr;
    rc = check_hip(hipMalloc(&d_counts, nb * 8), "hipMalloc");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&d_hops, 8), "hipMalloc");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&d_hit, nw * 4), "hipMalloc");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&d_hp, nw * 4), "hipMalloc");
    if (rc == FRW_OK) rc = check_hip(hipMalloc(&d_off, nw * 8), "hipMalloc");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_counts, 0, nb * 8), "memset");
    if (rc == FRW_OK) rc = check_hip(hipMemset(d_hops, 0, 8), "memset");
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
    for (void* x : {(void*)d_counts, (void*)d_hops, (void*)d_hit, (void*)d_hp, (void*)d_off}) (void)hipFree(x);
    return rc;
}

namespace {

int check_learned(const frw_tree* t, const frw_mlp* m, const frw_walk_params* p, float alpha)
{
    if (!t || !m || !p) return set_error(FRW_E_INVALID, "bad arguments");
    if (m->tree != t) return set_error(FRW_E_INVALID, "network built for another tree");
    if (p->master < 0 || p->master >= t->n_rects || p->n_walks <= 0 || p->first_walk < 0 ||
        p->first_walk + p->n_walks > ((int64_t)1 << 32))
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
