/*
 * CPU oracle, part 3: the learned square-size proposal (no counterpart in the
 * reference).  Test infrastructure only (see oracle.h).
 *
 * Restates, in plain FP32 loops, the definition the library's kernels
 * implement, so the GPU features, forward pass and gradients can be checked:
 *
 * Features of a point p with previous square half-size h (67 values):
 *   K0  = smallest k with 2^k > max(W, H) of the bounds, Lmax = min(K0, 15)
 *   e   = floor(log2(max(h, 1))), level l = clamp(K0 - e + 1, 0, Lmax),
 *   s   = 2^(K0 - l) (cell size), q = (p - bounds corner) / s, (ci, cj) = floor(q)
 *   f[v*8+u] = 1 if cell (ci-4+u, cj-4+v) meets a rect (closed rect against
 *              the half-open cell) or lies outside the 2^l x 2^l grid, else 0
 *   f[64] = l / 16, f[65], f[66] = q - floor(q) (x, y); all rounded to bf16
 *   (a point 16 or more cells outside the grid reads as all-occupied, 0, 0)
 * Target: log2(max(d*, s/16) / s) clamped to [-4, 4], d* = nearest-rect
 * distance.
 *
 * Verification of a square of half-size d centred at p (closed):
 *   Vmax = clamp(K0 - floor(log2(max(1, spacing / 128))), 0, min(K0, 15)),
 *          spacing = sqrt(W H / n_rects)
 *   v = clamp(K0 - floor(log2 d) + 4, 0, Vmax), s_v = 2^(K0 - v)
 *   certified when the square is inside the bounds (inclusive), its covering
 *   cells floor((corner - bounds corner) / s_v) lie in the 2^v grid, and no
 *   rect meets that cell block.
 *
 * Network: 67 (padded to 72) -> 64 -> 64 -> 64 -> 1, GELU (erf form) on the
 * hidden layers.  Weights and every activation fed to a layer are rounded to
 * bfloat16 (round to nearest even); sums and the output are FP32.  Gradients
 * of the mean of (y - t)^2 / 2, with the hidden-layer gradients rounded to
 * bfloat16 before they enter the weight-gradient products, as on the GPU.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define IN 67
#define INP 72
#define HID 64

float orc_bf16(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return x; /* inf / nan unchanged */
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7FFFu + lsb;
    u &= 0xFFFF0000u;
    float r;
    memcpy(&r, &u, 4);
    return r;
}

/* Rounding of weights and activations to bf16 can be switched off so the
 * gradient restatement can be checked against a float64 autodiff of the same
 * network; it is on by default (the GPU's arithmetic). */
static int g_round = 1;

void orc_mlp_set_rounding(int on) { g_round = on; }

static float R(float x) { return g_round ? orc_bf16(x) : x; }

static float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

static float gelu_grad(float z)
{
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

/* ---- features --------------------------------------------------------- */

static int k0_of(const orc_box* b)
{
    const double W = b->x1 - b->x0, H = b->y1 - b->y0;
    const double m = W > H ? W : H;
    int k = 0;
    while (ldexp(1.0, k) <= m) ++k;
    return k;
}

int orc_mlp_level(const orc_box* bounds, double h, double* cell)
{
    const int K0 = k0_of(bounds);
    const int Lmax = K0 < 15 ? K0 : 15;
    const int e = h >= 1.0 ? ilogb(h) : 0;
    int l = K0 - e + 1;
    if (l < 0) l = 0;
    if (l > Lmax) l = Lmax;
    if (cell) *cell = ldexp(1.0, K0 - l);
    return l;
}

/* Brute force over all rects: the window the GPU reads from its bitmaps. */
void orc_mlp_features(const double* xyxy, int n, const orc_box* bounds, double px, double py,
                      double h, float out[67])
{
    double s;
    const int l = orc_mlp_level(bounds, h, &s);
    const int64_t ncell = (int64_t)1 << l;
    const double qx = (px - bounds->x0) / s, qy = (py - bounds->y0) / s;
    const double fx = floor(qx), fy = floor(qy);
    out[64] = orc_bf16((float)l / 16.0f);
    if (!(fx > -16.0 && fy > -16.0 && fx < (double)(ncell + 16) && fy < (double)(ncell + 16))) {
        for (int k = 0; k < 64; ++k) out[k] = 1.0f;
        out[65] = out[66] = 0.0f;
        return;
    }
    out[65] = orc_bf16((float)(qx - fx));
    out[66] = orc_bf16((float)(qy - fy));
    const int64_t ci = (int64_t)fx;
    const int64_t cj = (int64_t)fy;
    for (int v = 0; v < 8; ++v)
        for (int u = 0; u < 8; ++u) {
            const int64_t i = ci - 4 + u, j = cj - 4 + v;
            float bit = 1.0f;
            if (i >= 0 && j >= 0 && i < ncell && j < ncell) {
                const double a = i * s, b = (i + 1) * s, c = j * s, d = (j + 1) * s;
                bit = 0.0f;
                for (int k = 0; k < n; ++k) {
                    const double* r = xyxy + 4 * k;
                    if (r[0] - bounds->x0 < b && r[2] - bounds->x0 >= a && r[1] - bounds->y0 < d &&
                        r[3] - bounds->y0 >= c) {
                        bit = 1.0f;
                        break;
                    }
                }
            }
            out[v * 8 + u] = bit;
        }
}

int orc_mlp_vmax(const orc_box* bounds, int n_rects)
{
    return orc_mlp_vmax_capped(bounds, n_rects, 15);
}

int orc_mlp_vmax_capped(const orc_box* bounds, int n_rects, int cap)
{
    const int K0 = k0_of(bounds);
    double W = bounds->x1 - bounds->x0, H = bounds->y1 - bounds->y0;
    if (W < 1) W = 1;
    if (H < 1) H = 1;
    const double spacing = sqrt(W * H / (double)(n_rects > 0 ? n_rects : 1));
    const double target = spacing / 128.0 > 1.0 ? spacing / 128.0 : 1.0;
    int v = K0 - (int)floor(log2(target));
    if (v > K0) v = K0;
    if (v > cap) v = cap;
    if (v < 0) v = 0;
    return v;
}

/* Brute force over all rects (no summed-area table). */
int orc_mlp_certify(const double* xyxy, int n, const orc_box* bounds, double px, double py, double d)
{
    if (!(d > 0.0 && d < 1e300)) return 0;
    const double x0 = px - d, y0 = py - d, x1 = px + d, y1 = py + d;
    if (!(y0 >= bounds->y0 && x0 >= bounds->x0 && y1 <= bounds->y1 && x1 <= bounds->x1)) return 0;
    const int K0 = k0_of(bounds), Vmax = orc_mlp_vmax(bounds, n);
    int v = K0 - ilogb(d) + 4;
    if (v < 0) v = 0;
    if (v > Vmax) v = Vmax;
    const double s = ldexp(1.0, K0 - v);
    const double nl = ldexp(1.0, v);
    const double a0 = floor((x0 - bounds->x0) / s), a1 = floor((x1 - bounds->x0) / s);
    const double b0 = floor((y0 - bounds->y0) / s), b1 = floor((y1 - bounds->y0) / s);
    if (!(a0 >= 0.0 && b0 >= 0.0 && a1 < nl && b1 < nl)) return 0;
    /* the cell block [a0 s, (a1+1) s) x [b0 s, (b1+1) s) against closed rects */
    for (int k = 0; k < n; ++k) {
        const double* r = xyxy + 4 * k;
        if (r[0] - bounds->x0 < (a1 + 1) * s && r[2] - bounds->x0 >= a0 * s &&
            r[1] - bounds->y0 < (b1 + 1) * s && r[3] - bounds->y0 >= b0 * s)
            return 0;
    }
    return 1;
}

/* ---- network ------------------------------------------------------------
 * Parameter block (FP32 master values, row-major [out][in]):
 *   W1[64][72], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], w4[64], b4[1] */
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + HID * INP)
#define OFF_W2 (OFF_B1 + HID)
#define OFF_B2 (OFF_W2 + HID * HID)
#define OFF_W3 (OFF_B2 + HID)
#define OFF_B3 (OFF_W3 + HID * HID)
#define OFF_W4 (OFF_B3 + HID)
#define OFF_B4 (OFF_W4 + HID)
#define N_PARAMS (OFF_B4 + 1)

int orc_mlp_param_count(void) { return N_PARAMS; }

static void layer(const float* W, const float* b, const float* in, int nin, float* z, float* a)
{
    for (int o = 0; o < HID; ++o) {
        float acc = 0.0f;
        for (int i = 0; i < nin; ++i) acc += R(W[o * nin + i]) * in[i];
        z[o] = acc + b[o];
        a[o] = R(gelu(z[o]));
    }
}

/* y[B] for inputs X[B][72] (values already bf16-exact). */
void orc_mlp_forward(const float* P, const float* X, int B, float* y)
{
    float z[HID], a1[HID], a2[HID], a3[HID], xin[INP];
    for (int s = 0; s < B; ++s) {
        for (int i = 0; i < INP; ++i) xin[i] = R(X[s * INP + i]);
        layer(P + OFF_W1, P + OFF_B1, xin, INP, z, a1);
        layer(P + OFF_W2, P + OFF_B2, a1, HID, z, a2);
        layer(P + OFF_W3, P + OFF_B3, a2, HID, z, a3);
        float acc = 0.0f;
        for (int j = 0; j < HID; ++j) acc += R(P[OFF_W4 + j]) * a3[j];
        y[s] = acc + P[OFF_B4];
    }
}

/* Loss and gradient G (same layout as P) of mean((y - t)^2)/2 over B samples. */
float orc_mlp_backward(const float* P, const float* X, const float* t, int B, float* G)
{
    memset(G, 0, sizeof(float) * N_PARAMS);
    float loss = 0.0f;
    float xin[INP], z1[HID], z2[HID], z3[HID], a1[HID], a2[HID], a3[HID];
    float d3[HID], d2[HID], d1[HID];
    for (int s = 0; s < B; ++s) {
        for (int i = 0; i < INP; ++i) xin[i] = R(X[s * INP + i]);
        layer(P + OFF_W1, P + OFF_B1, xin, INP, z1, a1);
        layer(P + OFF_W2, P + OFF_B2, a1, HID, z2, a2);
        layer(P + OFF_W3, P + OFF_B3, a2, HID, z3, a3);
        float y = 0.0f;
        for (int j = 0; j < HID; ++j) y += R(P[OFF_W4 + j]) * a3[j];
        y += P[OFF_B4];
        const float r = y - t[s];
        loss += 0.5f * r * r;
        const float dy = r / (float)B;
        G[OFF_B4] += dy;
        for (int j = 0; j < HID; ++j) {
            G[OFF_W4 + j] += dy * a3[j];
            d3[j] = R(dy * R(P[OFF_W4 + j]) * gelu_grad(z3[j]));
        }
        for (int o = 0; o < HID; ++o) {
            G[OFF_B3 + o] += d3[o];
            for (int i = 0; i < HID; ++i) G[OFF_W3 + o * HID + i] += d3[o] * a2[i];
        }
        for (int i = 0; i < HID; ++i) {
            float acc = 0.0f;
            for (int o = 0; o < HID; ++o) acc += d3[o] * R(P[OFF_W3 + o * HID + i]);
            d2[i] = R(acc * gelu_grad(z2[i]));
        }
        for (int o = 0; o < HID; ++o) {
            G[OFF_B2 + o] += d2[o];
            for (int i = 0; i < HID; ++i) G[OFF_W2 + o * HID + i] += d2[o] * a1[i];
        }
        for (int i = 0; i < HID; ++i) {
            float acc = 0.0f;
            for (int o = 0; o < HID; ++o) acc += d2[o] * R(P[OFF_W2 + o * HID + i]);
            d1[i] = R(acc * gelu_grad(z1[i]));
        }
        for (int o = 0; o < HID; ++o) {
            G[OFF_B1 + o] += d1[o];
            for (int i = 0; i < INP; ++i) G[OFF_W1 + o * INP + i] += d1[o] * xin[i];
        }
    }
    return loss / (float)B;
}
