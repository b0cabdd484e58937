/*
 * CPU oracle, part 4: the weighted (physical) FRW estimator of capacitance.
 * Test infrastructure only (see oracle.h).  The reference computes Green's-
 * function gradients but never uses them (randomWalk/CPU/getAvgPathLen.cpp:
 * 83-84); this restates, from the mathematics, the estimator the library's
 * weighted walk mode implements.
 *
 * Capacitance (per unit length, permittivity eps) between the master m and
 * conductor j, with m's Gaussian surface G (outward normal n, perimeter L):
 *   C_mj = Q_m for V_j = 1 = -eps * contour_G  n . grad phi_j dl,
 *   grad phi_j(r) = contour_S(r) grad_r P(r, r1) phi_j(r1) dl1,
 * S(r) the largest empty square centred at r and P its exit density.  One
 * sample: r uniform on G, r1 from P on S(r) (a Green-CDF panel, its centre),
 * weight w = -eps L (n . g_k) / (w_k 2h) with h the half-side of S, w_k the
 * panel's probability and g_k the gradient of that probability with respect
 * to the source for the unit square; phi_j(r1) is the walk from r1 (1 when it
 * ends on j).  So C_mj ~ mean of w over walks ending on j, and the weights of
 * all walks (escapes included) average to zero.
 *
 * Exit density of the unit square on a side, at arc parameter s (clockwise),
 * for a source at distance v from that side and at u along it:
 *   G(s; u, v) = sum_n 2 sin(n pi s) sin(n pi u) sinh(n pi (1 - v)) / sinh(n pi)
 * At the centre (u = v = 1/2):
 *   dG/du = A(s) = sum_{n even} n pi cos(n pi / 2) sin(n pi s) / cosh(n pi / 2)
 *   dG/dv = B(s) = -sum_{n odd} n pi sin(n pi / 2) sin(n pi s) / sinh(n pi / 2)
 * In y-down coordinates (top side y = 0, panels clockwise from the top-left
 * corner): top u = x0, v = y0; right u = y0, v = 1 - x0; bottom u = 1 - x0,
 * v = 1 - y0; left u = 1 - y0, v = x0.  Panel k (side k / 50, s = (k % 50 +
 * 1/2) / 50) has g_k = grad G(s_k) / (50 T), T the normalisation of the
 * panel probabilities (orc_green_cdf).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

static const double W_PI = 3.14159265358979323846;

/* Exit density of a side for a source at (u, v), truncated series. */
double orc_square_density(double s, double u, double v)
{
    double k = 0.0;
    for (int n = 1; n <= 60; ++n)
        k += 2.0 * sin(n * W_PI * s) * sin(n * W_PI * u) * (sinh(n * W_PI * (1.0 - v)) / sinh(n * W_PI));
    return k;
}

static double grad_a(double s)
{
    double a = 0.0;
    for (int n = 2; n <= 40; n += 2) {
        const double c = (n % 4 == 0) ? 1.0 : -1.0; /* cos(n pi / 2) */
        a += n * W_PI * c * sin(n * W_PI * s) / cosh(n * W_PI / 2.0);
    }
    return a;
}

static double grad_b(double s)
{
    double b = 0.0;
    for (int n = 1; n <= 39; n += 2) {
        const double sg = ((n & 3) == 1) ? 1.0 : -1.0; /* sin(n pi / 2) */
        b -= n * W_PI * sg * sin(n * W_PI * s) / sinh(n * W_PI / 2.0);
    }
    return b;
}

/* gx[k], gy[k]: gradient of panel k's probability with respect to the source
 * position, unit square, source at the centre, y down. */
void orc_green_grad(double gx[ORC_NSAMPLE], double gy[ORC_NSAMPLE])
{
    const int per = ORC_NSAMPLE / 4;
    double intg[ORC_NSAMPLE + 1];
    orc_green_cdf(intg);
    /* the normalisation the cdf applied: raw midpoint mass / normalised mass */
    double raw = 0.0;
    for (int j = 0; j < per; ++j) {
        const double s = (j + 0.5) / per;
        double k = 0.0;
        for (int n = 1; n <= 19; n += 2) k += (((n & 3) == 1) ? 1.0 : -1.0) * sin(n * W_PI * s) / cosh(n * W_PI / 2.0);
        raw += k / per;
    }
    const double T = 4.0 * raw;
    for (int k = 0; k < ORC_NSAMPLE; ++k) {
        const int side = k / per;
        const double s = ((k % per) + 0.5) / per;
        const double A = grad_a(s) / (per * T), B = grad_b(s) / (per * T);
        switch (side) {
        case 0: gx[k] = A;  gy[k] = B;  break;  /* top: u = x0, v = y0 */
        case 1: gx[k] = -B; gy[k] = A;  break;  /* right: u = y0, v = 1 - x0 */
        case 2: gx[k] = -A; gy[k] = -B; break;  /* bottom: u = 1 - x0, v = 1 - y0 */
        default: gx[k] = B; gy[k] = -A; break;  /* left: u = 1 - y0, v = x0 */
        }
    }
}

/* First-hop importance table (Yu et al. 2013, importance sampling of the
 * first hop): panel k drawn with probability q_k proportional to |g[k]|, as
 * u32 thresholds T[k] = floor(sum_{j<k} |g_j| / S * 2^32); q_k is the exact
 * draw probability (threshold difference). */
void orc_importance_table(const double g[ORC_NSAMPLE], uint32_t T[ORC_NSAMPLE], double q[ORC_NSAMPLE])
{
    double S = 0.0;
    for (int k = 0; k < ORC_NSAMPLE; ++k) S += fabs(g[k]);
    double cum = 0.0;
    for (int k = 0; k < ORC_NSAMPLE; ++k) {
        T[k] = (uint32_t)floor(cum / S * 4294967296.0);
        cum += fabs(g[k]);
    }
    for (int k = 0; k < ORC_NSAMPLE; ++k) {
        const double hi = k + 1 < ORC_NSAMPLE ? (double)T[k + 1] : 4294967296.0;
        q[k] = (hi - (double)T[k]) / 4294967296.0;
    }
}

/* Centre of panel k on the square (x0, y0) - (x0 + a, y0 + a), clockwise. */
void orc_panel_centre(double x0, double y0, double a, int k, double* px, double* py)
{
    const int per = ORC_NSAMPLE / 4;
    const double t = ((k % per) + 0.5) / per * a;
    switch (k / per) {
    case 0: *px = x0 + t;     *py = y0;         break;
    case 1: *px = x0 + a;     *py = y0 + t;     break;
    case 2: *px = x0 + a - t; *py = y0 + a;     break;
    default: *px = x0;        *py = y0 + a - t; break;
    }
}

/* Start of a weighted walk: the point on the Gaussian surface G from the
 * walk's 'GAUS' Philox word, its outward normal, and the perimeter. */
void orc_gauss_point(const orc_box* G, uint64_t seed, uint64_t walk, uint32_t job, double* rx, double* ry,
                     double* nx, double* ny, double* L)
{
    uint32_t ctr[4] = {(uint32_t)walk, job, 0u, 0x47415553u ^ (uint32_t)(walk >> 32)}; /* 'GAUS' ^ high word */
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    orc_philox(ctr, key, out);
    const double w = G->x1 - G->x0, h = G->y1 - G->y0;
    *L = 2.0 * (w + h);
    const double t = (double)out[0] * (1.0 / 4294967296.0) * (*L);
    if (t < w) {
        *rx = G->x0 + t; *ry = G->y0; *nx = 0.0; *ny = -1.0;
    } else if (t < w + h) {
        *rx = G->x1; *ry = G->y0 + (t - w); *nx = 1.0; *ny = 0.0;
    } else if (t < 2.0 * w + h) {
        *rx = G->x1 - (t - w - h); *ry = G->y1; *nx = 0.0; *ny = 1.0;
    } else {
        *rx = G->x0; *ry = G->y1 - (t - 2.0 * w - h); *nx = -1.0; *ny = 0.0;
    }
}

/* One weighted walk.  With importance sampling (tx/ty/qx/qy non-NULL) the
 * first panel comes from the |gx| table when the normal is along x, else from
 * the |gy| table, and the weight divides by that table's probability. */
int orc_walk_weighted_one(const orc_index* ix, const orc_box* G, const uint32_t T[ORC_NSAMPLE],
                          const double gx[ORC_NSAMPLE], const double gy[ORC_NSAMPLE],
                          const double pw[ORC_NSAMPLE], const uint32_t* tx, const uint32_t* ty,
                          const double* qx, const double* qy, uint64_t seed, uint32_t job, uint64_t walk,
                          int64_t max_hops, double eps, int64_t* hops, double* weight)
{
    const orc_box B = orc_index_bounds(ix);
    double rx, ry, nx, ny, L;
    orc_gauss_point(G, seed, walk, job, &rx, &ry, &nx, &ny, &L);
    /* first square: the largest empty square at r inside the bounds */
    double h = orc_nearest(ix, rx, ry, NULL);
    h = fmin(h, fmin(fmin(rx - B.x0, B.x1 - rx), fmin(ry - B.y0, B.y1 - ry)));
    *weight = 0.0;
    if (!(h > 0.0)) { *hops = 0; return ORC_HIT_STUCK; }
    const uint32_t u0 = orc_walk_u32(seed, walk, job, 0u);
    int k;
    double prob;
    if (tx) {
        const int along_x = nx != 0.0;
        k = orc_sample_index(along_x ? tx : ty, u0);
        prob = along_x ? qx[k] : qy[k];
    } else {
        k = orc_sample_index(T, u0);
        prob = pw[k];
    }
    double px, py;
    orc_panel_centre(rx - h, ry - h, 2.0 * h, k, &px, &py);
    *weight = -eps * L * (nx * gx[k] + ny * gy[k]) / (prob * 2.0 * h);
    int64_t n = 0;
    orc_box sq;
    for (;;) {
        if (!(px > B.x0 && py > B.y0 && px < B.x1 && py < B.y1)) { *hops = n; return ORC_HIT_ESCAPED; }
        sq = orc_q3(ix, px, py, NULL);
        const int hit = orc_q2(ix, px, py);
        n++;
        if (hit >= 0) { *hops = n; return hit; }
        if (n >= max_hops) { *hops = n; return ORC_HIT_STUCK; }
        const uint32_t u = orc_walk_u32(seed, walk, job, (uint32_t)n);
        orc_point_from_index(&sq, orc_sample_index(T, u), ORC_NSAMPLE, &px, &py);
    }
}

int64_t orc_walk_weighted_many(const orc_index* ix, int master, uint64_t seed, uint32_t job,
                               int64_t first_walk, int64_t n_walks, int64_t max_hops, double eps,
                               int importance, int32_t* hit_out, int32_t* hops_out, double* w_out)
{
    uint32_t T[ORC_NSAMPLE];
    orc_green_thresholds(T);
    double intg[ORC_NSAMPLE + 1], gx[ORC_NSAMPLE], gy[ORC_NSAMPLE], pw[ORC_NSAMPLE];
    orc_green_cdf(intg);
    orc_green_grad(gx, gy);
    /* the probability each panel is drawn with (threshold differences) */
    for (int k = 0; k < ORC_NSAMPLE; ++k) {
        const double hi = k + 1 < ORC_NSAMPLE ? (double)T[k + 1] : 4294967296.0;
        pw[k] = (hi - (double)T[k]) / 4294967296.0;
    }
    uint32_t tx[ORC_NSAMPLE], ty[ORC_NSAMPLE];
    double qx[ORC_NSAMPLE], qy[ORC_NSAMPLE];
    orc_importance_table(gx, tx, qx);
    orc_importance_table(gy, ty, qy);
    const orc_box G = orc_gauss_from(ix, master);
    int64_t total = 0;
    for (int64_t i = 0; i < n_walks; ++i) {
        int64_t h = 0;
        double w = 0.0;
        const int hit = orc_walk_weighted_one(ix, &G, T, gx, gy, pw, importance ? tx : NULL, importance ? ty : NULL,
                                              importance ? qx : NULL, importance ? qy : NULL, seed, job,
                                              (uint64_t)(first_walk + i), max_hops, eps, &h, &w);
        if (hit_out) hit_out[i] = hit;
        if (hops_out) hops_out[i] = (int32_t)h;
        if (w_out) w_out[i] = w;
        total += h;
    }
    return total;
}
