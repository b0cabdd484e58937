/* Sanitizer driver of the CPU oracle (test infrastructure): built with
 * AddressSanitizer and UndefinedBehaviorSanitizer by `make -C oracle san`
 * and run by the CPU test-suite.  Exercises every oracle entry point on a
 * layout file: parser and bounds, the three indices (brute force, the
 * reference tree, the reference tree with the full-depth rect query), the
 * queries, Gaussian surfaces, plain and weighted walks, the proposal
 * network's features, probe, forward and backward passes.  Exit status 0
 * and no sanitizer report is the pass condition.
 *   san_driver <layout> <master> <walks> */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

int main(int argc, char** argv)
{
    if (argc < 4) return 2;
    const int master = atoi(argv[2]);
    const int n_walks = atoi(argv[3]);
    int n_layers = 0;
    const int cap = 20000;
    double* xyxy = (double*)malloc(sizeof(double) * 4 * cap);
    const int n = orc_parse(argv[1], 0, xyxy, cap, &n_layers);
    if (n <= 0 || master >= n) return 3;
    orc_box b;
    orc_bounds(xyxy, n, &b);
    uint32_t T[ORC_NSAMPLE];
    orc_green_thresholds(T);
    double intg[ORC_NSAMPLE + 1];
    orc_green_cdf(intg);
    long long check = 0;
    for (int i = 0; i < 4096; ++i) {
        const uint32_t u = orc_walk_u32(7, (uint32_t)i, 0, (uint32_t)(i % 13));
        check += orc_sample_index(T, u) + orc_get_index_fp(intg, u / 4294967296.0);
    }
    const int kinds[3] = {ORC_EXACT, ORC_REFTREE, ORC_REFTREE_DFS};
    for (int k = 0; k < 3; ++k) {
        orc_index* ix = orc_index_new(xyxy, n, &b, kinds[k]);
        int64_t st[3];
        orc_index_stats(ix, st);
        const int m = 257;
        double* pts = (double*)malloc(sizeof(double) * 2 * m);
        double* boxes = (double*)malloc(sizeof(double) * 4 * m);
        uint8_t* q1 = (uint8_t*)malloc(m);
        int32_t* q2 = (int32_t*)malloc(sizeof(int32_t) * m);
        double* q3 = (double*)malloc(sizeof(double) * 4 * m);
        for (int i = 0; i < m; ++i) {
            const double fx = (double)orc_walk_u32(3, (uint32_t)i, 1, 0) / 4294967296.0;
            const double fy = (double)orc_walk_u32(3, (uint32_t)i, 1, 1) / 4294967296.0;
            pts[2 * i] = b.x0 + fx * (b.x1 - b.x0);
            pts[2 * i + 1] = b.y0 + fy * (b.y1 - b.y0);
            const double h = 1.0 + (i % 7) * 10.0;
            boxes[4 * i] = pts[2 * i] - h;
            boxes[4 * i + 1] = pts[2 * i + 1] - h;
            boxes[4 * i + 2] = pts[2 * i] + h;
            boxes[4 * i + 3] = pts[2 * i + 1] + h;
            int arg = -1, nq = 0;
            check += (long long)orc_nearest(ix, pts[2 * i], pts[2 * i + 1], &arg) + arg;
            orc_box sq = orc_q3(ix, pts[2 * i], pts[2 * i + 1], &nq);
            check += nq + (long long)sq.x0;
        }
        orc_queries_many(ix, pts, boxes, m, m, q1, q2, q3, 1);
        for (int i = 0; i < m; ++i) check += q1[i] + q2[i];
        orc_box g = orc_gauss_from(ix, master);
        int64_t hops = 0;
        check += orc_walk_one(ix, &g, T, 1, 0, 5, 1000000, &hops) + hops;
        int64_t* counts = (int64_t*)calloc((size_t)n + 2, sizeof(int64_t));
        int32_t* hit = (int32_t*)malloc(sizeof(int32_t) * n_walks);
        int32_t* hp = (int32_t*)malloc(sizeof(int32_t) * n_walks);
        check += orc_walk_many(ix, master, 1, 0, 0, n_walks, 1000000, counts, hit, hp, 1);
        double* w = (double*)malloc(sizeof(double) * n_walks);
        if (k != ORC_REFTREE) {
            check += orc_walk_weighted_many(ix, master, 1, 0, 0, n_walks / 4 + 1, 1000000, 1.0, 0, hit, hp, w);
            check += orc_walk_weighted_many(ix, master, 1, 0, 0, n_walks / 4 + 1, 1000000, 1.0, 1, hit, hp, w);
        }
        free(w);
        free(hit);
        free(hp);
        free(counts);
        free(pts);
        free(boxes);
        free(q1);
        free(q2);
        free(q3);
        orc_index_free(ix);
    }
    /* proposal network: features, probe, forward, backward */
    const int np = orc_mlp_param_count();
    float* params = (float*)malloc(sizeof(float) * np);
    for (int i = 0; i < np; ++i) params[i] = (float)((int)(orc_walk_u32(9, (uint32_t)i, 2, 0) % 2001) - 1000) / 5000.0f;
    const int B = 64;
    float* X = (float*)calloc((size_t)B * 72, sizeof(float));
    float* t = (float*)malloc(sizeof(float) * B);
    float* y = (float*)malloc(sizeof(float) * B);
    float* grad = (float*)malloc(sizeof(float) * np);
    for (int i = 0; i < B; ++i) {
        const double px = b.x0 + (b.x1 - b.x0) * (i + 0.5) / B, py = 0.5 * (b.y0 + b.y1);
        orc_mlp_features(xyxy, n, &b, px, py, 1.0 + 30.0 * i, X + 72 * i);
        check += orc_mlp_certify(xyxy, n, &b, px, py, 3.0 + i);
        t[i] = (float)(i % 5) - 2.0f;
    }
    check += orc_mlp_vmax(&b, n) + orc_mlp_vmax_capped(&b, n, 12);
    orc_mlp_forward(params, X, B, y);
    const float loss = orc_mlp_backward(params, X, t, B, grad);
    printf("ok %lld %.6g\n", check, (double)loss);
    free(params);
    free(X);
    free(t);
    free(y);
    free(grad);
    free(xyxy);
    return 0;
}
