/*
 * CPU oracle, part 1: RNG, Green CDF, geometry, layout reader.
 * Test infrastructure only (see oracle.h).  Built with -ffp-contract=off so
 * every FP64 expression rounds exactly as written, operation by operation.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) -------------------
 * Round: (c0,c1,c2,c3) -> (hi(M1*c2)^c1^k0, lo(M1*c2), hi(M0*c0)^c3^k1,
 * lo(M0*c0)); key bumped by the Weyl constants between rounds.  Replaces the
 * JKISS generator the reference re-seeds per walk
 * (randomWalk/CPU/mainFunctions.cpp:58-62). */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Walk stream: counter = (walk low word, job, hop/4, 'WALK' ^ walk high
 * word), key = seed; one 32-bit word per hop.  Walk ids below 2^32 leave the
 * domain word 'WALK'; ids up to 2^48 stay apart from the 'GAUS' start words,
 * whose high half differs. */
uint32_t orc_walk_u32(uint64_t seed, uint64_t walk, uint32_t job, uint32_t hop)
{
    uint32_t ctr[4] = { (uint32_t)walk, job, hop >> 2, 0x57414C4Bu ^ (uint32_t)(walk >> 32) };
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    uint32_t out[4];
    orc_philox(ctr, key, out);
    return out[hop & 3u];
}

/* ---- Green CDF -------------------------------------------------------------
 * Harmonic measure density on one side of the unit square seen from its
 * centre (separation of variables, side length a = 1):
 *     k(s) = sum_{n odd} sin(n*pi/2) * sin(n*pi*s) / cosh(n*pi/2)
 * truncated at n = 19 as the reference does; 50 midpoint panels per side, the
 * four sides concatenated clockwise, cumulative sum normalised to 1.  Same
 * behaviour as randomWalk/green/green.cpp:27-54,125-152 (precompute_unit_
 * square_green); re-derived, nothing taken from that file. */
static const double ORC_PI = 3.14159265358979323846;

static double green_side_density(double s)
{
    double k = 0.0;
    for (int n = 1; n <= 19; n += 2) {
        double sgn = ((n & 3) == 1) ? 1.0 : -1.0;
        k += sgn * sin(n * ORC_PI * s) / cosh(n * ORC_PI / 2.0);
    }
    return k;
}

void orc_green_cdf(double intg[ORC_NSAMPLE + 1])
{
    const int per_side = ORC_NSAMPLE / 4;
    double w[ORC_NSAMPLE / 4];
    for (int j = 0; j < per_side; ++j)
        w[j] = green_side_density((j + 0.5) / per_side) / per_side;
    intg[0] = 0.0;
    for (int i = 0; i < ORC_NSAMPLE; ++i) intg[i + 1] = intg[i] + w[i % per_side];
    double tot = intg[ORC_NSAMPLE];
    for (int i = 0; i <= ORC_NSAMPLE; ++i) intg[i] = intg[i] / tot;
}

/* T[i] = round-to-nearest-even(intg[i] * 2^32); an index is sampled from a
 * uniform u32 u as the largest i with T[i] <= u, i.e. intg[i] <= r < intg[i+1]
 * for r = u / 2^32 in [0,1) (randomWalk/CPU/mainFunctions.cpp:40-46, without
 * the out-of-range read at r = 1). */
void orc_green_thresholds(uint32_t T[ORC_NSAMPLE])
{
    double intg[ORC_NSAMPLE + 1];
    orc_green_cdf(intg);
    for (int i = 0; i < ORC_NSAMPLE; ++i)
        T[i] = (uint32_t)nearbyint(intg[i] * 4294967296.0);
}

int orc_sample_index(const uint32_t T[ORC_NSAMPLE], uint32_t u)
{
    int i = 0;
    while (i + 1 < ORC_NSAMPLE && T[i + 1] <= u) ++i;
    return i;
}

/* getIndex as written (linear scan over intg); -1 where the reference falls
 * off the end. */
int orc_get_index_fp(const double intg[ORC_NSAMPLE + 1], double r)
{
    for (int i = 0; i < ORC_NSAMPLE; ++i)
        if (intg[i] <= r && intg[i + 1] > r) return i;
    return -1;
}

/* ---- geometry (randomWalk/CPU/rectHost.cpp) ----------------------------- */

/* rectsCollision, rectHost.cpp:43-52: closed intervals, touching collides. */
int orc_collide(const orc_box* a, const orc_box* b)
{
    return a->x1 >= b->x0 && b->x1 >= a->x0 && a->y1 >= b->y0 && b->y1 >= a->y0;
}

/* rectContains(point), rectHost.cpp:54-63: closed. */
int orc_contains(const orc_box* r, double px, double py)
{
    return px >= r->x0 && py >= r->y0 && px <= r->x1 && py <= r->y1;
}

/* getPointFromNindex, rectHost.cpp:102-161, operation for operation:
 * integer-truncated perimeter, panel (index-1) offsets, the int cast on the
 * top side and the 1.5x corner-panel offsets are all kept. */
void orc_point_from_index(const orc_box* s, int index, int nsample, double* px, double* py)
{
    double perimeter = (double)(int)(2 * (s->x1 - s->x0) + 2 * (s->y1 - s->y0));
    double vector = perimeter / (double)nsample;
    double heigth = s->y1 - s->y0;
    double width = s->x1 - s->x0;
    double a = index * vector;          /* index*vector */
    double b = (index - 1) * vector;    /* (index-1)*vector */
    if (a < width) {
        *px = (double)(int)(b + vector / 2 + s->x0);
        *py = s->y0;
    } else if (b < width && a > width) {
        *px = s->x1;
        *py = s->y0;
    } else if (a < (width + heigth)) {
        *px = s->x1;
        if (b > width)
            *py = s->y0 + b - width + vector / 2;
        else
            *py = s->y0 + a - width + (a - width) / 2;
    } else if (b < (width + heigth) && a > (width + heigth)) {
        *px = s->x1;
        *py = s->y1;
    } else if (a < (2 * width + heigth)) {
        *py = s->y1;
        if (b > width + heigth)
            *px = s->x1 - (b - width - heigth + vector / 2);
        else
            *px = s->x1 - (a - width - heigth + (a - width - heigth) / 2);
    } else if (b < (2 * width + heigth) && a > (2 * width + heigth)) {
        *px = s->x0;
        *py = s->y1;
    } else {
        *px = s->x0;
        if (b > 2 * width + heigth)
            *py = s->y1 - (b - 2 * width - heigth + vector / 2);
        else
            *py = s->y1 - (a - 2 * width - heigth + (a - 2 * width - heigth) / 2);
    }
}

/* createGaussianSurface, rectHost.cpp:84-100. */
orc_box orc_gauss_surface(const orc_box* r, double fx, double fy)
{
    double middleX = (r->x0 + r->x1) / 2.;
    double middleY = (r->y0 + r->y1) / 2.;
    double vectorX = r->x0 - middleX;
    double vectorY = r->y0 - middleY;
    vectorX *= fx;
    vectorY *= fy;
    orc_box g;
    g.x0 = middleX + vectorX;
    g.y0 = middleY + vectorY;
    g.x1 = middleX - vectorX;
    g.y1 = middleY - vectorY;
    return g;
}

/* ---- layout reader (randomWalk/CPU/Parser.cpp:15-50, 89-125) --------------
 * Lines are read 49 characters at a time (getline into char[50]); a longer
 * line sets failbit and ends the read.  The first line containing "<<" is
 * skipped, every later one closes a layer; a line containing "rect" yields a
 * rect from the 4 numbers after its first token.  Rects after the last "<<"
 * are dropped.  Where the reference reads uninitialised or out-of-bounds
 * memory (fewer or more than 4 numbers) this restatement reports an error. */
#define ORC_MAX_LINE 50

static int parse_rect_line(const char* line, double c[4])
{
    const char* s = line;
    while (*s == ' ' || *s == '\t' || *s == '\r' || *s == '\v' || *s == '\f' || *s == '\n') ++s;
    while (*s && !(*s == ' ' || *s == '\t' || *s == '\r' || *s == '\v' || *s == '\f' || *s == '\n')) ++s;
    int k = 0;
    for (;;) {
        while (*s == ' ' || *s == '\t' || *s == '\r' || *s == '\v' || *s == '\f' || *s == '\n') ++s;
        if (!*s) break;
        if (!((*s >= '0' && *s <= '9') || *s == '-' || *s == '+' || *s == '.')) break;
        char* end = NULL;
        double v = strtod(s, &end);
        if (end == s) break;
        if (k >= 4) return -1;
        c[k++] = v;
        s = end;
    }
    return k == 4 ? 0 : -1;
}

int orc_parse(const char* path, int layer, double* xyxy, int cap, int* n_layers)
{
    FILE* f = fopen(path, "rb");
    if (!f) return -1;
    char line[ORC_MAX_LINE];
    int first = 1, cur_layer = 0, n_out = 0, cur_count = 0, err = 0;
    for (;;) {
        /* istream::getline(line, 50) */
        int len = 0, ch = 0, got_any = 0, fail = 0;
        for (;;) {
            ch = fgetc(f);
            if (ch == EOF) break;
            got_any = 1;
            if (ch == '\n') break;
            if (len == ORC_MAX_LINE - 1) { fail = 1; break; }
            line[len++] = (char)ch;
        }
        line[len] = 0;
        if (fail || (!got_any && ch == EOF)) break;
        if (strstr(line, "<<")) {
            if (first) { first = 0; continue; }
            if (cur_layer == layer) n_out = cur_count;
            cur_layer++;
            cur_count = 0;
            continue;
        }
        if (strstr(line, "rect")) {
            double c[4];
            if (parse_rect_line(line, c) != 0) { err = 1; break; }
            if (cur_layer == layer) {
                if (cur_count < cap) memcpy(xyxy + 4 * cur_count, c, sizeof(c));
            }
            cur_count++;
        }
        if (ch == EOF) break;
    }
    fclose(f);
    if (err) return -3;
    if (n_layers) *n_layers = cur_layer;
    if (layer < 0 || layer >= cur_layer) return -2;
    return n_out;
}

/* getLayerSize, Parser.cpp:51-87: bbox padded on all four sides by
 * int(0.01 * width) (the height pad is computed but the width pad is used). */
void orc_bounds(const double* r, int n, orc_box* out)
{
    double leftX = 0, rightX = 0, topY = 0, bottomY = 0;
    for (int i = 0; i < n; ++i) {
        const double* q = r + 4 * i;
        if (i == 0) { leftX = q[0]; topY = q[1]; rightX = q[2]; bottomY = q[3]; continue; }
        if (q[0] < leftX) leftX = q[0];
        if (q[1] < topY) topY = q[1];
        if (q[2] > rightX) rightX = q[2];
        if (q[3] > bottomY) bottomY = q[3];
    }
    int add_space_w = (int)((double)(rightX - leftX) * 0.01);
    out->x0 = leftX - add_space_w;
    out->y0 = topY - add_space_w;
    out->x1 = rightX + add_space_w;
    out->y1 = bottomY + add_space_w;
}
