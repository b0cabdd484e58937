// Host-side layout handling: the reference layout-file reader, padded bounds,
// synthetic layout generators, the Green CDF table and the result writer.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "frw_internal.h"

namespace frw {

// Padded absorbing bounds, Parser::getLayerSize (randomWalk/CPU/Parser.cpp:51-87):
// the bbox grown on all four sides by int(0.01 * width).
int4 padded_bounds(const std::vector<int4>& r)
{
    if (r.empty()) return make_int4(0, 0, 0, 0);
    double lx = r[0].x, ty = r[0].y, rx = r[0].z, by = r[0].w;
    for (size_t i = 1; i < r.size(); ++i) {
        lx = std::min(lx, (double)r[i].x);
        ty = std::min(ty, (double)r[i].y);
        rx = std::max(rx, (double)r[i].z);
        by = std::max(by, (double)r[i].w);
    }
    const int pad = (int)((double)(rx - lx) * 0.01);
    return make_int4((int)lx - pad, (int)ty - pad, (int)rx + pad, (int)by + pad);
}

// ---------------------------------------------------------------- reader
// Same line rules as randomWalk/CPU/Parser.cpp:15-50,89-125: lines are read
// into a 50-byte buffer (a longer line ends the read), the first line holding
// "<<" is skipped and each later one closes a layer, a line holding "rect"
// gives a rect from the four numbers after its first token.
namespace {

constexpr int kMaxLine = 50;

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f'; }

int parse_rect(const char* s, int4& out)
{
    while (is_space(*s)) ++s;
    while (*s && !is_space(*s)) ++s;  // header token
    double c[4];
    int k = 0;
    for (;;) {
        while (is_space(*s)) ++s;
        if (!*s) break;
        if (!((*s >= '0' && *s <= '9') || *s == '-' || *s == '+' || *s == '.')) break;
        char* end = nullptr;
        const double v = std::strtod(s, &end);
        if (end == s) break;
        if (k >= 4) return FRW_E_PARSE;
        c[k++] = v;
        s = end;
    }
    if (k != 4) return FRW_E_PARSE;
    for (int i = 0; i < 4; ++i)
        if (c[i] != std::floor(c[i]) || std::fabs(c[i]) > 1.0e9) return FRW_E_PARSE;
    out = make_int4((int)c[0], (int)c[1], (int)c[2], (int)c[3]);
    return FRW_OK;
}

}  // namespace

namespace {

// What one line of the file means to the reader.
struct LineEvent {
    int32_t kind;    // 0 rect, 1 layer delimiter, 2 line too long (ends the read), 3 malformed rect
    int64_t line;    // line number within the chunk (1-based)
    int4 r;
};

// Lines of buf[beg, end) (beg at a line start; end at a line start or the
// end of the file), in order.
void scan_chunk(const char* buf, size_t beg, size_t end, std::vector<LineEvent>* ev, int64_t* n_lines)
{
    char line[kMaxLine];
    int64_t ln = 0;
    size_t p = beg;
    while (p < end) {
        size_t q = p;
        while (q < end && buf[q] != '\n') ++q;
        const size_t len = q - p;
        ++ln;
        if (len >= (size_t)kMaxLine) {  // the 50-byte getline buffer overflows: the read ends
            ev->push_back(LineEvent{2, ln, make_int4(0, 0, 0, 0)});
            break;
        }
        std::memcpy(line, buf + p, len);
        line[len] = 0;
        if (std::strstr(line, "<<")) {
            ev->push_back(LineEvent{1, ln, make_int4(0, 0, 0, 0)});
        } else if (std::strstr(line, "rect")) {
            LineEvent e{0, ln, make_int4(0, 0, 0, 0)};
            if (parse_rect(line, e.r) != FRW_OK) e.kind = 3;
            ev->push_back(e);
        }
        p = q + 1;
    }
    *n_lines = ln;
}

}  // namespace

// The file is read whole and its lines classified in parallel chunks (split at
// line starts), then the events are applied in file order, so the result is
// the sequential reader's: same layers, same rects, same stopping line.
int read_layout(const char* path, int layer, std::vector<int4>* rects, int* n_layers)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return set_error(FRW_E_IO, std::string("cannot open layout file ") + path);
    std::vector<char> buf;
    {
        char tmp[1 << 16];
        size_t got;
        while ((got = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + got);
        const bool bad = std::ferror(f) != 0;
        std::fclose(f);
        if (bad) return set_error(FRW_E_IO, std::string("cannot read layout file ") + path);
    }
    const size_t n = buf.size();
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int T = (int)std::max<size_t>(1, std::min<size_t>({(size_t)std::min(hw, 16u), n >> 18}));
    std::vector<size_t> cut(T + 1, n);
    cut[0] = 0;
    for (int t = 1; t < T; ++t) {
        size_t c = std::max(cut[t - 1], n * (size_t)t / (size_t)T);
        while (c < n && c > 0 && buf[c - 1] != '\n') ++c;  // start of a line
        cut[t] = c;
    }
    std::vector<std::vector<LineEvent>> ev(T);
    std::vector<int64_t> lines(T, 0);
    if (T == 1) {
        scan_chunk(buf.data(), 0, n, &ev[0], &lines[0]);
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back(scan_chunk, buf.data(), cut[t], cut[t + 1], &ev[t], &lines[t]);
        for (std::thread& x : th) x.join();
    }
    bool first = true, stop = false;
    int cur = 0;
    int64_t base = 0;
    std::vector<int4> acc;
    for (int t = 0; t < T && !stop; ++t) {
        for (const LineEvent& e : ev[t]) {
            if (e.kind == 2) {
                stop = true;
                break;
            }
            if (e.kind == 3)
                return set_error(FRW_E_PARSE, "malformed rect line " + std::to_string(base + e.line) + " in " + path);
            if (e.kind == 1) {
                if (first) {
                    first = false;
                    continue;
                }
                if (cur == layer && rects) *rects = acc;
                acc.clear();
                ++cur;
                continue;
            }
            if (cur == layer) acc.push_back(e.r);
        }
        base += lines[t];
    }
    if (n_layers) *n_layers = cur;
    if (layer < 0 || layer >= cur)
        return set_error(FRW_E_LAYER, "layer " + std::to_string(layer) + " not in " + path);
    return FRW_OK;
}

int validate_rects(const std::vector<int4>& r)
{
    for (size_t i = 0; i < r.size(); ++i) {
        if (r[i].x > r[i].z || r[i].y > r[i].w)
            return set_error(FRW_E_PARSE, "rect " + std::to_string(i) + " has x0 > x1 or y0 > y1");
        const int lim = 1 << 28;
        if (std::abs(r[i].x) > lim || std::abs(r[i].y) > lim || std::abs(r[i].z) > lim || std::abs(r[i].w) > lim)
            return set_error(FRW_E_RANGE, "rect " + std::to_string(i) + " has a coordinate beyond 2^28");
    }
    return FRW_OK;
}

// ---------------------------------------------------------------- Green CDF
// Harmonic measure on a side of the unit square from its centre:
//   k(s) = sum_{n odd <= 19} sin(n pi / 2) sin(n pi s) / cosh(n pi / 2)
// at 50 midpoint panels per side, four sides clockwise, cumulated and
// normalised; thresholds T[i] = round-half-even(intg[i] * 2^32).
void green_thresholds(uint32_t T[kNSample])
{
    const double pi = 3.14159265358979323846;
    const int per_side = kNSample / 4;
    double w[kNSample / 4];
    for (int j = 0; j < per_side; ++j) {
        const double s = (j + 0.5) / per_side;
        double k = 0.0;
        for (int n = 1; n <= 19; n += 2) {
            const double sgn = ((n & 3) == 1) ? 1.0 : -1.0;
            k += sgn * std::sin(n * pi * s) / std::cosh(n * pi / 2.0);
        }
        w[j] = k / per_side;
    }
    double intg[kNSample + 1];
    intg[0] = 0.0;
    for (int i = 0; i < kNSample; ++i) intg[i + 1] = intg[i] + w[i % per_side];
    const double tot = intg[kNSample];
    for (int i = 0; i < kNSample; ++i) T[i] = (uint32_t)std::nearbyint((intg[i] / tot) * 4294967296.0);
}

// First-hop weight tables of the weighted estimator: the gradient of each
// panel's exit probability with respect to the source, for the unit square
// with the source at its centre (y down, panels clockwise from the top-left
// corner), and each panel's draw probability under the thresholds T.
// Exit density of a side for a source at u along it and v from it:
//   G(s; u, v) = sum_n 2 sin(n pi s) sin(n pi u) sinh(n pi (1 - v)) / sinh(n pi),
// whose centre derivatives are
//   dG/du = sum_{n even} n pi cos(n pi / 2) sin(n pi s) / cosh(n pi / 2),
//   dG/dv = -sum_{n odd} n pi sin(n pi / 2) sin(n pi s) / sinh(n pi / 2);
// (u, v) per side: top (x, y), right (y, 1 - x), bottom (1 - x, 1 - y),
// left (1 - y, x).  Panel values are midpoint densities / (50 T), T the
// normalisation of the panel probabilities.
// First-hop importance table: panel k drawn with probability q_k proportional
// to |g[k]| (one component of the Green gradient), as integer thresholds over
// u32 draws; q_k is the exact draw probability (threshold difference).
static void importance_table(const double* g, uint32_t* T, double* q)
{
    double S = 0.0;
    for (int k = 0; k < kNSample; ++k) S += std::fabs(g[k]);
    double cum = 0.0;
    for (int k = 0; k < kNSample; ++k) {
        T[k] = (uint32_t)std::floor(cum / S * 4294967296.0);
        cum += std::fabs(g[k]);
    }
    for (int k = 0; k < kNSample; ++k) {
        const double hi = k + 1 < kNSample ? (double)T[k + 1] : 4294967296.0;
        q[k] = (hi - (double)T[k]) / 4294967296.0;
    }
}

void green_grad_tables(double* wt, uint32_t* ithr)
{
    double* gx = wt;
    double* gy = wt + kNSample;
    double* pw = wt + 2 * kNSample;
    const double pi = 3.14159265358979323846;
    const int per = kNSample / 4;
    double raw = 0.0;
    for (int j = 0; j < per; ++j) {
        const double s = (j + 0.5) / per;
        double k = 0.0;
        for (int n = 1; n <= 19; n += 2) k += (((n & 3) == 1) ? 1.0 : -1.0) * std::sin(n * pi * s) / std::cosh(n * pi / 2.0);
        raw += k / per;
    }
    const double T = 4.0 * raw;
    for (int k = 0; k < kNSample; ++k) {
        const double s = ((k % per) + 0.5) / per;
        double a = 0.0, b = 0.0;
        for (int n = 2; n <= 40; n += 2)
            a += n * pi * ((n % 4 == 0) ? 1.0 : -1.0) * std::sin(n * pi * s) / std::cosh(n * pi / 2.0);
        for (int n = 1; n <= 39; n += 2)
            b -= n * pi * (((n & 3) == 1) ? 1.0 : -1.0) * std::sin(n * pi * s) / std::sinh(n * pi / 2.0);
        a /= per * T;
        b /= per * T;
        switch (k / per) {
        case 0: gx[k] = a;  gy[k] = b;  break;
        case 1: gx[k] = -b; gy[k] = a;  break;
        case 2: gx[k] = -a; gy[k] = -b; break;
        default: gx[k] = b; gy[k] = -a; break;
        }
    }
    uint32_t th[kNSample];
    green_thresholds(th);
    for (int k = 0; k < kNSample; ++k) {
        const double hi = k + 1 < kNSample ? (double)th[k + 1] : 4294967296.0;
        pw[k] = (hi - (double)th[k]) / 4294967296.0;
    }
    importance_table(gx, ithr, wt + 3 * kNSample);
    importance_table(gy, ithr + kNSample, wt + 4 * kNSample);
}

// ---------------------------------------------------------------- generator
namespace {

constexpr uint32_t kGenDomain = 0x47454E52u;  // 'GENR'

void philox_host(uint32_t c[4], uint64_t seed)
{
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// floor(u * range / 2^32): unbiased enough for layout synthesis, exact and
// identical on every host.
inline int64_t scale_u32(uint32_t u, int64_t range) { return (int64_t)(((unsigned __int128)u * (uint64_t)range) >> 32); }

struct Grid {
    int64_t cell;
    std::unordered_map<uint64_t, std::vector<int32_t>> m;
    static uint64_t key(int64_t cx, int64_t cy) { return ((uint64_t)(uint32_t)cx << 32) | (uint32_t)cy; }
    bool hits(const std::vector<int4>& rs, const int4& r) const
    {
        for (int64_t cx = r.x / cell; cx <= r.z / cell; ++cx)
            for (int64_t cy = r.y / cell; cy <= r.w / cell; ++cy) {
                auto it = m.find(key(cx, cy));
                if (it == m.end()) continue;
                for (int32_t j : it->second) {
                    const int4& q = rs[j];
                    if (q.z >= r.x && r.z >= q.x && q.w >= r.y && r.w >= q.y) return true;
                }
            }
        return false;
    }
    void add(const int4& r, int32_t id)
    {
        for (int64_t cx = r.x / cell; cx <= r.z / cell; ++cx)
            for (int64_t cy = r.y / cell; cy <= r.w / cell; ++cy) m[key(cx, cy)].push_back(id);
    }
};

}  // namespace

int generate_layout(const frw_gen_params* p, std::vector<int4>* out)
{
    if (p->n < 0 || p->box <= 0 || p->side_min < 1 || p->side_max < p->side_min || p->side_max >= p->box)
        return set_error(FRW_E_INVALID, "bad generator parameters");
    if (p->kind == FRW_GEN_CLUSTERED && (p->n_clusters <= 0 || p->sigma <= 0))
        return set_error(FRW_E_INVALID, "clustered layout needs n_clusters > 0 and sigma > 0");
    std::vector<int4>& rs = *out;
    rs.clear();
    rs.reserve((size_t)p->n);
    Grid g;
    g.cell = std::max<int64_t>(p->side_max + 1, 1);
    std::vector<double> cx, cy;
    if (p->kind == FRW_GEN_CLUSTERED) {
        for (int c = 0; c < p->n_clusters; ++c) {
            uint32_t w[4] = {(uint32_t)c, 0u, 2u, kGenDomain};
            philox_host(w, p->seed);
            cx.push_back((double)scale_u32(w[0], (int64_t)p->box + 1));
            cy.push_back((double)scale_u32(w[1], (int64_t)p->box + 1));
        }
    }
    const double frac = p->kind == FRW_GEN_CLUSTERED ? p->cluster_frac : 0.0;
    const int64_t span = (int64_t)p->side_max - p->side_min + 1;
    const uint64_t max_attempts = (uint64_t)std::max<int64_t>(p->n, 1) * 1000ull;
    uint64_t attempt = 0;
    for (int64_t j = 0; j < p->n; ++j) {
        const bool clustered = std::floor((j + 1) * frac) > std::floor(j * frac);
        for (;;) {
            if (attempt >= max_attempts)
                return set_error(FRW_E_RANGE, "layout too dense: rejection sampling did not finish");
            uint32_t w[4] = {(uint32_t)attempt, (uint32_t)(attempt >> 32), 0u, kGenDomain};
            philox_host(w, p->seed);
            const int64_t wd = p->side_min + scale_u32(w[0], span);
            const int64_t ht = p->side_min + scale_u32(w[1], span);
            int64_t x0, y0;
            if (clustered) {
                uint32_t z[4] = {(uint32_t)attempt, (uint32_t)(attempt >> 32), 1u, kGenDomain};
                philox_host(z, p->seed);
                const int c = (int)(w[2] % (uint32_t)p->n_clusters);
                // Box-Muller on two uniforms in (0, 1]
                const double u1 = ((double)z[0] + 1.0) / 4294967296.0;
                const double u2 = ((double)z[1]) / 4294967296.0;
                const double rad = std::sqrt(-2.0 * std::log(u1)) * p->sigma;
                const double th = 2.0 * 3.14159265358979323846 * u2;
                const double gx = cx[c] + rad * std::cos(th) - wd / 2.0;
                const double gy = cy[c] + rad * std::sin(th) - ht / 2.0;
                x0 = std::min<int64_t>(std::max<int64_t>((int64_t)std::llround(gx), 0), p->box - wd);
                y0 = std::min<int64_t>(std::max<int64_t>((int64_t)std::llround(gy), 0), p->box - ht);
            } else {
                x0 = scale_u32(w[2], (int64_t)p->box - wd + 1);
                y0 = scale_u32(w[3], (int64_t)p->box - ht + 1);
            }
            ++attempt;
            const int4 r = make_int4((int)x0, (int)y0, (int)(x0 + wd), (int)(y0 + ht));
            if (g.hits(rs, r)) continue;
            g.add(r, (int32_t)rs.size());
            rs.push_back(r);
            break;
        }
    }
    return FRW_OK;
}

int save_layout(const std::vector<int4>& rs, const char* path)
{
    FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(FRW_E_IO, std::string("cannot write ") + path);
    std::fprintf(f, "magic\ntech frw\ntimestamp 0\n<< metal3 >>\n");
    for (const int4& r : rs) std::fprintf(f, "rect %d %d %d %d\n", r.x, r.y, r.z, r.w);
    std::fprintf(f, "<< metal3 >>\n");
    std::fclose(f);
    return FRW_OK;
}

// Result file in the reference format (randomWalk/CPU/getAvgPathLen.cpp:13-33):
// one "x0 y0 x1 y1: P%" line per rect that was hit, and the escaped share.
// Percentages are float(count)/float(N)*100 printed like std::ostream (%g).
int write_results(const std::vector<int4>& rs, const int64_t* counts, int64_t n_walks,
                  const char* path)
{
    FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(FRW_E_IO, std::string("cannot write ") + path);
    std::fprintf(f, "[GPU]\n");
    const int64_t n = (int64_t)rs.size();
    if (counts[n] > 0)
        std::fprintf(f, "Ile scieżek wyszło poza granice warstwy: %g%%\n",
                     (double)((float)counts[n] / (float)n_walks * 100));
    for (int64_t i = 0; i < n; ++i) {
        if (counts[i] == 0) continue;
        std::fprintf(f, "%g %g %g %g: %g%%\n", (double)rs[i].x, (double)rs[i].y, (double)rs[i].z,
                     (double)rs[i].w, (double)(((float)counts[i] / (float)n_walks) * 100));
    }
    std::fclose(f);
    return FRW_OK;
}

}  // namespace frw
