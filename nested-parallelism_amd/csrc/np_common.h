// np_common.h -- definitions shared by host and device code of the SSSP path.
#pragma once
#include <stdint.h>
#include <math.h>

#if defined(__HIP__)
#define NP_HD __host__ __device__ inline
#else
#define NP_HD static inline
#endif

namespace np {

constexpr int INF = 1073741824;         // common/global.h:19
constexpr int UNDEFINED = 0x7FFFFFFF;   // apps/bfs-rec/src/gpu/bfs_rec.h:13
constexpr int NCLS = 32;                // out-degree classes
constexpr int GRID_CHUNK_SHIFT = 11;    // grid-scope work item = 2048 edges of one vertex
constexpr int LUT_SHIFT = 10;           // batch: instance lookup granularity (slots)

// Out-degree class: 0 for d == 0, else 1 + floor(log2 d).  A vertex of class c has
// 2^(c-1) <= d < 2^c, so "d >= 2^k" is exactly "class >= k + 1".
NP_HD int deg_class(int d) { return d <= 0 ? 0 : 32 - __builtin_clz((unsigned)d); }

// Grid-scope work items of one vertex of class c (chunks of 2^GRID_CHUNK_SHIFT edges).
NP_HD int grid_chunks(int c) { return c > GRID_CHUNK_SHIFT ? (1 << (c - GRID_CHUNK_SHIFT)) : 1; }

// Bin boundaries of one expansion, as class numbers:
//   lane bin [1, b1), wavefront bin [b1, b2), workgroup bin [b2, b3), grid bin [b3, 32).
struct Cfg { int level, klo, khi, b1, b2, b3; };

// Schedule (level, log2 thresholds) -> class boundaries.  The "mid" bin between the
// thresholds is always wavefront-scope; the level names the scope of the top bin.
NP_HD Cfg make_cfg(int level, int klo, int khi) {
    Cfg c;
    if (level < 0 || level > 2) level = 1;
    klo = klo < 0 ? 0 : (klo > 30 ? 30 : klo);
    khi = khi < 0 ? 0 : (khi > 30 ? 30 : khi);
    c.level = level; c.klo = klo; c.khi = khi;
    c.b1 = klo + 1;
    int bhi = khi + 1 > c.b1 ? khi + 1 : c.b1;
    if (level == 0) { c.b2 = NCLS; c.b3 = NCLS; }
    else if (level == 1) { c.b2 = bhi; c.b3 = NCLS; }
    else { c.b2 = bhi; c.b3 = bhi; }
    return c;
}

// ---------------------------------------------------------------------------------
// Counter-based random numbers: SplitMix64 (G. Steele, D. Lea, C. Flood, "Fast
// splittable pseudorandom number generators", OOPSLA 2014) finaliser over a keyed
// counter.  draw(key, i) is the i-th output of SplitMix64 seeded with key.
// ---------------------------------------------------------------------------------
constexpr uint64_t GAMMA = 0x9E3779B97F4A7C15ull;
NP_HD uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
NP_HD uint64_t stream_key(uint64_t seed, uint64_t stream) { return mix64(seed * GAMMA + stream + 1); }
NP_HD uint64_t draw(uint64_t key, uint64_t i) { return mix64(key + (i + 1) * GAMMA); }
// Uniform integer in [0, range) by multiply-high of a 32-bit draw.
NP_HD uint32_t uniform_below(uint32_t r32, uint32_t range) {
    return (uint32_t)(((uint64_t)r32 * range) >> 32);
}

// Bijective keyed mixing of [0, 2^bits): rounds of odd multiplication, addition and
// xor-shift, each a bijection modulo 2^bits.  Used to permute R-MAT vertex ids.
NP_HD uint32_t permute_id(uint32_t x, int bits, uint64_t key) {
    const uint32_t mask = bits >= 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    const int sh = bits / 2 + 1;
    for (int r = 0; r < 4; ++r) {
        uint64_t k = draw(key, (uint64_t)r);
        x = (x * ((uint32_t)k | 1u)) & mask;
        x = (x + (uint32_t)(k >> 32)) & mask;
        if (sh < bits) x ^= x >> sh;
        x &= mask;
    }
    return x;
}

// R-MAT edge e (Chakrabarti, Zhan, Faloutsos, SDM 2004): at each of `scale` levels one
// quadrant is chosen with probabilities a (0,0), b (0,1), c (1,0), d (1,1) as
// (src bit, dst bit).  ta/tab/tabc are a, a+b, a+b+c scaled to 2^32.
NP_HD void rmat_edge(uint64_t key, uint64_t e, int scale, uint64_t ta, uint64_t tab, uint64_t tabc,
                     uint32_t *src, uint32_t *dst) {
    uint32_t s = 0, t = 0;
    const uint64_t per = (uint64_t)(scale + 1) / 2;
    uint64_t r = 0;
    for (int l = 0; l < scale; ++l) {
        if ((l & 1) == 0) r = draw(key, e * per + (uint64_t)(l >> 1));
        uint64_t u = (l & 1) ? (r >> 32) : (r & 0xFFFFFFFFull);
        uint32_t sb, tb;
        if (u < ta) { sb = 0; tb = 0; }
        else if (u < tab) { sb = 0; tb = 1; }
        else if (u < tabc) { sb = 1; tb = 0; }
        else { sb = 1; tb = 1; }
        s = (s << 1) | sb;
        t = (t << 1) | tb;
    }
    *src = s; *dst = t;
}

// Stream ids of the generators.
enum : uint64_t { ST_RMAT = 1, ST_WEIGHT = 2, ST_PERM = 3, ST_ER = 4, ST_LONG = 5, ST_SRC = 6 };

// Erdos-Renyi edge e: uniform endpoints, a self-loop is redrawn (attempt counter), so
// exactly m edges remain.
NP_HD void er_edge(uint64_t key, uint64_t e, uint32_t n, uint32_t *src, uint32_t *dst) {
    for (uint64_t a = 0;; ++a) {
        uint64_t r = draw(key, e * 64 + a);
        uint32_t s = uniform_below((uint32_t)r, n), t = uniform_below((uint32_t)(r >> 32), n);
        if (s != t || n < 2 || a == 63) { *src = s; *dst = t; return; }
    }
}

// 2-D grid road-like graph: vertex r*cols+c, 4-neighbour bidirectional edges in the
// order right, left, down, up per vertex; then n_long random directed long-range edges
// (self-loops redrawn).  grid_deg is the grid out-degree of vertex v.
NP_HD int grid_deg(uint32_t v, int rows, int cols) {
    int r = (int)(v / (uint32_t)cols), c = (int)(v % (uint32_t)cols);
    return (c + 1 < cols) + (c > 0) + (r + 1 < rows) + (r > 0);
}
NP_HD uint32_t grid_nbr(uint32_t v, int rows, int cols, int k) {
    int r = (int)(v / (uint32_t)cols), c = (int)(v % (uint32_t)cols);
    int idx = 0;
    if (c + 1 < cols) { if (idx++ == k) return v + 1; }
    if (c > 0) { if (idx++ == k) return v - 1; }
    if (r + 1 < rows) { if (idx++ == k) return v + (uint32_t)cols; }
    if (r > 0) { if (idx++ == k) return v - (uint32_t)cols; }
    return v;
}

// Scheduler features of one frontier (class counts) at iteration `it` of an n-vertex
// graph: [log2(1+frontier)/24, log2(1+edges pending)/32, six class-group fractions,
// min(it,256)/64, log2(n)/32].  Edges pending are estimated from the classes
// (a class-c vertex has on average 0.75 * 2^c out-edges).
// The group sums are six scalars updated by selects (adding 0 leaves a sum unchanged), not an
// array indexed by the group, which a GPU thread would keep in scratch memory.
NP_HD void features(const int *count, int it, int n, float *x) {
    float tot = 0.f, edges = 0.f;
    float g0 = 0.f, g1 = 0.f, g2 = 0.f, g3 = 0.f, g4 = 0.f, g5 = 0.f;
    for (int c = 1; c < NCLS; ++c) {
        float k = (float)count[c];
        tot += k;
        edges += k * 0.75f * (float)(1u << (c < 31 ? c : 31));
        g0 += c <= 2 ? k : 0.f;
        g1 += c > 2 && c <= 4 ? k : 0.f;
        g2 += c > 4 && c <= 6 ? k : 0.f;
        g3 += c > 6 && c <= 8 ? k : 0.f;
        g4 += c > 8 && c <= 12 ? k : 0.f;
        g5 += c > 12 ? k : 0.f;
    }
    x[0] = log2f(1.f + tot) / 24.f;
    x[1] = log2f(1.f + edges) / 32.f;
    x[2] = tot > 0.f ? g0 / tot : 0.f;
    x[3] = tot > 0.f ? g1 / tot : 0.f;
    x[4] = tot > 0.f ? g2 / tot : 0.f;
    x[5] = tot > 0.f ? g3 / tot : 0.f;
    x[6] = tot > 0.f ? g4 / tot : 0.f;
    x[7] = tot > 0.f ? g5 / tot : 0.f;
    x[8] = (float)(it < 256 ? it : 256) / 64.f;
    x[9] = log2f((float)(n > 1 ? n : 1)) / 32.f;
}

}  // namespace np
