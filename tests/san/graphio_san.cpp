// AddressSanitizer / UndefinedBehaviorSanitizer run of the host-side graph code of the library
// (nested-parallelism_amd/csrc/np_graph_io.cpp: the DIMACS9/DIMACS10/SLNDC readers, the binary
// CSR cache, the host generators, source picking, scheduler features and heuristic), compiled
// with g++ and no device code.  Reads every text fixture under tests/golden/graphs, round-trips
// each through the CSR cache, feeds the readers and the cache damaged and hostile inputs (which
// must be refused, not crash), and runs the generators.  Built and run by tests/san/Makefile.
//   graphio_san <fixture dir> <scratch dir>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/np_sssp.h"

namespace np {
void set_error(const char *fmt, ...) {  // np_api.cpp's, which this build does not link
    static char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
}
}  // namespace np

static int failures = 0;
static void check(bool ok, const char *what, const std::string &ctx) {
    if (!ok) { std::fprintf(stderr, "graphio_san: %s (%s)\n", what, ctx.c_str()); ++failures; }
}

static bool csr_ok(int32_t n, int64_t m, const int32_t *row, const int32_t *col) {
    if (row[0] != 0 || row[n] != m) return false;
    for (int32_t v = 0; v < n; ++v) if (row[v + 1] < row[v]) return false;
    for (int64_t e = 0; e < m; ++e) if (col[e] < 0 || col[e] >= n) return false;
    return true;
}

static void write_file(const std::string &path, const std::string &text) {
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (fp) { std::fwrite(text.data(), 1, text.size(), fp); std::fclose(fp); }
}

static void read_and_cache(int fmt, const std::string &path, const std::string &tmp) {
    int32_t n = 0, src = 0, *row = nullptr, *col = nullptr, *w = nullptr;
    int64_t m = 0;
    int rc = np_read_graph(fmt, path.c_str(), &n, &m, &src, &row, &col, &w);
    check(rc == NP_OK, "fixture read", path);
    if (rc != NP_OK) return;
    check(csr_ok(n, m, row, col), "fixture CSR malformed", path);
    const std::string bin = tmp + "/cache.npcsr";
    check(np_csr_save(bin.c_str(), n, m, src, row, col, w) == NP_OK, "cache save", path);
    int32_t n2 = 0, s2 = 0, *r2 = nullptr, *c2 = nullptr, *w2 = nullptr;
    int64_t m2 = 0;
    check(np_csr_load(bin.c_str(), &n2, &m2, &s2, &r2, &c2, &w2) == NP_OK, "cache load", path);
    if (r2) {
        check(n2 == n && m2 == m && s2 == src && !std::memcmp(r2, row, sizeof(int32_t) * ((size_t)n + 1)) &&
                  (m == 0 || (!std::memcmp(c2, col, sizeof(int32_t) * (size_t)m) &&
                              !std::memcmp(w2, w, sizeof(int32_t) * (size_t)m))),
              "cache round trip differs", path);
        int32_t n3 = 0, s3 = 0, *r3 = nullptr, *c3 = nullptr, *w3 = nullptr;
        int64_t m3 = 0;
        check(np_read_graph(3, bin.c_str(), &n3, &m3, &s3, &r3, &c3, &w3) == NP_OK && n3 == n && m3 == m,
              "format 3 read", path);
        np_host_free(r3); np_host_free(c3); np_host_free(w3);
        // truncated and bit-flipped caches are refused
        FILE *fp = std::fopen(bin.c_str(), "rb");
        std::vector<char> bytes;
        if (fp) {
            int ch;
            while ((ch = std::fgetc(fp)) != EOF) bytes.push_back((char)ch);
            std::fclose(fp);
        }
        for (size_t cut : {(size_t)3, (size_t)20, bytes.size() / 2, bytes.size() - 1}) {
            if (cut >= bytes.size()) continue;
            write_file(bin, std::string(bytes.data(), cut));
            int32_t *a = nullptr, *b = nullptr, *c = nullptr;
            check(np_csr_load(bin.c_str(), &n3, &m3, &s3, &a, &b, &c) == NP_ERR_IO, "truncated cache accepted", path);
            np_host_free(a); np_host_free(b); np_host_free(c);
        }
        if (m > 0 && bytes.size() > 40) {
            std::vector<char> bad = bytes;
            int32_t huge = 0x7FFFFFF0;  // first col entry out of range
            std::memcpy(bad.data() + 32 + 4 * ((size_t)n + 1), &huge, 4);
            write_file(bin, std::string(bad.data(), bad.size()));
            int32_t *a = nullptr, *b = nullptr, *c = nullptr;
            check(np_csr_load(bin.c_str(), &n3, &m3, &s3, &a, &b, &c) == NP_ERR_IO, "bad col accepted", path);
            np_host_free(a); np_host_free(b); np_host_free(c);
            int64_t neg = -5;  // negative vertex count in the header
            bad = bytes;
            std::memcpy(bad.data() + 8, &neg, 8);
            write_file(bin, std::string(bad.data(), bad.size()));
            check(np_csr_load(bin.c_str(), &n3, &m3, &s3, &a, &b, &c) == NP_ERR_IO, "negative n accepted", path);
            np_host_free(a); np_host_free(b); np_host_free(c);
        }
    }
    np_host_free(r2); np_host_free(c2); np_host_free(w2);
    np_host_free(row); np_host_free(col); np_host_free(w);
}

static void hostile_text(const std::string &tmp) {
    // inputs the readers must refuse (or read without touching memory they do not own)
    const char *cases[][2] = {
        {"0", "p sp 3 2\na 1 2 5\na 4 1 2\n"},            // endpoint past n
        {"0", "p sp 3 2\na 0 2 5\na 1 2 2\n"},            // endpoint 0 (1-based ids)
        {"0", "p sp -4 2\na 1 2 5\n"},                    // negative n
        {"0", "p sp 3 5\na 1 2 5\n"},                     // fewer arcs than announced
        {"0", "a 1 2 5\n"},                               // no problem line
        {"0", ""},                                        // empty file
        {"1", "3 2\n2\n1 3\n"},                           // DIMACS10: fewer rows than n
        {"1", "2 1\n2 9\n1\n"},                           // DIMACS10: neighbour past n
        {"1", "-1 0\n"},                                  // DIMACS10: negative n
        {"2", "# Nodes: 3 Edges: 2\n1 2\n"},              // SLNDC: fewer edges than announced
        {"2", "# Nodes: 2 Edges: 3\n1 2\n2 3\n3 1\n"},    // SLNDC: more ids than announced nodes
        {"2", "1 2\n2 1\n"},                              // SLNDC: no header
    };
    for (auto &c : cases) {
        const std::string path = tmp + "/hostile.txt";
        write_file(path, c[1]);
        int32_t n = 0, src = 0, *row = nullptr, *col = nullptr, *w = nullptr;
        int64_t m = 0;
        int rc = np_read_graph(std::atoi(c[0]), path.c_str(), &n, &m, &src, &row, &col, &w);
        if (rc == NP_OK) check(csr_ok(n, m, row, col), "hostile input read into a malformed CSR", c[1]);
        np_host_free(row); np_host_free(col); np_host_free(w);
    }
    int32_t n = 0, src = 0, *row = nullptr, *col = nullptr, *w = nullptr;
    int64_t m = 0;
    check(np_read_graph(0, (tmp + "/does-not-exist").c_str(), &n, &m, &src, &row, &col, &w) == NP_ERR_IO,
          "missing file", "");
    check(np_read_graph(9, (tmp + "/hostile.txt").c_str(), &n, &m, &src, &row, &col, &w) != NP_OK, "bad format", "");
}

static void generators() {
    int32_t *row = nullptr, *col = nullptr, *w = nullptr, n = 0;
    int64_t m = 0;
    check(np_gen_rmat_host(10, 16, 0.57, 0.19, 0.19, 1, 255, 3, 1, &row, &col, &w, &n, &m) == NP_OK &&
              csr_ok(n, m, row, col), "rmat", "");
    std::vector<int32_t> srcs(8);
    check(np_pick_sources(row, n, 8, 3, srcs.data()) == NP_OK, "pick sources", "");
    for (int32_t s : srcs) check(s >= 0 && s < n && row[s + 1] > row[s], "picked source has no out-edge", "");
    np_host_free(row); np_host_free(col); np_host_free(w);
    row = col = w = nullptr;
    check(np_gen_er_host(3000, 24000, 1, 100, 1, &row, &col, &w) == NP_OK && csr_ok(3000, 24000, row, col), "er", "");
    np_host_free(row); np_host_free(col); np_host_free(w);
    row = col = w = nullptr;
    check(np_gen_grid_host(33, 17, -1, 1, 1000, 5, &row, &col, &w, &m) == NP_ERR_ARG, "negative long-range count", "");
    check(np_gen_grid_host(33, 17, 20, 1, 1000, 5, &row, &col, &w, &m) == NP_OK && csr_ok(33 * 17, m, row, col),
          "grid", "");
    np_host_free(row); np_host_free(col); np_host_free(w);
    int32_t count[NP_NCLASS] = {0};
    for (int c = 0; c < NP_NCLASS; ++c) count[c] = c * 1000;
    float x[NP_MLP_NIN];
    check(np_features(count, 3, 1 << 20, x) == NP_OK, "features", "");
    np_sched s;
    check(np_heuristic(count, 3, 1 << 20, &s) == NP_OK, "heuristic", "");
    np_sched_default(1, &s);
}

int main(int argc, char **argv) {
    if (argc < 3) { std::fprintf(stderr, "usage: graphio_san <fixture dir> <scratch dir>\n"); return 2; }
    const std::string dir = argv[1], tmp = argv[2];
    read_and_cache(0, dir + "/d9_small.gr", tmp);
    read_and_cache(0, dir + "/d9_rand.gr", tmp);
    read_and_cache(0, dir + "/d9_path.gr", tmp);
    read_and_cache(1, dir + "/d10_small.graph", tmp);
    read_and_cache(2, dir + "/slndc_small.txt", tmp);
    hostile_text(tmp);
    generators();
    std::printf("graphio_san: %s\n", failures ? "FAILED" : "ok");
    return failures ? 1 : 0;
}
