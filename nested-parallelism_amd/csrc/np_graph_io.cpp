// np_graph_io.cpp -- graph files and host-side generators (no device code).
//
// Readers follow common/global.cpp:29-190 token for token (the same fgets/strtok/atoi
// parsing, the same 0/1-based id handling, DIMACS10/SLNDC weights rand()%1000+1 drawn
// after srand(7) in the same order) and produce directly the CSR that convertCSR
// (common/global.cpp:192-235) builds from the adjacency lists: vertex i's out-edges in
// the reverse of their order in the file, row[n] = edge count.  The std::list stage is
// replaced by one edge array and a counting sort.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <string>

#include "np_common.h"
#include "../../include/np_sssp.h"

namespace np {
void set_error(const char *fmt, ...);
}

namespace {

constexpr size_t BUFF_SIZE = 1024000;  // common/global.h:15

struct EdgeList {
    std::vector<int> tail, head, w;
};

// Counting sort by tail; within a tail, edges in reverse input order (convertCSR pops
// from the back of each adjacency list, common/global.cpp:216-221).
int to_csr(int n, const EdgeList &el, int32_t **row, int32_t **col, int32_t **w) {
    const size_t m = el.tail.size();
    int32_t *r = (int32_t *)std::malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t *c = (int32_t *)std::malloc(sizeof(int32_t) * (m ? m : 1));
    int32_t *ww = (int32_t *)std::malloc(sizeof(int32_t) * (m ? m : 1));
    if (!r || !c || !ww) {
        std::free(r); std::free(c); std::free(ww);
        np::set_error("host allocation failed");
        return NP_ERR_NOMEM;
    }
    std::vector<long long> cur((size_t)n + 1, 0);
    for (size_t e = 0; e < m; ++e) cur[(size_t)el.tail[e] + 1]++;
    for (int i = 0; i < n; ++i) cur[(size_t)i + 1] += cur[(size_t)i];
    for (int i = 0; i <= n; ++i) r[i] = (int32_t)cur[(size_t)i];
    for (size_t k = m; k-- > 0;) {
        long long p = cur[(size_t)el.tail[k]]++;
        c[p] = el.head[k];
        ww[p] = el.w[k];
    }
    *row = r; *col = c; *w = ww;
    return NP_OK;
}

int read_dimacs9(FILE *fp, int *n_out, int *src_out, EdgeList &el) {
    std::vector<char> buff(BUFF_SIZE);
    int n = -1;
    long long m_hdr = 0;
    int source = 0;
    while (std::fgets(buff.data(), (int)BUFF_SIZE, fp) != nullptr) {
        char *b = buff.data();
        if (b[0] == 'c') continue;
        if (b[0] == 'p') {  // global.cpp:35-42
            char *p = std::strtok(b + 4, " ");
            if (!p) { np::set_error("DIMACS9: bad 'p' line"); return NP_ERR_IO; }
            n = std::atoi(p);
            p = std::strtok(nullptr, " ");
            if (!p) { np::set_error("DIMACS9: bad 'p' line"); return NP_ERR_IO; }
            m_hdr = std::atoll(p);
            if (n < 0 || m_hdr < 0) { np::set_error("DIMACS9: negative sizes"); return NP_ERR_IO; }
            el.tail.reserve((size_t)m_hdr); el.head.reserve((size_t)m_hdr); el.w.reserve((size_t)m_hdr);
        }
        if (b[0] == 's') {  // global.cpp:43-46
            char *p = std::strtok(b + 1, " ");
            source = p ? std::atoi(p) - 1 : -1;
        }
        if (b[0] == 'a') {  // global.cpp:47-60
            if (n < 0) { np::set_error("DIMACS9: arc before 'p' line"); return NP_ERR_IO; }
            char *p = std::strtok(b + 1, " ");
            char *q = p ? std::strtok(nullptr, " ") : nullptr;
            char *r = q ? std::strtok(nullptr, " ") : nullptr;
            if (!r) { np::set_error("DIMACS9: bad 'a' line"); return NP_ERR_IO; }
            int t = std::atoi(p) - 1, h = std::atoi(q) - 1, wt = std::atoi(r);
            if (t < 0 || t >= n || h < 0 || h >= n) { np::set_error("DIMACS9: vertex id out of range"); return NP_ERR_IO; }
            el.tail.push_back(t); el.head.push_back(h); el.w.push_back(wt);
        }
    }
    if (n < 0) { np::set_error("DIMACS9: no 'p' line"); return NP_ERR_IO; }
    if ((long long)el.tail.size() != m_hdr) {
        np::set_error("DIMACS9: %lld arcs read, header says %lld", (long long)el.tail.size(), m_hdr);
        return NP_ERR_IO;
    }
    if (source < 0 || (n > 0 && source >= n)) { np::set_error("DIMACS9: source out of range"); return NP_ERR_IO; }
    *n_out = n; *src_out = source;
    return NP_OK;
}

int read_dimacs10(FILE *fp, int *n_out, EdgeList &el) {
    std::vector<char> buff(BUFF_SIZE);
    char *b = buff.data();
    bool got = false;
    while (std::fgets(b, (int)BUFF_SIZE, fp) != nullptr) {  // global.cpp:71-74
        if (b[0] == '%') continue;
        got = true;
        break;
    }
    if (!got) { np::set_error("DIMACS10: no header"); return NP_ERR_IO; }
    char *p = std::strtok(b, " ");
    if (!p) { np::set_error("DIMACS10: bad header"); return NP_ERR_IO; }
    int n = std::atoi(p);
    p = std::strtok(nullptr, " ");
    if (!p) { np::set_error("DIMACS10: bad header"); return NP_ERR_IO; }
    long long m_hdr = 2ll * std::atoll(p);  // global.cpp:80
    if (n < 0 || m_hdr < 0) { np::set_error("DIMACS10: negative sizes"); return NP_ERR_IO; }
    el.tail.reserve((size_t)m_hdr); el.head.reserve((size_t)m_hdr); el.w.reserve((size_t)m_hdr);
    std::srand(7);  // global.cpp:84
    for (int i = 0; i < n; ++i) {
        if (std::fgets(b, (int)BUFF_SIZE, fp) == nullptr) {
            np::set_error("Parsing adjacencylist fails!");
            return NP_ERR_IO;
        }
        if (std::strlen(b) == BUFF_SIZE - 1) { np::set_error("DIMACS10: line %d too long", i); return NP_ERR_IO; }
        p = std::strtok(b, " ");
        while (p != nullptr && *p != '\n') {  // global.cpp:94-101
            int h = std::atoi(p) - 1;
            if (h < 0 || h >= n) { np::set_error("DIMACS10: vertex id out of range"); return NP_ERR_IO; }
            el.tail.push_back(i); el.head.push_back(h);
            el.w.push_back(std::rand() % 1000 + 1);
            p = std::strtok(nullptr, " ");
        }
    }
    if ((long long)el.tail.size() != m_hdr) {
        np::set_error("DIMACS10: %lld entries read, header says %lld", (long long)el.tail.size(), m_hdr);
        return NP_ERR_IO;
    }
    *n_out = n;
    return NP_OK;
}

int read_slndc(FILE *fp, int *n_out, EdgeList &el) {
    char b[256];
    if (!std::fgets(b, 256, fp) || !std::fgets(b, 256, fp) || !std::fgets(b, 256, fp)) {
        np::set_error("SLNDC: short header"); return NP_ERR_IO;
    }
    if (std::strlen(b) < 10) { np::set_error("SLNDC: bad node line"); return NP_ERR_IO; }
    char *p = std::strtok(b + 9, " ");  // global.cpp:125-128 ("# Nodes: N Edges: M")
    if (!p) { np::set_error("SLNDC: bad node line"); return NP_ERR_IO; }
    int n_hdr = std::atoi(p);
    if (n_hdr < 0) { np::set_error("SLNDC: bad node count"); return NP_ERR_IO; }
    if (!std::fgets(b, 256, fp)) { np::set_error("SLNDC: short header"); return NP_ERR_IO; }
    const long long lim = 2ll * n_hdr;
    std::vector<int> t_in, h_in;
    int tail, head;
    int maxNode = 0;
    for (;;) {  // global.cpp:136-150
        int k = std::fscanf(fp, "%d%d", &tail, &head);
        if (k == EOF) break;
        if (k != 2) { np::set_error("SLNDC: malformed edge list"); return NP_ERR_IO; }
        if (tail < 0 || head < 0 || tail >= lim || head >= lim) {
            np::set_error("SLNDC: vertex id %d/%d beyond twice the header count", tail, head);
            return NP_ERR_IO;
        }
        if (tail > maxNode) maxNode = tail;
        if (head > maxNode) maxNode = head;
        t_in.push_back(tail); h_in.push_back(head);
    }
    // ids present, compacted in increasing order (global.cpp:152-161)
    std::vector<int> map((size_t)maxNode + 1, -1);
    for (size_t e = 0; e < t_in.size(); ++e) { map[(size_t)t_in[e]] = 1; map[(size_t)h_in[e]] = 1; }
    int n = 0;
    for (int i = 0; i <= maxNode; ++i) {
        if (map[(size_t)i] == 1) map[(size_t)i] = n++;
        else map[(size_t)i] = -1;
    }
    if (t_in.empty()) n = 0;
    // adjacency per original tail id in file order, weights drawn tail by tail
    // (global.cpp:164-186)
    std::vector<long long> first((size_t)maxNode + 2, 0);
    for (size_t e = 0; e < t_in.size(); ++e) first[(size_t)t_in[e] + 1]++;
    for (int i = 0; i <= maxNode; ++i) first[(size_t)i + 1] += first[(size_t)i];
    std::vector<int> byTail(t_in.size());
    {
        std::vector<long long> cur(first.begin(), first.end() - 1);
        for (size_t e = 0; e < t_in.size(); ++e) byTail[(size_t)cur[(size_t)t_in[e]]++] = h_in[e];
    }
    el.tail.reserve(t_in.size()); el.head.reserve(t_in.size()); el.w.reserve(t_in.size());
    std::srand(7);
    for (int i = 0; i <= maxNode && !t_in.empty(); ++i) {
        if (map[(size_t)i] == -1) continue;
        for (long long k = first[(size_t)i]; k < first[(size_t)i + 1]; ++k) {
            el.tail.push_back(map[(size_t)i]);
            el.head.push_back(map[(size_t)byTail[(size_t)k]]);
            el.w.push_back(std::rand() % 1000 + 1);
        }
    }
    *n_out = n;
    return NP_OK;
}

}  // namespace

extern "C" {

void np_host_free(void *p) { std::free(p); }

// ---- binary CSR cache -------------------------------------------------------------------
// A parsed or generated graph saved once and read back without parsing.  Layout (native
// little-endian): 8-byte magic "NPCSR01\0", int64 n, int64 m, int32 source, int32 flags (bit 0:
// weights present), then row[n + 1], col[m] and, with bit 0, w[m], all int32.  A read checks
// the sizes against the file length and the CSR itself (row[0] = 0, row[n] = m, row
// nondecreasing, every col in [0, n)) and fails with NP_ERR_IO otherwise.
static const char kCsrMagic[8] = {'N', 'P', 'C', 'S', 'R', '0', '1', '\0'};

struct CsrHeader {
    char magic[8];
    int64_t n, m;
    int32_t src, flags;
};

static bool write_all(FILE *fp, const void *p, size_t bytes) { return bytes == 0 || std::fwrite(p, 1, bytes, fp) == bytes; }
static bool read_all(FILE *fp, void *p, size_t bytes) { return bytes == 0 || std::fread(p, 1, bytes, fp) == bytes; }

int np_csr_save(const char *path, int32_t n, int64_t m, int32_t src, const int32_t *row, const int32_t *col,
                const int32_t *w) {
    if (!path || n < 0 || m < 0 || !row || (m > 0 && !col)) { np::set_error("bad arguments"); return NP_ERR_ARG; }
    FILE *fp = std::fopen(path, "wb");
    if (!fp) { np::set_error("cannot open %s for writing", path); return NP_ERR_IO; }
    CsrHeader h;
    std::memcpy(h.magic, kCsrMagic, 8);
    h.n = n; h.m = m; h.src = src; h.flags = w ? 1 : 0;
    bool ok = write_all(fp, &h, sizeof h) && write_all(fp, row, sizeof(int32_t) * ((size_t)n + 1)) &&
              write_all(fp, col, sizeof(int32_t) * (size_t)m) && (!w || write_all(fp, w, sizeof(int32_t) * (size_t)m));
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) { np::set_error("write to %s failed", path); return NP_ERR_IO; }
    return NP_OK;
}

static int csr_load_stream(FILE *fp, int32_t *n, int64_t *m, int32_t *src, int32_t **row, int32_t **col,
                           int32_t **w) {
    CsrHeader h;
    if (!read_all(fp, &h, sizeof h) || std::memcmp(h.magic, kCsrMagic, 8) != 0) {
        np::set_error("not a binary CSR file"); return NP_ERR_IO;
    }
    if (h.n < 0 || h.n > 0x7FFFFFFE || h.m < 0 || h.m > 0x7FFFFFFF || (h.flags & ~1)) {
        np::set_error("binary CSR header out of range"); return NP_ERR_IO;
    }
    const size_t nb = sizeof(int32_t) * ((size_t)h.n + 1), mb = sizeof(int32_t) * (size_t)h.m;
    long here = std::ftell(fp);
    if (here >= 0 && std::fseek(fp, 0, SEEK_END) == 0) {
        const long end = std::ftell(fp);
        if (end < 0 || (unsigned long)(end - here) != nb + mb * ((h.flags & 1) ? 2 : 1)) {
            np::set_error("binary CSR file size does not match its header"); return NP_ERR_IO;
        }
        std::fseek(fp, here, SEEK_SET);
    }
    int32_t *r = (int32_t *)std::malloc(nb);
    int32_t *c = (int32_t *)std::malloc(mb ? mb : 4);
    int32_t *ww = (int32_t *)std::malloc(mb ? mb : 4);
    if (!r || !c || !ww) { std::free(r); std::free(c); std::free(ww); np::set_error("host allocation failed"); return NP_ERR_NOMEM; }
    bool ok = read_all(fp, r, nb) && read_all(fp, c, mb) && (!(h.flags & 1) || read_all(fp, ww, mb));
    if (ok && !(h.flags & 1)) for (int64_t e = 0; e < h.m; ++e) ww[e] = 1;  // unweighted: unit weights
    if (ok) ok = r[0] == 0 && (int64_t)r[h.n] == h.m;
    for (int64_t v = 0; ok && v < h.n; ++v) ok = r[v] <= r[v + 1];
    for (int64_t e = 0; ok && e < h.m; ++e) ok = c[e] >= 0 && (int64_t)c[e] < h.n;
    if (!ok) {
        std::free(r); std::free(c); std::free(ww);
        np::set_error("binary CSR file is truncated or not a valid CSR"); return NP_ERR_IO;
    }
    *n = (int32_t)h.n; *m = h.m; *src = h.src; *row = r; *col = c; *w = ww;
    return NP_OK;
}

int np_csr_load(const char *path, int32_t *n, int64_t *m, int32_t *src, int32_t **row, int32_t **col,
                int32_t **w) {
    if (!path || !n || !m || !src || !row || !col || !w) { np::set_error("null argument"); return NP_ERR_ARG; }
    FILE *fp = std::fopen(path, "rb");
    if (!fp) { np::set_error("cannot open %s", path); return NP_ERR_IO; }
    int rc = csr_load_stream(fp, n, m, src, row, col, w);
    std::fclose(fp);
    return rc;
}

int np_read_graph_stream(int32_t fmt, void *stream, int32_t *n, int64_t *m, int32_t *src, int32_t **row,
                         int32_t **col, int32_t **w) {
    FILE *fp = (FILE *)stream;
    if (!fp || !n || !m || !src || !row || !col || !w) { np::set_error("null argument"); return NP_ERR_ARG; }
    if (fmt == 3) return csr_load_stream(fp, n, m, src, row, col, w);
    EdgeList el;
    int nn = 0, s = 0, rc;
    switch (fmt) {
        case 0: rc = read_dimacs9(fp, &nn, &s, el); break;
        case 1: rc = read_dimacs10(fp, &nn, el); break;
        case 2: rc = read_slndc(fp, &nn, el); break;
        default: np::set_error("Wrong code for dataset"); rc = NP_ERR_ARG;
    }
    if (rc) return rc;
    if (el.tail.size() > 0x7FFFFFFFull) { np::set_error("more than 2^31-1 edges"); return NP_ERR_CAPACITY; }
    rc = to_csr(nn, el, row, col, w);
    if (rc) return rc;
    *n = nn; *m = (int64_t)el.tail.size(); *src = s;
    return NP_OK;
}

int np_read_graph(int32_t fmt, const char *path, int32_t *n, int64_t *m, int32_t *src, int32_t **row,
                  int32_t **col, int32_t **w) {
    if (!path) { np::set_error("null path"); return NP_ERR_ARG; }
    FILE *fp = std::fopen(path, fmt == 3 ? "rb" : "r");
    if (!fp) { np::set_error("cannot open %s", path); return NP_ERR_IO; }
    int rc = np_read_graph_stream(fmt, fp, n, m, src, row, col, w);
    std::fclose(fp);
    return rc;
}

// ---- host generators: same sampling as np_gen.hip, CSR by counting sort -------------
static int edges_to_csr(int n, long long m, const std::vector<uint32_t> &s,
                        const std::vector<uint32_t> &d, const std::vector<int> &wt, int32_t **row,
                        int32_t **col, int32_t **w) {
    int32_t *r = (int32_t *)std::malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t *c = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)(m ? m : 1));
    int32_t *ww = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)(m ? m : 1));
    if (!r || !c || !ww) { std::free(r); std::free(c); std::free(ww); np::set_error("host allocation failed"); return NP_ERR_NOMEM; }
    std::vector<long long> cur((size_t)n + 1, 0);
    for (long long e = 0; e < m; ++e) cur[(size_t)s[(size_t)e] + 1]++;
    for (int i = 0; i < n; ++i) cur[(size_t)i + 1] += cur[(size_t)i];
    for (int i = 0; i <= n; ++i) r[i] = (int32_t)cur[(size_t)i];
    for (long long e = m; e-- > 0;) {
        long long p = cur[(size_t)s[(size_t)e]]++;
        c[p] = (int32_t)d[(size_t)e];
        ww[p] = wt[(size_t)e];
    }
    *row = r; *col = c; *w = ww;
    return NP_OK;
}

void np_rmat_thresholds_host(double a, double b, double c, uint64_t *ta, uint64_t *tab, uint64_t *tabc) {
    const double S = 4294967296.0;
    double x = a * S, y = (a + b) * S, z = (a + b + c) * S;
    *ta = x >= S ? (uint64_t)S : (uint64_t)x;
    *tab = y >= S ? (uint64_t)S : (uint64_t)y;
    *tabc = z >= S ? (uint64_t)S : (uint64_t)z;
}

int np_gen_rmat_host(int32_t scale, int32_t edge_factor, double a, double b, double c, int32_t wmin,
                     int32_t wmax, uint64_t seed, int32_t permute, int32_t **row, int32_t **col,
                     int32_t **w, int32_t *n_out, int64_t *m_out) {
    if (scale < 1 || scale > 30 || edge_factor < 1 || wmax < wmin || !row || !col || !w) {
        np::set_error("bad R-MAT arguments"); return NP_ERR_ARG;
    }
    const int n = 1 << scale;
    const long long m = (long long)edge_factor * n;
    if (m > 0x7FFFFFFFll) { np::set_error("R-MAT: more than 2^31-1 edges"); return NP_ERR_CAPACITY; }
    uint64_t ta, tab, tabc;
    np_rmat_thresholds_host(a, b, c, &ta, &tab, &tabc);
    const uint64_t key = np::stream_key(seed, np::ST_RMAT), wkey = np::stream_key(seed, np::ST_WEIGHT),
                   pkey = np::stream_key(seed, np::ST_PERM);
    const uint32_t wr = (uint32_t)(wmax - wmin + 1);
    std::vector<uint32_t> s((size_t)m), d((size_t)m);
    std::vector<int> wt((size_t)m);
    for (long long e = 0; e < m; ++e) {
        uint32_t u, v;
        np::rmat_edge(key, (uint64_t)e, scale, ta, tab, tabc, &u, &v);
        if (permute) { u = np::permute_id(u, scale, pkey); v = np::permute_id(v, scale, pkey); }
        s[(size_t)e] = u; d[(size_t)e] = v;
        wt[(size_t)e] = wmin + (int)np::uniform_below((uint32_t)np::draw(wkey, (uint64_t)e), wr);
    }
    if (n_out) *n_out = n;
    if (m_out) *m_out = m;
    return edges_to_csr(n, m, s, d, wt, row, col, w);
}

int np_gen_er_host(int32_t n, int64_t m, int32_t wmin, int32_t wmax, uint64_t seed, int32_t **row,
                   int32_t **col, int32_t **w) {
    if (n < 1 || m < 0 || m > 0x7FFFFFFFll || wmax < wmin || !row || !col || !w) {
        np::set_error("bad ER arguments"); return NP_ERR_ARG;
    }
    const uint64_t key = np::stream_key(seed, np::ST_ER), wkey = np::stream_key(seed, np::ST_WEIGHT);
    const uint32_t wr = (uint32_t)(wmax - wmin + 1);
    std::vector<uint32_t> s((size_t)m), d((size_t)m);
    std::vector<int> wt((size_t)m);
    for (long long e = 0; e < m; ++e) {
        uint32_t u, v;
        np::er_edge(key, (uint64_t)e, (uint32_t)n, &u, &v);
        s[(size_t)e] = u; d[(size_t)e] = v;
        wt[(size_t)e] = wmin + (int)np::uniform_below((uint32_t)np::draw(wkey, (uint64_t)e), wr);
    }
    return edges_to_csr(n, m, s, d, wt, row, col, w);
}

int np_gen_grid_host(int32_t rows, int32_t cols, int64_t n_long, int32_t wmin, int32_t wmax,
                     uint64_t seed, int32_t **row, int32_t **col, int32_t **w, int64_t *m_out) {
    if (rows < 1 || cols < 1 || n_long < 0 || wmax < wmin || !row || !col || !w) {
        np::set_error("bad grid arguments"); return NP_ERR_ARG;
    }
    const long long n = (long long)rows * cols;
    const long long grid_m = 2ll * ((long long)rows * (cols - 1) + (long long)cols * (rows - 1));
    const long long m = grid_m + n_long;
    if (n > 0x7FFFFFFFll || m > 0x7FFFFFFFll) { np::set_error("grid too large"); return NP_ERR_CAPACITY; }
    const uint64_t lkey = np::stream_key(seed, np::ST_LONG), wkey = np::stream_key(seed, np::ST_WEIGHT);
    const uint32_t wr = (uint32_t)(wmax - wmin + 1);
    std::vector<uint32_t> s((size_t)m), d((size_t)m);
    std::vector<int> wt((size_t)m);
    long long e = 0;
    for (long long v = 0; v < n; ++v) {
        int dg = np::grid_deg((uint32_t)v, rows, cols);
        for (int k = 0; k < dg; ++k, ++e) {
            s[(size_t)e] = (uint32_t)v;
            d[(size_t)e] = np::grid_nbr((uint32_t)v, rows, cols, k);
            wt[(size_t)e] = wmin + (int)np::uniform_below((uint32_t)np::draw(wkey, (uint64_t)e), wr);
        }
    }
    for (long long j = 0; j < n_long; ++j, ++e) {
        uint32_t u, v;
        np::er_edge(lkey, (uint64_t)j, (uint32_t)n, &u, &v);
        s[(size_t)e] = u; d[(size_t)e] = v;
        wt[(size_t)e] = wmin + (int)np::uniform_below((uint32_t)np::draw(wkey, (uint64_t)e), wr);
    }
    if (m_out) *m_out = m;
    return edges_to_csr((int)n, m, s, d, wt, row, col, w);
}

int np_pick_sources(const int32_t *row, int32_t n, int32_t k, uint64_t seed, int32_t *srcs) {
    if (!row || !srcs || n < 1 || k < 0) { np::set_error("bad arguments"); return NP_ERR_ARG; }
    const uint64_t key = np::stream_key(seed, np::ST_SRC);
    int got = 0;
    const long long limit = 64ll * n + 4096;
    for (long long i = 0; got < k && i < limit; ++i) {
        int v = (int)np::uniform_below((uint32_t)np::draw(key, (uint64_t)i), (uint32_t)n);
        if (row[v + 1] - row[v] < 1) continue;
        bool dup = false;
        for (int j = 0; j < got; ++j) dup |= srcs[j] == v;
        if (!dup) srcs[got++] = v;
    }
    if (got < k) { np::set_error("only %d vertices with out-degree >= 1 found", got); return NP_ERR_ARG; }
    return NP_OK;
}

int np_features(const int32_t *count, int32_t it, int32_t n, float *x) {
    if (!count || !x) return NP_ERR_ARG;
    np::features(count, it, n, x);
    return NP_OK;
}

void np_sched_default(int32_t level, np_sched *s) {
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->level = (level >= 0 && level <= 2) ? level : 1;
    s->log2_t_lo = 6;  // THRESHOLD = 64, apps/sssp/src/gpu/sssp_kernel.cu:7
    s->log2_t_hi = 6;
    s->drive = NP_DRIVE_PERSISTENT;
    s->frontier = NP_FRONTIER_SCAN;
}

// The fixed-threshold heuristic: the plain schedule of the level a caller configured
// (reference default CONSOLIDATE_LEVEL 1, apps/sssp/src/gpu/Makefile:28), THRESHOLD 64.
int np_heuristic(const int32_t *count, int32_t it, int32_t n, np_sched *s) {
    (void)count; (void)it; (void)n;
    if (!s) return NP_ERR_ARG;
    int lv = s->level >= 0 && s->level <= 2 ? s->level : 1;
    s->level = lv;
    s->log2_t_lo = 6;
    s->log2_t_hi = 6;
    return NP_OK;
}

}  // extern "C"
