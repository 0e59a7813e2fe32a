/* AddressSanitizer / UndefinedBehaviorSanitizer run of the CPU oracle (oracle/np_oracle.c):
 * every oracle function on random multigraphs (self loops, duplicates, zero weights, vertices
 * without edges, empty graphs), with the shortest-path functions cross-checked against each
 * other.  Built and run by tests/san/Makefile; any sanitizer report aborts with a nonzero
 * status. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

long long orc_dijkstra_lazy(int n, const int *row, const int *col, const int *w, int src, int *dist);
long long orc_dijkstra_lazy_multi(int n, const int *row, const int *col, const int *w, const int *srcs, int nsrc,
                                  int *dist);
int orc_dijkstra_lazy_many(int n, const int *row, const int *col, const int *w, const int *srcs, int nsrc,
                           int *dist, int nthreads);
int orc_dijkstra_dheap(int n, const int *row, const int *col, const int *w, int src, int arity, int *dist);
int orc_bellman_ford(int n, const int *row, const int *col, const int *w, const int *srcs, int nsrc, int *dist);
int orc_bfs_levels(int n, const int *row, const int *col, int src, int *level);
int orc_check_sssp(int n, const int *row, const int *col, const int *w, const int *srcs, int nsrc,
                   const int *dist, int inf, int unit, int *bad_vertex);
void orc_reached_work(int n, const int *row, const int *dist, int inf, long long *n_r, long long *m_r);
int orc_mlp_nparam(void);
void orc_mlp_forward(const float *p, const float *x, int batch, float *out);
float orc_mlp_loss_grad(const float *p, const float *x, const int *label, const float *target, int batch,
                        float lam, float *grad);
int orc_bc(int n, const int *row, const int *col, const int *srcs, int nsrc, double *bc);
void orc_color(int n, const int *row, const int *col, int *color);
int orc_pagerank(int n, const int *row, const int *col, double damping, double total, double eps, int max_iters,
                 double *rank);
void orc_spmv(int n, const int *row, const int *col, const double *data, const double *x, double *y);

#define INF 1073741824
#define UNDEF 0x7FFFFFFF

static unsigned long long rs = 0x9E3779B97F4A7C15ull;
static unsigned rnd(unsigned k) {
    rs ^= rs << 13; rs ^= rs >> 7; rs ^= rs << 17;
    return k ? (unsigned)(rs % k) : 0u;
}

static int fail(const char *what, int t) {
    fprintf(stderr, "oracle_san: %s (case %d)\n", what, t);
    return 1;
}

static int one_case(int t, int n, int m, int wmin, int wmax) {
    int *row = calloc((size_t)n + 1, sizeof(int)), *col = malloc(sizeof(int) * (size_t)(m ? m : 1));
    int *w = malloc(sizeof(int) * (size_t)(m ? m : 1)), *cnt = calloc((size_t)n + 1, sizeof(int));
    int *src = malloc(sizeof(int) * (size_t)(m ? m : 1));
    for (int e = 0; e < m; ++e) { src[e] = (int)rnd((unsigned)n); cnt[src[e] + 1]++; }
    for (int v = 0; v < n; ++v) row[v + 1] = row[v] + cnt[v + 1];
    int *at = calloc((size_t)n + 1, sizeof(int));
    for (int e = 0; e < m; ++e) {
        int u = src[e], k = row[u] + at[u]++;
        col[k] = (int)rnd((unsigned)n);
        w[k] = wmin + (int)rnd((unsigned)(wmax - wmin) + 1u);
    }
    int *d1 = malloc(sizeof(int) * (size_t)n), *d2 = malloc(sizeof(int) * (size_t)n);
    int *d3 = malloc(sizeof(int) * (size_t)n), *lv = malloc(sizeof(int) * (size_t)n);
    int *many = malloc(sizeof(int) * (size_t)n * 3), *col_ = malloc(sizeof(int) * (size_t)n);
    double *bc = malloc(sizeof(double) * (size_t)n), *rank = malloc(sizeof(double) * (size_t)n);
    double *data = malloc(sizeof(double) * (size_t)(m ? m : 1)), *x = malloc(sizeof(double) * (size_t)n);
    double *y = malloc(sizeof(double) * (size_t)n);
    int rc = 0, bad = -1;
    int s = (int)rnd((unsigned)n), srcs3[3] = {s, (int)rnd((unsigned)n), (int)rnd((unsigned)n)};
    if (orc_dijkstra_lazy(n, row, col, w, s, d1) < 0) rc |= fail("dijkstra", t);
    if (orc_dijkstra_dheap(n, row, col, w, s, 2 + t % 3, d2) != 0) rc |= fail("dheap", t);
    if (orc_bellman_ford(n, row, col, w, &s, 1, d3) < 0) rc |= fail("bellman-ford", t);
    if (memcmp(d1, d2, sizeof(int) * (size_t)n) || memcmp(d1, d3, sizeof(int) * (size_t)n))
        rc |= fail("shortest-path functions disagree", t);
    if (wmin > 0 && orc_check_sssp(n, row, col, w, &s, 1, d1, INF, 0, &bad) != 0)
        rc |= fail("certificate rejects Dijkstra", t);
    if (orc_dijkstra_lazy_many(n, row, col, w, srcs3, 3, many, 2) != 0) rc |= fail("dijkstra_many", t);
    if (memcmp(many, d1, sizeof(int) * (size_t)n)) rc |= fail("dijkstra_many source 0", t);
    if (orc_dijkstra_lazy_multi(n, row, col, w, srcs3, 3, d2) < 0) rc |= fail("dijkstra_multi", t);
    if (orc_bfs_levels(n, row, col, s, lv) != 0) rc |= fail("bfs", t);
    if (orc_check_sssp(n, row, col, NULL, &s, 1, lv, UNDEF, 1, &bad) != 0) rc |= fail("certificate rejects BFS", t);
    long long nr = 0, mr = 0;
    orc_reached_work(n, row, d1, INF, &nr, &mr);
    if (orc_bc(n, row, col, srcs3, 3, bc) != 0) rc |= fail("bc", t);
    orc_color(n, row, col, col_);
    if (orc_pagerank(n, row, col, 0.5, 1000.0, 1e-13, 200, rank) < 0) rc |= fail("pagerank", t);
    for (int e = 0; e < m; ++e) data[e] = e % 13 + 0.5;
    for (int v = 0; v < n; ++v) x[v] = v % 17 + 0.5;
    orc_spmv(n, row, col, data, x, y);
    free(row); free(col); free(w); free(cnt); free(src); free(at); free(d1); free(d2); free(d3); free(lv);
    free(many); free(col_); free(bc); free(rank); free(data); free(x); free(y);
    return rc;
}

int main(void) {
    int rc = 0, t = 0;
    /* n, m, weight range (zero weights in some) */
    const int sizes[][4] = {{1, 0, 1, 9}, {2, 1, 1, 9}, {5, 30, 0, 0}, {40, 10, 1, 5}, {300, 3000, 1, 100},
                            {1000, 20000, 1, 255}, {2000, 1000, 0, 3}, {64, 4096, 1, 1}, {500, 5000, 0, 50}};
    for (size_t i = 0; i < sizeof sizes / sizeof sizes[0]; ++i)
        for (int r = 0; r < 3; ++r) rc |= one_case(t++, sizes[i][0], sizes[i][1], sizes[i][2], sizes[i][3]);
    const int np = orc_mlp_nparam();
    float *p = malloc(sizeof(float) * (size_t)np), *g = malloc(sizeof(float) * (size_t)np);
    float xx[4 * 10], tg[4 * 2], out[4 * 5];
    int lab[4] = {0, 1, 2, 1};
    for (int i = 0; i < np; ++i) p[i] = (float)((int)rnd(2001) - 1000) / 4000.0f;
    for (int i = 0; i < 40; ++i) xx[i] = (float)rnd(100) / 100.0f;
    for (int i = 0; i < 8; ++i) tg[i] = (float)rnd(12);
    orc_mlp_forward(p, xx, 4, out);
    (void)orc_mlp_loss_grad(p, xx, lab, tg, 4, 0.1f, g);
    free(p); free(g);
    printf("oracle_san: %d cases, %s\n", t, rc ? "FAILED" : "ok");
    return rc;
}
