/*
 * np_oracle.c -- CPU oracle for the SSSP / recursive-BFS nested-parallelism path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the C-ABI,
 * the drop-in entry points) links, loads or calls this file.  It is used by the
 * tests, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, and
 * there only as the checker / the timed CPU baseline, never as the thing measured.
 *
 * Contents and what pins each function (see tests/test_oracle.py):
 *   orc_dijkstra_lazy   restates apps/sssp/src/cpu/sssp.cpp:13-47 (SSSP_CPU) with the
 *                       lazy-deletion binary heap of common/heap.cpp:29-59,80-109.
 *                       Pinned by: the reference itself compiled from source
 *                       (oracle/_ref, identical distances AND identical pop count),
 *                       brute-force Bellman-Ford, closed forms (grid = Manhattan,
 *                       path, star), the shortest-path certificate.
 *   orc_dijkstra_dheap  Dijkstra with the decrease-key d-ary heap of
 *                       common/dheap.cpp:51-121.  Pinned by equality with the two above
 *                       and with a Dijkstra the harness drives over the reference's
 *                       own dheap class (oracle/_ref).
 *   orc_bellman_ford    brute force, O(n*m); pinned by closed forms.
 *   orc_bfs_levels      restates bfs_cpu, apps/bfs-rec/src/gpu/bfs_rec.cpp:100-119.
 *                       Pinned by the reference's bfs_cpu compiled in oracle/_ref and by
 *                       unit-weight Dijkstra.
 *   orc_check_sssp      shortest-path certificate (O(m)): dist[s]=0 for sources, every
 *                       edge relaxed, every reached non-source vertex has a tight in-edge,
 *                       unreached == INF.  With positive weights this proves the
 *                       distances exact.  Pinned by planted mistakes (tests flip one
 *                       distance / one edge and expect a reject).
 *   orc_reached_work    n_r and m_r (vertices reached, out-edges of reached vertices),
 *                       the units of the algorithmic-byte count.
 *   orc_mlp_*           fp32 restatement of the scheduler MLP forward/backward; pinned
 *                       by finite differences in the tests.
 *   orc_pagerank        the reference's PageRank iteration (pagerank_wrapper.cu:199-244,
 *                       pagerank_kernel.cu:83-107) in double.  Pinned by the linear system of
 *                       its fixed point solved with scipy (tests), closed forms (a directed
 *                       cycle is uniform, a two-way star by hand) and the total rank (1000).
 *   orc_spmv            y = A x over the CSR (apps/SpMV/src/gpu/SpMV_kernel.cu:77-92) in
 *                       double.  Pinned by scipy's CSR product and closed forms (tests).
 *   orc_color           the colors of the reference's graph-coloring rounds
 *                       (apps/graph-color/src/gpu/graph_color_kernel.cu:14-35):
 *                       color[v] = 1 + max color of the out-neighbours with a smaller id.
 *                       Pinned by a simulation of the rounds themselves in random thread
 *                       orders (tests), closed forms (paths, complete digraph) and, on
 *                       symmetric graphs, the proper-coloring invariant.
 *   orc_bc              Brandes betweenness, restating apps/bc/src/cpu/bc_kernel.cpp:20-95
 *                       and bc.cpp:152-185 (and the GPU's bc_fp/bc_bp kernels): path counts
 *                       over edges into the next BFS level, dependencies back from the
 *                       deepest level, bc[v] += delta[v] for v != source; in double.  Pinned
 *                       by the reference's bc_cpu compiled in oracle/_ref, closed forms
 *                       (directed path, star, complete graph) and the definition
 *                       sum over s != v != t of sigma_st(v) / sigma_st by brute force.
 *
 * Conventions follow common/global.h:19 (INF = 1073741824) and
 * apps/bfs-rec/src/gpu/bfs_rec.h:13 (UNDEFINED = 0x7FFFFFFF).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_INF 1073741824
#define ORC_UNDEFINED 0x7FFFFFFF

/* ------------------------------------------------------------------------- */
/* Lazy binary heap of (id,key) items: common/heap.cpp.                       */
/* insert = push_back + siftup (strict '>' with the parent, heap.cpp:80-90);   */
/* remove = take front, move back to front, pop_back, siftdown (heap.cpp:50-59,*/
/* 93-109: pick the right child only when left.key > right.key, swap only when */
/* child.key < parent.key).  No capacity cap (the reference exits at 102.4M).  */
/* ------------------------------------------------------------------------- */
typedef struct { int id; int key; } orc_item;
typedef struct { orc_item *a; long long n, cap; } orc_heap;

static int heap_push(orc_heap *h, int id, int key) {
    if (h->n == h->cap) {
        long long nc = h->cap ? h->cap * 2 : 1024;
        orc_item *na = (orc_item *)realloc(h->a, (size_t)nc * sizeof(orc_item));
        if (!na) return -1;
        h->a = na; h->cap = nc;
    }
    long long i = h->n++;
    h->a[i].id = id; h->a[i].key = key;
    while (i > 0) {
        long long p = (i - 1) >> 1;
        if (h->a[p].key > h->a[i].key) {
            orc_item t = h->a[p]; h->a[p] = h->a[i]; h->a[i] = t; i = p;
        } else break;
    }
    return 0;
}

static orc_item heap_pop(orc_heap *h) {
    orc_item top = h->a[0];
    h->a[0] = h->a[h->n - 1];
    h->n--;
    long long i = 0;
    for (;;) {
        long long l = 2 * i + 1, r = 2 * i + 2, c;
        if (l >= h->n) break;
        c = l;
        if (r < h->n && h->a[l].key > h->a[r].key) c = r;
        if (h->a[c].key < h->a[i].key) {
            orc_item t = h->a[c]; h->a[c] = h->a[i]; h->a[i] = t; i = c;
        } else break;
    }
    return top;
}

/* SSSP_CPU (apps/sssp/src/cpu/sssp.cpp:13-47), generalised to a source set: every
 * source is inserted with key 0 (one source = the reference exactly).  A vertex is
 * final on its first pop; edges are pushed only into vertices still at INF.
 * Returns the reference's printed "Iteration number" (pops + 1), or -1 on OOM. */
long long orc_dijkstra_lazy_multi(int n, const int *row, const int *col, const int *w,
                                  const int *srcs, int nsrc, int *dist) {
    orc_heap h = {0, 0, 0};
    for (int i = 0; i < n; ++i) dist[i] = ORC_INF;
    for (int s = 0; s < nsrc; ++s)
        if (heap_push(&h, srcs[s], 0)) { free(h.a); return -1; }
    long long iteration = 1;
    while (h.n) {
        orc_item it = heap_pop(&h);
        int u = it.id;
        if (dist[u] == ORC_INF) {
            dist[u] = it.key;
            for (int e = row[u]; e < row[u + 1]; ++e) {
                int v = col[e];
                if (dist[v] == ORC_INF)
                    if (heap_push(&h, v, dist[u] + w[e])) { free(h.a); return -1; }
            }
        }
        iteration++;
    }
    free(h.a);
    return iteration;
}

long long orc_dijkstra_lazy(int n, const int *row, const int *col, const int *w, int src,
                            int *dist) {
    return orc_dijkstra_lazy_multi(n, row, col, w, &src, 1, dist);
}

/* Many independent sources on one graph, one sequential Dijkstra per OpenMP thread
 * (how the CPU baseline is run).  dist is [nsrc][n].  Returns 0 or -1. */
int orc_dijkstra_lazy_many(int n, const int *row, const int *col, const int *w,
                           const int *srcs, int nsrc, int *dist, int nthreads) {
    int bad = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
    for (int s = 0; s < nsrc; ++s) {
        if (orc_dijkstra_lazy(n, row, col, w, srcs[s], dist + (size_t)s * n) < 0) bad |= 1;
    }
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------- */
/* d-ary heap with decrease-key over items 1..N (common/dheap.cpp:51-121):     */
/* p(x)=(x+d-2)/d, left(x)=d(x-1)+2, right(x)=dx+1; siftup moves the item up   */
/* while the parent key is strictly larger; siftdown follows minchild (first   */
/* strictly smaller child wins) while its key is strictly smaller.             */
/* ------------------------------------------------------------------------- */
typedef struct { long long N; long long n; int d; long long *h; long long *pos; long long *kvec; } orc_dheap;

static long long dh_minchild(orc_dheap *H, long long x) {
    long long minc = H->d * (x - 1) + 2;
    if (minc > H->n) return 0;
    for (long long y = minc + 1; y <= H->d * x + 1 && y <= H->n; y++)
        if (H->kvec[H->h[y]] < H->kvec[H->h[minc]]) minc = y;
    return minc;
}
static void dh_siftup(orc_dheap *H, long long i, long long x) {
    long long px = (x + (H->d - 2)) / H->d;
    while (x > 1 && H->kvec[H->h[px]] > H->kvec[i]) {
        H->h[x] = H->h[px]; H->pos[H->h[x]] = x;
        x = px; px = (x + (H->d - 2)) / H->d;
    }
    H->h[x] = i; H->pos[i] = x;
}
static void dh_siftdown(orc_dheap *H, long long i, long long x) {
    long long cx = dh_minchild(H, x);
    while (cx != 0 && H->kvec[H->h[cx]] < H->kvec[i]) {
        H->h[x] = H->h[cx]; H->pos[H->h[x]] = x;
        x = cx; cx = dh_minchild(H, x);
    }
    H->h[x] = i; H->pos[i] = x;
}
static void dh_insert(orc_dheap *H, long long i, long long k) { H->kvec[i] = k; H->n++; dh_siftup(H, i, H->n); }
static void dh_remove(orc_dheap *H, long long i) {
    long long j = H->h[H->n--];
    if (i != j && H->kvec[j] <= H->kvec[i]) dh_siftup(H, j, H->pos[i]);
    else if (i != j && H->kvec[j] > H->kvec[i]) dh_siftdown(H, j, H->pos[i]);
    H->pos[i] = 0;
}
static long long dh_deletemin(orc_dheap *H) { if (H->n == 0) return 0; long long i = H->h[1]; dh_remove(H, i); return i; }
static void dh_changekey(orc_dheap *H, long long i, long long k) {
    long long ki = H->kvec[i]; H->kvec[i] = k;
    if (k < ki) dh_siftup(H, i, H->pos[i]);
    else if (k > ki) dh_siftdown(H, i, H->pos[i]);
}

/* Decrease-key Dijkstra over the d-heap; vertex v is heap item v+1. Returns 0/-1. */
int orc_dijkstra_dheap(int n, const int *row, const int *col, const int *w, int src, int arity,
                       int *dist) {
    orc_dheap H;
    H.N = n; H.n = 0; H.d = arity < 2 ? 2 : arity;
    H.h = (long long *)malloc(sizeof(long long) * (size_t)(n + 1));
    H.pos = (long long *)calloc((size_t)n + 1, sizeof(long long));
    H.kvec = (long long *)malloc(sizeof(long long) * (size_t)(n + 1));
    char *done = (char *)calloc((size_t)n + 1, 1);
    if (!H.h || !H.pos || !H.kvec || !done) { free(H.h); free(H.pos); free(H.kvec); free(done); return -1; }
    for (int i = 0; i < n; ++i) dist[i] = ORC_INF;
    dh_insert(&H, (long long)src + 1, 0);
    while (H.n) {
        long long it = dh_deletemin(&H);
        int u = (int)(it - 1);
        long long du = H.kvec[it];
        dist[u] = (int)du;
        done[u] = 1;
        for (int e = row[u]; e < row[u + 1]; ++e) {
            int v = col[e];
            if (done[v]) continue;
            long long alt = du + w[e];
            if (H.pos[v + 1] != 0) {
                if (alt < H.kvec[v + 1]) dh_changekey(&H, v + 1, alt);
            } else {
                dh_insert(&H, (long long)v + 1, alt);
            }
        }
    }
    free(H.h); free(H.pos); free(H.kvec); free(done);
    return 0;
}

/* Brute-force Bellman-Ford (rounds until no change, at most n).  Tiny graphs only. */
int orc_bellman_ford(int n, const int *row, const int *col, const int *w, const int *srcs, int nsrc,
                     int *dist) {
    for (int i = 0; i < n; ++i) dist[i] = ORC_INF;
    for (int s = 0; s < nsrc; ++s) dist[srcs[s]] = 0;
    for (int round = 0; round <= n; ++round) {
        int changed = 0;
        for (int u = 0; u < n; ++u) {
            if (dist[u] == ORC_INF) continue;
            for (int e = row[u]; e < row[u + 1]; ++e) {
                long long alt = (long long)dist[u] + w[e];
                if (alt < dist[col[e]]) { dist[col[e]] = (int)alt; changed = 1; }
            }
        }
        if (!changed) return round;
    }
    return -1; /* negative cycle: not a valid input for this path */
}

/* bfs_cpu (apps/bfs-rec/src/gpu/bfs_rec.cpp:100-119): FIFO BFS from src, a neighbour is
 * (re)labelled when UNDEFINED or when its level exceeds level(node)+1 (unsigned compare). */
int orc_bfs_levels(int n, const int *row, const int *col, int src, int *level) {
    int *q = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1) * 2);
    long long cap = (long long)(n > 0 ? n : 1) * 2, head = 0, tail = 0, cnt = 0;
    if (!q) return -1;
    for (int i = 0; i < n; ++i) level[i] = ORC_UNDEFINED;
    level[src] = 0;
    q[tail] = src; tail = (tail + 1) % cap; cnt = 1;
    while (cnt) {
        int node = q[head]; head = (head + 1) % cap; cnt--;
        unsigned next = (unsigned)level[node] + 1u;
        for (int e = row[node]; e < row[node + 1]; ++e) {
            int nb = col[e];
            if (level[nb] == ORC_UNDEFINED || (unsigned)level[nb] > next) {
                level[nb] = (int)next;
                if (cnt == cap) { free(q); return -1; }
                q[tail] = nb; tail = (tail + 1) % cap; cnt++;
            }
        }
    }
    free(q);
    return 0;
}

/* Shortest-path certificate.  inf is the unreached sentinel (ORC_INF for SSSP,
 * ORC_UNDEFINED for BFS levels).  unit!=0 ignores w (all weights 1).
 * Returns 0 when dist is exactly the shortest-path distance from the source set;
 * otherwise a positive code and *bad_vertex:
 *   1 a source not at 0, 2 a value out of range, 3 an edge not relaxed,
 *   4 a reached vertex without a tight in-edge, 5 a non-positive weight on a reached
 *   edge (certificate not applicable), 6 out of memory. */
int orc_check_sssp(int n, const int *row, const int *col, const int *w, const int *srcs, int nsrc,
                   const int *dist, int inf, int unit, int *bad_vertex) {
    int code = 0; int badv = -1;
    unsigned char *tight = (unsigned char *)calloc((size_t)(n > 0 ? n : 1), 1);
    if (!tight) return 6;
    for (int s = 0; s < nsrc; ++s) {
        if (dist[srcs[s]] != 0) { code = 1; badv = srcs[s]; goto out; }
        tight[srcs[s]] = 1;
    }
#pragma omp parallel for schedule(dynamic, 4096)
    for (int u = 0; u < n; ++u) {
        int du = dist[u];
        if (du == inf) continue;
        if (du < 0 || du >= inf) {
#pragma omp critical
            { if (!code || code > 2) { code = 2; badv = u; } }
            continue;
        }
        for (int e = row[u]; e < row[u + 1]; ++e) {
            int v = col[e];
            int we = unit ? 1 : w[e];
            if (we <= 0) {
#pragma omp critical
                { if (!code) { code = 5; badv = u; } }
                continue;
            }
            long long alt = (long long)du + we;
            if ((long long)dist[v] > alt) {
#pragma omp critical
                { if (!code || code > 3) { code = 3; badv = v; } }
            } else if ((long long)dist[v] == alt) {
                __atomic_store_n(&tight[v], 1, __ATOMIC_RELAXED);
            }
        }
    }
    if (code) goto out;
    for (int v = 0; v < n; ++v) {
        if (dist[v] != inf && !tight[v]) { code = 4; badv = v; break; }
    }
out:
    free(tight);
    if (bad_vertex) *bad_vertex = badv;
    return code;
}

/* n_r = vertices with dist != inf; m_r = sum of their out-degrees. */
void orc_reached_work(int n, const int *row, const int *dist, int inf, long long *n_r, long long *m_r) {
    long long a = 0, b = 0;
#pragma omp parallel for reduction(+ : a, b)
    for (int v = 0; v < n; ++v)
        if (dist[v] != inf) { a++; b += row[v + 1] - row[v]; }
    *n_r = a; *m_r = b;
}

/* ------------------------------------------------------------------------- */
/* Scheduler MLP restatement (fp32): x[10] -> 32 (ReLU) -> 32 (ReLU) -> 5 outputs */
/* (3 level logits + 2 threshold regressions).  Parameter vector layout:        */
/*   W1[32][10] b1[32] W2[32][32] b2[32] W3[5][32] b3[5]                         */
/* Loss per sample: cross-entropy(softmax(logits), label) + lam * sum_j          */
/* (reg_j - target_j)^2, averaged over the batch.                                */
/* ------------------------------------------------------------------------- */
#define MLP_IN 10
#define MLP_H 32
#define MLP_OUT 5
#define MLP_NPARAM (MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_OUT * MLP_H + MLP_OUT)

int orc_mlp_nparam(void) { return MLP_NPARAM; }

void orc_mlp_forward(const float *p, const float *x, int batch, float *out) {
    const float *W1 = p, *b1 = W1 + MLP_H * MLP_IN, *W2 = b1 + MLP_H, *b2 = W2 + MLP_H * MLP_H;
    const float *W3 = b2 + MLP_H, *b3 = W3 + MLP_OUT * MLP_H;
    for (int s = 0; s < batch; ++s) {
        float h1[MLP_H], h2[MLP_H];
        const float *xs = x + (size_t)s * MLP_IN;
        for (int j = 0; j < MLP_H; ++j) {
            float a = b1[j];
            for (int i = 0; i < MLP_IN; ++i) a += W1[j * MLP_IN + i] * xs[i];
            h1[j] = a > 0.f ? a : 0.f;
        }
        for (int j = 0; j < MLP_H; ++j) {
            float a = b2[j];
            for (int i = 0; i < MLP_H; ++i) a += W2[j * MLP_H + i] * h1[i];
            h2[j] = a > 0.f ? a : 0.f;
        }
        for (int j = 0; j < MLP_OUT; ++j) {
            float a = b3[j];
            for (int i = 0; i < MLP_H; ++i) a += W3[j * MLP_H + i] * h2[i];
            out[(size_t)s * MLP_OUT + j] = a;
        }
    }
}

/* Mean loss over the batch and its gradient w.r.t. every parameter (grad zeroed here). */
float orc_mlp_loss_grad(const float *p, const float *x, const int *label, const float *target,
                        int batch, float lam, float *grad) {
    const float *W1 = p, *b1 = W1 + MLP_H * MLP_IN, *W2 = b1 + MLP_H, *b2 = W2 + MLP_H * MLP_H;
    const float *W3 = b2 + MLP_H, *b3 = W3 + MLP_OUT * MLP_H;
    float *gW1 = grad, *gb1 = gW1 + MLP_H * MLP_IN, *gW2 = gb1 + MLP_H, *gb2 = gW2 + MLP_H * MLP_H;
    float *gW3 = gb2 + MLP_H, *gb3 = gW3 + MLP_OUT * MLP_H;
    double loss = 0.0;
    memset(grad, 0, sizeof(float) * MLP_NPARAM);
    float inv = batch > 0 ? 1.0f / (float)batch : 0.f;
    for (int s = 0; s < batch; ++s) {
        float a1[MLP_H], h1[MLP_H], a2[MLP_H], h2[MLP_H], o[MLP_OUT], go[MLP_OUT], gh2[MLP_H], gh1[MLP_H];
        const float *xs = x + (size_t)s * MLP_IN;
        for (int j = 0; j < MLP_H; ++j) {
            float a = b1[j];
            for (int i = 0; i < MLP_IN; ++i) a += W1[j * MLP_IN + i] * xs[i];
            a1[j] = a; h1[j] = a > 0.f ? a : 0.f;
        }
        for (int j = 0; j < MLP_H; ++j) {
            float a = b2[j];
            for (int i = 0; i < MLP_H; ++i) a += W2[j * MLP_H + i] * h1[i];
            a2[j] = a; h2[j] = a > 0.f ? a : 0.f;
        }
        for (int j = 0; j < MLP_OUT; ++j) {
            float a = b3[j];
            for (int i = 0; i < MLP_H; ++i) a += W3[j * MLP_H + i] * h2[i];
            o[j] = a;
        }
        /* softmax cross-entropy over the 3 level logits */
        float mx = o[0]; if (o[1] > mx) mx = o[1]; if (o[2] > mx) mx = o[2];
        float e0 = expf(o[0] - mx), e1 = expf(o[1] - mx), e2 = expf(o[2] - mx);
        float z = e0 + e1 + e2;
        float pr[3] = {e0 / z, e1 / z, e2 / z};
        int lb = label[s];
        loss += -(double)logf(pr[lb] > 1e-30f ? pr[lb] : 1e-30f);
        for (int j = 0; j < 3; ++j) go[j] = (pr[j] - (j == lb ? 1.f : 0.f)) * inv;
        for (int j = 0; j < 2; ++j) {
            float d = o[3 + j] - target[(size_t)s * 2 + j];
            loss += (double)(lam * d * d);
            go[3 + j] = 2.f * lam * d * inv;
        }
        for (int j = 0; j < MLP_OUT; ++j) {
            gb3[j] += go[j];
            for (int i = 0; i < MLP_H; ++i) gW3[j * MLP_H + i] += go[j] * h2[i];
        }
        for (int i = 0; i < MLP_H; ++i) {
            float g = 0.f;
            for (int j = 0; j < MLP_OUT; ++j) g += W3[j * MLP_H + i] * go[j];
            gh2[i] = a2[i] > 0.f ? g : 0.f;
        }
        for (int j = 0; j < MLP_H; ++j) {
            gb2[j] += gh2[j];
            for (int i = 0; i < MLP_H; ++i) gW2[j * MLP_H + i] += gh2[j] * h1[i];
        }
        for (int i = 0; i < MLP_H; ++i) {
            float g = 0.f;
            for (int j = 0; j < MLP_H; ++j) g += W2[j * MLP_H + i] * gh2[j];
            gh1[i] = a1[i] > 0.f ? g : 0.f;
        }
        for (int j = 0; j < MLP_H; ++j) {
            gb1[j] += gh1[j];
            for (int i = 0; i < MLP_IN; ++i) gW1[j * MLP_IN + i] += gh1[j] * xs[i];
        }
    }
    return (float)(loss * inv);
}

/* Brandes betweenness summed over srcs[0..nsrc) (nsrc = 0: every vertex, as bc_cpu's loop
 * over i does, bc.cpp:164): directed, unit edge lengths, unnormalised; a parallel edge is a
 * distinct shortest path (sigma gains once per edge, bc_kernel.cpp:42-44).  Forward: queue BFS
 * in level order, sigma[w] += sigma[u] for every edge u -> w with level[w] = level[u] + 1.
 * Backward in reverse BFS order: delta[v] += sigma[v] / sigma[w] * (1 + delta[w]) over the same
 * edges (bc_kernel.cpp:53-65), then bc[v] += delta[v] unless v is the source (:79-95). */
int orc_bc(int n, const int *row, const int *col, const int *srcs, int nsrc, double *bc) {
    int *lvl = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *order = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    double *sig = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double *del = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!lvl || !order || !sig || !del) { free(lvl); free(order); free(sig); free(del); return -1; }
    for (int v = 0; v < n; ++v) bc[v] = 0.0;
    const int ns = nsrc > 0 ? nsrc : n;
    for (int k = 0; k < ns; ++k) {
        const int s = nsrc > 0 ? srcs[k] : k;
        if (s < 0 || s >= n) { free(lvl); free(order); free(sig); free(del); return -2; }
        for (int v = 0; v < n; ++v) { lvl[v] = -1; sig[v] = 0.0; del[v] = 0.0; }
        lvl[s] = 0; sig[s] = 1.0;
        int head = 0, tail = 0;
        order[tail++] = s;
        while (head < tail) {
            const int u = order[head++];
            for (int e = row[u]; e < row[u + 1]; ++e) {
                const int w = col[e];
                if (lvl[w] < 0) { lvl[w] = lvl[u] + 1; order[tail++] = w; }
                if (lvl[w] == lvl[u] + 1) sig[w] += sig[u];
            }
        }
        for (int i = tail - 1; i >= 0; --i) {
            const int v = order[i];
            for (int e = row[v]; e < row[v + 1]; ++e) {
                const int w = col[e];
                if (lvl[w] == lvl[v] + 1) del[v] += sig[v] / sig[w] * (1.0 + del[w]);
            }
            if (v != s) bc[v] += del[v];
        }
    }
    free(lvl); free(order); free(sig); free(del);
    return 0;
}

/* The reference's graph coloring (graph_color_wrapper.cu:598-635): round c colors every
 * uncolored v none of whose out-neighbours u < v is uncolored or colored c.  In increasing id
 * order every smaller neighbour already has its final color, so one pass gives it:
 * color[v] = 1 + max color[u] over the edges v -> u with u < v (self loops have u == v). */
void orc_color(int n, const int *row, const int *col, int *color) {
    for (int v = 0; v < n; ++v) {
        int c = 0;
        for (int e = row[v]; e < row[v + 1]; ++e)
            if (col[e] < v && color[col[e]] > c) c = color[col[e]];
        color[v] = c + 1;
    }
}

/* PageRank as PAGERANK_GPU iterates it: from rank = total/n, each iteration
 *   new[v] = damping * sum over edges p -> v of rank[p]/outdeg[p]
 *            + (1 - damping) * total / n + damping * (sum of dangling ranks) / n,
 * until no vertex changes by more than eps (or max_iters).  Sums are pushed along out-edges in
 * vertex order (the order only moves the result by rounding).  Returns the iterations run. */
int orc_pagerank(int n, const int *row, const int *col, double damping, double total, double eps,
                 int max_iters, double *rank) {
    double *nr = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!nr) return -1;
    for (int v = 0; v < n; ++v) rank[v] = total / n;
    int it = 0;
    while (it < max_iters) {
        double dang = 0.0;
        for (int v = 0; v < n; ++v)
            if (row[v + 1] == row[v]) dang += rank[v];
        const double add = (1.0 - damping) * total / n + damping * dang / n;
        for (int v = 0; v < n; ++v) nr[v] = 0.0;
        for (int u = 0; u < n; ++u) {
            const int d = row[u + 1] - row[u];
            if (d == 0) continue;
            const double c = rank[u] / d;
            for (int e = row[u]; e < row[u + 1]; ++e) nr[col[e]] += c;
        }
        double mx = 0.0;
        for (int v = 0; v < n; ++v) {
            const double r = damping * nr[v] + add;
            const double dd = fabs(r - rank[v]);
            if (dd > mx) mx = dd;
            rank[v] = r;
        }
        ++it;
        if (!(mx > eps)) break;
    }
    free(nr);
    return it;
}

/* y[v] = sum over row v of data[e] * x[col[e]] (csr_spmv_kernel, SpMV_kernel.cu:77-92). */
void orc_spmv(int n, const int *row, const int *col, const double *data, const double *x, double *y) {
    for (int v = 0; v < n; ++v) {
        double s = 0.0;
        for (int e = row[v]; e < row[v + 1]; ++e) s += data[e] * x[col[e]];
        y[v] = s;
    }
}
