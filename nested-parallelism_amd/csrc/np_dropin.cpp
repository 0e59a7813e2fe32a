// np_dropin.cpp -- the reference's C++ entry points (include/np_dropin.h) over the C ABI.
//
// Globals, readers and output helpers keep the reference's names, types and effects
// (common/global.cpp); SSSP_GPU / BFS_REC_GPU keep their contract: read the global CSR
// and `source`, write graph.costArray / graph.levelArray, accumulate the phase timers
// (apps/sssp/src/gpu/sssp_wrapper.cu:548-583, apps/bfs-rec/src/gpu/bfs_rec_wrapper.cu:
// 231-269).  Error behaviour follows the reference: a reader failure prints and exits
// with status 0 (the EXIT macro, common/global.h:21-24); a device failure prints
// "Error: ..." and exits with status 1 (cudaCheckError, sssp_wrapper.cu:50-56).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <sys/time.h>

#include "np_dropin.h"
#include "np_sssp.h"

#ifndef CONSOLIDATE_LEVEL
#define CONSOLIDATE_LEVEL 1      // apps/sssp/src/gpu/Makefile:28
#endif
#ifndef BFS_CONSOLIDATE_LEVEL
#define BFS_CONSOLIDATE_LEVEL 2  // apps/bfs-rec/src/gpu/bfs_rec_wrapper.cu:20-22
#endif

struct _GRAPH_ graph;
int noNodeTotal = 0;
int noEdgeTotal = 0;
int source = 0;
FILE *fp = NULL;
int VERBOSE;
int DEBUG;

CONF config;
double start_time = 0;
double end_time = 0;
double init_time = 0;
double d_malloc_time = 0;
double h2d_memcpy_time = 0;
double ker_exe_time = 0;
double d2h_memcpy_time = 0;

namespace {

int32_t *s_row = nullptr, *s_col = nullptr, *s_w = nullptr;  // CSR staged by readInput*
np_ctx *s_ctx = nullptr;
int s_ctx_dev = -1;

void drop_staged() {
    np_host_free(s_row); np_host_free(s_col); np_host_free(s_w);
    s_row = s_col = s_w = nullptr;
}

int read_common(int fmt) {
    drop_staged();
    int32_t n = 0, src = 0;
    int64_t m = 0;
    if (!fp) {
        fprintf(stderr, "info: %s:%d: %s", __FILE__, __LINE__, "NULL file pointer\n");
        exit(0);
    }
    int rc = np_read_graph_stream(fmt, fp, &n, &m, &src, &s_row, &s_col, &s_w);
    if (rc != NP_OK) {
        fprintf(stderr, "info: %s:%d: %s\n", __FILE__, __LINE__, np_last_error());
        exit(0);
    }
    noNodeTotal = n;
    noEdgeTotal = (int)m;
    if (fmt == 0) source = src;  // only DIMACS9 carries a source line
    return 0;
}

void fail(const char *what) {
    printf("Error: %s: %s\n", what, np_last_error());
    exit(1);
}

np_ctx *ctx_for_device() {
    if (s_ctx && s_ctx_dev == config.device_num) return s_ctx;
    if (s_ctx) np_destroy(s_ctx);
    s_ctx = nullptr;
    if (np_create(config.device_num, &s_ctx) != NP_OK) fail("np_create");
    s_ctx_dev = config.device_num;
    return s_ctx;
}

int env_level(int dflt) {
    const char *e = getenv("NP_CONSOLIDATE_LEVEL");
    if (e && *e) {
        int v = atoi(e);
        if (v >= 0 && v <= 2) return v;
    }
    return dflt;
}

// Schedules for config.solution (apps/sssp/src/gpu/sssp.cpp:31-38, sssp_wrapper.cu:553-570).
// Every solution relaxes with the same arithmetic; they differ in how frontier vertices are
// mapped to workers.  Solutions 0-5 keep the reference's shape -- a frontier queue per
// iteration, a lane per light vertex, heavy vertices (out-degree >= THRESHOLD 64) handed on --
// and run on the host loop with one readback per iteration, as sssp_wrapper.cu does:
//   0 thread queue            NP_SCHEME_THREAD        (sssp_tqueue_gpu, :151-191)
//   1 dual queue              NP_SCHEME_DUAL_QUEUE    (sssp_dual_queue_gpu, :193-293)
//   2 shared delayed buffer   NP_SCHEME_LDS_BUFFER    (:295-331)
//   3 global delayed buffer   NP_SCHEME_GLOBAL_BUFFER (:333-396)
//   4 nested launch per heavy vertex   NP_SCHEME_WAVE_NESTED (sssp_np_naive_gpu, :398-433)
//   5 nested launch per block          NP_SCHEME_WG_NESTED   (sssp_np_opt_gpu, :435-473)
//   6 consolidation           NP_SCHEME_CONSOLIDATE at CONSOLIDATE_LEVEL (:475-546)
bool sssp_schedule(int solution, np_sched *s) {
    static const int scheme[6] = {NP_SCHEME_THREAD, NP_SCHEME_DUAL_QUEUE, NP_SCHEME_LDS_BUFFER,
                                  NP_SCHEME_GLOBAL_BUFFER, NP_SCHEME_WAVE_NESTED, NP_SCHEME_WG_NESTED};
    if (solution == 6) { np_sched_default(env_level(CONSOLIDATE_LEVEL), s); return true; }
    if (solution < 0 || solution > 5) return false;
    np_sched_default(1, s);
    s->scheme = scheme[solution];
    s->drive = NP_DRIVE_HOSTLOOP;
    return true;
}

// Schedules for BFS_REC_GPU's config.solution (apps/bfs-rec/src/gpu/bfs_rec.cpp:31-34,
// bfs_rec_wrapper.cu:241-262):
//   0 flat                    NP_SCHEME_BFS_FLAT: a scan of every vertex per level
//   1 naive recursion         a workgroup per frontier vertex (the reference's one child
//                             block per vertex), improved children appended directly
//                             (the push frontier: no scan, as a nested launch needs none)
//   2 hierarchical recursion  NP_SCHEME_BFS_HIER: two levels per expansion
//   3 consolidation           NP_SCHEME_CONSOLIDATE at CONSOLIDATE_LEVEL
bool bfs_schedule(int solution, np_sched *s) {
    switch (solution) {
        case 0: np_sched_default(1, s); s->scheme = NP_SCHEME_BFS_FLAT; s->drive = NP_DRIVE_HOSTLOOP; return true;
        case 1:
            np_sched_default(1, s);
            s->log2_t_lo = 0; s->log2_t_hi = 0;   // every vertex of out-degree >= 1 in the workgroup bin
            s->frontier = NP_FRONTIER_PUSH; s->drive = NP_DRIVE_HOSTLOOP;
            return true;
        case 2: np_sched_default(1, s); s->scheme = NP_SCHEME_BFS_HIER; s->drive = NP_DRIVE_HOSTLOOP; return true;
        case 3: np_sched_default(env_level(BFS_CONSOLIDATE_LEVEL), s); return true;
        default: return false;
    }
}

np_graph *upload_graph(np_ctx *c) {
    np_graph *g = nullptr;
    double t0 = gettime();
    int rc = np_graph_upload(c, noNodeTotal, noEdgeTotal, graph.vertexArray, graph.edgeArray,
                             graph.weightArray, &g);
    double t1 = gettime();
    if (rc != NP_OK) fail("graph upload");
    h2d_memcpy_time += t1 - t0;
    return g;
}

}  // namespace

double gettime() {
    struct timeval t;
    gettimeofday(&t, NULL);
    return t.tv_sec + t.tv_usec * 1e-6;
}

double gettime_ms() {
    struct timeval t;
    gettimeofday(&t, NULL);
    return t.tv_sec * 1e+3 + t.tv_usec * 1e-3;
}

int readInputDIMACS9() { return read_common(0); }
int readInputDIMACS10() { return read_common(1); }
int readInputSLNDC() { return read_common(2); }
int readInput() { return 0; }  // common/readInput.cpp:3-9

int convertCSR() {
    const int n = noNodeTotal, m = noEdgeTotal;
    graph.vertexArray = new int[n + 1]();
    graph.costArray = new int[n]();
    graph.levelArray = new int[n]();
    graph.edgeArray = new int[m]();
    graph.weightArray = new int[m]();
    graph.frontier = new char[n]();
    graph.update = new char[n]();
    graph.visited = new char[n]();
    if (s_row) {
        memcpy(graph.vertexArray, s_row, sizeof(int) * ((size_t)n + 1));
        if (m) {
            memcpy(graph.edgeArray, s_col, sizeof(int) * (size_t)m);
            memcpy(graph.weightArray, s_w, sizeof(int) * (size_t)m);
        }
    }
    graph.vertexArray[n] = noEdgeTotal;
    if (VERBOSE) {
        fprintf(stderr, "Node no: %d\n", noNodeTotal);
        fprintf(stderr, "Edge no: %d\n", noEdgeTotal);
    }
    drop_staged();
    return 0;
}

float *bc = nullptr;  // common/global.cpp:15; the caller's main allocates n floats

int outputBC(FILE *file) {
    if (!file) file = stdout;
    fprintf(stderr, "------- BC value ------\n");
    for (int i = 0; i < noNodeTotal; ++i) fprintf(file, "Node %d: %f\n", i + 1, bc[i]);
    return 0;
}

int outputRank(FILE *file) {
    if (!file) file = stdout;
    fprintf(stderr, "------- Rank value ------\n");
    for (int i = 0; i < noNodeTotal; ++i) fprintf(file, "Node %d: %0.10lf\n", i + 1, (double)graph.rankArray[i]);
    return 0;
}

// invertGraph (common/global.cpp:237-276): the in-edge CSR in graph.childVertexArray /
// graph.rEdgeArray, each vertex's parents in decreasing source order (the reference pushes them
// in increasing order and pops from the back).  PAGERANK_GPU builds its own on the device;
// this keeps the host arrays a caller of the reference may read.
int invertGraph() {
    delete[] graph.childVertexArray;
    delete[] graph.rEdgeArray;
    const int n = noNodeTotal;
    graph.childVertexArray = new int[n + 1]();
    graph.rEdgeArray = new int[noEdgeTotal > 0 ? noEdgeTotal : 1]();
    for (int u = 0; u < n; ++u)
        for (int e = graph.vertexArray[u]; e < graph.vertexArray[u + 1]; ++e) graph.childVertexArray[graph.edgeArray[e] + 1]++;
    for (int i = 0; i < n; ++i) graph.childVertexArray[i + 1] += graph.childVertexArray[i];
    std::vector<int> fill(graph.childVertexArray + 1, graph.childVertexArray + n + 1);  // ends, filled downwards
    for (int u = 0; u < n; ++u)
        for (int e = graph.vertexArray[u]; e < graph.vertexArray[u + 1]; ++e) {
            const int t = graph.edgeArray[e];
            graph.rEdgeArray[--fill[t]] = u;
        }
    return 0;
}

int outputColor(FILE *file) {
    if (!file) file = stdout;
    fprintf(stderr, "------- Color value ------\n");
    for (int i = 0; i < noNodeTotal; ++i) fprintf(file, "Node %d: %d\n", i + 1, graph.colorArray[i]);
    return 0;
}

int outputCost(FILE *file) {
    if (!file) file = stdout;
    for (int i = 0; i < noNodeTotal; ++i)
        fprintf(file, "From Node %d -> %d is %d\n", source + 1, i + 1, graph.costArray[i]);
    return 0;
}

int clear() {
    delete[] graph.vertexArray; graph.vertexArray = nullptr;
    delete[] graph.edgeArray; graph.edgeArray = nullptr;
    delete[] graph.weightArray; graph.weightArray = nullptr;
    delete[] graph.frontier; graph.frontier = nullptr;
    return 0;
}

void SSSP_GPU() {
    start_time = gettime();
    np_ctx *c = ctx_for_device();
    end_time = gettime();
    init_time += end_time - start_time;
    if (DEBUG) fprintf(stderr, "Choose CUDA device: %d\n", config.device_num);
    if (noNodeTotal <= 0) return;  // an empty graph: nothing to solve
    if (source < 0 || source >= noNodeTotal) {
        fprintf(stderr, "Error: source %d outside the graph's %d vertices\n", source + 1, noNodeTotal);
        exit(1);
    }
    // host-side initialisation of prepare_gpu (sssp_wrapper.cu:92-96)
    for (int i = 0; i < noNodeTotal; i++) graph.costArray[i] = NP_INF;
    if (graph.update) graph.update[source] = 1;
    graph.costArray[source] = 0;
    np_graph *g = upload_graph(c);
    np_sched s;
    np_stats st;
    memset(&st, 0, sizeof st);
    start_time = gettime();
    if (sssp_schedule(config.solution, &s)) {
        if (np_sssp(c, g, source, &s, graph.costArray, &st) != NP_OK) fail("SSSP");
        if (config.solution == 6) printf("iteration: %d\n", st.iterations);
    }
    end_time = gettime();
    const double kernel_s = st.kernel_ms * 1e-3;
    ker_exe_time += kernel_s;
    d2h_memcpy_time += (end_time - start_time) - kernel_s > 0 ? (end_time - start_time) - kernel_s : 0;
    if (DEBUG) fprintf(stderr, "Iteration: %d\n", st.iterations);
    np_graph_free(g);
}

void BFS_REC_GPU() {
    if (noNodeTotal <= 0) return;  // an empty graph: nothing to solve
    start_time = gettime_ms();
    np_ctx *c = ctx_for_device();
    end_time = gettime_ms();
    init_time += end_time - start_time;
    np_graph *g = nullptr;
    start_time = gettime_ms();
    if (np_graph_upload(c, noNodeTotal, noEdgeTotal, graph.vertexArray, graph.edgeArray,
                        graph.weightArray, &g) != NP_OK)
        fail("graph upload");
    end_time = gettime_ms();
    h2d_memcpy_time += end_time - start_time;
    np_stats st;
    memset(&st, 0, sizeof st);
    double t0 = gettime_ms();
    np_sched s;
    if (bfs_schedule(config.solution, &s)) {
        if (np_bfs(c, g, source, &s, graph.levelArray, &st) != NP_OK) fail("BFS");
        if (DEBUG) printf("===> GPU #%d - BFS levels in %d iterations.\n", config.solution + 1, st.iterations);
    }
    double t1 = gettime_ms();
    ker_exe_time += st.kernel_ms;
    d2h_memcpy_time += (t1 - t0) - st.kernel_ms > 0 ? (t1 - t0) - st.kernel_ms : 0;
    np_graph_free(g);
}

// BC_GPU (apps/bc/src/gpu/bc_wrapper.cu:783-809): every vertex is a source.  Solutions 0-7
// differ only in how frontier work is assigned to threads; each runs np_bc (7 is the
// consolidated one).  Other solution numbers compute nothing and leave bc zero, as the
// reference's switch does.  The timing lines are the reference's.
void BC_GPU() {
    np_ctx *c = ctx_for_device();
    np_graph *g = upload_graph(c);
    np_stats st;
    memset(&st, 0, sizeof st);
    start_time = gettime();
    if (config.solution >= 0 && config.solution <= 7) {
        if (np_bc(c, g, nullptr, 0, bc, &st) != NP_OK) fail("BC");
    } else {
        for (int i = 0; i < noNodeTotal; ++i) bc[i] = 0.f;
    }
    end_time = gettime();
    fprintf(stdout, "Execution time:\t\t%lf\n", st.kernel_ms * 1e-3);
    fprintf(stdout, "cudaMemcpy to CPU:\t\t%lf\n", (end_time - start_time) - st.kernel_ms * 1e-3);
    np_graph_free(g);
}

// GRAPH_COLOR_GPU (apps/graph-color/src/gpu/graph_color_wrapper.cu:662-699): solutions 0-12
// differ in how the rounds assign vertices to threads and all give the same colors; each runs
// np_color (dependency counting).  Other solution numbers compute nothing, as the reference's
// switch does.  The caller's main allocates graph.colorArray (graph_color.cpp:99-105).
void GRAPH_COLOR_GPU() {
    np_ctx *c = ctx_for_device();
    np_graph *g = upload_graph(c);
    np_stats st;
    memset(&st, 0, sizeof st);
    start_time = gettime();
    if (config.solution >= 0 && config.solution <= 12) {
        if (np_color(c, g, 0, graph.colorArray, &st) != NP_OK) fail("graph coloring");
        if (DEBUG) fprintf(stderr, "Graph Coloring ends in %d iterations.\n", st.iterations);
    }
    end_time = gettime();
    fprintf(stderr, "Execution time:\t\t%lf\n", end_time - start_time);
    np_graph_free(g);
}

// PAGERANK_GPU (apps/pagerank/src/gpu/pagerank_wrapper.cu:631-661): solutions 0-6 differ in how
// the pull is assigned to threads and converge to the same ranks; each runs np_pagerank.  The
// caller's main allocates graph.rankArray (pagerank.cpp:103-107).
void PAGERANK_GPU() {
    np_ctx *c = ctx_for_device();
    np_graph *g = upload_graph(c);
    np_stats st;
    memset(&st, 0, sizeof st);
    start_time = gettime();
    if (config.solution >= 0 && config.solution <= 6) {
        if (np_pagerank(c, g, 0, graph.rankArray, &st) != NP_OK) fail("PageRank");
        if (DEBUG) fprintf(stderr, "PageRank ends in %d iterations.\n", st.iterations);
    }
    end_time = gettime();
    ker_exe_time += st.kernel_ms * 1e-3;
    d2h_memcpy_time += (end_time - start_time) - st.kernel_ms * 1e-3 > 0 ? (end_time - start_time) - st.kernel_ms * 1e-3 : 0;
    np_graph_free(g);
}

// SpMV_GPU (apps/SpMV/src/gpu/SpMV_wrapper.cu:97-113, 314-345): allocates and fills the caller's
// data (i % 13 + 0.5 per edge) and x (i % 17 + 0.5 per vertex) as the reference's prepare_gpu
// does, then y = A x over the global CSR (solutions 0-6 all compute this product).  The arrays
// are the main program's globals; they are weak references here, so a program that defines
// none of them (the other drivers) still loads, and SpMV_GPU then stops with an error.
extern double *data __attribute__((weak));
extern double *x __attribute__((weak));
extern double *y __attribute__((weak));

void SpMV_GPU() {
    if (!&data || !&x || !&y) fail("SpMV_GPU needs the globals double *data, *x, *y of its main program");
    np_ctx *c = ctx_for_device();
    np_graph *g = upload_graph(c);
    delete[] data; delete[] x; delete[] y;
    data = new double[noEdgeTotal > 0 ? noEdgeTotal : 1];
    x = new double[noNodeTotal > 0 ? noNodeTotal : 1];
    y = new double[noNodeTotal > 0 ? noNodeTotal : 1]();
    for (int i = 0; i < noEdgeTotal; ++i) data[i] = i % 13 + 0.5;
    for (int i = 0; i < noNodeTotal; ++i) x[i] = i % 17 + 0.5;
    np_stats st;
    memset(&st, 0, sizeof st);
    start_time = gettime();
    if (config.solution >= 0 && config.solution <= 6) {
        if (np_spmv(c, g, data, x, y, &st) != NP_OK) fail("SpMV");
    }
    end_time = gettime();
    ker_exe_time += st.kernel_ms * 1e-3;
    d2h_memcpy_time += (end_time - start_time) - st.kernel_ms * 1e-3 > 0 ? (end_time - start_time) - st.kernel_ms * 1e-3 : 0;
    np_graph_free(g);
}
