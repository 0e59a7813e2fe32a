// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A thin extern "C" harness linked against the reference's own CPU sources, compiled
// from where they lie (common/global.cpp, common/heap.cpp, common/dheap.cpp,
// apps/sssp/src/cpu/sssp.cpp; apps/bfs-rec/src/gpu/bfs_rec.cpp for the BFS build,
// apps/bc/src/cpu/bc.cpp for the BC build) by
// oracle/Makefile into oracle/_ref/.  It only sets the reference's globals and calls
// its functions; it re-implements none of them.  The outputs (distances, pop counts,
// CSR arrays of the reference readers) pin the C restatement in np_oracle.c and the
// product's graph readers.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unistd.h>

#include "global.h"   // reference: struct _GRAPH_ graph, noNodeTotal, readInput*, convertCSR

#ifdef NP_REF_SSSP
#include "dheap.h"    // reference: class dheap (common/dheap.h:55-78)
int SSSP_CPU();       // reference: apps/sssp/src/cpu/sssp.cpp:13
#endif
#ifdef NP_REF_BFS
void bfs_cpu(int *levelArray);   // reference: apps/bfs-rec/src/gpu/bfs_rec.cpp:100
#endif
#ifdef NP_REF_BC
#include <omp.h>
void bc_cpu();                   // reference: apps/bc/src/cpu/bc.cpp:152
extern int *sigma;               // reference globals: apps/bc/src/cpu/bc.h:31-37
extern float *delta;
extern char *p;
extern int *p_idx, *p_data;
#endif

static void free_graph() {
    delete[] graph.vertexArray; graph.vertexArray = nullptr;
    delete[] graph.costArray;   graph.costArray = nullptr;
    delete[] graph.levelArray;  graph.levelArray = nullptr;
    delete[] graph.edgeArray;   graph.edgeArray = nullptr;
    delete[] graph.weightArray; graph.weightArray = nullptr;
    delete[] graph.frontier;    graph.frontier = nullptr;
    delete[] graph.update;      graph.update = nullptr;
    delete[] graph.visited;     graph.visited = nullptr;
}

static void set_graph(int n, int m, const int *row, const int *col, const int *w) {
    free_graph();
    noNodeTotal = n; noEdgeTotal = m;
    graph.vertexArray = new int[n + 1];
    graph.edgeArray = new int[m > 0 ? m : 1];
    graph.weightArray = new int[m > 0 ? m : 1];
    graph.costArray = new int[n > 0 ? n : 1];
    graph.levelArray = new int[n > 0 ? n : 1];
    std::memcpy(graph.vertexArray, row, sizeof(int) * (n + 1));
    if (m) std::memcpy(graph.edgeArray, col, sizeof(int) * m);
    if (m && w) std::memcpy(graph.weightArray, w, sizeof(int) * m);
}

extern "C" {

// Read a graph file with the reference reader (fmt 0 DIMACS9, 1 DIMACS10, 2 SLNDC,
// apps/sssp/src/gpu/sssp.cpp:120-125) followed by convertCSR (common/global.cpp:192).
int ref_read_graph(int fmt, const char *path, int *n, int *m, int *src) {
    free_graph();
    source = 0;
    fp = std::fopen(path, "r");
    if (!fp) return -1;
    int rc = 0;
    switch (fmt) {
        case 0: readInputDIMACS9(); break;
        case 1: readInputDIMACS10(); break;
        case 2: readInputSLNDC(); break;
        default: rc = -2;
    }
    std::fclose(fp); fp = nullptr;
    if (rc) return rc;
    convertCSR();
    *n = noNodeTotal; *m = noEdgeTotal; *src = source;
    return 0;
}

void ref_copy_csr(int *row, int *col, int *w) {
    std::memcpy(row, graph.vertexArray, sizeof(int) * (noNodeTotal + 1));
    std::memcpy(col, graph.edgeArray, sizeof(int) * noEdgeTotal);
    std::memcpy(w, graph.weightArray, sizeof(int) * noEdgeTotal);
    free_graph();
}

// outputCost (common/global.cpp:369-375) into a file, from the given distances.
int ref_output_cost(const char *path, int n, int src, const int *dist) {
    std::FILE *f = std::fopen(path, "w");
    if (!f) return -1;
    int *save = graph.costArray;
    int save_n = noNodeTotal, save_src = source;
    graph.costArray = const_cast<int *>(dist);
    noNodeTotal = n; source = src;
    outputCost(f);
    graph.costArray = save; noNodeTotal = save_n; source = save_src;
    std::fclose(f);
    return 0;
}

#ifdef NP_REF_SSSP
// The reference CPU Dijkstra SSSP_CPU() on the given CSR.  *iterations receives the
// "Iteration number" it prints (captured from stdout).
int ref_sssp_cpu(int n, int m, const int *row, const int *col, const int *w, int src, int *dist,
                 long long *iterations) {
    set_graph(n, m, row, col, w);
    source = src;
    std::fflush(stdout);
    std::FILE *tmp = std::tmpfile();
    int saved = dup(1);
    if (tmp) dup2(fileno(tmp), 1);
    SSSP_CPU();
    std::fflush(stdout);
    if (tmp) dup2(saved, 1);
    close(saved);
    long long it = -1;
    if (tmp) {
        std::rewind(tmp);
        char line[256];
        while (std::fgets(line, sizeof line, tmp))
            if (std::sscanf(line, "Iteration number: %lld", &it) == 1) break;
        std::fclose(tmp);
    }
    if (iterations) *iterations = it;
    std::memcpy(dist, graph.costArray, sizeof(int) * n);
    free_graph();
    return 0;
}

// Decrease-key Dijkstra driven over the reference's own dheap class
// (common/dheap.cpp:51-121).  Vertex v is heap item v+1.
int ref_dheap_dijkstra(int n, const int *row, const int *col, const int *w, int src, int arity,
                       int *dist) {
    dheap h((unsigned long)n, arity);
    char *done = new char[n]();
    for (int i = 0; i < n; ++i) dist[i] = INF;
    h.insert((item)src + 1, 0);
    while (!h.empty()) {
        item it = h.deletemin();
        int u = (int)(it - 1);
        long du = h.key(it);
        dist[u] = (int)du;
        done[u] = 1;
        for (int e = row[u]; e < row[u + 1]; ++e) {
            int v = col[e];
            if (done[v]) continue;
            long alt = du + w[e];
            if (h.member((item)v + 1)) {
                if (alt < h.key((item)v + 1)) h.changekey((item)v + 1, alt);
            } else {
                h.insert((item)v + 1, alt);
            }
        }
    }
    delete[] done;
    return 0;
}
#endif

#ifdef NP_REF_BFS
// The reference bfs_cpu() (apps/bfs-rec/src/gpu/bfs_rec.cpp:100-119), initialised the
// way its main() does (:194-196).
int ref_bfs_cpu(int n, int m, const int *row, const int *col, int src, int *level) {
    set_graph(n, m, row, col, nullptr);
    source = src;
    for (int i = 0; i < n; ++i) level[i] = 0x7FFFFFFF;
    level[src] = 0;
    bfs_cpu(level);
    free_graph();
    return 0;
}
#endif

#ifdef NP_REF_BC
// The reference bc_cpu() (apps/bc/src/cpu/bc.cpp:152-185 with bc_kernel.cpp): Brandes from every
// vertex, float bc.  Its outer OpenMP loop over sources shares sigma/delta/p/levelArray between
// threads, so it is run on one thread.  p is an n x n byte matrix: small graphs only.
int ref_bc_cpu(int n, int m, const int *row, const int *col, float *out) {
    set_graph(n, m, row, col, nullptr);
    graph.frontier = new char[n > 0 ? n : 1]();
    graph.update = new char[n > 0 ? n : 1]();
    omp_set_num_threads(1);
    bc_cpu();
    std::memcpy(out, bc, sizeof(float) * n);
    delete[] bc; bc = nullptr;
    delete[] delta; delta = nullptr;
    delete[] sigma; sigma = nullptr;
    delete[] p; p = nullptr;
    delete[] p_idx; p_idx = nullptr;
    delete[] p_data; p_data = nullptr;
    free_graph();
    return 0;
}
#endif

}  // extern "C"
