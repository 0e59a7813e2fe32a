// np_cli.cpp -- command-line drivers with the reference's flags and output, over the
// drop-in entry points.  Built three times: -DNP_CLI_SSSP gives gpu-sssp
// (apps/sssp/src/gpu/sssp.cpp:17-156), -DNP_CLI_BFS gives gpu-bfs-rec
// (apps/bfs-rec/src/gpu/bfs_rec.cpp:17-217), -DNP_CLI_BC gives gpu-bc
// (apps/bc/src/gpu/bc.cpp:7-140), -DNP_CLI_COLOR gives gpu-graph-color
// (apps/graph-color/src/gpu/graph_color.cpp:7-160), -DNP_CLI_PR gives gpu-pagerank
// (apps/pagerank/src/gpu/pagerank.cpp:7-190), -DNP_CLI_SPMV gives gpu-spmv
// (apps/SpMV/src/gpu/SpMV.cpp:7-165).  Flags: -h -v -d -f <fmt> -s <solution>
// -e <device> -i <file> (and -t, listed by the reference and ignored by its GPU binary).
// Extra flag -o <file>: write the costs (levels, BC values) after the runs.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <queue>

#include "np_dropin.h"

#define N 10  // runs per invocation, apps/sssp/src/gpu/sssp.cpp:3

static const char *g_out = nullptr;

#ifdef NP_CLI_SPMV
double *data = nullptr, *x = nullptr, *y = nullptr;  // apps/SpMV/src/gpu/SpMV.cpp:9-11
#endif

static void usage() {
    fprintf(stderr, "\n");
#if defined(NP_CLI_SSSP)
    fprintf(stderr, "Usage:  gpu-sssp [option]\n");
#elif defined(NP_CLI_BC)
    fprintf(stderr, "Usage:  gpu-bc [option]\n");
#elif defined(NP_CLI_COLOR)
    fprintf(stderr, "Usage:  gpu-graph-color [option]\n");
#elif defined(NP_CLI_PR)
    fprintf(stderr, "Usage:  gpu-pagerank [option]\n");
#elif defined(NP_CLI_SPMV)
    fprintf(stderr, "Usage:  gpu-spmv [option]\n");
#else
    fprintf(stderr, "Usage:  gpu-bfs-rec [option]\n");
#endif
    fprintf(stderr, "\nOptions:\n");
    fprintf(stderr, "    --help,-h      print this message\n");
    fprintf(stderr, "    --verbose,-v   basic verbosity level\n");
    fprintf(stderr, "    --debug,-d     enhanced verbosity level\n");
    fprintf(stderr, "\nOther:\n");
    fprintf(stderr, "    --import,-i <graph_file>           import graph file\n");
    fprintf(stderr, "    --thread,-t <number of threads>    specify number of threads\n");
    fprintf(stderr, "    --format,-f <number>               specify the input format\n");
    fprintf(stderr, "                 0 - DIMACS9\n");
    fprintf(stderr, "                 1 - DIMACS10\n");
    fprintf(stderr, "                 2 - SLNDC\n");
    fprintf(stderr, "    --solution,-s <number>             specify the solution\n");
#ifdef NP_CLI_SSSP
    fprintf(stderr, "                 0 - Unordered + thread queue\n");
    fprintf(stderr, "                 1 - dual queue\n");
    fprintf(stderr, "                 2 - shared delayed buffer\n");
    fprintf(stderr, "                 3 - global dedayed buffer\n");
    fprintf(stderr, "                 4 - multiple dynamic parallelism per block\n");
    fprintf(stderr, "                 5 - single dynamic parallelism per block\n");
    fprintf(stderr, "                 6 - workload consolidation with dynamic parallelism\n");
#elif defined(NP_CLI_BC)
    fprintf(stderr, "                 0 - thread bitmap\n");
    fprintf(stderr, "                 1 - thread queue\n");
    fprintf(stderr, "                 2 - dual queue\n");
    fprintf(stderr, "                 3 - shared delayed buffer\n");
    fprintf(stderr, "                 4 - global delayed buffer\n");
    fprintf(stderr, "                 5 - naive dp\n");
    fprintf(stderr, "                 6 - optimized dp\n");
    fprintf(stderr, "                 7 - workload consolidation for dynamic parallelism\n");
#elif defined(NP_CLI_COLOR)
    fprintf(stderr, "                 0-6   bitmap working set (no pruning ... optimized dp)\n");
    fprintf(stderr, "                 7     bitmap + workload consolidation\n");
    fprintf(stderr, "                 8-11  queue working set (delayed buffers, dp)\n");
    fprintf(stderr, "                 12    queue + workload consolidation\n");
#elif defined(NP_CLI_PR)
    fprintf(stderr, "                 0 - pull, 1 - dual queue, 2 - shared delayed buffer,\n");
    fprintf(stderr, "                 3 - global delayed buffer, 4 - naive dp, 5 - optimized dp,\n");
    fprintf(stderr, "                 6 - workload consolidation\n");
#elif defined(NP_CLI_SPMV)
    fprintf(stderr, "                 0 - thread per row, 1 - dual queue, 2 - shared delayed buffer,\n");
    fprintf(stderr, "                 3 - global delayed buffer, 4 - naive dp, 5 - optimized dp,\n");
    fprintf(stderr, "                 6 - workload consolidation\n");
#else
    fprintf(stderr, "                 0 - GPU flat\n");
    fprintf(stderr, "                 1 - GPU naive dynamic parallelism\n");
    fprintf(stderr, "                 2 - GPU hierachical dynamic parallelism\n");
    fprintf(stderr, "                 3 - GPU consolidated dynamic parallelism\n");
#endif
    fprintf(stderr, "    --device,-e <number>               select the device\n");
    fprintf(stderr, "    --output,-o <file>                 write costs (outputCost) after the runs\n");
}

static void print_conf() {
    fprintf(stderr, "\nCONFIGURATION:\n");
    if (config.graph_file) {
        fprintf(stderr, "- Graph file: %s\n", config.graph_file);
        fprintf(stderr, "- Graph format: %d\n", config.data_set_format);
    }
    fprintf(stderr, "- Solution: %d\n", config.solution);
#ifdef NP_CLI_BFS
    fprintf(stderr, "- GPU Device: %d\n", config.device_num);
#endif
    if (config.verbose && config.debug) fprintf(stderr, "- verbose mode\n");
    if (config.debug) fprintf(stderr, "- debug mode\n");
}

static void init_conf() {
    config.data_set_format = 0;
    config.solution = 0;
    config.device_num = 0;
    config.verbose = false;
    config.debug = false;
}

static int parse_arguments(int argc, char **argv) {
    int i = 1;
    if (argc < 2) { usage(); return 0; }
    while (i < argc) {
        if (!strcmp(argv[i], "-h") || !strcmp(argv[i], "--help")) { usage(); return 0; }
        else if (!strcmp(argv[i], "-v") || !strcmp(argv[i], "--verbose")) VERBOSE = config.verbose = 1;
        else if (!strcmp(argv[i], "-d") || !strcmp(argv[i], "--debug")) DEBUG = config.debug = 1;
        else if (!strcmp(argv[i], "-f") || !strcmp(argv[i], "--format")) { if (++i < argc) config.data_set_format = atoi(argv[i]); }
        else if (!strcmp(argv[i], "-s") || !strcmp(argv[i], "--solution")) { if (++i < argc) config.solution = atoi(argv[i]); }
        else if (!strcmp(argv[i], "-e") || !strcmp(argv[i], "--device")) { if (++i < argc) config.device_num = atoi(argv[i]); }
        else if (!strcmp(argv[i], "-t") || !strcmp(argv[i], "--thread")) { ++i; }
        else if (!strcmp(argv[i], "-o") || !strcmp(argv[i], "--output")) { if (++i < argc) g_out = argv[i]; }
        else if (!strcmp(argv[i], "-i") || !strcmp(argv[i], "--import")) {
            ++i;
            if (i == argc) { fprintf(stderr, "Graph file name missing.\n"); return 0; }
            config.graph_file = argv[i];
        }
        ++i;
    }
    return 1;
}

static int read_graph() {
    switch (config.data_set_format) {
        case 0: readInputDIMACS9(); break;
        case 1: readInputDIMACS10(); break;
        case 2: readInputSLNDC(); break;
        default: fprintf(stderr, "Wrong code for dataset\n"); return 0;
    }
    return 1;
}

#ifdef NP_CLI_BFS
// Host-side validation of the driver, as the reference's main does after its runs
// (bfs_cpu + validateArrays, apps/bfs-rec/src/gpu/bfs_rec.cpp:100-139,194-204).
// It checks the device result; it never replaces it.
static void validation_bfs(int *level) {
    const unsigned UNDEF = 0x7FFFFFFF;
    for (int i = 0; i < noNodeTotal; ++i) level[i] = (int)UNDEF;
    if (source < 0 || source >= noNodeTotal) return;  // empty graph: nothing reached
    level[source] = 0;
    std::queue<int> q;
    q.push(source);
    while (!q.empty()) {
        int node = q.front(); q.pop();
        unsigned next = (unsigned)level[node] + 1;
        for (int e = graph.vertexArray[node]; e < graph.vertexArray[node + 1]; e++) {
            int nb = graph.edgeArray[e];
            if ((unsigned)level[nb] == UNDEF || (unsigned)level[nb] > next) { level[nb] = (int)next; q.push(nb); }
        }
    }
}
#endif

int main(int argc, char *argv[]) {
    init_conf();
    if (!parse_arguments(argc, argv)) return 0;
    print_conf();
    if (config.graph_file != NULL) {
        fp = fopen(config.graph_file, "r");
        if (fp == NULL) { fprintf(stderr, "Error: NULL file pointer.\n"); return 1; }
    } else {
        return 0;
    }
    double t0 = gettime();
    if (!read_graph()) return 0;
    double t1 = gettime();
    if (VERBOSE) fprintf(stderr, "Source node is: %d\n", source + 1);
    if (VERBOSE) fprintf(stderr, "Read data:\t\t%lf\n", t1 - t0);
    t0 = gettime();
    convertCSR();
    t1 = gettime();
    if (VERBOSE) fprintf(stderr, "AdjList to CSR:\t\t%lf\n", t1 - t0);
#if defined(NP_CLI_SPMV)
    SpMV_GPU();  // N = 1 run (SpMV.cpp:3)
    if (VERBOSE) {
        fprintf(stdout, "CUDA runtime initialization:\t\t%lf\n", init_time);
        fprintf(stdout, "CUDA H2D cudaMemcpy:\t\t%lf\n", h2d_memcpy_time);
        fprintf(stdout, "CUDA kernel execution:\t\t%lf\n", ker_exe_time);
        fprintf(stdout, "CUDA D2H cudaMemcpy:\t\t%lf\n", d2h_memcpy_time);
    }
    if (DEBUG)
        for (int i = 0; i < noNodeTotal; ++i) {
            fprintf(stderr, "x[%d] = %f\n", i, x[i]);
            fprintf(stderr, "y[%d] = %f\n", i, y[i]);
        }
    if (g_out) {
        FILE *f = fopen(g_out, "w");
        if (f) {
            for (int i = 0; i < noNodeTotal; ++i) fprintf(f, "y[%d] = %.17g\n", i, y[i]);
            fclose(f);
        }
    }
    delete[] data; delete[] x; delete[] y;
    data = x = y = nullptr;
#elif defined(NP_CLI_PR)
    invertGraph();
    graph.rankArray = new float[noNodeTotal > 0 ? noNodeTotal : 1];
    for (int i = 0; i < noNodeTotal; ++i) graph.rankArray[i] = 1000.0f / noNodeTotal;  // pagerank.cpp:103-107
    PAGERANK_GPU();  // N = 1 run (pagerank.cpp:3)
    if (VERBOSE) {
        fprintf(stdout, "CUDA runtime initialization:\t\t%lf\n", init_time);
        fprintf(stdout, "CUDA H2D cudaMemcpy:\t\t%lf\n", h2d_memcpy_time);
        fprintf(stdout, "CUDA kernel execution:\t\t%lf\n", ker_exe_time);
        fprintf(stdout, "CUDA D2H cudaMemcpy:\t\t%lf\n", d2h_memcpy_time);
    }
    if (DEBUG) outputRank(stdout);
    if (g_out) {
        FILE *f = fopen(g_out, "w");
        if (f) { outputRank(f); fclose(f); }
    }
    delete[] graph.rankArray;
    graph.rankArray = nullptr;
    delete[] graph.childVertexArray;
    graph.childVertexArray = nullptr;
    delete[] graph.rEdgeArray;
    graph.rEdgeArray = nullptr;
#elif defined(NP_CLI_COLOR)
    graph.colorArray = new int[noNodeTotal > 0 ? noNodeTotal : 1]();
    GRAPH_COLOR_GPU();
    if (DEBUG) outputColor(stdout);
    if (g_out) {
        FILE *f = fopen(g_out, "w");
        if (f) { outputColor(f); fclose(f); }
    }
    delete[] graph.colorArray;
    graph.colorArray = nullptr;
#elif defined(NP_CLI_BC)
    bc = new float[noNodeTotal > 0 ? noNodeTotal : 1];
    BC_GPU();
    if (DEBUG) outputBC(stdout);
    if (g_out) {
        FILE *f = fopen(g_out, "w");
        if (f) { outputBC(f); fclose(f); }
    }
    delete[] bc;
    bc = nullptr;
#elif defined(NP_CLI_SSSP)
    for (int i = 0; i < N; ++i) SSSP_GPU();
    if (VERBOSE) {
        fprintf(stdout, "CUDA runtime initialization:\t\t%lf\n", init_time / N);
        fprintf(stdout, "CUDA cudaMalloc:\t\t%lf\n", d_malloc_time / N);
        fprintf(stdout, "CUDA H2D cudaMemcpy:\t\t%lf\n", h2d_memcpy_time / N);
        fprintf(stdout, "CUDA kernel execution:\t\t%lf\n", ker_exe_time / N);
        fprintf(stdout, "CUDA D2H cudaMemcpy:\t\t%lf\n", d2h_memcpy_time / N);
    }
    if (g_out) {
        FILE *f = fopen(g_out, "w");
        if (f) { outputCost(f); fclose(f); }
    }
#else
    printf("\n===MAIN=== :: [num_nodes,num_edges] = %u, %u\n", noNodeTotal, noEdgeTotal);
    for (int i = 0; i < N; ++i) {
        for (int k = 0; k < noNodeTotal; ++k) graph.levelArray[k] = 0x7FFFFFFF;
        if (source >= 0 && source < noNodeTotal) graph.levelArray[source] = 0;
        BFS_REC_GPU();
    }
    int *cpu = new int[noNodeTotal > 0 ? noNodeTotal : 1];
    validation_bfs(cpu);
    int flag = 1;
    for (int node = 0; node < noNodeTotal; node++) {
        if (graph.levelArray[node] != cpu[node]) {
            printf("ERROR: validation error at %d: %s !\n", node, "GPU bfs rec");
            printf("Node %d : %d v.s. %d\n", node, graph.levelArray[node], cpu[node]);
            flag = 0;
            break;
        }
    }
    if (flag) printf("PASS: %s !\n", "GPU bfs rec");
    delete[] cpu;
    if (VERBOSE) {
        fprintf(stdout, "===MAIN=== :: CUDA runtime init:\t\t%.2lf ms.\n", init_time / N);
        fprintf(stdout, "===MAIN=== :: CUDA device cudaMalloc:\t\t%.2lf ms.\n", d_malloc_time / N);
        fprintf(stdout, "===MAIN=== :: CUDA H2D cudaMemcpy:\t\t%.2lf ms.\n", h2d_memcpy_time / N);
        fprintf(stdout, "===MAIN=== :: CUDA kernel execution:\t\t%.2lf ms.\n", ker_exe_time / N);
        fprintf(stdout, "===MAIN=== :: CUDA D2H cudaMemcpy:\t\t%.2lf ms.\n", d2h_memcpy_time / N);
    }
    if (g_out) {
        FILE *f = fopen(g_out, "w");
        if (f) {
            for (int i = 0; i < noNodeTotal; ++i) fprintf(f, "Node %d is in level %d\n", i + 1, graph.levelArray[i]);
            fclose(f);
        }
    }
#endif
    fclose(fp);
    clear();
    return 0;
}
