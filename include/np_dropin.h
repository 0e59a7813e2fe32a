/*
 * np_dropin.h -- the reference's own C++ entry points and globals, implemented by
 * libnpdropin.so over the C ABI in np_sssp.h.  A program written against the
 * reference's common/global.h + apps/sssp/src/gpu/sssp.h (or apps/bfs-rec/src/gpu/
 * bfs_rec.h) links this library instead of compiling sssp_wrapper.cu / bfs_rec_wrapper.cu
 * and common/global.cpp.
 *
 * Each declaration replaces the reference symbol of the same name:
 *   struct _GRAPH_, graph, noNodeTotal, noEdgeTotal, source, fp, VERBOSE, DEBUG
 *                                   common/global.h:28-42,48-58 (same member order/types)
 *   readInputDIMACS9/10/SLNDC       common/global.cpp:29-64, 66-105, 107-190
 *   readInput                       common/readInput.cpp:3-9 (a stub returning 0)
 *   convertCSR                      common/global.cpp:192-235
 *   outputCost                      common/global.cpp:369-375
 *   clear                           common/global.cpp:405-413
 *   gettime, gettime_ms             common/global.cpp:17-27
 *   CONF config, timing doubles     apps/sssp/src/gpu/sssp.h:13-30 (bfs_rec.h:15-32)
 *   SSSP_GPU                        apps/sssp/src/gpu/sssp_wrapper.cu:548-583
 *   BFS_REC_GPU                     apps/bfs-rec/src/gpu/bfs_rec_wrapper.cu:231-269
 *   BC_GPU, float *bc               apps/bc/src/gpu/bc_wrapper.cu:783-809, common/global.cpp:15
 *   outputBC                        common/global.cpp:377-383
 *   GRAPH_COLOR_GPU, outputColor    apps/graph-color/src/gpu/graph_color_wrapper.cu:662-699,
 *                                   common/global.cpp:393-399 (colors in graph.colorArray)
 *   PAGERANK_GPU, outputRank,       apps/pagerank/src/gpu/pagerank_wrapper.cu:631-661,
 *   invertGraph                     common/global.cpp:237-276, 385-391 (ranks in graph.rankArray)
 *   SpMV_GPU                        apps/SpMV/src/gpu/SpMV_wrapper.cu:97-113, 314-345: fills the
 *                                   caller's double *data, *x, *y (defined by its main, as in
 *                                   apps/SpMV/src/gpu/SpMV.cpp:9-11; referenced weakly here, so
 *                                   programs without them still link) and computes y = A x
 * The consolidation level is the CONSOLIDATE_LEVEL macro of the reference
 * (sssp_wrapper.cu:14-16, default 1 as in apps/sssp/src/gpu/Makefile:28; bfs_rec_wrapper.cu:
 * 20-22, default 2) given when libnpdropin.so is built, overridable at run time with the
 * environment variable NP_CONSOLIDATE_LEVEL.
 */
#ifndef NP_DROPIN_H
#define NP_DROPIN_H

#include <stdio.h>

struct _GRAPH_ {
    int *vertexArray;
    int *costArray;
    int *levelArray;
    int *edgeArray;
    int *weightArray;
    char *frontier;
    char *visited;
    char *update;
    int *childVertexArray;
    int *rEdgeArray;
    int *outdegreeArray;
    float *rankArray;
    int *colorArray;
};

typedef struct conf {
    bool verbose;
    bool debug;
    int data_set_format;
    int solution;
    int device_num;
    char *graph_file;
} CONF;

extern struct _GRAPH_ graph;
extern int noNodeTotal;
extern int noEdgeTotal;
extern int source;
extern FILE *fp;
extern int VERBOSE;
extern int DEBUG;

extern CONF config;
extern double start_time;
extern double end_time;
extern double init_time;
extern double d_malloc_time;
extern double h2d_memcpy_time;
extern double ker_exe_time;
extern double d2h_memcpy_time;

double gettime();
double gettime_ms();
int readInputDIMACS9();
int readInputDIMACS10();
int readInputSLNDC();
int readInput();
int convertCSR();
int outputCost(FILE *file);
int clear();

void SSSP_GPU();
void BFS_REC_GPU();

extern float *bc;
int outputBC(FILE *file);
void BC_GPU();

int outputColor(FILE *file);
void GRAPH_COLOR_GPU();

int invertGraph();
int outputRank(FILE *file);
void PAGERANK_GPU();

void SpMV_GPU();

#endif /* NP_DROPIN_H */
