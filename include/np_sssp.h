/*
 * np_sssp.h -- C ABI of the MI355X (gfx950) SSSP / recursive-BFS nested-parallelism path.
 *
 * Plain pointers and sizes only.  Every entry point returns an int status (NP_OK or a
 * negative NP_ERR_* code) and records a message readable with np_last_error() (per
 * host thread).  The library owns one HIP stream per context; a context is not
 * reentrant, separate contexts are.
 *
 * Graph layout is the reference CSR of common/global.h:28-42 as produced by
 * convertCSR() (common/global.cpp:192-235): row[n+1] (vertexArray, row[n] = m),
 * col[m] (edgeArray), w[m] (weightArray), all int32.  Distances are int32 with
 * unreached = NP_INF (= INF, common/global.h:19); BFS levels use NP_UNDEFINED
 * (apps/bfs-rec/src/gpu/bfs_rec.h:13).
 *
 * Reference interfaces replaced (the reference has no C ABI: its host code calls the
 * CUDA wrappers directly, so these stand where those calls were):
 *   np_graph_upload          prepare_gpu()'s cudaMalloc + H2D of the CSR
 *                            (apps/sssp/src/gpu/sssp_wrapper.cu:58-140)
 *   np_sssp                  sssp_np_consolidate_gpu() + the D2H of costArray
 *                            (sssp_wrapper.cu:475-546, :576), i.e. SSSP_GPU() with
 *                            config.solution == 6 (sssp_wrapper.cu:548-583)
 *   np_bfs                   bfs_rec_dp_cons_gpu() + D2H of levelArray
 *                            (apps/bfs-rec/src/gpu/bfs_rec_wrapper.cu:170-229, :264)
 *   np_sched.level           the CONSOLIDATE_LEVEL macro (sssp_wrapper.cu:14-16,503-531;
 *                            bfs_rec_wrapper.cu:20-22,180-214), now a runtime value
 *   np_sched.log2_t_lo       THRESHOLD (apps/sssp/src/gpu/sssp_kernel.cu:7), as a power of two
 *   np_graph_free/np_destroy clean_gpu() (sssp_wrapper.cu:142-149)
 *   np_read_graph            readInputDIMACS9/DIMACS10/SLNDC + convertCSR
 *                            (common/global.cpp:29-235)
 * Entry points without a counterpart in the reference (batching, the learned
 * scheduler, the generators, the autotuner) say so in their comment.
 */
#ifndef NP_SSSP_H
#define NP_SSSP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NP_INF 1073741824      /* common/global.h:19 */
#define NP_UNDEFINED 0x7FFFFFFF /* apps/bfs-rec/src/gpu/bfs_rec.h:13 */

enum {
    NP_OK = 0,
    NP_ERR_ARG = -1,           /* bad argument (NULL, source out of range, bad sizes) */
    NP_ERR_HIP = -2,           /* a HIP runtime call failed */
    NP_ERR_NOMEM = -3,         /* device or host allocation failed */
    NP_ERR_CAPACITY = -4,      /* a size limit of the path was exceeded */
    NP_ERR_NOT_CONVERGED = -5, /* iteration cap reached before the frontier emptied */
    NP_ERR_TIMEOUT = -6,       /* a grid barrier's bounded spin expired (kernel exited) */
    NP_ERR_UNSUPPORTED = -7,   /* no usable gfx950 device / feature */
    NP_ERR_IO = -8,            /* file could not be read or is malformed */
    NP_ERR_BOUNDS = -9         /* bounds-checked build (libnpsssp_bounds.so) only: a frontier-queue or
                                  delayed-buffer write fell outside its segment and was skipped */
};

/* Consolidation level (CONSOLIDATE_LEVEL: 0 warp, 1 block, 2 grid). */
enum { NP_LEVEL_WAVE = 0, NP_LEVEL_WORKGROUP = 1, NP_LEVEL_GRID = 2, NP_LEVEL_LEARNED = -1 };
/* How the iteration loop is driven. */
enum {
    NP_DRIVE_PERSISTENT = 0, /* one cooperative launch runs every iteration (grid barriers) */
    NP_DRIVE_HOSTLOOP = 1,   /* kernels per iteration, 4-byte frontier readback per iteration
                                (the reference's loop shape, sssp_wrapper.cu:490-541) */
    NP_DRIVE_GRAPH = 2       /* iterations captured in a hipGraph, replayed in chunks */
};
/* How the next frontier is formed. */
enum {
    NP_FRONTIER_SCAN = 0, /* update bitmap compacted after each expansion
                             (unorder_generateQueue_kernel, sssp_kernel.cu:648-658) */
    NP_FRONTIER_PUSH = 1  /* relaxers append improved vertices directly (no O(n) scan) */
};

/* Schedule of one solve.  Bins by out-degree d of a frontier vertex:
 *   d <  2^log2_t_lo                        lane bin: wavefront-wide balanced gather
 *   2^log2_t_lo <= d < 2^log2_t_hi          wavefront bin: one 64-lane wavefront per vertex
 *   d >= 2^log2_t_hi (and >= 2^log2_t_lo)   level 0: wavefront; level 1: one workgroup per
 *                                           vertex; level 2: grid-wide, split in edge chunks
 * The reference's fixed heuristic is log2_t_lo = log2_t_hi = 6 (THRESHOLD 64). */
typedef struct np_sched {
    int32_t level;       /* NP_LEVEL_*; NP_LEVEL_LEARNED = the trained MLP picks per iteration */
    int32_t log2_t_lo;   /* 0..30 */
    int32_t log2_t_hi;   /* 0..30 */
    int32_t drive;       /* NP_DRIVE_* */
    int32_t frontier;    /* NP_FRONTIER_* */
    int32_t wgs_per_cu;  /* persistent grid: workgroups per CU (0 = occupancy maximum) */
    int32_t graph_chunk; /* NP_DRIVE_GRAPH: iterations per captured chunk (0 = 16) */
    int32_t max_iters;   /* 0 = no cap beyond n+1 */
    int32_t trace;       /* nonzero: record per-iteration trace (np_trace_*) */
    int32_t scheme;      /* NP_SCHEME_*: 0 = the consolidation bins above; others are the
                            reference's other schedules (config.solution) */
    int32_t delta;       /* scan frontier: priority window.  0 = every flagged vertex is
                            expanded (the reference's unordered frontier); > 0 = only
                            vertices with dist < T are expanded, T rising to
                            min(pending) + delta when none is left.  Changes the order of
                            the relaxations only, never a distance. */
    int32_t pack_hold;   /* source-packed batches without a priority window: c in 1..31 holds back,
                            for one compaction pass, a vertex of out-degree >= 2^(c-1) that fewer
                            than 3/4 of its pack's sources have reached yet, so that the sources
                            arriving in the next iteration are expanded with the pending ones in
                            one read of its edges.  0 = never.  Order only, never a distance. */
    int32_t reserved[4];
} np_sched;


/* np_sched.scheme: how queued vertices are mapped to workers.  Schemes 1-8 are the
 * reference's other schedules behind config.solution (apps/sssp/src/gpu/sssp_wrapper.cu:
 * 553-570, apps/bfs-rec/src/gpu/bfs_rec_wrapper.cu:241-262); a vertex is heavy when its
 * out-degree is >= 2^log2_t_lo (THRESHOLD 64, apps/sssp/src/gpu/sssp_kernel.cu:7).  They run on
 * the host-loop drive (a persistent request runs host-loop) or the graph drive, with the scan
 * frontier and no priority window, on single-source solves (np_sssp, np_bfs); level, the
 * learned scheduler and log2_t_hi do not apply to them. */
enum {
    NP_SCHEME_CONSOLIDATE = 0,   /* degree bins of one launch (solution 6; bfs-rec 3) */
    NP_SCHEME_THREAD = 1,        /* a lane walks all edges of one vertex (solution 0,
                                    unorder_threadQueue_kernel, sssp_kernel.cu:34-55) */
    NP_SCHEME_DUAL_QUEUE = 2,    /* light vertices a lane each, then heavy ones a workgroup each,
                                    two launches (solution 1, sssp_wrapper.cu:193-293) */
    NP_SCHEME_LDS_BUFFER = 3,    /* light a lane each; heavy into a per-workgroup LDS buffer that
                                    the whole workgroup then walks (solution 2, shared delayed
                                    buffer, sssp_kernel.cu:112-172) */
    NP_SCHEME_GLOBAL_BUFFER = 4, /* light a lane each; heavy appended to a global buffer that a
                                    second launch drains, a wavefront per vertex (solution 3,
                                    sssp_kernel.cu:179-209 + sssp_process_buffer) */
    NP_SCHEME_WAVE_NESTED = 5,   /* light a lane each; each wavefront then walks its own heavy
                                    vertices, 64 lanes each (solution 4, a 64-thread nested launch
                                    per heavy vertex, sssp_kernel.cu:237-286) */
    NP_SCHEME_WG_NESTED = 6,     /* light a lane each; heavy into a per-workgroup list walked by its
                                    wavefronts, 64 lanes per vertex (solution 5, one nested launch
                                    per block, sssp_kernel.cu:289-327) */
    NP_SCHEME_BFS_FLAT = 7,      /* np_bfs only: level-synchronous scan of every vertex per level
                                    (bfs-rec solution 0, bfs_kernel_flat, bfs_rec_kernel.cu:41-58) */
    NP_SCHEME_BFS_HIER = 8       /* np_bfs only: two levels per expansion, lanes over a vertex's
                                    children and the wavefront over each lowered child's children
                                    (bfs-rec solution 2, bfs_kernel_dp_hier, bfs_rec_kernel.cu:95-140) */
};

/* Fills the plain fixed-threshold schedule of the reference at `level` (0,1,2), the
 * persistent drive and the scan frontier. */
void np_sched_default(int32_t level, np_sched *s);

typedef struct np_stats {
    int32_t iterations;     /* frontier expansions performed */
    int32_t status;         /* same as the return code */
    double kernel_ms;       /* device time of the solve, hipEvents on the context stream */
    int64_t frontier_total; /* sum of frontier sizes over iterations */
    int32_t drive;          /* NP_DRIVE_* that ran: a persistent solve runs host-loop when the
                               process gets no cooperative queue (one holder per device) */
    int32_t pad;
    int64_t pairs_total;    /* sum over iterations of queued (vertex, source) pairs: equals
                               frontier_total except in source-packed batch parts */
    int64_t edges_total;    /* source-packed parts of traced solves (sched.trace) on the
                               host-loop and graph drives: sum over iterations of the
                               out-degrees of the queued vertices (edge rows read); 0 elsewhere */
    int64_t items_total;    /* same solves: sum of out-degree x queued sources ((edge, source)
                               relaxations attempted); 0 elsewhere */
    int64_t reserved[2];
} np_stats;

typedef struct np_ctx np_ctx;
typedef struct np_graph np_graph;
typedef struct np_batch np_batch;

const char *np_last_error(void);
const char *np_version(void);

/* Context on HIP device `device` (cudaSetDevice in prepare_gpu, sssp_wrapper.cu:67). */
int np_create(int device, np_ctx **out);
/* Make the library launch on `stream` (a hipStream_t) instead of its own. */
int np_set_stream(np_ctx *ctx, void *stream);
void *np_get_stream(np_ctx *ctx);
int np_destroy(np_ctx *ctx);
/* Number of compute units and the persistent-kernel occupancy (diagnostics). */
int np_device_info(np_ctx *ctx, int32_t *num_cus, int32_t *max_wgs_per_cu);

/* Context options (no counterpart in the reference).
 * NP_CTX_RELABEL: solver vertex numbering of graphs created afterwards.
 *   0: the caller's ids; 1: always renumber by decreasing in-degree; 2 (default): renumber
 *   when the n/16 most-targeted vertices receive at least 40% of the edges.  Renumbering
 *   packs the distances that most relaxations touch into few cache lines; sources and
 *   results stay in the caller's ids (the mapping is applied on the device). */
#define NP_CTX_RELABEL 1
/* NP_CTX_PACK: 1 (default) batch instances that share a graph are solved source-packed
 * (np_batch_create); 0 every instance goes to the concatenated part. */
#define NP_CTX_PACK 2
/* NP_CTX_PACK_MAX: the most sources one pack holds, 2..32 (default 32); a graph's k sources
 * go to ceil(k / max) packs of near-equal size.  Smaller packs read narrower distance rows per
 * edge (fewer unused bytes when few of a row's sources are pending) and run side by side on the
 * graph drive. */
#define NP_CTX_PACK_MAX 3
int np_set_option(np_ctx *ctx, int32_t key, int32_t value);
/* 1 when the graph is solved in a renumbered layout, else 0; skew (optional) receives the
 * share of edges that its n/16 most-targeted vertices receive. */
int np_graph_relabeled(const np_graph *g, double *skew);

/* Device-memory plumbing for callers without their own HIP runtime (no counterpart in
 * the reference; its wrapper calls cudaMalloc/cudaMemcpy directly). */
int np_dev_alloc(np_ctx *ctx, uint64_t bytes, void **d_ptr);
int np_dev_free(np_ctx *ctx, void *d_ptr);
/* kind 0: host -> device, 1: device -> host, 2: device -> device; synchronous on the
 * context stream. */
int np_dev_copy(np_ctx *ctx, void *dst, const void *src, uint64_t bytes, int32_t kind);
/* Waits for all work queued on the context stream. */
int np_sync(np_ctx *ctx);

/* Checks a host CSR without a device: row[0] = 0, row non-decreasing, row[n] = m, 0 <= col < n
 * (NP_ERR_ARG otherwise); with weights (w non-null) every weight >= 0 (NP_ERR_ARG: shortest
 * paths need them) and < NP_INF (NP_ERR_CAPACITY: dist[u] + w must stay in int32, the
 * reference's int arithmetic, apps/sssp/src/gpu/sssp_kernel.cu:46-52).  *max_w (optional)
 * receives the largest weight.  np_graph_upload applies it; device-resident graphs
 * (np_graph_wrap_device, the generators) get the same checks on the device.  Solves whose
 * distances could reach NP_INF ((n - 1) x max weight >= NP_INF) check after the solve that
 * none did and fail with NP_ERR_CAPACITY otherwise: the reference reports such a vertex as
 * unreached.  No counterpart in the reference (it checks nothing). */
int np_csr_validate(int32_t n, int64_t m, const int32_t *row, const int32_t *col, const int32_t *w,
                    int32_t *max_w);
/* Copy a host CSR to the device (prepare_gpu's H2D, sssp_wrapper.cu:100-113). */
int np_graph_upload(np_ctx *ctx, int32_t n, int64_t m, const int32_t *row, const int32_t *col,
                    const int32_t *w, np_graph **out);
/* Use device-resident CSR arrays in place (no copy; the caller keeps them alive).
 * No counterpart in the reference (it always re-uploads). */
int np_graph_wrap_device(np_ctx *ctx, int32_t n, int64_t m, const int32_t *d_row,
                         const int32_t *d_col, const int32_t *d_w, np_graph **out);
int np_graph_info(const np_graph *g, int32_t *n, int64_t *m);
int np_graph_download(np_ctx *ctx, const np_graph *g, int32_t *row, int32_t *col, int32_t *w);
/* Device addresses of the CSR arrays (row, col, w). */
int np_graph_device_ptrs(const np_graph *g, const int32_t **row, const int32_t **col,
                         const int32_t **w);
int np_graph_free(np_graph *g);

/* Single-source shortest paths (SSSP_GPU with solution 6).  dist_out: host int32[n]. */
int np_sssp(np_ctx *ctx, np_graph *g, int32_t src, const np_sched *sched, int32_t *dist_out,
            np_stats *stats);
/* Same, distances written to device memory d_dist (int32[n]); nothing is copied to the
 * host and the call returns once the solve has finished on the device. */
int np_sssp_device(np_ctx *ctx, np_graph *g, int32_t src, const np_sched *sched, int32_t *d_dist,
                   np_stats *stats);
/* np_sssp_device without any device allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  The solver state (bitmaps, frontier queues, counters, the renumbered
 * graph's distances) lives in a file-scope device array of the library sized for graphs of
 * at most NP_NOALLOC_MAX_VERTICES vertices (about 32.5 MiB, held by every GPU the library
 * runs on until the process ends); larger graphs fail with NP_ERR_CAPACITY.  The graph
 * must have been created on ctx (its CSR and degree classes are the caller's inputs).
 * Results are identical to np_sssp_device's.  The graph drive runs as the host loop
 * (capturing a graph allocates); stats.drive says which drive ran.  Every call on one GPU
 * shares that GPU's array, which keeps the last call's contents: calls that overlap on
 * one GPU (two contexts, two threads) race and must be serialised by the caller. */
#define NP_NOALLOC_MAX_VERTICES 2097152
int np_sssp_device_noalloc(np_ctx *ctx, np_graph *g, int32_t src, const np_sched *sched, int32_t *d_dist,
                           np_stats *stats);
/* Recursive BFS levels (BFS_REC_GPU with solution 3): unit weights, level[src] = 0,
 * unreached = NP_UNDEFINED.  level_out: host int32[n]. */
int np_bfs(np_ctx *ctx, np_graph *g, int32_t src, const np_sched *sched, int32_t *level_out,
           np_stats *stats);
/* np_bfs with the levels left in device memory (d_level: device int32[n]), as np_sssp_device
 * is to np_sssp.  No separate reference counterpart: BFS_REC_GPU (bfs_rec_wrapper.cu:231-269)
 * always copies levelArray back; this form lets callers keep levels resident. */
int np_bfs_device(np_ctx *ctx, np_graph *g, int32_t src, const np_sched *sched, int32_t *d_level,
                  np_stats *stats);

/* Batched solves -- no counterpart in the reference.  `count` instances; instance i
 * solves graph graphs[i] from source srcs[i] (graphs may repeat: many sources on one
 * graph; or differ: many graphs of uneven size).  unit != 0 solves BFS levels instead of
 * distances.  Instances are laid out in two kinds of parts, each solved by one launch:
 *   packs         2..32 sources of a graph that appears more than once share one state
 *                 whose distances are stored source-minor (the distances of all its
 *                 sources at a vertex are adjacent), so an edge is read once for every
 *                 source pending at its tail (NP_CTX_PACK, default on)
 *   concatenated  every other instance: the graphs laid end to end in one slot space
 *                 with an instance-tagged worklist.
 * On the graph drive the parts run side by side, each on its own stream (without a trace);
 * on the other drives one after another.  Packs always use the scan frontier
 * (np_sched.frontier applies to the concatenated part); their priority window (delta) admits a (vertex, source) pair when the source's
 * distance at the vertex is below the window.  Every other field applies to both. */
/* The batch reads its graphs at every solve: free it before any graph it names. */
int np_batch_create(np_ctx *ctx, int32_t count, np_graph *const *graphs, const int32_t *srcs,
                    int32_t unit, np_batch **out);
/* stats: iterations and frontier_total summed over the parts; kernel_ms = device time of the
 * whole solve. */
int np_batch_solve(np_ctx *ctx, np_batch *b, const np_sched *sched, np_stats *stats);
/* Host copy of instance i's distances (int32[n_i]). */
int np_batch_get(np_ctx *ctx, np_batch *b, int32_t i, int32_t *dist_out);
/* Device address of instance i's distances, in the caller's vertex ids (valid until the
 * next solve or free of the batch). */
int np_batch_device_dist(np_batch *b, int32_t i, const int32_t **d_dist);
/* Number of packs and of instances in the concatenated part. */
int np_batch_layout(const np_batch *b, int32_t *packs, int32_t *concatenated);
int np_batch_free(np_batch *b);

/* Work counted for throughput: n_r = vertices with a finite distance, m_r = sum of
 * their out-degrees (device reduction over d_dist, which holds inf for unreached). */
int np_reached_work(np_ctx *ctx, np_graph *g, const int32_t *d_dist, int32_t inf,
                    int64_t *n_r, int64_t *m_r);

/* ---- per-iteration trace (sched.trace != 0) -- no counterpart in the reference ---- */
#define NP_NCLASS 32 /* out-degree classes: 0 for d == 0, else 1 + floor(log2 d) */
typedef struct np_trace_rec {
    int32_t iteration;
    int32_t level;            /* configuration used for this iteration */
    int32_t log2_t_lo, log2_t_hi;
    int32_t count[NP_NCLASS]; /* frontier size per out-degree class */
    int64_t t_begin, t_end;   /* device clock (100 MHz) at the start/end of the expansion */
    int64_t pairs, edges;     /* source-packed parts: queued (vertex, source) pairs and (host-loop
                                 and graph drives) the queued vertices' out-degrees, summed over
                                 iterations up to and including this one; 0 elsewhere */
} np_trace_rec;
/* Copies up to `cap` records of the last solve on ctx; *count = records available. */
int np_trace_get(np_ctx *ctx, np_trace_rec *recs, int32_t cap, int32_t *count);

/* ---- learned scheduler -- no counterpart in the reference --------------------------
 * fp32 MLP 10 -> 32 -> 32 -> 5 (3 level logits, 2 regressions: log2_t_lo, log2_t_hi),
 * ReLU hidden layers; parameters W1[32][10] b1[32] W2[32][32] b2[32] W3[5][32] b3[5]
 * (NP_MLP_NPARAM floats).  Features per iteration (np_features documents them). */
#define NP_MLP_NIN 10
#define NP_MLP_NPARAM (32 * 10 + 32 + 32 * 32 + 32 + 5 * 32 + 5)
int np_model_set(np_ctx *ctx, const float *params); /* host params -> device */
int np_model_get(np_ctx *ctx, float *params);
/* Feature vector of one frontier (class counts) at iteration `it` of a graph with n
 * vertices (the same function the kernels evaluate, for host-side training sets). */
int np_features(const int32_t *count, int32_t it, int32_t n, float *x);
/* Fixed-threshold heuristic applied when the network is absent or its output is not
 * finite: the learned scheduler's fallback, and the schedule it is measured against. */
int np_heuristic(const int32_t *count, int32_t it, int32_t n, np_sched *s);
/* One full-batch training run on the GPU with the project's forward/backward kernels
 * (Adam).  x[batch][10], label[batch] in {0,1,2}, target[batch][2] (log2 thresholds).
 * Updates the context's model in place; writes the loss per epoch into loss_out. */
int np_model_train(np_ctx *ctx, const float *x, const int32_t *label, const float *target,
                   int32_t batch, int32_t epochs, float lr, float lam, float *loss_out);
/* Loss and gradient (host outputs) of the current model on a batch -- for gradient checks. */
int np_model_loss_grad(np_ctx *ctx, const float *x, const int32_t *label, const float *target,
                       int32_t batch, float lam, float *loss, float *grad);
/* Forward pass on the GPU: out[batch][5]. */
int np_model_forward(np_ctx *ctx, const float *x, int32_t batch, float *out);

/* ---- single-iteration stepping, for timing labels -- no counterpart in the reference --
 * Starts a stepped solve; each np_step_run performs one expansion with the given
 * schedule (level/thresholds) from a saved state, returning its device time, so that
 * every candidate schedule is timed on the same frontier. */
int np_step_begin(np_ctx *ctx, np_graph *g, int32_t src, np_stats *stats);
/* Frontier class counts of the current state; returns NP_OK and *done = 1 when empty. */
int np_step_frontier(np_ctx *ctx, int32_t *count, int32_t *done);
/* Time one expansion with `sched` from the current state without advancing it. */
int np_step_try(np_ctx *ctx, const np_sched *sched, int32_t reps, double *ms);
/* Advance the state by one expansion with `sched`. */
int np_step_advance(np_ctx *ctx, const np_sched *sched, double *ms);
int np_step_end(np_ctx *ctx, int32_t *dist_out);

/* ---- graph generators (device) -- no counterpart in the reference ------------------
 * Deterministic from `seed` (SplitMix64-keyed counters).  CSR per-vertex order is the
 * reverse of generation order, as convertCSR produces for a file read in that order. */
int np_gen_rmat(np_ctx *ctx, int32_t scale, int32_t edge_factor, double a, double b, double c,
                int32_t wmin, int32_t wmax, uint64_t seed, int32_t permute, np_graph **out);
int np_gen_er(np_ctx *ctx, int32_t n, int64_t m, int32_t wmin, int32_t wmax, uint64_t seed,
              np_graph **out);
int np_gen_grid(np_ctx *ctx, int32_t rows, int32_t cols, int64_t n_long, int32_t wmin,
                int32_t wmax, uint64_t seed, np_graph **out);
/* Host versions of the same generators (identical CSR), into malloc'd arrays freed with
 * np_host_free. */
int np_gen_rmat_host(int32_t scale, int32_t edge_factor, double a, double b, double c,
                     int32_t wmin, int32_t wmax, uint64_t seed, int32_t permute, int32_t **row,
                     int32_t **col, int32_t **w, int32_t *n, int64_t *m);
int np_gen_er_host(int32_t n, int64_t m, int32_t wmin, int32_t wmax, uint64_t seed, int32_t **row,
                   int32_t **col, int32_t **w);
int np_gen_grid_host(int32_t rows, int32_t cols, int64_t n_long, int32_t wmin, int32_t wmax,
                     uint64_t seed, int32_t **row, int32_t **col, int32_t **w, int64_t *m);
/* `k` distinct sources with out-degree >= 1 drawn with `seed` (Graph500 practice). */
int np_pick_sources(const int32_t *row, int32_t n, int32_t k, uint64_t seed, int32_t *srcs);
void np_host_free(void *p);

/* ---- graph files: readInputDIMACS9/10/SLNDC + convertCSR (common/global.cpp:29-235) --
 * fmt 0 DIMACS9, 1 DIMACS10, 2 SLNDC (apps/sssp/src/gpu/sssp.cpp:120-125).  Same CSR as the
 * reference (per-vertex order reversed, DIMACS10/SLNDC weights rand()%1000+1 after
 * srand(7) in file order), built without adjacency lists.  *src gets the DIMACS9 's'
 * line (1-based in the file) or 0.  fmt 3 = the binary CSR cache below (no counterpart in
 * the reference).  Arrays are freed with np_host_free. */
int np_read_graph(int32_t fmt, const char *path, int32_t *n, int64_t *m, int32_t *src,
                  int32_t **row, int32_t **col, int32_t **w);
/* Same, from an open FILE* (passed as void*), as the reference readers use the global fp. */
int np_read_graph_stream(int32_t fmt, void *fp, int32_t *n, int64_t *m, int32_t *src,
                         int32_t **row, int32_t **col, int32_t **w);

/* ---- betweenness centrality: bc_np_consolidate_gpu + BC_GPU ----------------------------
 * (apps/bc/src/gpu/bc_wrapper.cu:678-809, kernels bc_fp_kernel.cu:429-489, bc_bp_kernel.cu:4-36)
 * Brandes' dependencies summed over the sources srcs[0..nsrc) (nsrc = 0: every vertex, as
 * BC_GPU's loop does), unnormalised, on the directed graph with unit edge lengths; parallel
 * edges count as distinct shortest paths, as the reference's per-edge path-count update does.
 * bc_out[n] receives float values (the reference's float *bc), accumulated in double.  Sources
 * run in batches of up to 32 on the same frontier machinery as the source-packed SSSP; forward
 * levels go out in runs as long as the previous batch's depth + 1, with one readback of the
 * runs' (first entry, count) pairs per run.  stats: iterations = levels summed over batches,
 * frontier_total = (vertex, source mask) entries, kernel_ms = device time of the whole call. */
int np_bc(np_ctx *ctx, np_graph *g, const int32_t *srcs, int32_t nsrc, float *bc_out, np_stats *stats);

/* ---- graph coloring: GRAPH_COLOR_GPU --------------------------------------------------
 * (apps/graph-color/src/gpu/graph_color_wrapper.cu:598-699, kernels graph_color_kernel.cu:
 * 14-35, 779-834)  The reference's rounds give every vertex v
 *     color[v] = 1 + max { color[u] : v -> u, u < v }   (1 when there is no such u),
 * whatever the thread timing; color_out[n] receives these int32 values (graph.colorArray).
 * method 0: dependency counting over the transposed smaller-neighbour lists, O(m) work;
 * method 1: the reference's rounds as they are, O(rounds x m) (the baseline).
 * stats: iterations = rounds, kernel_ms = device time of the whole call. */
int np_color(np_ctx *ctx, np_graph *g, int32_t method, int32_t *color_out, np_stats *stats);

/* ---- PageRank: PAGERANK_GPU --------------------------------------------------------------
 * (apps/pagerank/src/gpu/pagerank_wrapper.cu:199-244, 565-661, kernels pagerank_kernel.cu:60-107)
 * Pull iterations over the in-edges from rank = 1000/n:
 *   new[v] = 0.5 * sum_{p -> v} rank[p]/outdeg[p] + 0.5 * 1000/n + 0.5 * (rank of dangling)/n
 * until no rank changes by more than 1e-13 (damping, total and epsilon of the reference), in
 * float64; rank_out[n] receives float values (graph.rankArray, FLOAT_T = float).  max_iters = 0:
 * no cap beyond 100000 (NP_ERR_NOT_CONVERGED past it).  stats.iterations = iterations run.
 * Its device buffers (in-edge CSR, row bins, ranks) come from a workspace the context keeps
 * and grows to the largest call (shared with np_spmv, freed by np_destroy), so repeated calls
 * allocate nothing; calls on one context must not overlap (one context per thread). */
int np_pagerank(np_ctx *ctx, np_graph *g, int32_t max_iters, float *rank_out, np_stats *stats);

/* ---- SpMV: SpMV_GPU (apps/SpMV/src/gpu/SpMV_wrapper.cu:314-345, SpMV_kernel.cu:77-92) ------
 * y[v] = sum over the graph's row v of data[e] * x[col[e]] (float64, the reference's FLOAT_T);
 * host arrays data[m], x[n], y[n].  kernel_ms covers the product only (not the copies).  Device
 * copies of data, x, y and the row bins come from the context's workspace (see np_pagerank). */
int np_spmv(np_ctx *ctx, np_graph *g, const double *data, const double *x, double *y, np_stats *stats);

/* ---- binary CSR cache -- no counterpart in the reference ----------------------------
 * A graph read or generated once is saved and read back without parsing (the reference
 * re-parses text on every run).  Layout, native little-endian: magic "NPCSR01\0", int64 n,
 * int64 m, int32 source, int32 flags (bit 0: weights present), row[n+1], col[m], w[m] (int32).
 * np_csr_load checks the sizes against the file and the CSR itself (row[0] = 0, row[n] = m,
 * nondecreasing, col in [0, n)) and returns NP_ERR_IO otherwise; without weights it returns
 * unit weights.  w may be NULL for np_csr_save.  Arrays are freed with np_host_free. */
int np_csr_save(const char *path, int32_t n, int64_t m, int32_t src, const int32_t *row,
                const int32_t *col, const int32_t *w);
int np_csr_load(const char *path, int32_t *n, int64_t *m, int32_t *src, int32_t **row,
                int32_t **col, int32_t **w);

/* ---- autotuner -- no counterpart in the reference ----------------------------------
 * Times the candidate schedules (drive persistent or graph x level x log2_t_lo x
 * log2_t_hi x wgs_per_cu) on `g` from the given sources, then the best of them with a
 * priority window of 2, 4, 8, 16 and 64 times the mean edge weight; when a window wins, level
 * and thresholds are searched again under it, then half and twice the window.  Returns the
 * fastest in *best (frontier mode and the other fields kept from `base`); *best_ms
 * (optional) = its mean device time per source. */
int np_autotune(np_ctx *ctx, np_graph *g, const int32_t *srcs, int32_t nsrc, const np_sched *base,
                np_sched *best, double *best_ms);
/* The same search on a batch: each candidate times one np_batch_solve (all parts);
 * *best_ms = the fastest solve.  The batch holds the last candidate's results after. */
int np_batch_autotune(np_ctx *ctx, np_batch *b, const np_sched *base, np_sched *best, double *best_ms);

#ifdef __cplusplus
}
#endif
#endif /* NP_SSSP_H */
