// np_internal.h -- internal interfaces between the C-ABI layer (np_api.cpp) and the
// HIP translation units (np_kernels.hip, np_gen.hip).  Not installed, not stable.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "np_common.h"
#include "../../include/np_sssp.h"

namespace np {

constexpr int BAR_LEAVES = 32;
struct alignas(128) BarLeaf { unsigned long long n; unsigned long long pad[15]; };

// Device-side control block of one solve.
struct Ctrl {
    BarLeaf leaf[BAR_LEAVES]; // grid barrier: cumulative arrivals per leaf (own cache lines)
    unsigned long long bar;   // grid barrier: cumulative leaf completions
    int error;                // NP_ERR_* raised by a kernel (0 = none)
    int done;                 // frontier emptied
    int iters;                // expansions performed
    int it_base;              // graph drive: iteration index of the captured chunk's first slot
    long long frontier_total; // sum of frontier sizes
    int cfg_level, cfg_klo, cfg_khi, cfg_pad;  // schedule decided for the next expansion
    int thresh;               // priority window: vertices with dist < thresh are expanded
    int pass;                 // compaction passes so far (host-loop / graph drives)
    int skip;                 // host-loop / graph drives: this slot only raised the window
    int defer_cnt[3];         // per compaction pass (rotating): flagged vertices kept back
    int defer_min[3];         // ... and the least distance among them
    long long pairs_total;    // source-packed: sum over iterations of queued (vertex, source) pairs
    long long edges_total;    // source-packed: sum over iterations of the queued vertices' out-degrees
    long long items_total;    // source-packed: ... of out-degree x queued sources ((edge, source) items)
};

// Everything an expansion / compaction touches.  Indices are "slots": the vertex ids of
// a single solve, or instance-tagged vertex ids of a batch (slot = base_i + v).
struct DevState {
    const int *row;           // row offsets, indexed VB_i + v (graphs concatenated)
    const int *col;           // local target ids
    const int *w;             // weights (unused for unit-weight solves)
    const unsigned char *cls; // out-degree class per slot
    int *dist;                // per slot
    unsigned *bm[2];          // update bitmaps per slot (scan uses bm[0]; push ping-pongs)
    int *queue[2];            // frontier, segmented by class (seg_off)
    const int *seg_off;       // [NCLS+1] class segment offsets inside a queue
    int *cnt;                 // [3][NCLS] frontier class counters, rotating by iteration
    const int *lut;           // batch: slot >> LUT_SHIFT -> instance
    const int2 *inst;         // batch: per instance {slot base, row base}
    long long nslots;
    long long nwords;         // bitmap words
    int inf;                  // unreached value written at init
    Ctrl *ctrl;
    // source-packed batch (kshift >= 0): 2^kshift sources of one graph per vertex row,
    // dist[(v << kshift) + i]; mflag[v] = sources pending at v; bm[0] = vertices with a
    // nonzero mflag; qmask[] = source mask of each queue entry
    int kshift;               // -1: one distance per slot
    unsigned *mflag;
    unsigned *qmask[2];
    // NP_SCHEME_GLOBAL_BUFFER: heavy vertices of one expansion and the buffer's fill
    // counters [3], rotating by iteration
    int *sbuf;
    int *sbcnt;
    // source-packed, host-loop and graph drives: a counting kernel after each compaction
    // accumulates Ctrl::edges_total / items_total (set when the solve records a trace)
    int count_work;
    // source-packed, no priority window: a flagged vertex of out-degree class >= hold_cls that
    // fewer than hold_min of the pack's sources have reached is held back one compaction pass
    // (once per flagging; bm[1] marks the held vertices); hold_cls 0 = never
    int hold_cls;
    int hold_min;
};

// Largest number of sources packed into one distance row.
constexpr int PACK_MAX = 32;

struct SchedDev {
    int learned;            // evaluate the MLP each iteration
    int level, klo, khi;    // fixed schedule (the fallback when the MLP is unusable)
    int push;               // NP_FRONTIER_PUSH
    int delta;              // priority window (0 = off)
    int max_iters;
    int trace_cap;
    int n_feat;             // vertex count used by the features
    const float *mlp;       // NP_MLP_NPARAM device floats
    np_trace_rec *trace;
    int scheme;             // NP_SCHEME_* (0: the consolidation bins)
};

// ---- kernels.hip: solve drivers ----------------------------------------------------
// Launch shapes are chosen by the callers in np_api.cpp.
hipError_t launch_fill_i32(int *p, long long n, int v, hipStream_t s);
hipError_t launch_fill_u32(unsigned *p, long long n, unsigned v, hipStream_t s);
hipError_t launch_degree_class(const int *row, int n, unsigned char *cls, int *hist, hipStream_t s);
// out3[0] = CSR violations, out3[1] = negative weights, out3[2] = largest weight (w may be
// null: no weight checks); out3 must hold zeros.
hipError_t launch_validate_csr(const int *row, const int *col, const int *w, int n, long long m, int *out3,
                               hipStream_t s);
// *out = largest finite distance of the state's buffer (0 when none).
hipError_t launch_max_finite(const DevState &st, int *out, hipStream_t s);
// *bad = reached edges u -> v with dist[v] = inf and dist[u] + w >= inf (a distance that the
// int32 range below inf cannot hold); batch: concatenated slots.
hipError_t launch_dist_escape(const DevState &st, bool batch, int *bad, hipStream_t s);
hipError_t launch_init_solve(const DevState &st, const int *d_src_slots, int nsrc, int push, int delta,
                             hipStream_t s);
hipError_t launch_reached_work(const int *row, const int *dist, int n, int inf,
                               unsigned long long *out2, hipStream_t s);
hipError_t launch_rebase_row(const int *row, int n, int add, int *out, hipStream_t s);
hipError_t launch_fill_lut(int *lut, long long first, long long count, int inst, hipStream_t s);
hipError_t launch_copy_cls(const unsigned char *src, unsigned char *dst, long long n,
                           hipStream_t s);

// Source-packed state: dist rows set to inf, source i of srcs[] at lane i (nsrc <= 2^kshift);
// delta > 0 starts a priority window at T = delta.
hipError_t launch_init_packed(const DevState &st, const int *d_srcs, int nsrc, int delta, hipStream_t s);
// out[v] = dist[(map(v) << kshift) + lane], map = old2new or identity (old2new null).
hipError_t launch_unpack(const int *dist, int kshift, int lane, const int *old2new, int n, int *out,
                         hipStream_t s);

// Persistent drive: one cooperative launch.  `wgs` workgroups of 256 threads.  A state
// with kshift >= 0 runs the source-packed kernels (here and in compact/expand).
hipError_t launch_persistent(const DevState &st, const SchedDev &sp, bool batch, bool unit,
                             int wgs, hipStream_t s);
int persistent_max_wgs_per_cu(bool batch, bool unit, bool push, bool packed);

// Host-loop / graph drive, one iteration = [compact] + decide + expand.
hipError_t launch_compact(const DevState &st, int it_off, bool batch, int wgs, hipStream_t s);
hipError_t launch_decide(const DevState &st, const SchedDev &sp, int it_off, hipStream_t s);
hipError_t launch_expand(const DevState &st, const SchedDev &sp, int it_off, bool batch, bool unit,
                         int wgs, hipStream_t s);
hipError_t launch_advance(const DevState &st, int k, hipStream_t s);
// NP_SCHEME_BFS_FLAT: one level-synchronous pass over every vertex (levels in st.dist);
// *changed (device) is set when a level was written.
hipError_t launch_bfs_flat(const DevState &st, int level, int *changed, int wgs, hipStream_t s);
// Stepping: force the next expansion's schedule (writes ctrl->cfg_*).
hipError_t launch_set_cfg(const DevState &st, int level, int klo, int khi, hipStream_t s);

// ---- kernels.hip: scheduler MLP ------------------------------------------------------
hipError_t launch_mlp_forward(const float *p, const float *x, int batch, float *out,
                              hipStream_t s);
// Per-sample forward+backward then per-parameter reduction: grad (NP_MLP_NPARAM) and the
// mean loss (loss[0]).  work: scratch of mlp_work_floats(batch) floats.
size_t mlp_work_floats(int batch);
hipError_t launch_mlp_loss_grad(const float *p, const float *x, const int *label,
                                const float *target, int batch, float lam, float *work,
                                float *grad, float *loss, hipStream_t s);
hipError_t launch_adam(float *p, const float *grad, float *m, float *v, int n, float lr,
                       float b1, float b2, float eps, int step, hipStream_t s);

// ---- bc.hip: betweenness centrality, up to 32 sources per batch -----------------------
constexpr int BC_UNSEEN = 0x7FFFFFFF;
struct BcDev {
    const int *row, *col;  // CSR (original ids)
    int n, ks;             // slots (v << ks) + j for batch source j < 2^ks
    int *lvl;              // BFS level per slot, BC_UNSEEN when not reached
    double *sig, *del;     // shortest-path counts and dependencies per slot
    unsigned *nmask;       // [n] sources that first reached v in the last expansion
};
hipError_t launch_bc_reset(const BcDev &b, hipStream_t s);
hipError_t launch_bc_sources(const BcDev &b, const int *srcs, int nsrc, hipStream_t s);
// Flags -> entries (v, mask) at ent_v/ent_m[*count...] with chunk counts ent_c[*count...].
// Flagged vertices -> entries at ent_v/ent_m + *eoff, counted in *count; the first chunk of
// each at cst[0 .. count), chunks reserved on *ctot (zero before the level).
hipError_t launch_bc_compact(const BcDev &b, int *ent_v, unsigned *ent_m, const int *eoff, long long *cst,
                             int *count, unsigned long long *ctot, hipStream_t s);
// Level d's first entry and count into lv[2d], lv[2d + 1]; *eoff moves past them, *count = 0.
hipError_t launch_bc_advance(int *eoff, int *count, int *lv, int d, hipStream_t s);
// Chunk ranges of ne entries (first chunk at cst[o]) reserved on *ctot (zero before).
hipError_t launch_bc_chunks(const BcDev &b, const int *ent_v, const unsigned *ent_m, int ne, long long *cst,
                            unsigned long long *ctot, hipStream_t s);
// Chunk owners of ne entries at ent_v/ent_m + *eoff (*ne_dev when given; ne then bounds it); a
// range past cap sets *err (nothing is written past cap).
hipError_t launch_bc_owner(const BcDev &b, const int *ent_v, const unsigned *ent_m, const int *eoff,
                           const long long *cst, int ne, const int *ne_dev, int *owner, long long cap, int *err,
                           hipStream_t s);
// Expansion of the chunks of the entries at ent_v + *eoff (eoff null: ent_v); chunk total
// *ctot, read on the device, at most cap.
hipError_t launch_bc_expand(const BcDev &b, bool fwd, const int *ent_v, const unsigned *ent_m, const int *eoff,
                            const long long *cst, const int *owner, const unsigned long long *ctot, long long cap,
                            int d, int num_cus, hipStream_t s);
hipError_t launch_bc_accum(const BcDev &b, const int *ent_v, const unsigned *ent_m, int ne, double *bc,
                           hipStream_t s);
hipError_t launch_bc_to_float(const double *bc, int n, float *out, hipStream_t s);

// ---- color.hip: graph coloring (longest decreasing-id path labels) -----------------------
hipError_t launch_color_count(const int *row, const int *col, int n, int *dep, long long *ncnt, hipStream_t s);
hipError_t launch_color_fill(const int *row, const int *col, int n, const long long *nrow, int *cursor, int *ncol,
                             hipStream_t s);
hipError_t launch_color_init(const int *dep, int n, int *color, int *q, int *qn, hipStream_t s);
// One dependency-counting round (graph-capturable: no host values depend on the frontier).
hipError_t launch_color_step(const int *q_in, const int *qn_in, int *q_out, int *qn_out, int *qn_zero,
                             const long long *nrow, const int *ncol, int *dep, int *color, int *rounds, int num_cus,
                             hipStream_t s);
hipError_t launch_color_round(const int *row, const int *col, int n, int *color, int c, int *nonstop, hipStream_t s);
hipError_t color_scan(void *temp, size_t *temp_bytes, const long long *in, long long *out, int n, hipStream_t s);

// ---- pagerank.hip: PageRank over the in-edge CSR ------------------------------------------
// Rows by length: < PR_HEAVY entries PR_GROUP lanes each; < PR_HUGE one wavefront each (med);
// longer rows (huge) in PR_CHUNK-entry chunks, one wavefront per chunk, partial sums in part[].
constexpr int PR_GROUP = 8;
constexpr int PR_HEAVY = 32;
constexpr int PR_CHUNK = 2048;
constexpr int PR_HUGE = 2 * PR_CHUNK;
struct RowBins {
    int *med, *nmed;              // medium rows [n]; count
    int *huge, *hc0, *nhuge;      // huge rows and their first chunk [m / PR_HUGE + 1]; count
    int *crow, *ck, *nchunk;      // chunk -> row, chunk index [pr_chunk_cap(m)]; count
    double *part;                 // per-chunk partial sums [pr_chunk_cap(m)]
};
long long pr_chunk_cap(long long m);
// In-edge CSR: src[e] = source of edge e; sort (col, src) pairs; offsets from the sorted targets.
// src[e] = source of edge e for the m edges of the CSR row (heads: m ints of scratch; temp null
// queries *temp_bytes).
hipError_t pr_sources(void *temp, size_t *temp_bytes, const int *row, int n, long long m, int *heads, int *src,
                      hipStream_t s);
hipError_t pr_sort(void *temp, size_t *temp_bytes, const unsigned *key_in, unsigned *key_out, const int *val_in,
                   int *val_out, long long m, int end_bit, hipStream_t s);
hipError_t launch_pr_offsets(const unsigned *key, long long m, int n, long long *irow, hipStream_t s);
hipError_t launch_pr_init(double *rank, int n, double r0, hipStream_t s);
// Bins over irow (counts *nmed, *nhuge, *nchunk zero on entry).
hipError_t launch_pr_bins(const long long *irow, int n, long long m, const RowBins &rb, hipStream_t s);
// Iteration i, rank -> nrank, decided on the device: ctl[0] (done) is set when iteration i - 1
// changed no vertex by more than eps, and then nothing runs; else ctl[1] = i + 1.  Sums of
// iteration i in dsum2[i & 1], dmax2[i & 1] (cleared by the iteration itself); contrib: scratch.
hipError_t launch_pr_iteration(int i, int *ctl, const int *row, const long long *irow, const int *pcol,
                               const RowBins &rb, const double *rank, double *nrank, double *contrib, int n,
                               double damping, double rw, double eps, double *dsum2, unsigned long long *dmax2,
                               int num_cus, hipStream_t s);
hipError_t launch_pr_to_float(const double *rank, int n, float *out, hipStream_t s);
// SpMV y = A x over the graph's CSR, with row bins built by launch_spmv_bins.
hipError_t launch_spmv_bins(const int *row, int n, long long m, const RowBins &rb, hipStream_t s);
hipError_t launch_spmv(const int *row, const int *col, const double *data, const double *x, double *y, int n,
                       const RowBins &rb, int num_cus, hipStream_t s);

// ---- gen.hip: generators and CSR construction ---------------------------------------
struct GenOut { int *row; int *col; int *w; };
// Builds CSR (reverse generation order per vertex) from an edge list in generation order.
hipError_t build_csr_device(const uint32_t *src, const uint32_t *dst, const int *w, int n,
                            long long m, GenOut *out, hipStream_t s);
hipError_t gen_rmat_edges(int scale, long long m, uint64_t seed, double a, double b, double c,
                          int permute, int wmin, int wmax, uint32_t *src, uint32_t *dst, int *w,
                          hipStream_t s);
hipError_t gen_er_edges(int n, long long m, uint64_t seed, int wmin, int wmax, uint32_t *src,
                        uint32_t *dst, int *w, hipStream_t s);
hipError_t gen_grid_edges(int rows, int cols, long long n_long, uint64_t seed, int wmin, int wmax,
                          const long long *grid_off, uint32_t *src, uint32_t *dst, int *w,
                          hipStream_t s);

hipError_t grid_offsets(int rows, int cols, long long *grid_off, hipStream_t s);

// Solver layout renumbered by decreasing in-degree.  `skew` = share of the edges that the
// top n/16 vertices receive; the layout is built (active) only when skew >= min_skew.
struct Relabel {
    bool active = false;
    double skew = 0.0;
    int *row = nullptr, *col = nullptr, *w = nullptr;  // renumbered CSR
    int *old2new = nullptr;                             // original id -> solver id
};
hipError_t relabel_by_indegree(const int *row, const int *col, const int *w, int n, long long m,
                               double min_skew, Relabel *out, hipStream_t s);
// out[v] = src[old2new[v]]
hipError_t launch_unpermute(const int *src, const int *old2new, int n, int *out, hipStream_t s);
// Thresholds of rmat_edge from probabilities.
void rmat_thresholds(double a, double b, double c, uint64_t *ta, uint64_t *tab, uint64_t *tabc);

}  // namespace np
