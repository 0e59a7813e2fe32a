// np_api.cpp -- the C ABI (include/np_sssp.h) over the gfx950 kernels.
//
// Reference mapping: np_graph_upload is prepare_gpu()'s allocation + H2D
// (apps/sssp/src/gpu/sssp_wrapper.cu:58-140); np_sssp is sssp_np_consolidate_gpu()
// (:475-546) followed by the D2H of costArray (:576); np_bfs is bfs_rec_dp_cons_gpu()
// (apps/bfs-rec/src/gpu/bfs_rec_wrapper.cu:170-229).  Unlike the reference, the device
// graph stays resident between solves and the solver state is allocated once per graph.
#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <vector>
#include <algorithm>
#include <string>
#include <functional>
#include <utility>
#include <mutex>

#include "np_internal.h"

namespace np {

static thread_local char g_err[512] = "";

static const char *BOUNDS_MSG = "bounds-checked build: a frontier-queue or delayed-buffer write fell outside its "
                                "segment (skipped; the distances are not valid)";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

static int hip_fail(hipError_t e, const char *what, int line) {
    set_error("%s failed (line %d): %s", what, line, hipGetErrorString(e));
    return NP_ERR_HIP;
}
#define HIPCHK(x)                                                      \
    do {                                                               \
        hipError_t _e = (x);                                           \
        if (_e != hipSuccess) return hip_fail(_e, #x, __LINE__);       \
    } while (0)

template <class T>
static hipError_t dalloc(T **p, size_t count) {
    *p = nullptr;
    return hipMalloc((void **)p, sizeof(T) * (count ? count : 1));
}
template <class T>
static void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

static constexpr int TRACE_CAP = 1 << 15;

}  // namespace np

using namespace np;

// ------------------------------------------------------------------------------------
// solver state shared by single-graph solves and batches
// ------------------------------------------------------------------------------------
struct Space {
    long long nslots = 0, nwords = 0;
    long long qcap = 0;
    int hist[NCLS] = {0};
    int *dist = nullptr;          // owned distance buffer (np_sssp copies out of it)
    unsigned *bm[2] = {nullptr, nullptr};
    int *queue[2] = {nullptr, nullptr};
    int *cnt = nullptr;
    int *seg_off = nullptr;
    unsigned char *cls = nullptr; // owned for batches; borrowed from the graph otherwise
    bool own_cls = false;
    int *lut = nullptr;
    int2 *inst = nullptr;
    int *src_slots = nullptr;
    int src_cap = 0;
    unsigned *mflag = nullptr;    // source-packed: pending sources per vertex
    unsigned *qmask[2] = {nullptr, nullptr};
    int *sbuf = nullptr, *sbcnt = nullptr;  // NP_SCHEME_GLOBAL_BUFFER (allocated on first use)
    // a batch part run concurrently with the others: its own control block, stream, event
    Ctrl *pctrl = nullptr;
    Ctrl *h_pctrl = nullptr;
    hipStream_t pstream = nullptr;
    hipEvent_t pev = nullptr;
    // captured chunk of the graph drive
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    SchedDev gsp{};
    DevState gst{};
    int gchunk = 0, gwgs = 0;
    bool gbatch = false, gunit = false;
    // largest edge weight of the graph(s) solved in this space, and whether a distance could
    // reach INF: (largest vertex count - 1) x max_w >= INF (checked after every solve)
    int max_w = 0;
    bool guard = false;

    void release() {
        dfree(dist); dfree(bm[0]); dfree(bm[1]); dfree(queue[0]); dfree(queue[1]);
        dfree(cnt); dfree(seg_off);
        if (own_cls) dfree(cls);
        cls = nullptr;
        dfree(lut); dfree(inst); dfree(src_slots);
        dfree(mflag); dfree(qmask[0]); dfree(qmask[1]);
        dfree(sbuf); dfree(sbcnt);
        dfree(pctrl);
        if (h_pctrl) (void)hipHostFree(h_pctrl);
        if (pstream) (void)hipStreamDestroy(pstream);
        if (pev) (void)hipEventDestroy(pev);
        h_pctrl = nullptr; pstream = nullptr; pev = nullptr;
        drop_graph();
    }
    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr; graph = nullptr;
    }
};

struct np_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int num_cus = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Ctrl *ctrl = nullptr;
    Ctrl *h_ctrl = nullptr;          // pinned readback
    np_trace_rec *trace = nullptr;
    int last_iters = 0;
    float *model = nullptr;          // NP_MLP_NPARAM floats on the device
    bool has_model = false;
    float *adam_m = nullptr, *adam_v = nullptr;
    int adam_step = 0;
    unsigned long long *d_work2 = nullptr;
    // stepping
    np_graph *step_g = nullptr;
    int step_it = 0;
    int *snap_dist = nullptr;
    unsigned *snap_bm = nullptr;
    long long snap_n = 0;
    int relabel = 2;  // NP_CTX_RELABEL: 0 never, 1 always, 2 when in-degree is skewed
    int pack = 1;     // NP_CTX_PACK: batch instances sharing a graph go to source-packed states
    int pack_max = PACK_MAX;  // NP_CTX_PACK_MAX: most sources per pack
    // grow-only device workspace for PageRank and SpMV (their buffers are carved from it, so
    // repeated calls allocate nothing once it is large enough) and a small pinned readback
    char *ws = nullptr;
    size_t ws_bytes = 0;
    unsigned long long *h_small = nullptr;  // 4 words
    int *d_guard = nullptr;                 // 4 ints: CSR validation, distance-range check
    // Live objects made on this context.  np_destroy releases what they hold on the device
    // and detaches them: their handles stay valid for np_graph_free / np_batch_free, which
    // then only delete the host object, and every other call rejects them.
    std::vector<np_graph *> graphs;
    std::vector<np_batch *> batches;
    std::vector<void *> allocs;             // np_dev_alloc
    unsigned char *scratch = nullptr;       // this device's copy of the static scratch array
};

template <class T>
static void unlink_from(std::vector<T *> &v, T *p) {
    auto it = std::find(v.begin(), v.end(), p);
    if (it != v.end()) v.erase(it);
}

struct np_graph {
    np_ctx *ctx = nullptr;
    int n = 0;
    long long m = 0;
    int *row = nullptr, *col = nullptr, *w = nullptr;  // the caller's CSR (original ids)
    bool owns = false;
    Relabel rl;                                         // solver layout, when renumbered
    const int *srow = nullptr, *scol = nullptr, *sw = nullptr;  // CSR the kernels run on
    unsigned char *cls = nullptr;                        // degree class per solver id
    int hist[NCLS] = {0};
    Space sp;  // single-source solver state (lazily allocated)
    bool sp_ready = false;
    int *out = nullptr;  // un-permuted distances (renumbered graphs)
    int max_w = 0;       // largest edge weight (validated: 0 <= w < INF)

    // solver id of original vertex v (one 4-byte read when renumbered)
    int solver_id(int v, hipStream_t s) const {
        if (!rl.active) return v;
        int r = -1;
        if (hipMemcpyAsync(&r, rl.old2new + v, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return -1;
        return r;
    }
};

// Up to PACK_MAX sources of one graph in one source-packed state.
struct Pack {
    np_graph *g = nullptr;
    int kshift = 0;          // dist rows of 2^kshift sources
    std::vector<int> inst;   // batch instance of each lane
    Space sp;
};

struct np_batch {
    np_ctx *ctx = nullptr;
    int count = 0;
    int unit = 0;
    // per instance
    std::vector<int> nverts;
    std::vector<np_graph *> graphs;
    std::vector<int> pack_of, lane_of;  // pack and lane, or -1 (concatenated part)
    std::vector<int> base;              // slot base in the concatenated part, or -1
    std::vector<long long> out_off;     // offset of the instance's result in `out`
    std::vector<Pack *> packs;
    // concatenated part
    int nconcat = 0;
    int *row = nullptr, *col = nullptr, *w = nullptr;  // concatenated graphs (or borrowed)
    bool owns_graph = false;
    Space sp;
    int *out = nullptr;                 // results in original ids (unpacked / mapped back)
    long long out_total = 0;
};

// ------------------------------------------------------------------------------------
// helpers
// ------------------------------------------------------------------------------------
static int ctx_sync(np_ctx *c) {
    HIPCHK(hipStreamSynchronize(c->stream));
    return NP_OK;
}

static void seg_from_hist(const int *hist, int *seg, long long *qcap) {
    long long acc = 0;
    for (int c = 0; c < NCLS; ++c) {
        seg[c] = (int)acc;
        if (c > 0) acc += hist[c];  // class-0 vertices are never queued
    }
    seg[NCLS] = (int)acc;
    *qcap = acc;
}

// Allocates the frontier/bitmap/counter buffers of a space whose hist/nslots are set.
static int space_alloc(np_ctx *c, Space &s, bool want_dist) {
    int seg[NCLS + 1];
    seg_from_hist(s.hist, seg, &s.qcap);
    s.nwords = (s.nslots + 31) / 32;
    if (want_dist) HIPCHK(dalloc(&s.dist, (size_t)s.nslots));
    HIPCHK(dalloc(&s.bm[0], (size_t)s.nwords));
    HIPCHK(dalloc(&s.bm[1], (size_t)s.nwords));
    HIPCHK(dalloc(&s.queue[0], (size_t)s.qcap));
    HIPCHK(dalloc(&s.queue[1], (size_t)s.qcap));
    HIPCHK(dalloc(&s.cnt, 3 * NCLS));
    HIPCHK(dalloc(&s.seg_off, NCLS + 1));
    HIPCHK(hipMemcpyAsync(s.seg_off, seg, sizeof(seg), hipMemcpyHostToDevice, c->stream));
    return ctx_sync(c);
}

static int ensure_src(np_ctx *c, Space &s, int nsrc) {
    if (s.src_cap >= nsrc) return NP_OK;
    dfree(s.src_slots);
    HIPCHK(dalloc(&s.src_slots, (size_t)nsrc));
    s.src_cap = nsrc;
    (void)c;
    return NP_OK;
}

static DevState make_state(np_ctx *c, Space &s, const int *row, const int *col, const int *w,
                           int *dist, int inf) {
    DevState st{};
    st.row = row; st.col = col; st.w = w;
    st.cls = s.cls;
    st.dist = dist;
    st.bm[0] = s.bm[0]; st.bm[1] = s.bm[1];
    st.queue[0] = s.queue[0]; st.queue[1] = s.queue[1];
    st.seg_off = s.seg_off;
    st.cnt = s.cnt;
    st.lut = s.lut;
    st.inst = s.inst;
    st.nslots = s.nslots;
    st.nwords = s.nwords;
    st.inf = inf;
    st.ctrl = c->ctrl;
    st.kshift = -1;
    st.sbuf = s.sbuf;
    st.sbcnt = s.sbcnt;
    return st;
}

static int validate_sched(const np_sched *in, np_sched *out) {
    if (in) *out = *in;
    else np_sched_default(1, out);
    if (out->level < -1 || out->level > 2) { set_error("level must be -1..2"); return NP_ERR_ARG; }
    if (out->log2_t_lo < 0 || out->log2_t_lo > 30 || out->log2_t_hi < 0 || out->log2_t_hi > 30) {
        set_error("log2 thresholds must be 0..30"); return NP_ERR_ARG;
    }
    if (out->drive < 0 || out->drive > 2) { set_error("bad drive"); return NP_ERR_ARG; }
    if (out->frontier < 0 || out->frontier > 1) { set_error("bad frontier mode"); return NP_ERR_ARG; }
    if (out->wgs_per_cu < 0 || out->graph_chunk < 0 || out->max_iters < 0 || out->delta < 0) {
        set_error("negative schedule field"); return NP_ERR_ARG;
    }
    if (out->scheme < NP_SCHEME_CONSOLIDATE || out->scheme > NP_SCHEME_BFS_HIER) {
        set_error("scheme must be 0..8"); return NP_ERR_ARG;
    }
    if (out->pack_hold < 0 || out->pack_hold >= NCLS) { set_error("pack_hold must be 0..%d", NCLS - 1); return NP_ERR_ARG; }
    return NP_OK;
}

static SchedDev make_schedd(np_ctx *c, const np_sched &s, int n_feat) {
    SchedDev d{};
    d.learned = (s.level == NP_LEVEL_LEARNED) ? 1 : 0;
    d.level = s.level >= 0 ? s.level : 1;  // learned: fallback is the plain block level
    d.klo = s.log2_t_lo;
    d.khi = s.log2_t_hi;
    d.push = s.frontier == NP_FRONTIER_PUSH;
    d.delta = d.push ? 0 : s.delta;
    // With non-negative weights the frontier empties within n+1 expansions (every vertex
    // is final once all shortest paths of <= k edges are), so n+2 only stops inputs
    // with negative cycles.
    d.max_iters = s.max_iters > 0 ? s.max_iters : (n_feat < 0x7FFFFFFD ? n_feat + 2 : 0x7FFFFFFF);
    d.trace_cap = s.trace ? TRACE_CAP : 0;
    d.trace = s.trace ? c->trace : nullptr;
    d.n_feat = n_feat;
    d.mlp = (d.learned && c->has_model) ? c->model : nullptr;
    if (d.learned && !c->has_model) d.learned = 0;  // no network: the heuristic runs
    d.scheme = s.scheme;
    if (d.scheme != NP_SCHEME_CONSOLIDATE) {
        // the reference's other schedules: fixed heavy threshold, scan frontier, no window
        d.learned = 0; d.mlp = nullptr; d.level = 1; d.push = 0; d.delta = 0;
    }
    return d;
}

static int persistent_wgs(np_ctx *c, const np_sched &s, bool batch, bool unit, bool packed) {
    int occ = persistent_max_wgs_per_cu(batch, unit, s.frontier == NP_FRONTIER_PUSH, packed);
    if (occ < 1) occ = 1;
    int per = s.wgs_per_cu > 0 ? std::min(s.wgs_per_cu, occ) : occ;
    return per * c->num_cus;
}

static int read_ctrl(np_ctx *c) {
    HIPCHK(hipMemcpyAsync(c->h_ctrl, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, c->stream));
    return ctx_sync(c);
}

// Whether a solve over graphs of up to n vertices and largest weight max_w needs the
// distance-range check.
static void set_guard(Space &s, long long n, int max_w) {
    s.max_w = std::max(s.max_w, max_w);
    s.guard = s.guard || (n - 1) * (long long)max_w >= (long long)INF;
}

// The distance-range check after a solve of SSSP distances (not BFS levels, which stay
// below n).  The reference's relaxation `if (cost[v] > alt) atomicMin` (sssp_kernel.cu:46-52)
// never stores alt >= INF = 2^30, so a vertex whose distance would reach INF is reported as
// unreached.  When (vertices - 1) x (largest weight) >= INF this can happen: the largest
// finite distance is read back, and only when it is within one weight of INF are the edges
// swept for a reached tail whose target stayed at INF (NP_ERR_CAPACITY).  Weights are below
// INF (checked at graph creation), so dist[u] + w never overflows int32.
static int distance_guard(np_ctx *c, const Space &s, const DevState &st, bool batch) {
    if (!s.guard) return NP_OK;
    int *h = (int *)c->h_small;
    HIPCHK(launch_max_finite(st, c->d_guard, c->stream));
    HIPCHK(hipMemcpyAsync(h, c->d_guard, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if ((long long)h[0] + s.max_w < (long long)INF) return NP_OK;
    HIPCHK(launch_dist_escape(st, batch, c->d_guard, c->stream));
    HIPCHK(hipMemcpyAsync(h, c->d_guard, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (h[0] > 0) {
        set_error("a shortest distance reaches INF = 2^30 (%d relaxations out of the int32 range of the reference)", h[0]);
        return NP_ERR_CAPACITY;
    }
    return NP_OK;
}

// Captures (or reuses) the graph drive's chunk of sch.graph_chunk iterations for this state.
static int ensure_graph(np_ctx *c, Space &s, const DevState &st, const np_sched &sch, const SchedDev &sd,
                        bool batch, bool unit) {
    const int K = sch.graph_chunk > 0 ? sch.graph_chunk : 16;
    const int wgs = c->num_cus * 8;
    bool same = s.gexec && s.gchunk == K && s.gwgs == wgs && s.gbatch == batch && s.gunit == unit &&
                std::memcmp(&s.gsp, &sd, sizeof(SchedDev)) == 0 &&
                std::memcmp(&s.gst, &st, sizeof(DevState)) == 0;
    if (same) return NP_OK;
    s.drop_graph();
    hipStream_t cs = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        for (int j = 0; j < K && e == hipSuccess; ++j) {
            if (!sd.push) e = launch_compact(st, j, batch, wgs, cs);
            if (e == hipSuccess) e = launch_decide(st, sd, j, cs);
            if (e == hipSuccess) e = launch_expand(st, sd, j, batch, unit, wgs, cs);
        }
        if (e == hipSuccess) e = launch_advance(st, K, cs);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(cs, &g);
        if (e == hipSuccess) e = e2;
        if (e == hipSuccess) {
            s.graph = g;
            e = hipGraphInstantiate(&s.gexec, g, nullptr, nullptr, 0);
        } else if (g) {
            (void)hipGraphDestroy(g);
        }
    }
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) { s.drop_graph(); return hip_fail(e, "graph capture", __LINE__); }
    s.gchunk = K; s.gwgs = wgs; s.gbatch = batch; s.gunit = unit; s.gsp = sd; s.gst = st;
    return NP_OK;
}

// Runs a solve whose state was initialised; fills stats.
static int run_drive(np_ctx *c, Space &s, const DevState &st, const np_sched &sch, const SchedDev &sd,
                     bool batch, bool unit, np_stats *stats) {
    int rc = NP_OK;
    float ms = 0.f;
    int drive = sch.drive;
    // the reference's other schedules are per-iteration launches
    if (sd.scheme != NP_SCHEME_CONSOLIDATE && drive == NP_DRIVE_PERSISTENT) drive = NP_DRIVE_HOSTLOOP;
    if (drive == NP_DRIVE_PERSISTENT) {
        int wgs = persistent_wgs(c, sch, batch, unit, st.kshift >= 0);
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        hipError_t e = launch_persistent(st, sd, batch, unit, wgs, c->stream);
        if (e == hipErrorIllegalState || e == hipErrorCooperativeLaunchTooLarge) {
            // No cooperative queue for this process (another process on the device holds
            // the cooperative dispatch resource) or the grid no longer fits: nothing was
            // enqueued, so the same state is driven one kernel per step instead.
            (void)hipGetLastError();
            drive = NP_DRIVE_HOSTLOOP;
        } else {
            if (e != hipSuccess) return hip_fail(e, "persistent launch", __LINE__);
            HIPCHK(hipEventRecord(c->ev1, c->stream));
            if ((rc = read_ctrl(c))) return rc;
        }
    }
    if (drive == NP_DRIVE_HOSTLOOP) {
        const int wgs = c->num_cus * 8;
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        for (int it = 0;; ++it) {
            if (!sd.push) HIPCHK(launch_compact(st, it, batch, wgs, c->stream));
            HIPCHK(launch_decide(st, sd, it, c->stream));
            HIPCHK(launch_expand(st, sd, it, batch, unit, wgs, c->stream));
            // one small readback per iteration, as the reference's loop (sssp_wrapper.cu:539)
            if ((rc = read_ctrl(c))) return rc;
            if (c->h_ctrl->done || c->h_ctrl->error) break;
        }
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        if ((rc = read_ctrl(c))) return rc;
    } else if (drive == NP_DRIVE_GRAPH) {
        if ((rc = ensure_graph(c, s, st, sch, sd, batch, unit))) return rc;
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        for (;;) {
            HIPCHK(hipGraphLaunch(s.gexec, c->stream));
            if ((rc = read_ctrl(c))) return rc;
            if (c->h_ctrl->done || c->h_ctrl->error) break;
        }
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        if ((rc = read_ctrl(c))) return rc;
    }
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    const Ctrl &h = *c->h_ctrl;
    c->last_iters = h.iters;
    rc = h.error ? h.error : NP_OK;
    if (rc == NP_ERR_TIMEOUT) set_error("a grid barrier timed out; the kernel exited early");
    if (rc == NP_ERR_NOT_CONVERGED) set_error("iteration cap reached");
    if (rc == NP_ERR_BOUNDS) set_error("%s", BOUNDS_MSG);
    if (rc == NP_OK && !unit) rc = distance_guard(c, s, st, batch);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->iterations = h.iters;
        stats->status = rc;
        stats->kernel_ms = ms;
        stats->drive = drive;
        stats->pairs_total = st.kshift >= 0 ? h.pairs_total : h.frontier_total;
        stats->frontier_total = h.frontier_total;
        stats->edges_total = h.edges_total;
        stats->items_total = h.items_total;
    }
    return rc;
}

static int graph_space(np_ctx *c, np_graph *g) {
    if (g->sp_ready) return NP_OK;
    Space &s = g->sp;
    s.nslots = g->n;
    std::memcpy(s.hist, g->hist, sizeof(s.hist));
    s.cls = g->cls;
    s.own_cls = false;
    set_guard(s, g->n, g->max_w);
    int rc = space_alloc(c, s, true);
    if (rc) { s.release(); return rc; }
    if ((rc = ensure_src(c, s, 1))) return rc;
    g->sp_ready = true;
    return NP_OK;
}

// BFS_REC_GPU solution 0 (bfs_flat_gpu, bfs_rec_wrapper.cu:111-133): one flat pass per level
// with a one-byte "changed" readback, until a level writes nothing.  The level array holds
// UNDEFINED and 0 at the source (launch_init_solve).
static int run_bfs_flat(np_ctx *c, const DevState &st, int n, np_stats *stats) {
    const int wgs = c->num_cus * 8;
    int *changed = &c->d_guard[3];
    int *h = (int *)c->h_small;
    int level = 0;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (;; ++level) {
        if (level >= n) { set_error("BFS deeper than n levels"); return NP_ERR_NOT_CONVERGED; }
        HIPCHK(launch_fill_i32(changed, 1, 0, c->stream));
        HIPCHK(launch_bfs_flat(st, level, changed, wgs, c->stream));
        HIPCHK(hipMemcpyAsync(h, changed, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (!h[0]) break;
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_iters = level + 1;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->iterations = level + 1;
        stats->kernel_ms = ms;
        stats->drive = NP_DRIVE_HOSTLOOP;
    }
    return NP_OK;
}

static int solve_single(np_ctx *c, np_graph *g, int src, const np_sched *sched, int *d_dist_user,
                        int *h_dist, bool unit, np_stats *stats) {
    if (!c || !g) { set_error("null context or graph"); return NP_ERR_ARG; }
    if (g->ctx != c) { set_error("graph belongs to another context or its context was destroyed"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    if (src < 0 || src >= g->n) { set_error("source %d out of range [0,%d)", src, g->n); return NP_ERR_ARG; }
    np_sched sch;
    int rc = validate_sched(sched, &sch);
    if (rc) return rc;
    if (!unit && (sch.scheme == NP_SCHEME_BFS_FLAT || sch.scheme == NP_SCHEME_BFS_HIER)) {
        set_error("scheme %d computes BFS levels (np_bfs) only", sch.scheme);
        return NP_ERR_ARG;
    }
    if ((rc = graph_space(c, g))) return rc;
    Space &s = g->sp;
    if (sch.scheme == NP_SCHEME_GLOBAL_BUFFER && !s.sbuf) {
        // at most every queued vertex is heavy in one expansion: the queue's capacity
        HIPCHK(dalloc(&s.sbuf, (size_t)std::max(1ll, s.qcap)));
        HIPCHK(dalloc(&s.sbcnt, 3));
    }
    const int ssrc = g->solver_id(src, c->stream);
    if (ssrc < 0) { set_error("source mapping failed"); return NP_ERR_HIP; }
    // a renumbered graph is solved in its own buffer, then mapped back to original ids
    int *dist = (d_dist_user && !g->rl.active) ? d_dist_user : s.dist;
    const int inf = unit ? UNDEFINED : INF;
    DevState st = make_state(c, s, g->srow, g->scol, unit ? nullptr : g->sw, dist, inf);
    SchedDev sd = make_schedd(c, sch, g->n);
    HIPCHK(hipMemcpyAsync(s.src_slots, &ssrc, sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_init_solve(st, s.src_slots, 1, sd.push, sd.delta, c->stream));
    rc = sch.scheme == NP_SCHEME_BFS_FLAT ? run_bfs_flat(c, st, g->n, stats) : run_drive(c, s, st, sch, sd, false, unit, stats);
    if (rc) return rc;
    const int *result = dist;
    if (g->rl.active) {
        int *target = d_dist_user;
        if (!target) {
            if (!g->out) HIPCHK(dalloc(&g->out, (size_t)g->n));
            target = g->out;
        }
        HIPCHK(launch_unpermute(dist, g->rl.old2new, g->n, target, c->stream));
        result = target;
    }
    if (h_dist) {
        HIPCHK(hipMemcpyAsync(h_dist, result, sizeof(int) * (size_t)g->n, hipMemcpyDeviceToHost, c->stream));
    }
    return ctx_sync(c);
}

// Validates a device CSR (structure, and weights: none negative, every one below INF so
// that dist[u] + w stays in int32), renumbers it when skewed, classes its vertices by
// out-degree and registers it with the context.
static int finish_graph(np_ctx *c, np_graph *g) {
    int *hist = nullptr;
    int chk[3] = {0, 0, 0};
    HIPCHK(dalloc(&hist, NCLS + 3));
    hipError_t e = launch_fill_i32(hist, NCLS + 3, 0, c->stream);
    if (e == hipSuccess) e = launch_validate_csr(g->row, g->col, g->w, g->n, g->m, hist + NCLS, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(chk, hist + NCLS, sizeof(chk), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { dfree(hist); return hip_fail(e, "CSR validation", __LINE__); }
    if (chk[0]) {
        dfree(hist);
        set_error("invalid CSR: %d violations of row[0]=0, row non-decreasing, row[n]=m, 0<=col<n", chk[0]);
        return NP_ERR_ARG;
    }
    if (chk[1]) {
        dfree(hist);
        set_error("%d negative edge weights: shortest paths need weights >= 0", chk[1]);
        return NP_ERR_ARG;
    }
    if (chk[2] >= INF) {
        dfree(hist);
        set_error("edge weight %d >= INF = 2^30: dist[u] + w could overflow the int32 distances", chk[2]);
        return NP_ERR_CAPACITY;
    }
    g->max_w = chk[2];
    g->srow = g->row; g->scol = g->col; g->sw = g->w;
    if (c->relabel != 0) {
        const double min_skew = c->relabel == 1 ? -1.0 : 0.4;
        e = relabel_by_indegree(g->row, g->col, g->w, g->n, g->m, min_skew, &g->rl, c->stream);
        if (e != hipSuccess) { dfree(hist); return hip_fail(e, "in-degree renumbering", __LINE__); }
        if (g->rl.active) { g->srow = g->rl.row; g->scol = g->rl.col; g->sw = g->rl.w; }
    }
    e = dalloc(&g->cls, (size_t)g->n);
    if (e == hipSuccess) e = launch_fill_i32(hist, NCLS, 0, c->stream);
    if (e == hipSuccess) e = launch_degree_class(g->srow, g->n, g->cls, hist, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(g->hist, hist, sizeof(g->hist), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(hist);
    if (e != hipSuccess) return hip_fail(e, "degree classes", __LINE__);
    c->graphs.push_back(g);
    return NP_OK;
}


static void init_model_params(float *p) {
    // He-style deterministic initialisation from SplitMix64 draws.
    uint64_t key = stream_key(12345, 99);
    const int fan[3] = {NP_MLP_NIN, 32, 32};
    const int sizes[3] = {32 * NP_MLP_NIN, 32 * 32, 5 * 32};
    const int bias[3] = {32, 32, 5};
    int off = 0, k = 0;
    for (int l = 0; l < 3; ++l) {
        float sc = std::sqrt(2.0f / (float)fan[l]);
        for (int i = 0; i < sizes[l]; ++i, ++k) {
            double u = (double)(draw(key, (uint64_t)k) >> 11) * (1.0 / 9007199254740992.0);
            p[off++] = (float)((2.0 * u - 1.0) * sc);
        }
        for (int i = 0; i < bias[l]; ++i) p[off++] = 0.f;
    }
}

// Device resources of a graph / batch (the host object stays).
static void graph_release(np_graph *g) {
    g->sp.release();
    g->sp_ready = false;
    if (g->owns) { dfree(g->row); dfree(g->col); dfree(g->w); }
    dfree(g->rl.row); dfree(g->rl.col); dfree(g->rl.w); dfree(g->rl.old2new);
    g->rl.active = false;
    dfree(g->cls);
    dfree(g->out);
}

static void batch_release(np_batch *b) {
    b->sp.release();
    for (Pack *pk : b->packs) { pk->sp.release(); delete pk; }
    b->packs.clear();
    if (b->owns_graph) { dfree(b->row); dfree(b->col); dfree(b->w); }
    dfree(b->out);
    b->nconcat = 0;
}

// ------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------
extern "C" {

const char *np_last_error(void) { return g_err; }
#ifndef NP_BOUNDS_CHECK
#define NP_BOUNDS_CHECK 0
#endif
const char *np_version(void) { return NP_BOUNDS_CHECK ? "np-sssp gfx950 1.0 bounds-checked" : "np-sssp gfx950 1.0"; }


// ------------------------------------------------------------------------------------
// allocation-free single-source solves (np_sssp_device_noalloc)
// ------------------------------------------------------------------------------------
// Solver state of one solve on a graph of at most NOALLOC_MAX_N vertices, carved from a
// file-scope device array: distances (renumbered graphs), two update bitmaps, two frontier
// queues, the delayed buffer of NP_SCHEME_GLOBAL_BUFFER, class counters and offsets and the
// source slot.  Every GPU the library runs on holds one copy until the process ends; calls
// on one GPU share it, so calls that overlap on one GPU race.
static constexpr long long NOALLOC_MAX_N = NP_NOALLOC_MAX_VERTICES;
static constexpr size_t NA_ALIGN = 256;
static constexpr size_t na_round(size_t b) { return (b + NA_ALIGN - 1) / NA_ALIGN * NA_ALIGN; }
static constexpr size_t noalloc_bytes(long long n) {
    return 4 * na_round(4 * (size_t)n)                    // dist, queue[0], queue[1], sbuf
           + 2 * na_round(4 * (size_t)((n + 31) / 32))   // bm[0], bm[1]
           + na_round(4 * 3 * NCLS) + na_round(4 * (NCLS + 1)) + na_round(4) + na_round(4 * 3);
}
static constexpr size_t NOALLOC_BYTES = noalloc_bytes(NOALLOC_MAX_N);
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static constexpr int NA_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[NA_MAX_DEVICES] = {nullptr};

// The device address of this device's `scratch` (the device must be current).  In host
// code the array's name is only a host-side stand-in; kernels get this address.
static int scratch_setup(int device, unsigned char **out) {
    *out = nullptr;
    if (device >= NA_MAX_DEVICES) return NP_OK;  // no allocation-free solves on such a device
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch symbol address", __LINE__);
        g_scratch_ptr[device] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[device];
    return NP_OK;
}

int np_create(int device, np_ctx **out) {
    if (!out) { set_error("null out"); return NP_ERR_ARG; }
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("no HIP device visible");
        return NP_ERR_UNSUPPORTED;
    }
    if (device < 0 || device >= ndev) { set_error("device %d out of range", device); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("device %d is %s; this library is built for gfx950 only", device, prop.gcnArchName);
        return NP_ERR_UNSUPPORTED;
    }
    if (!prop.cooperativeLaunch) { set_error("cooperative launch unsupported"); return NP_ERR_UNSUPPORTED; }
    np_ctx *c = new np_ctx();
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    c->own_stream = true;
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) e = dalloc(&c->ctrl, 1);
    if (e == hipSuccess) e = hipHostMalloc((void **)&c->h_ctrl, sizeof(Ctrl), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&c->h_small, 4 * sizeof(unsigned long long), hipHostMallocDefault);
    if (e == hipSuccess) e = dalloc(&c->trace, TRACE_CAP);
    if (e == hipSuccess) e = dalloc(&c->model, NP_MLP_NPARAM);
    if (e == hipSuccess) e = dalloc(&c->d_work2, 4);
    if (e == hipSuccess) e = dalloc(&c->d_guard, 4);
    if (e != hipSuccess) {
        int rc = hip_fail(e, "context setup", __LINE__);
        np_destroy(c);
        return rc;
    }
    std::memset(c->h_ctrl, 0, sizeof(Ctrl));
    int rc = scratch_setup(device, &c->scratch);
    if (rc) { np_destroy(c); return rc; }
    *out = c;
    return NP_OK;
}

int np_set_stream(np_ctx *c, void *stream) {
    if (!c || !stream) { set_error("null argument"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    c->stream = (hipStream_t)stream;
    c->own_stream = false;
    return NP_OK;
}

void *np_get_stream(np_ctx *c) { return c ? (void *)c->stream : nullptr; }

int np_set_option(np_ctx *c, int32_t key, int32_t value) {
    if (!c) { set_error("null context"); return NP_ERR_ARG; }
    switch (key) {
    case NP_CTX_RELABEL:
        if (value < 0 || value > 2) { set_error("NP_CTX_RELABEL takes 0, 1 or 2"); return NP_ERR_ARG; }
        c->relabel = value;
        return NP_OK;
    case NP_CTX_PACK_MAX:
        if (value < 2 || value > PACK_MAX) { set_error("NP_CTX_PACK_MAX takes 2..%d", PACK_MAX); return NP_ERR_ARG; }
        c->pack_max = value;
        return NP_OK;
    case NP_CTX_PACK:
        if (value < 0 || value > 1) { set_error("NP_CTX_PACK takes 0 or 1"); return NP_ERR_ARG; }
        c->pack = value;
        return NP_OK;
    default:
        set_error("unknown option");
        return NP_ERR_ARG;
    }
}

int np_graph_relabeled(const np_graph *g, double *skew) {
    if (!g) { set_error("null graph"); return NP_ERR_ARG; }
    if (skew) *skew = g->rl.skew;
    return g->rl.active ? 1 : 0;
}

int np_destroy(np_ctx *c) {
    if (!c) return NP_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    // batches first: they read their graphs
    for (np_batch *b : c->batches) { batch_release(b); b->ctx = nullptr; }
    for (np_graph *g : c->graphs) { graph_release(g); g->ctx = nullptr; }
    for (void *p : c->allocs) (void)hipFree(p);
    c->batches.clear(); c->graphs.clear(); c->allocs.clear();
    c->step_g = nullptr;
    dfree(c->ctrl); dfree(c->trace); dfree(c->model); dfree(c->adam_m); dfree(c->adam_v);
    dfree(c->d_work2); dfree(c->d_guard); dfree(c->snap_dist); dfree(c->snap_bm);
    if (c->h_ctrl) (void)hipHostFree(c->h_ctrl);
    if (c->h_small) (void)hipHostFree(c->h_small);
    if (c->ws) (void)hipFree(c->ws);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return NP_OK;
}

int np_device_info(np_ctx *c, int32_t *num_cus, int32_t *max_wgs_per_cu) {
    if (!c) { set_error("null context"); return NP_ERR_ARG; }
    if (num_cus) *num_cus = c->num_cus;
    if (max_wgs_per_cu) *max_wgs_per_cu = persistent_max_wgs_per_cu(false, false, false, false);
    return NP_OK;
}

int np_dev_alloc(np_ctx *c, uint64_t bytes, void **d_ptr) {
    if (!c || !d_ptr) { set_error("null argument"); return NP_ERR_ARG; }
    *d_ptr = nullptr;
    HIPCHK(hipSetDevice(c->device));
    hipError_t e = hipMalloc(d_ptr, bytes ? bytes : 1);
    if (e != hipSuccess) { set_error("hipMalloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e)); return NP_ERR_NOMEM; }
    c->allocs.push_back(*d_ptr);
    return NP_OK;
}

int np_dev_free(np_ctx *c, void *d_ptr) {
    if (!c) { set_error("null context"); return NP_ERR_ARG; }
    if (d_ptr) {
        auto it = std::find(c->allocs.begin(), c->allocs.end(), d_ptr);
        if (it == c->allocs.end()) { set_error("pointer was not allocated by np_dev_alloc on this context"); return NP_ERR_ARG; }
        c->allocs.erase(it);
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(d_ptr));
    }
    return NP_OK;
}

int np_dev_copy(np_ctx *c, void *dst, const void *src, uint64_t bytes, int32_t kind) {
    if (!c || (!dst && bytes) || (!src && bytes)) { set_error("null argument"); return NP_ERR_ARG; }
    if (kind < 0 || kind > 2) { set_error("bad copy kind"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    if (!bytes) return NP_OK;
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, k, c->stream));
    return ctx_sync(c);
}

int np_sync(np_ctx *c) {
    if (!c) { set_error("null context"); return NP_ERR_ARG; }
    return ctx_sync(c);
}

static int new_graph(np_ctx *c, int32_t n, int64_t m, np_graph **out) {
    if (!c || !out) { set_error("null argument"); return NP_ERR_ARG; }
    if (n < 0 || m < 0) { set_error("negative sizes"); return NP_ERR_ARG; }
    if (m > 0x7FFFFFFFll) { set_error("m = %lld exceeds the int32 CSR of the reference", (long long)m); return NP_ERR_CAPACITY; }
    *out = nullptr;
    return NP_OK;
}

int np_csr_validate(int32_t n, int64_t m, const int32_t *row, const int32_t *col, const int32_t *w,
                    int32_t *max_w) {
    if (n < 0 || m < 0 || !row || (m > 0 && !col)) { set_error("bad CSR arguments"); return NP_ERR_ARG; }
    if (m > 0x7FFFFFFFll) { set_error("m = %lld exceeds the int32 CSR of the reference", (long long)m); return NP_ERR_CAPACITY; }
    if (row[0] != 0 || row[n] != m) { set_error("row[0] must be 0 and row[n] must be m"); return NP_ERR_ARG; }
    for (int32_t v = 0; v < n; ++v)
        if (row[v + 1] < row[v]) { set_error("row decreases at vertex %d", v); return NP_ERR_ARG; }
    int32_t mx = 0;
    for (int64_t e = 0; e < m; ++e) {
        if (col[e] < 0 || col[e] >= n) { set_error("col[%lld] = %d outside [0,%d)", (long long)e, col[e], n); return NP_ERR_ARG; }
        if (w) {
            if (w[e] < 0) { set_error("negative weight %d at edge %lld: shortest paths need weights >= 0", w[e], (long long)e); return NP_ERR_ARG; }
            mx = std::max(mx, w[e]);
        }
    }
    if (mx >= INF) {
        set_error("edge weight %d >= INF = 2^30: dist[u] + w could overflow the int32 distances", mx);
        return NP_ERR_CAPACITY;
    }
    if (max_w) *max_w = mx;
    return NP_OK;
}

int np_graph_upload(np_ctx *c, int32_t n, int64_t m, const int32_t *row, const int32_t *col,
                    const int32_t *w, np_graph **out) {
    int rc = new_graph(c, n, m, out);
    if (rc) return rc;
    if (!row || (m > 0 && (!col || !w))) { set_error("null CSR array"); return NP_ERR_ARG; }
    if ((rc = np_csr_validate(n, m, row, col, w, nullptr))) return rc;
    HIPCHK(hipSetDevice(c->device));
    np_graph *g = new np_graph();
    g->ctx = c; g->n = n; g->m = m; g->owns = true;
    hipError_t e = dalloc(&g->row, (size_t)n + 1);
    if (e == hipSuccess) e = dalloc(&g->col, (size_t)m);
    if (e == hipSuccess) e = dalloc(&g->w, (size_t)m);
    if (e == hipSuccess) e = hipMemcpyAsync(g->row, row, sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && m) e = hipMemcpyAsync(g->col, col, sizeof(int) * (size_t)m, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && m) e = hipMemcpyAsync(g->w, w, sizeof(int) * (size_t)m, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) { rc = hip_fail(e, "graph upload", __LINE__); np_graph_free(g); return rc; }
    if ((rc = finish_graph(c, g))) { np_graph_free(g); return rc; }
    *out = g;
    return NP_OK;
}

int np_graph_wrap_device(np_ctx *c, int32_t n, int64_t m, const int32_t *d_row, const int32_t *d_col,
                         const int32_t *d_w, np_graph **out) {
    int rc = new_graph(c, n, m, out);
    if (rc) return rc;
    if (!d_row || (m > 0 && (!d_col || !d_w))) { set_error("null CSR array"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    np_graph *g = new np_graph();
    g->ctx = c; g->n = n; g->m = m; g->owns = false;
    g->row = const_cast<int *>(d_row); g->col = const_cast<int *>(d_col); g->w = const_cast<int *>(d_w);
    if ((rc = finish_graph(c, g))) { np_graph_free(g); return rc; }
    *out = g;
    return NP_OK;
}

int np_graph_info(const np_graph *g, int32_t *n, int64_t *m) {
    if (!g) { set_error("null graph"); return NP_ERR_ARG; }
    if (n) *n = g->n;
    if (m) *m = g->m;
    return NP_OK;
}

int np_graph_download(np_ctx *c, const np_graph *g, int32_t *row, int32_t *col, int32_t *w) {
    if (!c || !g) { set_error("null argument"); return NP_ERR_ARG; }
    if (g->ctx != c) { set_error("graph belongs to another context or its context was destroyed"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    if (row) HIPCHK(hipMemcpyAsync(row, g->row, sizeof(int) * ((size_t)g->n + 1), hipMemcpyDeviceToHost, c->stream));
    if (col && g->m) HIPCHK(hipMemcpyAsync(col, g->col, sizeof(int) * (size_t)g->m, hipMemcpyDeviceToHost, c->stream));
    if (w && g->m) HIPCHK(hipMemcpyAsync(w, g->w, sizeof(int) * (size_t)g->m, hipMemcpyDeviceToHost, c->stream));
    return ctx_sync(c);
}

int np_graph_device_ptrs(const np_graph *g, const int32_t **row, const int32_t **col, const int32_t **w) {
    if (!g) { set_error("null graph"); return NP_ERR_ARG; }
    if (row) *row = g->row;
    if (col) *col = g->col;
    if (w) *w = g->w;
    return NP_OK;
}

int np_graph_free(np_graph *g) {
    if (!g) return NP_OK;
    if (np_ctx *c = g->ctx) {  // a graph whose context was destroyed holds nothing on the device
        (void)hipSetDevice(c->device);
        if (c->step_g == g) c->step_g = nullptr;
        unlink_from(c->graphs, g);
        graph_release(g);
    }
    delete g;
    return NP_OK;
}

int np_sssp(np_ctx *c, np_graph *g, int32_t src, const np_sched *sched, int32_t *dist_out, np_stats *stats) {
    if (!dist_out) { set_error("null dist_out"); return NP_ERR_ARG; }
    return solve_single(c, g, src, sched, nullptr, dist_out, false, stats);
}

int np_sssp_device(np_ctx *c, np_graph *g, int32_t src, const np_sched *sched, int32_t *d_dist, np_stats *stats) {
    if (!d_dist) { set_error("null d_dist"); return NP_ERR_ARG; }
    return solve_single(c, g, src, sched, d_dist, nullptr, false, stats);
}

int np_sssp_device_noalloc(np_ctx *c, np_graph *g, int32_t src, const np_sched *sched, int32_t *d_dist,
                           np_stats *stats) {
    if (!c || !g) { set_error("null context or graph"); return NP_ERR_ARG; }
    if (!d_dist) { set_error("null d_dist"); return NP_ERR_ARG; }
    if (g->ctx != c) { set_error("graph belongs to another context or its context was destroyed"); return NP_ERR_ARG; }
    if (!c->scratch) { set_error("no scratch array on device %d", c->device); return NP_ERR_UNSUPPORTED; }
    if ((long long)g->n > NOALLOC_MAX_N) {
        set_error("graph of %d vertices exceeds the %lld vertices of the allocation-free solve's scratch", g->n,
                  NOALLOC_MAX_N);
        return NP_ERR_CAPACITY;
    }
    if (src < 0 || src >= g->n) { set_error("source %d out of range [0,%d)", src, g->n); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    np_sched sch;
    int rc = validate_sched(sched, &sch);
    if (rc) return rc;
    if (sch.scheme == NP_SCHEME_BFS_FLAT || sch.scheme == NP_SCHEME_BFS_HIER) {
        set_error("scheme %d computes BFS levels (np_bfs) only", sch.scheme);
        return NP_ERR_ARG;
    }
    if (sch.drive == NP_DRIVE_GRAPH) sch.drive = NP_DRIVE_HOSTLOOP;  // capturing a graph allocates
    // the solver state of graph_space(), carved from this device's scratch (nothing is freed:
    // the local Space only names the pieces)
    Space s;
    s.nslots = g->n;
    std::memcpy(s.hist, g->hist, sizeof(s.hist));
    s.cls = g->cls;
    set_guard(s, g->n, g->max_w);
    int seg[NCLS + 1];
    seg_from_hist(s.hist, seg, &s.qcap);
    s.nwords = (s.nslots + 31) / 32;
    unsigned char *at = c->scratch;
    auto carve = [&](size_t bytes) { unsigned char *p = at; at += na_round(bytes); return p; };
    const size_t n4 = 4 * (size_t)g->n;
    s.dist = (int *)carve(n4);
    s.queue[0] = (int *)carve(n4);
    s.queue[1] = (int *)carve(n4);
    s.sbuf = (int *)carve(n4);
    s.bm[0] = (unsigned *)carve(4 * (size_t)s.nwords);
    s.bm[1] = (unsigned *)carve(4 * (size_t)s.nwords);
    s.cnt = (int *)carve(4 * 3 * NCLS);
    s.seg_off = (int *)carve(4 * (NCLS + 1));
    s.src_slots = (int *)carve(4);
    s.sbcnt = (int *)carve(4 * 3);
    if ((size_t)(at - c->scratch) > NOALLOC_BYTES) { set_error("scratch layout exceeds its array"); return NP_ERR_CAPACITY; }
    const int ssrc = g->solver_id(src, c->stream);
    if (ssrc < 0) { set_error("source mapping failed"); return NP_ERR_HIP; }
    int *dist = g->rl.active ? s.dist : d_dist;
    DevState st = make_state(c, s, g->srow, g->scol, g->sw, dist, INF);
    SchedDev sd = make_schedd(c, sch, g->n);
    HIPCHK(hipMemcpyAsync(s.seg_off, seg, sizeof(seg), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(s.src_slots, &ssrc, sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_init_solve(st, s.src_slots, 1, sd.push, sd.delta, c->stream));
    rc = run_drive(c, s, st, sch, sd, false, false, stats);
    if (rc) { (void)ctx_sync(c); return rc; }
    if (g->rl.active) HIPCHK(launch_unpermute(dist, g->rl.old2new, g->n, d_dist, c->stream));
    return ctx_sync(c);
}

int np_bfs(np_ctx *c, np_graph *g, int32_t src, const np_sched *sched, int32_t *level_out, np_stats *stats) {
    if (!level_out) { set_error("null level_out"); return NP_ERR_ARG; }
    return solve_single(c, g, src, sched, nullptr, level_out, true, stats);
}

int np_bfs_device(np_ctx *c, np_graph *g, int32_t src, const np_sched *sched, int32_t *d_level, np_stats *stats) {
    if (!d_level) { set_error("null d_level"); return NP_ERR_ARG; }
    return solve_single(c, g, src, sched, d_level, nullptr, true, stats);
}

int np_reached_work(np_ctx *c, np_graph *g, const int32_t *d_dist, int32_t inf, int64_t *n_r, int64_t *m_r) {
    if (!c || !g || !d_dist) { set_error("null argument"); return NP_ERR_ARG; }
    if (g->ctx != c) { set_error("graph belongs to another context or its context was destroyed"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    unsigned long long h[2] = {0, 0};
    HIPCHK(launch_reached_work(g->row, d_dist, g->n, inf, c->d_work2, c->stream));
    HIPCHK(hipMemcpyAsync(h, c->d_work2, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    int rc = ctx_sync(c);
    if (rc) return rc;
    if (n_r) *n_r = (int64_t)h[0];
    if (m_r) *m_r = (int64_t)h[1];
    return NP_OK;
}

// ---- betweenness centrality ----------------------------------------------------------
// bc_np_consolidate_gpu + BC_GPU (apps/bc/src/gpu/bc_wrapper.cu:678-809): sources in batches of
// up to 32; per batch a level-synchronous forward BFS with path counts records every level's
// entries, then the backward sweep replays them from the deepest level up.
struct BcBufs {
    int *lvl = nullptr, *ent_v = nullptr, *owner = nullptr, *srcs = nullptr, *err = nullptr;
    int *ectl = nullptr, *lv = nullptr, *h_lv = nullptr;
    unsigned *ent_m = nullptr, *nmask = nullptr;
    double *sig = nullptr, *del = nullptr, *acc = nullptr;
    long long *cst = nullptr, *h_total = nullptr;
    unsigned long long *ctot = nullptr;
    float *out = nullptr;
    ~BcBufs() {
        dfree(lvl); dfree(ent_v); dfree(owner); dfree(srcs); dfree(err); dfree(ent_m); dfree(nmask);
        dfree(sig); dfree(del); dfree(acc); dfree(cst); dfree(ctot); dfree(out); dfree(ectl); dfree(lv);
        if (h_total) (void)hipHostFree(h_total);
        if (h_lv) (void)hipHostFree(h_lv);
    }
};

int np_bc(np_ctx *c, np_graph *g, const int32_t *srcs, int32_t nsrc, float *bc_out, np_stats *stats) {
    if (!c || !g || g->ctx != c || !bc_out || nsrc < 0 || (nsrc > 0 && !srcs)) {
        set_error("bad arguments"); return NP_ERR_ARG;
    }
    HIPCHK(hipSetDevice(c->device));
    const int n = g->n;
    std::vector<int> S;
    if (nsrc == 0) {  // every vertex, as BC_GPU does (bc_wrapper.cu:692)
        S.resize((size_t)n);
        for (int i = 0; i < n; ++i) S[(size_t)i] = i;
    } else {
        S.assign(srcs, srcs + nsrc);
        for (int i = 0; i < nsrc; ++i)
            if (S[(size_t)i] < 0 || S[(size_t)i] >= n) { set_error("source %d out of range", S[(size_t)i]); return NP_ERR_ARG; }
    }
    np_stats st{};
    st.drive = NP_DRIVE_HOSTLOOP;
    if (n == 0) { if (stats) *stats = st; return NP_OK; }
    const int P = (int)std::min<size_t>(32, S.size() ? S.size() : 1);
    int ks = 0;
    while ((1 << ks) < P) ++ks;
    const long long cap_e = (long long)n * P;  // entries of one batch: at most one per (vertex, source)
    if (cap_e > 0x7FFFFFFFll) { set_error("n x batch sources exceeds 2^31 entries"); return NP_ERR_CAPACITY; }
    const long long cap_owner = g->m / 2 + (long long)n + 1;  // chunks of one level
    if (cap_owner > 0x7FFFFFFFll) { set_error("too many chunks for one level"); return NP_ERR_CAPACITY; }
    BcBufs bb;
    const size_t slots = (size_t)n << ks;
    HIPCHK(dalloc(&bb.lvl, slots));
    HIPCHK(dalloc(&bb.sig, slots));
    HIPCHK(dalloc(&bb.del, slots));
    HIPCHK(dalloc(&bb.nmask, (size_t)n));
    HIPCHK(dalloc(&bb.ent_v, (size_t)cap_e));
    HIPCHK(dalloc(&bb.ent_m, (size_t)cap_e));
    HIPCHK(dalloc(&bb.cst, (size_t)n + 1));
    HIPCHK(dalloc(&bb.owner, (size_t)cap_owner));
    HIPCHK(dalloc(&bb.err, 1));
    HIPCHK(hipMemsetAsync(bb.err, 0, sizeof(int), c->stream));
    HIPCHK(dalloc(&bb.srcs, 32));
    HIPCHK(dalloc(&bb.acc, (size_t)n));
    HIPCHK(dalloc(&bb.out, (size_t)n));
    HIPCHK(hipHostMalloc((void **)&bb.h_total, 2 * sizeof(long long)));
    // Forward levels go out several at a time with the entry counts kept on the device; the host
    // reads each run's (first entry, count) pairs once.  The first run is as long as the previous
    // source batch's depth + 1 (depths barely vary on one graph), at most BC_RUN; runs after it
    // have 2, 4, 8, ... levels (deep graphs), at most BC_RUN; levels past the last are empty and
    // cost a few small launches.
    constexpr int BC_RUN = 64;
    HIPCHK(dalloc(&bb.ectl, 2));  // first entry of the current level, its entry count
    HIPCHK(dalloc(&bb.lv, 2 * ((size_t)n + BC_RUN + 1)));
    HIPCHK(hipHostMalloc((void **)&bb.h_lv, 2 * BC_RUN * sizeof(int)));
    int run = 8;
    // chunk counters, one per forward level (ctot[d]) and per backward level (ctot[NB + li]),
    // zeroed once per source batch
    const size_t NB = (size_t)n + BC_RUN + 1;
    HIPCHK(dalloc(&bb.ctot, 2 * NB));
    HIPCHK(hipMemsetAsync(bb.acc, 0, sizeof(double) * (size_t)n, c->stream));
    BcDev b{g->row, g->col, n, ks, bb.lvl, bb.sig, bb.del, bb.nmask};
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    std::vector<std::pair<long long, int>> levels;
    for (size_t b0 = 0; b0 < S.size(); b0 += (size_t)P) {
        const int np_ = (int)std::min<size_t>((size_t)P, S.size() - b0);
        HIPCHK(hipMemcpyAsync(bb.srcs, S.data() + b0, sizeof(int) * (size_t)np_, hipMemcpyHostToDevice, c->stream));
        HIPCHK(launch_bc_reset(b, c->stream));
        HIPCHK(launch_bc_sources(b, bb.srcs, np_, c->stream));
        levels.clear();
        long long E = 0;
        HIPCHK(hipMemsetAsync(bb.ectl, 0, 2 * sizeof(int), c->stream));
        HIPCHK(hipMemsetAsync(bb.ctot, 0, 2 * NB * sizeof(unsigned long long), c->stream));
        for (int d0 = 0, last = -1, k = run, grow = 2; last < 0;
             d0 += k, k = grow, grow = std::min(BC_RUN, 2 * grow)) {
            if (d0 > n) { set_error("more than n BFS levels"); return NP_ERR_NOT_CONVERGED; }
            for (int d = d0; d < d0 + k; ++d) {
                // entry count and chunk total stay on the device: compaction (reserving each
                // entry's chunks), owners, expansion, then the level's record
                HIPCHK(launch_bc_compact(b, bb.ent_v, bb.ent_m, bb.ectl, bb.cst, bb.ectl + 1, bb.ctot + d, c->stream));
                HIPCHK(launch_bc_owner(b, bb.ent_v, bb.ent_m, bb.ectl, bb.cst, n, bb.ectl + 1, bb.owner, cap_owner,
                                       bb.err, c->stream));
                HIPCHK(launch_bc_expand(b, true, bb.ent_v, bb.ent_m, bb.ectl, bb.cst, bb.owner, bb.ctot + d, cap_owner,
                                        d, c->num_cus, c->stream));
                HIPCHK(launch_bc_advance(bb.ectl, bb.ectl + 1, bb.lv, d, c->stream));
            }
            HIPCHK(hipMemcpyAsync(bb.h_lv, bb.lv + 2 * (size_t)d0, 2 * (size_t)k * sizeof(int), hipMemcpyDeviceToHost,
                                  c->stream));
            int rc = ctx_sync(c);
            if (rc) return rc;
            for (int j = 0; j < k; ++j) {
                const int ne = bb.h_lv[2 * j + 1];
                if (ne <= 0) { last = d0 + j; break; }
                levels.push_back({(long long)bb.h_lv[2 * j], ne});
                E += ne;
            }
        }
        run = std::min<int>(BC_RUN, (int)levels.size() + 1);
        // the deepest level has no successors (dependency 0): its slots only take their factor
        // 1 / sigma (k_bc_accum adds zeros to bc); level 0 holds the sources
        if (levels.size() >= 2) {
            const long long E0 = levels.back().first;
            HIPCHK(launch_bc_accum(b, bb.ent_v + E0, bb.ent_m + E0, levels.back().second, bb.acc, c->stream));
        }
        for (int li = (int)levels.size() - 2; li >= 1; --li) {
            const long long E0 = levels[(size_t)li].first;
            const int ne = levels[(size_t)li].second;
            unsigned long long *ct = bb.ctot + NB + li;
            HIPCHK(launch_bc_chunks(b, bb.ent_v + E0, bb.ent_m + E0, ne, bb.cst, ct, c->stream));
            HIPCHK(launch_bc_owner(b, bb.ent_v + E0, bb.ent_m + E0, nullptr, bb.cst, ne, nullptr, bb.owner, cap_owner,
                                   bb.err, c->stream));
            HIPCHK(launch_bc_expand(b, false, bb.ent_v + E0, bb.ent_m + E0, nullptr, bb.cst, bb.owner, ct, cap_owner,
                                    li, c->num_cus, c->stream));
            HIPCHK(launch_bc_accum(b, bb.ent_v + E0, bb.ent_m + E0, ne, bb.acc, c->stream));
        }
        st.iterations += (int)levels.size();
        st.frontier_total += E;
    }
    HIPCHK(launch_bc_to_float(bb.acc, n, bb.out, c->stream));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipMemcpyAsync(bc_out, bb.out, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(bb.h_total, bb.err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    int rc = ctx_sync(c);
    if (rc) return rc;
    int err = 0;
    std::memcpy(&err, bb.h_total, sizeof(int));
    if (err) { set_error("a level exceeded its chunk capacity"); return NP_ERR_CAPACITY; }
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    st.kernel_ms = ms;
    if (stats) *stats = st;
    return NP_OK;
}

// ---- graph coloring --------------------------------------------------------------------
// GRAPH_COLOR_GPU (apps/graph-color/src/gpu/graph_color_wrapper.cu:598-699): method 0 = the
// dependency-counting frontier (O(m)), method 1 = the reference's rounds (O(rounds x m)).
struct ColorBufs {
    int *dep = nullptr, *cursor = nullptr, *ncol = nullptr, *q[3] = {nullptr, nullptr, nullptr}, *qn = nullptr,
        *color = nullptr, *h = nullptr;
    long long *ncnt = nullptr, *nrow = nullptr;
    void *temp = nullptr;
    hipStream_t cs = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    ~ColorBufs() {
        dfree(dep); dfree(cursor); dfree(ncol); dfree(q[0]); dfree(q[1]); dfree(q[2]); dfree(qn); dfree(color);
        dfree(ncnt); dfree(nrow);
        if (temp) (void)hipFree(temp);
        if (h) (void)hipHostFree(h);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cs) (void)hipStreamDestroy(cs);
    }
};
static constexpr int COLOR_CHUNK = 30;  // rounds per captured graph (a multiple of the 3 rotating queues)

int np_color(np_ctx *c, np_graph *g, int32_t method, int32_t *color_out, np_stats *stats) {
    if (!c || !g || g->ctx != c || !color_out || method < 0 || method > 1) { set_error("bad arguments"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    const int n = g->n;
    np_stats st{};
    st.drive = NP_DRIVE_HOSTLOOP;
    if (n == 0) { if (stats) *stats = st; return NP_OK; }
    ColorBufs cb;
    HIPCHK(dalloc(&cb.color, (size_t)n));
    HIPCHK(hipHostMalloc((void **)&cb.h, 2 * sizeof(int)));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    int rounds = 0;
    long long frontier_total = 0;
    if (method == 1) {
        HIPCHK(dalloc(&cb.qn, 1));
        HIPCHK(hipMemsetAsync(cb.color, 0, sizeof(int) * (size_t)n, c->stream));
        for (int col = 1;; ++col) {
            HIPCHK(hipMemsetAsync(cb.qn, 0, sizeof(int), c->stream));
            HIPCHK(launch_color_round(g->row, g->col, n, cb.color, col, cb.qn, c->stream));
            HIPCHK(hipMemcpyAsync(cb.h, cb.qn, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            int rc = ctx_sync(c);
            if (rc) return rc;
            ++rounds;
            if (cb.h[0] == 0) break;  // nothing colored: every vertex has its color
            if (col > n) { set_error("coloring did not finish in n rounds"); return NP_ERR_NOT_CONVERGED; }
        }
    } else {
        HIPCHK(dalloc(&cb.dep, (size_t)n));
        HIPCHK(dalloc(&cb.cursor, (size_t)n));
        HIPCHK(dalloc(&cb.ncol, (size_t)(g->m > 0 ? g->m : 1)));
        for (int k = 0; k < 3; ++k) HIPCHK(dalloc(&cb.q[k], (size_t)n));
        HIPCHK(dalloc(&cb.qn, 4));  // three rotating frontier counts + the round counter
        HIPCHK(dalloc(&cb.ncnt, (size_t)n + 1));
        HIPCHK(dalloc(&cb.nrow, (size_t)n + 1));
        size_t temp_bytes = 0;
        HIPCHK(color_scan(nullptr, &temp_bytes, cb.ncnt, cb.nrow, n, c->stream));
        HIPCHK(hipMalloc(&cb.temp, temp_bytes ? temp_bytes : 1));
        HIPCHK(hipMemsetAsync(cb.ncnt, 0, sizeof(long long) * ((size_t)n + 1), c->stream));
        HIPCHK(hipMemsetAsync(cb.cursor, 0, sizeof(int) * (size_t)n, c->stream));
        HIPCHK(hipMemsetAsync(cb.qn, 0, 4 * sizeof(int), c->stream));
        HIPCHK(launch_color_count(g->row, g->col, n, cb.dep, cb.ncnt, c->stream));
        size_t tb = temp_bytes;
        HIPCHK(color_scan(cb.temp, &tb, cb.ncnt, cb.nrow, n, c->stream));
        HIPCHK(launch_color_fill(g->row, g->col, n, cb.nrow, cb.cursor, cb.ncol, c->stream));
        HIPCHK(launch_color_init(cb.dep, n, cb.color, cb.q[0], cb.qn, c->stream));
        // COLOR_CHUNK rounds captured once: round r reads queue r % 3 and fills (r + 1) % 3
        HIPCHK(hipStreamCreateWithFlags(&cb.cs, hipStreamNonBlocking));
        HIPCHK(hipStreamBeginCapture(cb.cs, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        for (int r = 0; r < COLOR_CHUNK && e == hipSuccess; ++r)
            e = launch_color_step(cb.q[r % 3], cb.qn + r % 3, cb.q[(r + 1) % 3], cb.qn + (r + 1) % 3,
                                  cb.qn + (r + 2) % 3, cb.nrow, cb.ncol, cb.dep, cb.color, cb.qn + 3, c->num_cus, cb.cs);
        hipError_t e2 = hipStreamEndCapture(cb.cs, &cb.graph);
        HIPCHK(e);
        HIPCHK(e2);
        HIPCHK(hipGraphInstantiate(&cb.gexec, cb.graph, nullptr, nullptr, 0));
        for (long long done = 0;; done += COLOR_CHUNK) {
            if (done > (long long)n + COLOR_CHUNK) { set_error("coloring did not finish in n rounds"); return NP_ERR_NOT_CONVERGED; }
            HIPCHK(hipGraphLaunch(cb.gexec, c->stream));
            HIPCHK(hipMemcpyAsync(cb.h, cb.qn, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));  // count 0
            int rc = ctx_sync(c);
            if (rc) return rc;
            if (cb.h[0] == 0) break;  // the frontier of the next chunk's first round is empty
        }
        HIPCHK(hipMemcpyAsync(cb.h, cb.qn + 3, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        int rc = ctx_sync(c);
        if (rc) return rc;
        rounds = cb.h[0];
        frontier_total = n;  // every vertex is in exactly one round's frontier
        st.drive = NP_DRIVE_GRAPH;
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipMemcpyAsync(color_out, cb.color, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    int rc = ctx_sync(c);
    if (rc) return rc;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    st.kernel_ms = ms;
    st.iterations = rounds;
    st.frontier_total = frontier_total;
    if (stats) *stats = st;
    return NP_OK;
}

// ---- PageRank and SpMV ----------------------------------------------------------------------
}  // extern "C"

// Buffers carved from the context's workspace: a first pass with base == nullptr sizes the
// layout, the second hands out pointers (each 256-byte aligned).
struct Carve {
    char *base = nullptr;
    size_t off = 0;
    template <class T>
    T *take(size_t count) {
        off = (off + 255) & ~(size_t)255;
        T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off += sizeof(T) * (count ? count : 1);
        return p;
    }
};

// The context's workspace, grown to at least `bytes` (its old contents are dropped).
static hipError_t ctx_workspace(np_ctx *c, size_t bytes, char **p) {
    if (bytes > c->ws_bytes) {
        if (c->ws) (void)hipFree(c->ws);
        c->ws = nullptr;
        c->ws_bytes = 0;
        hipError_t e = hipMalloc((void **)&c->ws, bytes);
        if (e != hipSuccess) return e;
        c->ws_bytes = bytes;
    }
    *p = c->ws;
    return hipSuccess;
}

// Row-bin storage (medium rows, huge rows, their chunks) for n rows holding m entries; the
// three counts are adjacent, cleared by bins_clear.
static void bins_carve(Carve &cv, RowBins &rb, int n, long long m) {
    const size_t nh = (size_t)(m / PR_HUGE) + 1, nc = (size_t)pr_chunk_cap(m);
    rb.med = cv.take<int>((size_t)n);
    rb.huge = cv.take<int>(nh);
    rb.hc0 = cv.take<int>(nh);
    rb.crow = cv.take<int>(nc);
    rb.ck = cv.take<int>(nc);
    rb.part = cv.take<double>(nc);
    rb.nmed = cv.take<int>(3);
    rb.nhuge = rb.nmed ? rb.nmed + 1 : nullptr;
    rb.nchunk = rb.nmed ? rb.nmed + 2 : nullptr;
}
static hipError_t bins_clear(const RowBins &rb, hipStream_t s) { return hipMemsetAsync(rb.nmed, 0, 3 * sizeof(int), s); }

// PAGERANK_GPU (apps/pagerank/src/gpu/pagerank_wrapper.cu:199-244, 565-661): damping 0.5, total
// rank 1000, until no vertex changes by more than 1e-13 (pagerank_wrapper.cu:5-6, pagerank.h:15).
struct PrBufs {
    long long *irow;
    int *src, *pcol;
    unsigned *tkey;
    RowBins rb;
    double *rank, *nrank, *contrib, *dsum;
    unsigned long long *dmax;
    int *ctl;
    float *out;
    void *temp;
};
static void pr_carve(Carve &cv, PrBufs &pb, int n, long long m, size_t temp_bytes) {
    pb.irow = cv.take<long long>((size_t)n + 1);
    pb.src = cv.take<int>((size_t)m);
    pb.pcol = cv.take<int>((size_t)m);
    pb.tkey = cv.take<unsigned>((size_t)m);
    bins_carve(cv, pb.rb, n, m);
    pb.rank = cv.take<double>((size_t)n);
    pb.nrank = cv.take<double>((size_t)n);
    pb.contrib = cv.take<double>((size_t)n);
    pb.dsum = cv.take<double>(2);
    pb.dmax = cv.take<unsigned long long>(2);
    pb.ctl = cv.take<int>(2);
    pb.out = cv.take<float>((size_t)n);
    pb.temp = cv.take<char>(temp_bytes);
}

extern "C" {

int np_pagerank(np_ctx *c, np_graph *g, int32_t max_iters, float *rank_out, np_stats *stats) {
    if (!c || !g || g->ctx != c || !rank_out || max_iters < 0) { set_error("bad arguments"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    const int n = g->n;
    const double damping = 0.5, total = 1000.0, eps = 1.0e-13;
    np_stats st{};
    st.drive = NP_DRIVE_HOSTLOOP;
    if (n == 0) { if (stats) *stats = st; return NP_OK; }
    const long long m = g->m;
    int end_bit = 1;  // key bits covering vertex ids 0..n-1
    while (end_bit < 31 && (1ll << end_bit) < n) ++end_bit;
    const unsigned *tkey_in = reinterpret_cast<const unsigned *>(g->col);
    size_t temp_bytes = 0, scan_bytes = 0;
    HIPCHK(pr_sort(nullptr, &temp_bytes, tkey_in, nullptr, nullptr, nullptr, m, end_bit, c->stream));
    HIPCHK(pr_sources(nullptr, &scan_bytes, g->row, n, m, nullptr, nullptr, c->stream));
    temp_bytes = std::max(temp_bytes, scan_bytes);
    PrBufs pb{};
    Carve cv;
    pr_carve(cv, pb, n, m, temp_bytes);
    HIPCHK(ctx_workspace(c, cv.off, &cv.base));
    cv.off = 0;
    pr_carve(cv, pb, n, m, temp_bytes);
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    HIPCHK(bins_clear(pb.rb, c->stream));
    if (m > 0) {
        size_t sb = scan_bytes;
        HIPCHK(pr_sources(pb.temp, &sb, g->row, n, m, pb.pcol, pb.src, c->stream));  // pcol: scratch until the sort
        size_t tb = temp_bytes;
        HIPCHK(pr_sort(pb.temp, &tb, tkey_in, pb.tkey, pb.src, pb.pcol, m, end_bit, c->stream));
    }
    HIPCHK(launch_pr_offsets(pb.tkey, m, n, pb.irow, c->stream));
    HIPCHK(launch_pr_bins(pb.irow, n, m, pb.rb, c->stream));
    HIPCHK(launch_pr_init(pb.rank, n, total / n, c->stream));
    HIPCHK(hipMemsetAsync(pb.ctl, 0, 2 * sizeof(int), c->stream));
    const double rw = (1.0 - damping) * total / n;
    const int cap = max_iters > 0 ? max_iters : 100000;
    // Iterations go out PR_BATCH at a time; the device ends the run (see launch_pr_iteration) and
    // the host reads the iteration count once per batch instead of the change once per iteration.
    constexpr int PR_BATCH = 8;
    double *r[2] = {pb.rank, pb.nrank};
    int launched = 0, it = 0;
    for (;;) {
        if (launched >= cap) { set_error("PageRank did not converge in %d iterations", cap); return NP_ERR_NOT_CONVERGED; }
        const int k = std::min(PR_BATCH, cap - launched);
        for (int j = 0; j < k; ++j, ++launched)
            HIPCHK(launch_pr_iteration(launched, pb.ctl, g->row, pb.irow, pb.pcol, pb.rb, r[launched & 1],
                                       r[(launched + 1) & 1], pb.contrib, n, damping, rw, eps, pb.dsum, pb.dmax,
                                       c->num_cus, c->stream));
        unsigned long long *h = c->h_small;
        HIPCHK(hipMemcpyAsync(h, pb.ctl, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(h + 1, pb.dmax, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
        int rc = ctx_sync(c);
        if (rc) return rc;
        int hc[2];
        std::memcpy(hc, h, sizeof hc);
        if (hc[0]) { it = hc[1]; break; }  // iteration hc[1] changed no vertex by more than EPSILON
        if (launched >= cap) {             // the last allowed iteration ran: its own change decides
            double dlast;
            std::memcpy(&dlast, h + 1 + ((launched - 1) & 1), sizeof dlast);
            if (!(dlast > eps)) { it = launched; break; }
        }
    }
    HIPCHK(launch_pr_to_float(r[it & 1], n, pb.out, c->stream));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipMemcpyAsync(rank_out, pb.out, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    int rc = ctx_sync(c);
    if (rc) return rc;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    st.kernel_ms = ms;
    st.iterations = it;
    st.frontier_total = (long long)it * n;
    if (stats) *stats = st;
    return NP_OK;
}

// ---- SpMV: SpMV_GPU (apps/SpMV/src/gpu/SpMV_wrapper.cu:314-345) --------------------------
int np_spmv(np_ctx *c, np_graph *g, const double *data, const double *x, double *y, np_stats *stats) {
    if (!c || !g || g->ctx != c || !x || !y || (g->m > 0 && !data)) { set_error("bad arguments"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    const int n = g->n;
    const long long m = g->m;
    np_stats st{};
    st.drive = NP_DRIVE_HOSTLOOP;
    if (n == 0) { if (stats) *stats = st; return NP_OK; }
    double *d_data = nullptr, *d_x = nullptr, *d_y = nullptr;
    RowBins rb{};
    Carve cv;
    for (int pass = 0; pass < 2; ++pass) {  // size, then carve from the context's workspace
        if (pass) {
            HIPCHK(ctx_workspace(c, cv.off, &cv.base));
            cv.off = 0;
        }
        d_data = cv.take<double>((size_t)m);
        d_x = cv.take<double>((size_t)n);
        d_y = cv.take<double>((size_t)n);
        bins_carve(cv, rb, n, m);
    }
    if (m > 0) HIPCHK(hipMemcpyAsync(d_data, data, sizeof(double) * (size_t)m, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d_x, x, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(bins_clear(rb, c->stream));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    HIPCHK(launch_spmv_bins(g->row, n, m, rb, c->stream));
    HIPCHK(launch_spmv(g->row, g->col, d_data, d_x, d_y, n, rb, c->num_cus, c->stream));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipMemcpyAsync(y, d_y, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    int rc = ctx_sync(c);
    if (rc) return rc;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    st.kernel_ms = ms;
    st.iterations = 1;
    if (stats) *stats = st;
    return NP_OK;
}

int np_trace_get(np_ctx *c, np_trace_rec *recs, int32_t cap, int32_t *count) {
    if (!c || !count) { set_error("null argument"); return NP_ERR_ARG; }
    int k = std::min(c->last_iters, TRACE_CAP);
    *count = k;
    if (recs && cap > 0) {
        int t = std::min(k, cap);
        HIPCHK(hipMemcpyAsync(recs, c->trace, sizeof(np_trace_rec) * (size_t)t, hipMemcpyDeviceToHost, c->stream));
        return ctx_sync(c);
    }
    return NP_OK;
}

// ---- batches -----------------------------------------------------------------------
// A batch holds two kinds of parts:
//   packs         2..32 sources of one graph in one source-packed state (dist rows of
//                 2^kshift sources per vertex), one solve per pack
//   concatenated  every other instance: graphs laid end to end in one slot space, one
//                 solve for all of them
// Instances of a graph that appears more than once go to packs (NP_CTX_PACK, default on).
static int build_pack(np_ctx *c, np_batch *b, np_graph *g, const std::vector<int> &ids, const int32_t *srcs) {
    Pack *pk = new Pack();
    b->packs.push_back(pk);
    pk->g = g;
    pk->inst = ids;
    const int k = (int)ids.size();
    int ks = 0;
    while ((1 << ks) < k) ++ks;
    pk->kshift = ks;
    Space &s = pk->sp;
    s.nslots = g->n;
    std::memcpy(s.hist, g->hist, sizeof(s.hist));
    s.cls = g->cls;
    s.own_cls = false;
    set_guard(s, g->n, g->max_w);
    int rc = space_alloc(c, s, false);
    if (rc) return rc;
    HIPCHK(dalloc(&s.dist, (size_t)g->n << ks));
    HIPCHK(dalloc(&s.mflag, (size_t)g->n));
    HIPCHK(dalloc(&s.qmask[0], (size_t)s.qcap));
    HIPCHK(dalloc(&s.qmask[1], (size_t)s.qcap));
    if ((rc = ensure_src(c, s, k))) return rc;
    std::vector<int> sid(k);
    for (int j = 0; j < k; ++j) {
        sid[j] = g->solver_id(srcs[ids[j]], c->stream);
        if (sid[j] < 0) { set_error("source mapping failed"); return NP_ERR_HIP; }
        b->pack_of[ids[j]] = (int)b->packs.size() - 1;
        b->lane_of[ids[j]] = j;
    }
    HIPCHK(hipMemcpyAsync(s.src_slots, sid.data(), sizeof(int) * (size_t)k, hipMemcpyHostToDevice, c->stream));
    return ctx_sync(c);
}

static int build_concat(np_ctx *c, np_batch *b, const std::vector<int> &ids, np_graph *const *graphs,
                        const int32_t *srcs) {
    std::vector<np_graph *> uniq;
    std::vector<int> gi(ids.size());
    for (size_t t = 0; t < ids.size(); ++t) {
        np_graph *g = graphs[ids[t]];
        int k = -1;
        for (size_t j = 0; j < uniq.size(); ++j) if (uniq[j] == g) k = (int)j;
        if (k < 0) { k = (int)uniq.size(); uniq.push_back(g); }
        gi[t] = k;
    }
    Space &s = b->sp;
    std::vector<long long> VB(uniq.size()), EB(uniq.size());
    // every graph runs in its solver layout (renumbered when skewed)
    if (uniq.size() == 1) {
        b->row = const_cast<int *>(uniq[0]->srow); b->col = const_cast<int *>(uniq[0]->scol);
        b->w = const_cast<int *>(uniq[0]->sw);
        b->owns_graph = false;
        VB[0] = 0; EB[0] = 0;
    } else {
        long long nv = 0, ne = 0;
        for (size_t j = 0; j < uniq.size(); ++j) { VB[j] = nv; EB[j] = ne; nv += uniq[j]->n + 1; ne += uniq[j]->m; }
        if (ne > 0x7FFFFFFFll || nv > 0x7FFFFFFFll) { set_error("batch exceeds 2^31 edges"); return NP_ERR_CAPACITY; }
        b->owns_graph = true;
        HIPCHK(dalloc(&b->row, (size_t)nv));
        HIPCHK(dalloc(&b->col, (size_t)ne));
        HIPCHK(dalloc(&b->w, (size_t)ne));
        for (size_t j = 0; j < uniq.size(); ++j) {
            np_graph *g = uniq[j];
            HIPCHK(launch_rebase_row(g->srow, g->n, (int)EB[j], b->row + VB[j], c->stream));
            if (g->m) {
                HIPCHK(hipMemcpyAsync(b->col + EB[j], g->scol, sizeof(int) * (size_t)g->m, hipMemcpyDeviceToDevice, c->stream));
                HIPCHK(hipMemcpyAsync(b->w + EB[j], g->sw, sizeof(int) * (size_t)g->m, hipMemcpyDeviceToDevice, c->stream));
            }
        }
    }
    // slots: instance i owns [base_i, base_i + pad(n_i)), pad to 2^LUT_SHIFT
    const long long G = 1ll << LUT_SHIFT;
    const int cnt = (int)ids.size();
    long long acc = 0;
    std::vector<int2> inst(cnt);
    for (int t = 0; t < cnt; ++t) {
        np_graph *g = uniq[gi[t]];
        b->base[ids[t]] = (int)acc;
        inst[t] = make_int2((int)acc, (int)VB[gi[t]]);
        for (int k = 0; k < NCLS; ++k) s.hist[k] += g->hist[k];
        set_guard(s, g->n, g->max_w);
        acc += ((long long)g->n + G - 1) / G * G;
        if (acc > 0x7FFFFFFFll - G) { set_error("batch exceeds 2^31 vertex slots"); return NP_ERR_CAPACITY; }
    }
    s.nslots = acc > 0 ? acc : G;
    int rc = space_alloc(c, s, true);
    if (rc) return rc;
    HIPCHK(dalloc(&s.cls, (size_t)s.nslots));
    s.own_cls = true;
    HIPCHK(dalloc(&s.lut, (size_t)(s.nslots / G + 1)));
    HIPCHK(dalloc(&s.inst, (size_t)cnt));
    HIPCHK(launch_fill_u32((unsigned *)s.cls, (s.nslots + 3) / 4, 0u, c->stream));
    for (int t = 0; t < cnt; ++t) {
        np_graph *g = uniq[gi[t]];
        HIPCHK(launch_copy_cls(g->cls, s.cls + b->base[ids[t]], g->n, c->stream));
        HIPCHK(launch_fill_lut(s.lut, b->base[ids[t]] / G, ((long long)g->n + G - 1) / G, t, c->stream));
    }
    HIPCHK(hipMemcpyAsync(s.inst, inst.data(), sizeof(int2) * (size_t)cnt, hipMemcpyHostToDevice, c->stream));
    if ((rc = ensure_src(c, s, cnt))) return rc;
    std::vector<int> slots(cnt);
    for (int t = 0; t < cnt; ++t) {
        const int sid = graphs[ids[t]]->solver_id(srcs[ids[t]], c->stream);
        if (sid < 0) { set_error("source mapping failed"); return NP_ERR_HIP; }
        slots[t] = b->base[ids[t]] + sid;
    }
    HIPCHK(hipMemcpyAsync(s.src_slots, slots.data(), sizeof(int) * (size_t)cnt, hipMemcpyHostToDevice, c->stream));
    b->nconcat = cnt;
    return ctx_sync(c);
}

int np_batch_create(np_ctx *c, int32_t count, np_graph *const *graphs, const int32_t *srcs, int32_t unit,
                    np_batch **out) {
    if (!c || !graphs || !srcs || !out || count < 1) { set_error("bad batch arguments"); return NP_ERR_ARG; }
    *out = nullptr;
    HIPCHK(hipSetDevice(c->device));
    std::vector<np_graph *> uniq;
    std::vector<std::vector<int>> by_graph;
    for (int i = 0; i < count; ++i) {
        np_graph *g = graphs[i];
        if (!g || g->ctx != c) { set_error("instance %d: bad graph", i); return NP_ERR_ARG; }
        if (srcs[i] < 0 || srcs[i] >= g->n) { set_error("instance %d: source out of range", i); return NP_ERR_ARG; }
        int k = -1;
        for (size_t j = 0; j < uniq.size(); ++j) if (uniq[j] == g) k = (int)j;
        if (k < 0) { k = (int)uniq.size(); uniq.push_back(g); by_graph.emplace_back(); }
        by_graph[k].push_back(i);
    }
    np_batch *b = new np_batch();
    b->ctx = c; b->count = count; b->unit = unit ? 1 : 0;
    b->nverts.resize(count); b->graphs.resize(count); b->base.assign(count, -1);
    b->pack_of.assign(count, -1); b->lane_of.assign(count, -1); b->out_off.resize(count);
    long long off = 0;
    for (int i = 0; i < count; ++i) {
        b->graphs[i] = graphs[i]; b->nverts[i] = graphs[i]->n;
        b->out_off[i] = off; off += graphs[i]->n;
    }
    b->out_total = off;
    std::vector<int> rest;
    int rc = NP_OK;
    for (size_t j = 0; j < uniq.size() && rc == NP_OK; ++j) {
        const std::vector<int> &ids = by_graph[j];
        const int k = (int)ids.size();
        // a pack's distance rows are indexed with 32-bit offsets: n x (sources, rounded up to a
        // power of two) stays within 2^31
        int cap = c->pack_max;
        while (cap > 1 && (long long)uniq[j]->n * cap > (1ll << 31)) cap >>= 1;
        if (!c->pack || k < 2 || cap < 2) { rest.insert(rest.end(), ids.begin(), ids.end()); continue; }
        const int npk = (k + cap - 1) / cap;  // packs of near-equal size
        for (int p = 0, at = 0; p < npk && rc == NP_OK; ++p) {
            const int sz = k / npk + (p < k % npk ? 1 : 0);
            rc = build_pack(c, b, uniq[j], std::vector<int>(ids.begin() + at, ids.begin() + at + sz), srcs);
            at += sz;
        }
    }
    if (rc == NP_OK && !rest.empty()) {
        std::sort(rest.begin(), rest.end());
        rc = build_concat(c, b, rest, graphs, srcs);
    }
    if (rc) { np_batch_free(b); return rc; }
    c->batches.push_back(b);
    *out = b;
    return NP_OK;
}

// One part of a batch solve.
struct PartJob {
    Space *s;
    DevState st;
    SchedDev sd;
    bool batch;
    bool done;
};

static int part_resources(Space &s) {
    if (s.pstream) return NP_OK;
    HIPCHK(dalloc(&s.pctrl, 1));
    HIPCHK(hipHostMalloc((void **)&s.h_pctrl, sizeof(Ctrl), hipHostMallocDefault));
    HIPCHK(hipStreamCreateWithFlags(&s.pstream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&s.pev, hipEventDisableTiming));
    return NP_OK;
}

// Graph drive over several parts at once: each part replays its captured chunk on its own
// stream with its own control block, so the parts' kernels overlap on the device; the host
// polls every part after each round of chunks.  Device time is taken from a fork event on
// the context stream to the join of every part stream.
static int run_parts_concurrent(np_ctx *c, std::vector<PartJob> &jobs, const np_sched &sch, bool unit,
                                np_stats *acc) {
    int rc = NP_OK;
    for (PartJob &j : jobs) {
        if ((rc = ensure_graph(c, *j.s, j.st, sch, j.sd, j.batch, unit))) return rc;
        HIPCHK(hipStreamSynchronize(j.s->pstream));  // initialisation done
    }
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (PartJob &j : jobs) HIPCHK(hipStreamWaitEvent(j.s->pstream, c->ev0, 0));
    size_t active = jobs.size();
    while (active > 0) {
        for (PartJob &j : jobs) {
            if (j.done) continue;
            HIPCHK(hipGraphLaunch(j.s->gexec, j.s->pstream));
            HIPCHK(hipMemcpyAsync(j.s->h_pctrl, j.s->pctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, j.s->pstream));
        }
        for (PartJob &j : jobs) {
            if (j.done) continue;
            HIPCHK(hipStreamSynchronize(j.s->pstream));
            const Ctrl &h = *j.s->h_pctrl;
            if (h.done || h.error) { j.done = true; --active; }
        }
    }
    for (PartJob &j : jobs) {
        HIPCHK(hipEventRecord(j.s->pev, j.s->pstream));
        HIPCHK(hipStreamWaitEvent(c->stream, j.s->pev, 0));
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    acc->kernel_ms = ms;
    acc->drive = NP_DRIVE_GRAPH;
    for (PartJob &j : jobs) {
        const Ctrl &h = *j.s->h_pctrl;
        acc->iterations += h.iters;
        acc->frontier_total += h.frontier_total;
        acc->pairs_total += j.st.kshift >= 0 ? h.pairs_total : h.frontier_total;
        acc->edges_total += h.edges_total;
        acc->items_total += h.items_total;
        if (h.error && rc == NP_OK) rc = h.error;
    }
    if (rc == NP_ERR_NOT_CONVERGED) set_error("iteration cap reached");
    if (rc == NP_ERR_TIMEOUT) set_error("a grid barrier timed out; the kernel exited early");
    if (rc == NP_ERR_BOUNDS) set_error("%s", BOUNDS_MSG);
    for (PartJob &j : jobs)
        if (rc == NP_OK && !unit) rc = distance_guard(c, *j.s, j.st, j.batch);
    return rc;
}

int np_batch_solve(np_ctx *c, np_batch *b, const np_sched *sched, np_stats *stats) {
    if (!c || !b || b->ctx != c) { set_error("bad batch"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    np_sched sch;
    int rc = validate_sched(sched, &sch);
    if (rc) return rc;
    if (sch.scheme != NP_SCHEME_CONSOLIDATE) {
        set_error("batches run the consolidation bins (scheme 0) only");
        return NP_ERR_ARG;
    }
    const bool unit = b->unit != 0;
    const int inf = unit ? UNDEFINED : INF;
    np_stats acc;
    std::memset(&acc, 0, sizeof acc);
    acc.drive = -1;
    auto add = [&](const np_stats &p) {
        acc.iterations += p.iterations;
        acc.kernel_ms += p.kernel_ms;
        acc.frontier_total += p.frontier_total;
        acc.pairs_total += p.pairs_total;
        acc.edges_total += p.edges_total;
        acc.items_total += p.items_total;
        if (acc.drive < 0) acc.drive = p.drive;
    };
    // packs: scan frontier
    np_sched psch = sch;
    psch.frontier = NP_FRONTIER_SCAN;
    const int nparts = (int)b->packs.size() + (b->nconcat > 0 ? 1 : 0);
    // the graph drive runs the parts side by side, each on its own stream
    const bool together = sch.drive == NP_DRIVE_GRAPH && nparts > 1 && !sch.trace;
    std::vector<PartJob> jobs;
    for (Pack *pk : b->packs) {
        np_graph *g = pk->g;
        Space &s = pk->sp;
        DevState st = make_state(c, s, g->srow, g->scol, unit ? nullptr : g->sw, s.dist, inf);
        st.kshift = pk->kshift;
        st.mflag = s.mflag;
        st.qmask[0] = s.qmask[0]; st.qmask[1] = s.qmask[1];
        st.count_work = sch.trace ? 1 : 0;
        if (psch.pack_hold > 0 && psch.delta <= 0) {
            // hold a heavy vertex while fewer than 3/4 of the pack's sources have reached it
            st.hold_cls = psch.pack_hold;
            st.hold_min = (3 * (int)pk->inst.size() + 3) / 4;
        }
        SchedDev sd = make_schedd(c, psch, g->n);
        if (psch.max_iters <= 0) {
            // Without a window every source's frontier empties within n+1 expansions; with
            // one, every expansion takes the least pending (vertex, source) pair, whose
            // distance is final, for the first time: at most n * sources expansions.
            const long long cap = (long long)g->n * (long long)pk->inst.size() + 2;
            sd.max_iters = cap > 0x7FFFFFFFll ? 0x7FFFFFFF : (int)cap;
        }
        if (together) {
            if ((rc = part_resources(s))) return rc;
            st.ctrl = s.pctrl;
            HIPCHK(launch_init_packed(st, s.src_slots, (int)pk->inst.size(), sd.delta, s.pstream));
            jobs.push_back(PartJob{&s, st, sd, false, false});
            continue;
        }
        HIPCHK(launch_init_packed(st, s.src_slots, (int)pk->inst.size(), sd.delta, c->stream));
        np_stats ps;
        if ((rc = run_drive(c, s, st, psch, sd, false, unit, &ps))) return rc;
        add(ps);
    }
    if (b->nconcat > 0) {
        Space &s = b->sp;
        DevState st = make_state(c, s, b->row, b->col, unit ? nullptr : b->w, s.dist, inf);
        int nmax = 1;
        long long nsum = 2;
        for (int i = 0; i < b->count; ++i)
            if (b->pack_of[i] < 0) { nmax = std::max(nmax, b->nverts[i]); nsum += b->nverts[i]; }
        SchedDev sd = make_schedd(c, sch, nmax);
        if (sch.max_iters <= 0 && sd.delta > 0)  // a window orders the instances' vertices together
            sd.max_iters = nsum > 0x7FFFFFFFll ? 0x7FFFFFFF : (int)nsum;
        if (together) {
            if ((rc = part_resources(s))) return rc;
            st.ctrl = s.pctrl;
            HIPCHK(launch_init_solve(st, s.src_slots, b->nconcat, sd.push, sd.delta, s.pstream));
            jobs.push_back(PartJob{&s, st, sd, true, false});
        } else {
            HIPCHK(launch_init_solve(st, s.src_slots, b->nconcat, sd.push, sd.delta, c->stream));
            np_stats ps;
            if ((rc = run_drive(c, s, st, sch, sd, true, unit, &ps))) return rc;
            add(ps);
        }
    }
    if (together) {
        // packs and the concatenated part share one schedule apart from the frontier mode
        rc = run_parts_concurrent(c, jobs, psch, unit, &acc);
        if (stats) { *stats = acc; stats->status = rc; }
        return rc;
    }
    if (stats) *stats = acc;
    return NP_OK;
}

// Instance i's distances in original vertex ids: unpacked from its pack, or mapped back
// when its graph is renumbered, into the batch's result buffer.
static int batch_result(np_ctx *c, np_batch *b, int i, const int **res) {
    const np_graph *g = b->graphs[i];
    const int *old2new = g->rl.active ? g->rl.old2new : nullptr;
    const int pk = b->pack_of[i];
    if (pk < 0 && !old2new) {
        *res = b->sp.dist + b->base[i];
        return NP_OK;
    }
    if (!b->out) HIPCHK(dalloc(&b->out, (size_t)b->out_total));
    int *dst = b->out + b->out_off[i];
    if (pk >= 0) {
        const Pack *p = b->packs[pk];
        HIPCHK(launch_unpack(p->sp.dist, p->kshift, b->lane_of[i], old2new, g->n, dst, c->stream));
    } else {
        HIPCHK(launch_unpermute(b->sp.dist + b->base[i], old2new, g->n, dst, c->stream));
    }
    *res = dst;
    return NP_OK;
}

int np_batch_get(np_ctx *c, np_batch *b, int32_t i, int32_t *dist_out) {
    if (!c || !b || !dist_out || i < 0 || i >= b->count) { set_error("bad argument"); return NP_ERR_ARG; }
    if (b->ctx != c) { set_error("batch belongs to another context or its context was destroyed"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    const int *d = nullptr;
    int rc = batch_result(c, b, i, &d);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(dist_out, d, sizeof(int) * (size_t)b->nverts[i], hipMemcpyDeviceToHost, c->stream));
    return ctx_sync(c);
}

int np_batch_device_dist(np_batch *b, int32_t i, const int32_t **d_dist) {
    if (!b || !d_dist || i < 0 || i >= b->count) { set_error("bad argument"); return NP_ERR_ARG; }
    if (!b->ctx) { set_error("the batch's context was destroyed"); return NP_ERR_ARG; }
    int rc = batch_result(b->ctx, b, i, d_dist);
    if (rc) return rc;
    return ctx_sync(b->ctx);
}

int np_batch_layout(const np_batch *b, int32_t *packs, int32_t *concatenated) {
    if (!b) { set_error("null batch"); return NP_ERR_ARG; }
    if (packs) *packs = (int32_t)b->packs.size();
    if (concatenated) *concatenated = b->nconcat;
    return NP_OK;
}

int np_batch_free(np_batch *b) {
    if (!b) return NP_OK;
    if (np_ctx *c = b->ctx) {  // a batch whose context was destroyed holds nothing on the device
        (void)hipSetDevice(c->device);
        unlink_from(c->batches, b);
        batch_release(b);
    }
    delete b;
    return NP_OK;
}

// ---- learned scheduler -------------------------------------------------------------
int np_model_set(np_ctx *c, const float *params) {
    if (!c || !params) { set_error("null argument"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    for (int i = 0; i < NP_MLP_NPARAM; ++i)
        if (!std::isfinite(params[i])) { set_error("parameter %d is not finite", i); return NP_ERR_ARG; }
    HIPCHK(hipMemcpyAsync(c->model, params, sizeof(float) * NP_MLP_NPARAM, hipMemcpyHostToDevice, c->stream));
    int rc = ctx_sync(c);
    if (rc) return rc;
    c->has_model = true;
    c->adam_step = 0;
    return NP_OK;
}

int np_model_get(np_ctx *c, float *params) {
    if (!c || !params) { set_error("null argument"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    if (!c->has_model) { set_error("no model set"); return NP_ERR_ARG; }
    HIPCHK(hipMemcpyAsync(params, c->model, sizeof(float) * NP_MLP_NPARAM, hipMemcpyDeviceToHost, c->stream));
    return ctx_sync(c);
}

struct TrainBufs {
    float *x = nullptr, *t = nullptr, *work = nullptr, *grad = nullptr, *loss = nullptr;
    int *lab = nullptr;
    ~TrainBufs() { dfree(x); dfree(t); dfree(work); dfree(grad); dfree(loss); dfree(lab); }
};

static int upload_train(np_ctx *c, TrainBufs &tb, const float *x, const int32_t *label, const float *target,
                        int batch, int nloss) {
    HIPCHK(dalloc(&tb.x, (size_t)batch * NP_MLP_NIN));
    HIPCHK(dalloc(&tb.t, (size_t)batch * 2));
    HIPCHK(dalloc(&tb.lab, (size_t)batch));
    HIPCHK(dalloc(&tb.work, mlp_work_floats(batch)));
    HIPCHK(dalloc(&tb.grad, NP_MLP_NPARAM));
    HIPCHK(dalloc(&tb.loss, (size_t)(nloss > 0 ? nloss : 1)));
    HIPCHK(hipMemcpyAsync(tb.x, x, sizeof(float) * (size_t)batch * NP_MLP_NIN, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(tb.t, target, sizeof(float) * (size_t)batch * 2, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(tb.lab, label, sizeof(int) * (size_t)batch, hipMemcpyHostToDevice, c->stream));
    return NP_OK;
}

static int check_train_args(np_ctx *c, const float *x, const int32_t *label, const float *target, int batch) {
    if (!c || !x || !label || !target || batch < 1) { set_error("bad training arguments"); return NP_ERR_ARG; }
    for (int i = 0; i < batch; ++i)
        if (label[i] < 0 || label[i] > 2) { set_error("label %d out of range", i); return NP_ERR_ARG; }
    return NP_OK;
}

static int ensure_model(np_ctx *c) {
    if (c->has_model) return NP_OK;
    std::vector<float> p(NP_MLP_NPARAM);
    init_model_params(p.data());
    return np_model_set(c, p.data());
}

int np_model_loss_grad(np_ctx *c, const float *x, const int32_t *label, const float *target, int32_t batch,
                       float lam, float *loss, float *grad) {
    int rc = check_train_args(c, x, label, target, batch);
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    if ((rc = ensure_model(c))) return rc;
    TrainBufs tb;
    if ((rc = upload_train(c, tb, x, label, target, batch, 1))) return rc;
    HIPCHK(launch_mlp_loss_grad(c->model, tb.x, tb.lab, tb.t, batch, lam, tb.work, tb.grad, tb.loss, c->stream));
    if (grad) HIPCHK(hipMemcpyAsync(grad, tb.grad, sizeof(float) * NP_MLP_NPARAM, hipMemcpyDeviceToHost, c->stream));
    if (loss) HIPCHK(hipMemcpyAsync(loss, tb.loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    return ctx_sync(c);
}

int np_model_train(np_ctx *c, const float *x, const int32_t *label, const float *target, int32_t batch,
                   int32_t epochs, float lr, float lam, float *loss_out) {
    int rc = check_train_args(c, x, label, target, batch);
    if (rc) return rc;
    if (epochs < 1) { set_error("epochs must be >= 1"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    if ((rc = ensure_model(c))) return rc;
    if (!c->adam_m) {
        HIPCHK(dalloc(&c->adam_m, NP_MLP_NPARAM));
        HIPCHK(dalloc(&c->adam_v, NP_MLP_NPARAM));
    }
    if (c->adam_step == 0) {
        HIPCHK(launch_fill_i32((int *)c->adam_m, NP_MLP_NPARAM, 0, c->stream));
        HIPCHK(launch_fill_i32((int *)c->adam_v, NP_MLP_NPARAM, 0, c->stream));
    }
    TrainBufs tb;
    if ((rc = upload_train(c, tb, x, label, target, batch, epochs))) return rc;
    for (int ep = 0; ep < epochs; ++ep) {
        HIPCHK(launch_mlp_loss_grad(c->model, tb.x, tb.lab, tb.t, batch, lam, tb.work, tb.grad, tb.loss + ep, c->stream));
        c->adam_step++;
        HIPCHK(launch_adam(c->model, tb.grad, c->adam_m, c->adam_v, NP_MLP_NPARAM, lr, 0.9f, 0.999f, 1e-8f,
                           c->adam_step, c->stream));
    }
    if (loss_out) HIPCHK(hipMemcpyAsync(loss_out, tb.loss, sizeof(float) * (size_t)epochs, hipMemcpyDeviceToHost, c->stream));
    return ctx_sync(c);
}

int np_model_forward(np_ctx *c, const float *x, int32_t batch, float *out) {
    if (!c || !x || !out || batch < 1) { set_error("bad arguments"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    int rc = ensure_model(c);
    if (rc) return rc;
    float *dx = nullptr, *dout = nullptr;
    HIPCHK(dalloc(&dx, (size_t)batch * NP_MLP_NIN));
    hipError_t e = dalloc(&dout, (size_t)batch * 5);
    if (e == hipSuccess) e = hipMemcpyAsync(dx, x, sizeof(float) * (size_t)batch * NP_MLP_NIN, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = launch_mlp_forward(c->model, dx, batch, dout, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(float) * (size_t)batch * 5, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(dx); dfree(dout);
    if (e != hipSuccess) return hip_fail(e, "model forward", __LINE__);
    return NP_OK;
}

// ---- stepping ----------------------------------------------------------------------
static int step_state(np_ctx *c, DevState &st, SchedDev &sd) {
    np_graph *g = c->step_g;
    if (!g) { set_error("no stepped solve in progress"); return NP_ERR_ARG; }
    st = make_state(c, g->sp, g->srow, g->scol, g->sw, g->sp.dist, INF);
    np_sched d;
    np_sched_default(1, &d);
    sd = make_schedd(c, d, g->n);
    return NP_OK;
}

static int step_compact(np_ctx *c) {
    DevState st; SchedDev sd;
    int rc = step_state(c, st, sd);
    if (rc) return rc;
    HIPCHK(launch_compact(st, c->step_it, false, c->num_cus * 8, c->stream));
    return ctx_sync(c);
}

int np_step_begin(np_ctx *c, np_graph *g, int32_t src, np_stats *stats) {
    if (!c || !g || g->ctx != c) { set_error("bad arguments"); return NP_ERR_ARG; }
    if (src < 0 || src >= g->n) { set_error("source out of range"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    int rc = graph_space(c, g);
    if (rc) return rc;
    if (c->snap_n < g->n) {
        dfree(c->snap_dist); dfree(c->snap_bm);
        HIPCHK(dalloc(&c->snap_dist, (size_t)g->n));
        HIPCHK(dalloc(&c->snap_bm, (size_t)(g->n + 31) / 32));
        c->snap_n = g->n;
    }
    const int sid = g->solver_id(src, c->stream);
    if (sid < 0) { set_error("source mapping failed"); return NP_ERR_HIP; }
    c->step_g = g;
    c->step_it = 0;
    DevState st; SchedDev sd;
    if ((rc = step_state(c, st, sd))) return rc;
    HIPCHK(hipMemcpyAsync(g->sp.src_slots, &sid, sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_init_solve(st, g->sp.src_slots, 1, 0, 0, c->stream));
    HIPCHK(launch_set_cfg(st, 1, 6, 6, c->stream));
    if (stats) std::memset(stats, 0, sizeof(*stats));
    return step_compact(c);
}

int np_step_frontier(np_ctx *c, int32_t *count, int32_t *done) {
    if (!c || !count || !done) { set_error("null argument"); return NP_ERR_ARG; }
    np_graph *g = c->step_g;
    if (!g) { set_error("no stepped solve in progress"); return NP_ERR_ARG; }
    HIPCHK(hipMemcpyAsync(count, g->sp.cnt + (c->step_it % 3) * NCLS, sizeof(int) * NCLS, hipMemcpyDeviceToHost, c->stream));
    int rc = ctx_sync(c);
    if (rc) return rc;
    long long tot = 0;
    for (int k = 1; k < NCLS; ++k) tot += count[k];
    *done = tot == 0;
    return NP_OK;
}

static int step_expand_timed(np_ctx *c, const np_sched *s, double *ms) {
    DevState st; SchedDev sd;
    int rc = step_state(c, st, sd);
    if (rc) return rc;
    np_sched sch;
    if ((rc = validate_sched(s, &sch))) return rc;
    int lv = sch.level >= 0 ? sch.level : 1;
    HIPCHK(launch_set_cfg(st, lv, sch.log2_t_lo, sch.log2_t_hi, c->stream));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    HIPCHK(launch_expand(st, sd, c->step_it, false, false, c->num_cus * 8, c->stream));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, c->ev0, c->ev1));
    if (ms) *ms = f;
    return NP_OK;
}

int np_step_try(np_ctx *c, const np_sched *s, int32_t reps, double *ms) {
    if (!c || !c->step_g) { set_error("no stepped solve in progress"); return NP_ERR_ARG; }
    np_graph *g = c->step_g;
    const size_t nw = (size_t)(g->n + 31) / 32;
    HIPCHK(hipMemcpyAsync(c->snap_dist, g->sp.dist, sizeof(int) * (size_t)g->n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->snap_bm, g->sp.bm[0], sizeof(unsigned) * nw, hipMemcpyDeviceToDevice, c->stream));
    double best = 1e30;
    for (int r = 0; r < (reps > 0 ? reps : 1); ++r) {
        double t = 0;
        int rc = step_expand_timed(c, s, &t);
        if (rc) return rc;
        best = std::min(best, t);
        HIPCHK(hipMemcpyAsync(g->sp.dist, c->snap_dist, sizeof(int) * (size_t)g->n, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(g->sp.bm[0], c->snap_bm, sizeof(unsigned) * nw, hipMemcpyDeviceToDevice, c->stream));
    }
    if (ms) *ms = best;
    return ctx_sync(c);
}

int np_step_advance(np_ctx *c, const np_sched *s, double *ms) {
    if (!c || !c->step_g) { set_error("no stepped solve in progress"); return NP_ERR_ARG; }
    int rc = step_expand_timed(c, s, ms);
    if (rc) return rc;
    c->step_it++;
    return step_compact(c);
}

int np_step_end(np_ctx *c, int32_t *dist_out) {
    if (!c || !c->step_g) { set_error("no stepped solve in progress"); return NP_ERR_ARG; }
    np_graph *g = c->step_g;
    if (dist_out) {
        const int *d = g->sp.dist;
        if (g->rl.active) {
            if (!g->out) HIPCHK(dalloc(&g->out, (size_t)g->n));
            HIPCHK(launch_unpermute(g->sp.dist, g->rl.old2new, g->n, g->out, c->stream));
            d = g->out;
        }
        HIPCHK(hipMemcpyAsync(dist_out, d, sizeof(int) * (size_t)g->n, hipMemcpyDeviceToHost, c->stream));
        int rc = ctx_sync(c);
        if (rc) return rc;
    }
    c->step_g = nullptr;
    return NP_OK;
}

// ---- autotuner ---------------------------------------------------------------------
// Mean of up to 64K weights of g (the first ones of the solver layout): the scale of the
// priority-window candidates.
static int sample_mean_weight(np_ctx *c, const np_graph *g, double *wmean) {
    *wmean = 0.0;
    const long long k = g->m < 65536 ? g->m : 65536;
    if (k <= 0 || !g->sw) return NP_OK;
    std::vector<int> h((size_t)k);
    HIPCHK(hipMemcpyAsync(h.data(), g->sw, sizeof(int) * (size_t)k, hipMemcpyDeviceToHost, c->stream));
    int rc = ctx_sync(c);
    if (rc) return rc;
    double sum = 0;
    for (int x : h) sum += x;
    *wmean = sum / (double)k;
    return NP_OK;
}

// Priority windows tried on top of the best schedule of the grid search.
static std::vector<int> window_candidates(double wmean) {
    std::vector<int> out;
    if (wmean <= 0) return out;
    for (double f : {2.0, 4.0, 8.0, 16.0, 64.0}) {
        const double d = f * wmean;
        if (d >= 1 && d < 1e9) out.push_back((int)d);
    }
    return out;
}

// Window stages of both autotuners.  `eval` times one schedule (ms).  The windows are tried on
// the best schedule of the grid search; when one wins, the level and bin thresholds are searched
// again under it on the same drive (a window changes which frontier sizes occur), and then half
// and twice the window.
static int tune_window(const std::function<int(const np_sched &, double *)> &eval, double wmean,
                       np_sched *bs, double *bestt) {
    const np_sched grid_best = *bs;
    for (int delta : window_candidates(wmean)) {
        np_sched s = grid_best;
        s.delta = delta;
        double t;
        if (int rc = eval(s, &t)) return rc;
        if (t < *bestt) { *bestt = t; *bs = s; }
    }
    if (bs->delta <= 0) return NP_OK;
    const np_sched win = *bs;
    for (int lv = 1; lv <= 2; ++lv)
        for (int klo : {4, 5, 6, 7})
            for (int khi : {8, 10, 12, 14, 30}) {
                np_sched s = win;
                s.level = lv; s.log2_t_lo = klo; s.log2_t_hi = khi;
                if (s.level == win.level && klo == win.log2_t_lo && khi == win.log2_t_hi) continue;
                double t;
                if (int rc = eval(s, &t)) return rc;
                if (t < *bestt) { *bestt = t; *bs = s; }
            }
    const np_sched th = *bs;
    for (int delta : {th.delta / 2, th.delta * 2}) {
        if (delta < 1 || delta >= 1000000000) continue;
        np_sched s = th;
        s.delta = delta;
        double t;
        if (int rc = eval(s, &t)) return rc;
        if (t < *bestt) { *bestt = t; *bs = s; }
    }
    return NP_OK;
}

int np_autotune(np_ctx *c, np_graph *g, const int32_t *srcs, int32_t nsrc, const np_sched *base,
                np_sched *best, double *best_ms) {
    if (!c || !g || g->ctx != c || !srcs || nsrc < 1 || !best) { set_error("bad arguments"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    np_sched b0;
    int rc = validate_sched(base, &b0);
    if (rc) return rc;
    b0.scheme = NP_SCHEME_CONSOLIDATE;  // the search is over the consolidation bins
    int occ = persistent_max_wgs_per_cu(false, false, b0.frontier == NP_FRONTIER_PUSH, false);
    std::vector<int> wgs_opts = {0};
    if (occ > 2) wgs_opts.push_back(2);
    if (occ > 4) wgs_opts.push_back(4);
    const int klos[] = {3, 4, 5, 6, 7};
    const int khis[] = {6, 8, 10, 12, 14, 30};
    double bestt = 1e30;
    np_sched bs = b0;
    std::vector<int> scratch((size_t)g->n);
    int *d_dist = nullptr;
    HIPCHK(dalloc(&d_dist, (size_t)g->n));
    for (int drive : {(int)NP_DRIVE_PERSISTENT, (int)NP_DRIVE_GRAPH})
    for (int lv = 0; lv <= 2; ++lv)
        for (int klo : klos)
            for (int khi : khis) {
                if (lv == 0 && khi != khis[0]) continue;  // level 0 has no upper bin
                if (khi < klo) continue;
                for (int wg : wgs_opts) {
                    if (drive != NP_DRIVE_PERSISTENT && wg != 0) continue;  // one grid size per kernel
                    np_sched s = b0;
                    s.drive = drive;
                    s.level = lv; s.log2_t_lo = klo; s.log2_t_hi = khi; s.wgs_per_cu = wg;
                    s.trace = 0;
                    double tot = 0;
                    for (int i = 0; i < nsrc && rc == NP_OK; ++i) {
                        np_stats st;
                        rc = solve_single(c, g, srcs[i], &s, d_dist, nullptr, false, &st);
                        tot += st.kernel_ms;
                    }
                    if (rc) { dfree(d_dist); return rc; }
                    if (tot < bestt) { bestt = tot; bs = s; }
                }
            }
    // priority window on top of the best scan-frontier schedule
    double wmean = 0;
    if (bs.frontier == NP_FRONTIER_SCAN && (rc = sample_mean_weight(c, g, &wmean)) != NP_OK) { dfree(d_dist); return rc; }
    if (bs.frontier == NP_FRONTIER_SCAN) {
        auto eval = [&](const np_sched &s, double *ms) {
            double tot = 0;
            for (int i = 0; i < nsrc; ++i) {
                np_stats st;
                if (int e = solve_single(c, g, srcs[i], &s, d_dist, nullptr, false, &st)) return e;
                tot += st.kernel_ms;
            }
            *ms = tot;
            return (int)NP_OK;
        };
        if ((rc = tune_window(eval, wmean, &bs, &bestt))) { dfree(d_dist); return rc; }
    }
    dfree(d_dist);
    *best = bs;
    if (best_ms) *best_ms = bestt / nsrc;
    return NP_OK;
}

int np_batch_autotune(np_ctx *c, np_batch *b, const np_sched *base, np_sched *best, double *best_ms) {
    if (!c || !b || b->ctx != c || !best) { set_error("bad arguments"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    np_sched b0;
    int rc = validate_sched(base, &b0);
    if (rc) return rc;
    b0.scheme = NP_SCHEME_CONSOLIDATE;
    int occ = persistent_max_wgs_per_cu(true, b->unit != 0, false, !b->packs.empty());
    std::vector<int> wgs_opts = {0};
    if (occ > 2) wgs_opts.push_back(2);
    if (occ > 4) wgs_opts.push_back(4);
    const int klos[] = {3, 4, 5, 6, 7};
    const int khis[] = {6, 8, 10, 12, 14, 30};
    // One solve per candidate in the searches; a single timing varies by a few percent, so the
    // closest candidates are timed twice more and the least of their three timings decides.
    std::vector<std::pair<double, np_sched>> tried;  // single timings
    auto once = [&](const np_sched &s, double *ms) {
        np_stats st;
        if (int e = np_batch_solve(c, b, &s, &st)) return e;
        *ms = st.kernel_ms;
        tried.emplace_back(st.kernel_ms, s);
        return (int)NP_OK;
    };
    auto confirm = [&](size_t top, np_sched *win, double *wint) {
        std::stable_sort(tried.begin(), tried.end(),
                         [](const std::pair<double, np_sched> &x, const std::pair<double, np_sched> &y) {
                             return x.first < y.first;
                         });
        *wint = 1e30;
        for (size_t i = 0; i < tried.size() && i < top; ++i) {
            double m = tried[i].first;
            for (int r = 0; r < 2; ++r) {
                np_stats st;
                if (int e = np_batch_solve(c, b, &tried[i].second, &st)) return e;
                m = std::min(m, (double)st.kernel_ms);
            }
            if (m < *wint) { *wint = m; *win = tried[i].second; }
        }
        tried.clear();
        tried.emplace_back(*wint, *win);  // confirmed: competes with the next stage's timings
        return (int)NP_OK;
    };
    double bestt = 1e30;
    np_sched bs = b0;
    for (int drive : {(int)NP_DRIVE_PERSISTENT, (int)NP_DRIVE_GRAPH})
    for (int lv = 0; lv <= 2; ++lv)
        for (int klo : klos)
            for (int khi : khis) {
                if (lv == 0 && khi != khis[0]) continue;  // level 0 has no upper bin
                if (khi < klo) continue;
                for (int wg : wgs_opts) {
                    if (drive != NP_DRIVE_PERSISTENT && wg != 0) continue;  // one grid size per kernel
                    np_sched s = b0;
                    s.drive = drive;
                    s.level = lv; s.log2_t_lo = klo; s.log2_t_hi = khi; s.wgs_per_cu = wg;
                    s.trace = 0;
                    double t;
                    if ((rc = once(s, &t))) return rc;
                }
            }
    if ((rc = confirm(4, &bs, &bestt))) return rc;
    // packs without a window: holding back heavy vertices that few sources have reached yet
    // (out-degree >= 64 .. 2048) on top of the best schedule
    if (!b->packs.empty()) {
        const np_sched base_bs = bs;
        for (int hold : {7, 8, 9, 10, 11, 12}) {
            np_sched s = base_bs;
            s.pack_hold = hold;
            double t;
            if ((rc = once(s, &t))) return rc;
        }
        if ((rc = confirm(3, &bs, &bestt))) return rc;
    }
    // priority window on top of the best schedule, scaled by the weights of the graph
    // that carries the most instances
    double wmean = 0;
    const np_graph *g0 = b->packs.empty() ? b->graphs[0] : b->packs[0]->g;
    if ((rc = sample_mean_weight(c, g0, &wmean))) return rc;
    if (!b->unit) {
        if ((rc = tune_window(once, wmean, &bs, &bestt))) return rc;
        if ((rc = confirm(3, &bs, &bestt))) return rc;
    }
    if (bs.delta > 0) bs.pack_hold = 0;  // a window turns the hold off: report what runs
    *best = bs;
    if (best_ms) *best_ms = bestt;
    return NP_OK;
}

// ---- device generators -------------------------------------------------------------
struct EdgeBufs {
    uint32_t *s = nullptr, *d = nullptr;
    int *w = nullptr;
    ~EdgeBufs() { dfree(s); dfree(d); dfree(w); }
};

static int finish_generated(np_ctx *c, int n, long long m, EdgeBufs &eb, np_graph **out) {
    GenOut go{};
    hipError_t e = build_csr_device(eb.s, eb.d, eb.w, n, m, &go, c->stream);
    if (e != hipSuccess) return hip_fail(e, "CSR build", __LINE__);
    np_graph *g = new np_graph();
    g->ctx = c; g->n = n; g->m = m; g->owns = true;
    g->row = go.row; g->col = go.col; g->w = go.w;
    int rc = finish_graph(c, g);
    if (rc) { np_graph_free(g); return rc; }
    *out = g;
    return NP_OK;
}

int np_gen_rmat(np_ctx *c, int32_t scale, int32_t edge_factor, double a, double b, double cc, int32_t wmin,
                int32_t wmax, uint64_t seed, int32_t permute, np_graph **out) {
    if (!c || !out || scale < 1 || scale > 30 || edge_factor < 1 || wmax < wmin) {
        set_error("bad R-MAT arguments"); return NP_ERR_ARG;
    }
    const int n = 1 << scale;
    const long long m = (long long)edge_factor * n;
    if (m > 0x7FFFFFFFll) { set_error("more than 2^31-1 edges"); return NP_ERR_CAPACITY; }
    HIPCHK(hipSetDevice(c->device));
    EdgeBufs eb;
    HIPCHK(dalloc(&eb.s, (size_t)m)); HIPCHK(dalloc(&eb.d, (size_t)m)); HIPCHK(dalloc(&eb.w, (size_t)m));
    HIPCHK(gen_rmat_edges(scale, m, seed, a, b, cc, permute, wmin, wmax, eb.s, eb.d, eb.w, c->stream));
    return finish_generated(c, n, m, eb, out);
}

int np_gen_er(np_ctx *c, int32_t n, int64_t m, int32_t wmin, int32_t wmax, uint64_t seed, np_graph **out) {
    if (!c || !out || n < 1 || m < 0 || m > 0x7FFFFFFFll || wmax < wmin) { set_error("bad ER arguments"); return NP_ERR_ARG; }
    HIPCHK(hipSetDevice(c->device));
    EdgeBufs eb;
    HIPCHK(dalloc(&eb.s, (size_t)m)); HIPCHK(dalloc(&eb.d, (size_t)m)); HIPCHK(dalloc(&eb.w, (size_t)m));
    HIPCHK(gen_er_edges(n, m, seed, wmin, wmax, eb.s, eb.d, eb.w, c->stream));
    return finish_generated(c, n, m, eb, out);
}

int np_gen_grid(np_ctx *c, int32_t rows, int32_t cols, int64_t n_long, int32_t wmin, int32_t wmax,
                uint64_t seed, np_graph **out) {
    if (!c || !out || rows < 1 || cols < 1 || n_long < 0 || wmax < wmin) { set_error("bad grid arguments"); return NP_ERR_ARG; }
    const long long n = (long long)rows * cols;
    const long long grid_m = 2ll * ((long long)rows * (cols - 1) + (long long)cols * (rows - 1));
    const long long m = grid_m + n_long;
    if (n > 0x7FFFFFFFll || m > 0x7FFFFFFFll) { set_error("grid too large"); return NP_ERR_CAPACITY; }
    HIPCHK(hipSetDevice(c->device));
    EdgeBufs eb;
    long long *off = nullptr;
    HIPCHK(dalloc(&eb.s, (size_t)m)); HIPCHK(dalloc(&eb.d, (size_t)m)); HIPCHK(dalloc(&eb.w, (size_t)m));
    HIPCHK(dalloc(&off, (size_t)n + 1));
    hipError_t e = grid_offsets(rows, cols, off, c->stream);
    if (e == hipSuccess) e = gen_grid_edges(rows, cols, n_long, seed, wmin, wmax, off, eb.s, eb.d, eb.w, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(off);
    if (e != hipSuccess) return hip_fail(e, "grid generation", __LINE__);
    return finish_generated(c, (int)n, m, eb, out);
}

}  // extern "C"
