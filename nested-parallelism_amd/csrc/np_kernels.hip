// np_kernels.hip -- gfx950 kernels of the SSSP / recursive-BFS nested-parallelism path.
//
// The reference (apps/sssp/src/gpu/sssp_kernel.cu) expands each frontier vertex in a
// thread and hands vertices of out-degree >= THRESHOLD to child grids launched from
// the device, consolidated per warp (consolidate_warp_dp_kernel, :330-404), per block
// (consolidate_block_dp_kernel, :407-481) or per grid (consolidate_grid_dp_kernel,
// :484-541).  HIP has no device-side launch, so here the frontier is kept as a queue
// segmented by out-degree class and every scope of consolidation is a mapping of those
// segments onto workers of the same launch:
//   lane bin       64 frontier vertices per wavefront, their edges spread evenly over the
//                  64 lanes (prefix sum of degrees + owner search), coalesced CSR reads
//   wavefront bin  one vertex per 64-lane wavefront (the reference's 64-thread child block)
//   workgroup bin  one vertex per 256-thread workgroup (block-level consolidation)
//   grid bin       a vertex split into 2048-edge chunks spread over every wavefront of
//                  the grid (grid-level consolidation)
// The relaxation is the reference's, bit for bit (sssp_kernel.cu:46-52):
//   alt = dist[u] + w;  if (dist[v] > alt) atomicMin(&dist[v], alt)  -- int32.
// A vertex whose distance improved is flagged in an update bitmap (the reference's
// update[] byte array, one bit per vertex) and, for the next iteration, either compacted
// from the bitmap (scan frontier, unorder_generateQueue_kernel :648-658) or appended by
// the relaxing wavefront (push frontier).  Batches of up to 32 sources of one graph run
// the same bins over (vertex, pending sources) work items ("source-packed batch" below).
#include <hip/hip_runtime.h>
#include "np_internal.h"

namespace np {

#define NP_AGENT __HIP_MEMORY_SCOPE_AGENT
static constexpr int WG = 256;
static constexpr int WAVES = WG / 64;
static constexpr unsigned long long SPIN_LIMIT = 1ull << 24;

// ------------------------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Data that one workgroup writes and another reads inside a persistent launch (queue
// entries, bitmap words, counters, distances) only moves through device-coherent
// accesses: relaxed agent-scope atomics, loads and stores (sc1 on gfx950: they bypass
// the per-CU cache and the non-coherent L2 state of the other XCDs).  The grid barrier
// then needs no L2 writeback or invalidate, which would also evict the read-only CSR.
__device__ __forceinline__ int ld_coherent(const int *p) {
    return __hip_atomic_load(const_cast<int *>(p), __ATOMIC_RELAXED, NP_AGENT);
}
__device__ __forceinline__ unsigned ld_coherent_u(const unsigned *p) {
    return __hip_atomic_load(const_cast<unsigned *>(p), __ATOMIC_RELAXED, NP_AGENT);
}
__device__ __forceinline__ void st_coherent(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, NP_AGENT);
}
__device__ __forceinline__ void st_coherent_u(unsigned *p, unsigned v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, NP_AGENT);
}

#ifndef NP_BOUNDS_CHECK
#define NP_BOUNDS_CHECK 0
#endif
// Bounds-checked build (NP_BOUNDS_CHECK=1, libnpsssp_bounds.so): every write into a frontier
// queue, its source masks or a scheme's delayed buffer checks its index against the segment it
// belongs to, [lo, hi); an index outside is not written and the solve fails with NP_ERR_BOUNDS.
// The normal build compiles the check out.
__device__ __forceinline__ bool seg_ok(Ctrl *ctrl, long long idx, long long lo, long long hi) {
#if NP_BOUNDS_CHECK
    if (idx >= lo && idx < hi) return true;
    atomicCAS(&ctrl->error, 0, (int)NP_ERR_BOUNDS);
    return false;
#else
    (void)ctrl; (void)idx; (void)lo; (void)hi;
    return true;
#endif
}

__device__ __forceinline__ int wave_incl_scan(int x) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(x, o);
        if (lane >= o) x += t;
    }
    return x;
}

template <bool BATCH>
__device__ __forceinline__ void decode(const DevState &st, int slot, int &B, int &VB) {
    if (BATCH) {
        int i = st.lut[slot >> LUT_SHIFT];
        int2 d = st.inst[i];
        B = d.x; VB = d.y;
    } else {
        B = 0; VB = 0;
    }
}

// Grid-wide barrier of a cooperative launch.  Every wave first waits for its own
// outstanding memory operations (s_waitcnt 0: its coherent stores and atomics have
// completed), the workgroup meets at s_barrier, then one lane arrives.  Arrivals form a
// two-level tree of cumulative 64-bit counters (no reset): workgroup b counts on leaf
// b % 32 (each on its own cache line), the last arrival of a leaf counts on the root,
// and every workgroup spins on the root with relaxed coherent loads.  The spin is
// bounded: on expiry the block raises NP_ERR_TIMEOUT and every block leaves its loop at
// the next barrier, so the grid always drains.
__device__ bool grid_sync(Ctrl *ctrl, unsigned nblocks, unsigned long long &gen) {
    __shared__ int s_ok;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        gen += 1;
        const unsigned nleaves = nblocks < (unsigned)BAR_LEAVES ? nblocks : (unsigned)BAR_LEAVES;
        const unsigned l = blockIdx.x % nleaves;
        const unsigned long long lsize = nblocks / nleaves + (l < nblocks % nleaves ? 1u : 0u);
        unsigned long long la =
            __hip_atomic_fetch_add(&ctrl->leaf[l].n, 1ull, __ATOMIC_RELAXED, NP_AGENT) + 1ull;
        const unsigned long long target = gen * (unsigned long long)nleaves;
        unsigned long long arrived = 0;
        if (la == gen * lsize)
            arrived = __hip_atomic_fetch_add(&ctrl->bar, 1ull, __ATOMIC_RELAXED, NP_AGENT) + 1ull;
        int ok = 1;
        if (arrived < target) {
            unsigned long long spins = 0;
            while (__hip_atomic_load(&ctrl->bar, __ATOMIC_RELAXED, NP_AGENT) < target) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > SPIN_LIMIT) { ok = 0; break; }
                if ((spins & 1023) == 0 && __hip_atomic_load(&ctrl->error, __ATOMIC_RELAXED, NP_AGENT)) {
                    ok = 0; break;
                }
            }
        }
        if (!ok) {
            int expected = 0;
            __hip_atomic_compare_exchange_strong(&ctrl->error, &expected, (int)NP_ERR_TIMEOUT,
                                                 __ATOMIC_RELAXED, __ATOMIC_RELAXED, NP_AGENT);
        }
        if (__hip_atomic_load(&ctrl->error, __ATOMIC_RELAXED, NP_AGENT)) ok = 0;
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// ------------------------------------------------------------------------------------
// relaxation
// ------------------------------------------------------------------------------------
struct PushCtx {
    unsigned *bm_next;  // bitmap receiving "improved" marks
    unsigned *bm_cur;   // push: bitmap of the frontier being consumed (cleared here)
    int *cnt_next;      // push: class counters of the next frontier
    int *q_next;        // push: next frontier queue
};

// Tail of one relaxation after the cheap check (alt < dist[tslot]) passed: the atomicMin of the
// reference, issued without a return value, and the update mark.  Marking on the check alone
// marks exactly the slots the atomic's old value would: when a concurrent atomicMin has already
// stored less than alt, that writer passed its own check in this same launch and marks the slot
// itself (every earlier store is visible to the check).  With a scanned frontier the mark is one
// more atomic without a return, so no lane waits on either; the compaction drops vertices without
// out-edges.  Pushing relaxers need the class of the slot and the bit's old value to append it
// once.  Returns true when the caller must append tslot (push only).
template <bool PUSH>
__device__ __forceinline__ bool relax_tail(const DevState &st, const PushCtx &pc, int tslot,
                                           int alt, int &c_out) {
    atomicMin(&st.dist[tslot], alt);
    const unsigned bit = 1u << (tslot & 31);
    if (!PUSH) {
        atomicOr(&pc.bm_next[tslot >> 5], bit);
        return false;
    }
    const int c = st.cls[tslot];
    if (c == 0) return false;  // no out-edges: nothing to propagate
    const unsigned prev = atomicOr(&pc.bm_next[tslot >> 5], bit);
    c_out = c;
    return (prev & bit) == 0;
}

// Wavefront-aggregated append into the class-segmented next frontier: one counter
// atomic per (wavefront, class) group.
__device__ __forceinline__ void push_wave(const DevState &st, const PushCtx &pc, bool want,
                                          int slot, int c) {
    unsigned long long mask = __ballot(want);
    while (mask) {
        int leader = __ffsll((long long)mask) - 1;
        int cl = __shfl(c, leader);
        unsigned long long same = __ballot(want && c == cl);
        int base = 0;
        if (lane_id() == leader) base = atomicAdd(&pc.cnt_next[cl], __popcll(same));
        base = __shfl(base, leader);
        if (want && c == cl) {
            int pos = base + __popcll(same & __lanemask_lt());
            if (seg_ok(st.ctrl, (long long)st.seg_off[cl] + pos, st.seg_off[cl], st.seg_off[cl + 1]))
                st_coherent(pc.q_next + st.seg_off[cl] + pos, slot);
            want = false;
        }
        mask &= ~same;
    }
}

constexpr int UNR = 4;

// GUARD: lanes past the end of their range issue no loads (per-iteration kernels, eight
// waves per SIMD: enough loads in flight already); otherwise every load is issued and only
// the atomics are guarded, so the unrolled items' loads are all in flight together
// (persistent kernels, four to five waves per SIMD).  Dead lanes then read a valid address.

// Edges [e0, e1) of one vertex (slot base B, distance du), worker t of `stride`.
// The loop trip count is uniform across the wavefront.
template <bool UNIT, bool PUSH, bool GUARD>
__device__ __forceinline__ void relax_range(const DevState &st, const PushCtx &pc, int B, int e0,
                                            int e1, int du, int t, int stride) {
    for (int base = e0; base < e1; base += UNR * stride) {
        int tv[UNR], alt[UNR], cur[UNR];
        bool ok[UNR];
#pragma unroll
        for (int k = 0; k < UNR; ++k) {
            int e = base + t + k * stride;
            ok[k] = e < e1;
            int ec = ok[k] ? e : e0;
            tv[k] = B + st.col[ec];
            alt[k] = du + (UNIT ? 1 : st.w[ec]);
        }
#pragma unroll
        for (int k = 0; k < UNR; ++k) cur[k] = (GUARD && !ok[k]) ? 0 : ld_coherent(st.dist + tv[k]);
#pragma unroll
        for (int k = 0; k < UNR; ++k) {
            bool want = false;
            int c = 0;
            if (ok[k] && alt[k] < cur[k]) want = relax_tail<PUSH>(st, pc, tv[k], alt[k], c);
            if (PUSH) push_wave(st, pc, want, tv[k], c);
        }
    }
}

// Lane bin: 64 frontier vertices (queue[qidx .. qidx+nvalid)), edges balanced over lanes.
template <bool BATCH, bool UNIT, bool PUSH, bool GUARD>
__device__ __forceinline__ void lane_tile(const DevState &st, const PushCtx &pc, const int *q,
                                          long long qidx, int nvalid) {
    const int lane = lane_id();
    const bool valid = lane < nvalid;
    int slot = valid ? ld_coherent(q + qidx + lane) : 0;
    int B = 0, VB = 0, s = 0, d = 0, du = 0;
    if (valid) {
        decode<BATCH>(st, slot, B, VB);
        int v = slot - B;
        s = st.row[VB + v];
        d = st.row[VB + v + 1] - s;
        du = ld_coherent(st.dist + slot);
        if (PUSH) st_coherent_u(pc.bm_cur + (slot >> 5), 0u);
    }
    const int incl = wave_incl_scan(d);
    const int total = __shfl(incl, 63);
    const int excl = incl - d;
    if (total == 0) return;
    // First edge of the tile (owner of index 0 = first lane with d > 0) and that lane's slot
    // base: the address the lanes whose index runs past the end read.  Edge and base come
    // from the same lane, so the target slot lies inside that lane's instance.
    const int first_lane = __ffsll((long long)__ballot(d > 0)) - 1;
    const int e_safe = __shfl(s, first_lane);
    const int b_safe = BATCH ? __shfl(B, first_lane) : 0;
    for (int base = 0; base < total; base += UNR * 64) {
        int tv[UNR], alt[UNR], cur[UNR];
        bool ok[UNR];
#pragma unroll
        for (int k = 0; k < UNR; ++k) {
            const int kk = base + lane + k * 64;
            ok[k] = kk < total;
            // owner: the last lane whose exclusive offset is <= kk
            int o = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1) {
                int t = __shfl(excl, o + step);
                if (t <= kk) o += step;
            }
            const int e = __shfl(s, o) + (kk - __shfl(excl, o));
            const int ou = __shfl(du, o);
            const int ob = BATCH ? __shfl(B, o) : 0;
            const int ec = ok[k] ? e : e_safe;
            tv[k] = (ok[k] ? ob : b_safe) + st.col[ec];
            alt[k] = ou + (UNIT ? 1 : st.w[ec]);
        }
#pragma unroll
        for (int k = 0; k < UNR; ++k) cur[k] = (GUARD && !ok[k]) ? 0 : ld_coherent(st.dist + tv[k]);
#pragma unroll
        for (int k = 0; k < UNR; ++k) {
            bool want = false;
            int c = 0;
            if (ok[k] && alt[k] < cur[k]) want = relax_tail<PUSH>(st, pc, tv[k], alt[k], c);
            if (PUSH) push_wave(st, pc, want, tv[k], c);
        }
    }
}

// ------------------------------------------------------------------------------------
// one expansion (all threads of every workgroup call it)
// ------------------------------------------------------------------------------------
// Per-workgroup item tables: exclusive prefix, by class, of the work items of each scope.
struct ItemTables {
    long long pre[4][NCLS + 1];  // 0 grid chunks, 1 workgroup items, 2 wavefront items, 3 lane tiles
    int cnt[NCLS];
    int seg[NCLS];
};

__device__ __forceinline__ int find_class(const long long *pre, long long idx) {
    int c = 0;
#pragma unroll
    for (int step = 16; step > 0; step >>= 1)
        if (c + step <= NCLS - 1 && pre[c + step] <= idx) c += step;
    return c;
}

// Thread 0 builds the tables from the frontier's class counts (already in tb.cnt).
__device__ __noinline__ void build_tables(ItemTables &tb, const Cfg &cfg) {
    long long acc[4] = {0, 0, 0, 0};
    for (int c = 0; c < NCLS; ++c) {
        for (int t = 0; t < 4; ++t) tb.pre[t][c] = acc[t];
        long long n = c == 0 ? 0 : tb.cnt[c];
        if (c >= cfg.b3) acc[0] += n * grid_chunks(c);
        else if (c >= cfg.b2) acc[1] += n;
        else if (c >= cfg.b1) acc[2] += n;
        else acc[3] += (n + 63) / 64;
    }
    for (int t = 0; t < 4; ++t) tb.pre[t][NCLS] = acc[t];
}

template <bool BATCH, bool UNIT, bool PUSH, bool GUARD>
__device__ void expansion(const DevState &st, const PushCtx &pc, const int *q, ItemTables &tb) {
    const long long G = tb.pre[0][NCLS], K = tb.pre[1][NCLS], W = tb.pre[2][NCLS],
                    T = tb.pre[3][NCLS];
    // workgroup items: the whole workgroup strides one vertex's edges
    for (long long k = blockIdx.x; k < K; k += gridDim.x) {
        int c = find_class(tb.pre[1], k);
        long long qi = tb.seg[c] + (k - tb.pre[1][c]);
        int slot = ld_coherent(q + qi);
        int B, VB;
        decode<BATCH>(st, slot, B, VB);
        int v = slot - B;
        int e0 = st.row[VB + v], e1 = st.row[VB + v + 1];
        int du = ld_coherent(st.dist + slot);
        if (PUSH && threadIdx.x == 0) st_coherent_u(pc.bm_cur + (slot >> 5), 0u);
        relax_range<UNIT, PUSH, GUARD>(st, pc, B, e0, e1, du, threadIdx.x, WG);
    }
    // wavefront-granular items: grid chunks, wavefront items, lane tiles, round-robin
    const long long NW = (long long)gridDim.x * WAVES;
    const long long gw = (long long)blockIdx.x * WAVES + (threadIdx.x >> 6);
    const long long U = G + W + T;
    for (long long u = gw; u < U; u += NW) {
        if (u < G) {
            int c = find_class(tb.pre[0], u);
            long long r = u - tb.pre[0][c];
            int sh = c > GRID_CHUNK_SHIFT ? c - GRID_CHUNK_SHIFT : 0;
            long long qi = tb.seg[c] + (r >> sh);
            int j = (int)(r & ((1ll << sh) - 1));
            int slot = ld_coherent(q + qi);
            int B, VB;
            decode<BATCH>(st, slot, B, VB);
            int v = slot - B;
            int s = st.row[VB + v], e = st.row[VB + v + 1];
            int e0 = s + (j << GRID_CHUNK_SHIFT);
            int e1 = e0 + (1 << GRID_CHUNK_SHIFT);
            if (e1 > e) e1 = e;
            if (PUSH && j == 0 && lane_id() == 0) st_coherent_u(pc.bm_cur + (slot >> 5), 0u);
            if (e0 < e1) {
                int du = ld_coherent(st.dist + slot);
                relax_range<UNIT, PUSH, GUARD>(st, pc, B, e0, e1, du, lane_id(), 64);
            }
        } else if (u < G + W) {
            long long i = u - G;
            int c = find_class(tb.pre[2], i);
            long long qi = tb.seg[c] + (i - tb.pre[2][c]);
            int slot = ld_coherent(q + qi);
            int B, VB;
            decode<BATCH>(st, slot, B, VB);
            int v = slot - B;
            int e0 = st.row[VB + v], e1 = st.row[VB + v + 1];
            int du = ld_coherent(st.dist + slot);
            if (PUSH && lane_id() == 0) st_coherent_u(pc.bm_cur + (slot >> 5), 0u);
            relax_range<UNIT, PUSH, GUARD>(st, pc, B, e0, e1, du, lane_id(), 64);
        } else {
            long long t = u - G - W;
            int c = find_class(tb.pre[3], t);
            long long ti = t - tb.pre[3][c];
            long long first = ti * 64;
            int nvalid = (int)((long long)tb.cnt[c] - first < 64 ? (long long)tb.cnt[c] - first : 64);
            lane_tile<BATCH, UNIT, PUSH, GUARD>(st, pc, q, tb.seg[c] + first, nvalid);
        }
    }
}

// ------------------------------------------------------------------------------------
// compaction (scan frontier): update bitmap -> class-segmented queue
// ------------------------------------------------------------------------------------
// With a priority window (T < INT_MAX) only flagged vertices with dist < T are queued and
// unflagged; the others keep their flag, and the pass adds how many were kept and the
// least distance among them to defer_cnt / defer_min.  T = INT_MAX queues every flag.
constexpr int NO_WINDOW = 0x7FFFFFFF;

// Single-source frontiers, and source-packed ones without a window, are sparse: a thread scans
// one bitmap word and visits its set bits, so empty words cost one load.  With a priority window
// a source-packed frontier is dense (deferred vertices stay flagged and are re-read every pass):
// there each thread owns CPT vertices of a tile of WG * CPT consecutive ones, so the lanes of a
// wave read the class bytes, masks and distance rows of consecutive vertices and the 32 lanes of
// one bitmap word rewrite it with one store.  Both count queue slots per class in LDS and
// reserve them once per class per tile.
constexpr int CPT = 4;

// One bitmap word per 32 lanes: the first lane of each half-wave stores the kept bits when the
// word changed.  v is this lane's vertex; its half-wave covers one word.
__device__ __forceinline__ void rewrite_word(unsigned *bm, long long v, long long nv, unsigned word, bool keep) {
    const unsigned nw = (unsigned)(__ballot(keep) >> (lane_id() & 32));
    if ((lane_id() & 31) == 0 && v < nv && nw != word) st_coherent_u(bm + (v >> 5), nw);
}

// LDS class counters of this tile -> global reservation (s_lb) and reset for the next tile.
__device__ __forceinline__ void reserve_classes(int *s_lc, int *s_lb, int *cnt_out) {
    __syncthreads();
    if (threadIdx.x < NCLS) {
        const int k = s_lc[threadIdx.x];
        s_lb[threadIdx.x] = k ? atomicAdd(&cnt_out[threadIdx.x], k) : 0;
        s_lc[threadIdx.x] = 0;
    }
    __syncthreads();
}

// Deferred-pair statistics of the workgroup -> defer_cnt / defer_min.
__device__ __forceinline__ void flush_deferred(int my_dc, int my_dm, int *s_dc, int *s_dm, int *defer_cnt,
                                               int *defer_min) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        my_dc += __shfl_xor(my_dc, o);
        const int x = __shfl_xor(my_dm, o);
        my_dm = x < my_dm ? x : my_dm;
    }
    if (lane_id() == 0 && my_dc) { atomicAdd(s_dc, my_dc); atomicMin(s_dm, my_dm); }
    __syncthreads();
    if (threadIdx.x == 0 && *s_dc) { atomicAdd(defer_cnt, *s_dc); atomicMin(defer_min, *s_dm); }
}

__device__ __forceinline__ void compaction_impl(Ctrl *ctrl, unsigned *bm, const unsigned char *cls, const int *seg_off,
                                                long long nwords, int *cnt_out, int *q_out, const int *dist,
                                                int T, int *defer_cnt, int *defer_min) {
    __shared__ int s_lc[NCLS], s_lb[NCLS];
    __shared__ int s_dc, s_dm;
    if (threadIdx.x == 0) { s_dc = 0; s_dm = NO_WINDOW; }
    int my_dc = 0, my_dm = NO_WINDOW;
    for (long long wb = (long long)blockIdx.x * WG; wb < nwords; wb += (long long)gridDim.x * WG) {
        if (threadIdx.x < NCLS) s_lc[threadIdx.x] = 0;
        __syncthreads();
        const long long wi = wb + threadIdx.x;
        const unsigned word0 = wi < nwords ? ld_coherent_u(bm + wi) : 0u;
        unsigned word = word0, take = 0, drop = 0;
        while (word) {
            int b = __ffs(word) - 1;
            word &= word - 1;
            const long long slot = wi * 32 + b;
            int c = cls[slot];
            if (!c) { drop |= 1u << b; continue; }  // no out-edges: nothing to propagate
            if (T != NO_WINDOW) {
                int d = ld_coherent(dist + slot);
                if (d >= T) { my_dc++; my_dm = d < my_dm ? d : my_dm; continue; }
            }
            take |= 1u << b;
            atomicAdd(&s_lc[c], 1);
        }
        __syncthreads();
        if (threadIdx.x < NCLS) {
            int k = s_lc[threadIdx.x];
            s_lb[threadIdx.x] = k ? atomicAdd(&cnt_out[threadIdx.x], k) : 0;
            s_lc[threadIdx.x] = 0;
        }
        __syncthreads();
        word = take;
        while (word) {
            int b = __ffs(word) - 1;
            word &= word - 1;
            long long slot = wi * 32 + b;
            int c = cls[slot];
            int p = atomicAdd(&s_lc[c], 1);
            if (seg_ok(ctrl, (long long)seg_off[c] + s_lb[c] + p, seg_off[c], seg_off[c + 1]))
                st_coherent(q_out + seg_off[c] + s_lb[c] + p, (int)slot);
        }
        if (take | drop) st_coherent_u(bm + wi, word0 & ~(take | drop));
        __syncthreads();
    }
    if (T != NO_WINDOW) flush_deferred(my_dc, my_dm, &s_dc, &s_dm, defer_cnt, defer_min);
}

// ------------------------------------------------------------------------------------
// schedule decision (all threads of a workgroup; the MLP is evaluated redundantly per
// workgroup so every workgroup reaches the same decision without another barrier)
// ------------------------------------------------------------------------------------
// The MLP's decision, or the fixed schedule (level, klo, khi) when its output is unusable.
// Called with scalars, not the SchedDev: a struct passed by reference to a function that is not
// inlined is first stored to scratch memory by every caller.
__device__ __noinline__ Cfg decide_mlp(const float *mlp, int n_feat, int flevel, int fklo, int fkhi,
                                       const int *cnt, int it) {
    static_assert(WG == 256 && NP_MLP_NIN <= 16, "decide() maps 32 outputs x 8 parts onto the workgroup");
    __shared__ float s_x[NP_MLP_NIN], s_h1[32], s_h2[32], s_o[5];
    const Cfg fixed = make_cfg(flevel, fklo, fkhi);
    const float *W1 = mlp, *b1 = W1 + 32 * NP_MLP_NIN, *W2 = b1 + 32, *b2 = W2 + 32 * 32;
    const float *W3 = b2 + 32, *b3 = W3 + 5 * 32;
    // Thread t computes part p = t % 8 of output j = t / 8 of each layer; the 8 parts of an output
    // are adjacent lanes, summed with three shuffles.  Every weight is loaded before the first
    // layer (W2 and W3 rows as float4, thread t at float 4t: one coalesced 4 KB read), so the
    // forward pass waits for global memory once, not once per weight.
    const int t = threadIdx.x, j = t >> 3, p = t & 7;
    const float w1a = W1[j * NP_MLP_NIN + p];
    const float w1b = p + 8 < NP_MLP_NIN ? W1[j * NP_MLP_NIN + p + 8] : 0.f;
    const float4 w2 = *(const float4 *)(W2 + 4 * t);
    const float4 w3 = t < 40 ? *(const float4 *)(W3 + 4 * t) : make_float4(0.f, 0.f, 0.f, 0.f);
    const float b1j = b1[j], b2j = b2[j], b3j = t < 40 ? b3[j] : 0.f;
    if (t == 0) features(cnt, it, n_feat, s_x);
    __syncthreads();
    float a = w1a * s_x[p] + (p + 8 < NP_MLP_NIN ? w1b * s_x[p + 8] : 0.f);
    a += __shfl_xor(a, 1);
    a += __shfl_xor(a, 2);
    a += __shfl_xor(a, 4);
    if (p == 0) s_h1[j] = a + b1j > 0.f ? a + b1j : 0.f;
    __syncthreads();
    a = w2.x * s_h1[4 * p] + w2.y * s_h1[4 * p + 1] + w2.z * s_h1[4 * p + 2] + w2.w * s_h1[4 * p + 3];
    a += __shfl_xor(a, 1);
    a += __shfl_xor(a, 2);
    a += __shfl_xor(a, 4);
    if (p == 0) s_h2[j] = a + b2j > 0.f ? a + b2j : 0.f;
    __syncthreads();
    a = w3.x * s_h2[4 * p] + w3.y * s_h2[4 * p + 1] + w3.z * s_h2[4 * p + 2] + w3.w * s_h2[4 * p + 3];
    a += __shfl_xor(a, 1);
    a += __shfl_xor(a, 2);
    a += __shfl_xor(a, 4);
    if (t < 40 && p == 0) s_o[j] = a + b3j;
    __syncthreads();
    float o0 = s_o[0], o1 = s_o[1], o2 = s_o[2], o3 = s_o[3], o4 = s_o[4];
    __syncthreads();
    if (!(isfinite(o0) && isfinite(o1) && isfinite(o2) && isfinite(o3) && isfinite(o4)) ||
        fabsf(o3) > 1e6f || fabsf(o4) > 1e6f)
        return fixed;  // fallback: the fixed-threshold heuristic
    int level = (o1 > o0) ? ((o2 > o1) ? 2 : 1) : ((o2 > o0) ? 2 : 0);
    return make_cfg(level, (int)rintf(o3), (int)rintf(o4));
}

__device__ __forceinline__ Cfg decide(const SchedDev &sp, const int *cnt, int it) {
    if (!sp.learned || !sp.mlp) return make_cfg(sp.level, sp.klo, sp.khi);
    return decide_mlp(sp.mlp, sp.n_feat, sp.level, sp.klo, sp.khi, cnt, it);
}

__device__ __noinline__ void record_trace_rec(np_trace_rec *r, int it, int level, int klo, int khi, const int *cnt,
                                              long long t0, Ctrl *ctrl) {
    r->pairs = __hip_atomic_load(&ctrl->pairs_total, __ATOMIC_RELAXED, NP_AGENT);
    r->edges = __hip_atomic_load(&ctrl->edges_total, __ATOMIC_RELAXED, NP_AGENT);
    r->iteration = it;
    r->level = level;
    r->log2_t_lo = klo;
    r->log2_t_hi = khi;
    for (int c = 0; c < NCLS; ++c) r->count[c] = cnt[c];
    r->t_begin = t0;
    r->t_end = t0;
}

__device__ __forceinline__ void record_trace(const SchedDev &sp, int it, const Cfg &cfg, const int *cnt,
                                             long long t0, Ctrl *ctrl) {
    if (!sp.trace || it >= sp.trace_cap) return;
    record_trace_rec(sp.trace + it, it, cfg.level, cfg.klo, cfg.khi, cnt, t0, ctrl);
}

// ------------------------------------------------------------------------------------
// persistent drive
// ------------------------------------------------------------------------------------
template <bool BATCH, bool UNIT, bool PUSH>
__global__ __launch_bounds__(WG) void k_persistent(DevState st, SchedDev sp) {
    __shared__ ItemTables tb;
    __shared__ int s_total;
    Ctrl *ctrl = st.ctrl;
    unsigned long long gen = 0;
    const unsigned nb = gridDim.x;
    long long t_prev = 0;
    // priority window: every workgroup computes the same sequence of thresholds
    int T = (!PUSH && sp.delta > 0) ? sp.delta : NO_WINDOW;
    int pass = 0;
    for (int it = 0;; ++it) {
        int *cnt_cur = st.cnt + (it % 3) * NCLS;
        int *cnt_zero = st.cnt + ((it + 2) % 3) * NCLS;
        int *q_cur = st.queue[it & 1];
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            long long now = (long long)wall_clock64();
            if (it > 0 && sp.trace && it - 1 < sp.trace_cap) sp.trace[it - 1].t_end = now;
            t_prev = now;
        }
        if (blockIdx.x == 0 && threadIdx.x < NCLS) st_coherent(cnt_zero + threadIdx.x, 0);
        for (;;) {
            int *dc = ctrl->defer_cnt + pass % 3, *dm = ctrl->defer_min + pass % 3;
            if (!PUSH) {
                // the statistics slot of the next pass was last read before the previous barrier
                if (blockIdx.x == 0 && threadIdx.x == 0) {
                    st_coherent(ctrl->defer_cnt + (pass + 1) % 3, 0);
                    st_coherent(ctrl->defer_min + (pass + 1) % 3, NO_WINDOW);
                }
                compaction_impl(st.ctrl, st.bm[0], st.cls, st.seg_off, st.nwords, cnt_cur, q_cur, st.dist, T, dc, dm);
                if (!grid_sync(ctrl, nb, gen)) return;
                pass++;
            }
            if (threadIdx.x < NCLS) {
                tb.cnt[threadIdx.x] = ld_coherent(cnt_cur + threadIdx.x);
                tb.seg[threadIdx.x] = st.seg_off[threadIdx.x];
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                long long tot = 0;
                for (int c = 1; c < NCLS; ++c) tot += tb.cnt[c];
                s_total = tot > 0x7FFFFFFF ? 0x7FFFFFFF : (int)tot;
            }
            __syncthreads();
            if (s_total > 0 || T == NO_WINDOW) break;
            // nothing below the window: raise it to the least pending distance + delta
            if (ld_coherent(dc) == 0) break;
            const int mn = ld_coherent(dm);
            T = mn > NO_WINDOW - sp.delta ? NO_WINDOW : mn + sp.delta;
        }
        const int total = s_total;
        if (total == 0 || (sp.max_iters > 0 && it >= sp.max_iters)) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                ctrl->iters = it;
                ctrl->done = total == 0 ? 1 : 0;
                if (total != 0) ctrl->error = NP_ERR_NOT_CONVERGED;
            }
            return;
        }
        Cfg cfg = decide(sp, tb.cnt, it);
        if (threadIdx.x == 0) build_tables(tb, cfg);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            ctrl->frontier_total += total;
            record_trace(sp, it, cfg, tb.cnt, t_prev, ctrl);
        }
        __syncthreads();
        PushCtx pc;
        pc.bm_next = PUSH ? st.bm[(it + 1) & 1] : st.bm[0];
        pc.bm_cur = st.bm[it & 1];
        pc.cnt_next = st.cnt + ((it + 1) % 3) * NCLS;
        pc.q_next = st.queue[(it + 1) & 1];
        expansion<BATCH, UNIT, PUSH, false>(st, pc, q_cur, tb);
        if (!grid_sync(ctrl, nb, gen)) return;
    }
}

// ------------------------------------------------------------------------------------
// host-loop / graph drive kernels
// ------------------------------------------------------------------------------------
__global__ __launch_bounds__(WG) void k_compact(DevState st, int it_off) {
    Ctrl *ctrl = st.ctrl;
    if (ld_coherent(&ctrl->done) || ld_coherent(&ctrl->error)) return;
    const int it = ld_coherent(&ctrl->it_base) + it_off;
    const int pass = ld_coherent(&ctrl->pass);
    const int T = ld_coherent(&ctrl->thresh);
    if (blockIdx.x == 0 && threadIdx.x < NCLS) st.cnt[((it + 2) % 3) * NCLS + threadIdx.x] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctrl->defer_cnt[(pass + 1) % 3] = 0;
        ctrl->defer_min[(pass + 1) % 3] = NO_WINDOW;
    }
    compaction_impl(st.ctrl, st.bm[0], st.cls, st.seg_off, st.nwords, st.cnt + (it % 3) * NCLS, st.queue[it & 1],
                    st.dist, T, ctrl->defer_cnt + pass % 3, ctrl->defer_min + pass % 3);
}

__global__ __launch_bounds__(WG) void k_decide(DevState st, SchedDev sp, int it_off) {
    __shared__ int s_cnt[NCLS];
    Ctrl *ctrl = st.ctrl;
    if (ld_coherent(&ctrl->done) || ld_coherent(&ctrl->error)) return;
    const int it = ld_coherent(&ctrl->it_base) + it_off;
    if (threadIdx.x < NCLS) s_cnt[threadIdx.x] = ld_coherent(st.cnt + (it % 3) * NCLS + threadIdx.x);
    __syncthreads();
    long long tot = 0;
    for (int c = 1; c < NCLS; ++c) tot += s_cnt[c];
    const int pass = ld_coherent(&ctrl->pass);  // the compaction of this slot
    if (tot == 0 && !sp.push) {
        const int T = ld_coherent(&ctrl->thresh);
        if ((T != NO_WINDOW || st.hold_cls) && ld_coherent(ctrl->defer_cnt + pass % 3) > 0) {
            // nothing queued but pairs kept back: raise the window (held vertices need only the
            // next pass); this slot's expansion is skipped
            if (threadIdx.x == 0) {
                if (T != NO_WINDOW) {
                    const int mn = ld_coherent(ctrl->defer_min + pass % 3);
                    ctrl->thresh = mn > NO_WINDOW - sp.delta ? NO_WINDOW : mn + sp.delta;
                }
                ctrl->skip = 1;
                ctrl->pass = pass + 1;
            }
            return;
        }
    }
    if (tot == 0 || (sp.max_iters > 0 && ctrl->iters >= sp.max_iters)) {
        if (threadIdx.x == 0) {
            if (tot == 0) ctrl->done = 1;
            else ctrl->error = NP_ERR_NOT_CONVERGED;
        }
        return;
    }
    Cfg cfg = decide(sp, s_cnt, ctrl->iters);
    if (threadIdx.x == 0) {
        if (ctrl->cfg_pad == 0) {  // cfg_pad != 0: a forced schedule (stepping) is kept
            ctrl->cfg_level = cfg.level;
            ctrl->cfg_klo = cfg.klo;
            ctrl->cfg_khi = cfg.khi;
        } else {
            cfg = make_cfg(ctrl->cfg_level, ctrl->cfg_klo, ctrl->cfg_khi);
        }
        ctrl->frontier_total += tot;
        record_trace(sp, ctrl->iters, cfg, s_cnt, (long long)wall_clock64(), ctrl);
        ctrl->iters += 1;
        ctrl->skip = 0;
        if (!sp.push) ctrl->pass = pass + 1;
    }
}

template <bool BATCH, bool UNIT, bool PUSH>
__global__ __launch_bounds__(WG) void k_expand(DevState st, int it_off) {
    __shared__ ItemTables tb;
    Ctrl *ctrl = st.ctrl;
    if (ld_coherent(&ctrl->done) || ld_coherent(&ctrl->error) || ld_coherent(&ctrl->skip)) return;
    const int it = ld_coherent(&ctrl->it_base) + it_off;
    if (threadIdx.x < NCLS) {
        tb.cnt[threadIdx.x] = ld_coherent(st.cnt + (it % 3) * NCLS + threadIdx.x);
        tb.seg[threadIdx.x] = st.seg_off[threadIdx.x];
    }
    if (PUSH && blockIdx.x == 0 && threadIdx.x < NCLS) st.cnt[((it + 2) % 3) * NCLS + threadIdx.x] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        Cfg cfg = make_cfg(ld_coherent(&ctrl->cfg_level), ld_coherent(&ctrl->cfg_klo),
                           ld_coherent(&ctrl->cfg_khi));
        build_tables(tb, cfg);
    }
    __syncthreads();
    PushCtx pc;
    pc.bm_next = PUSH ? st.bm[(it + 1) & 1] : st.bm[0];
    pc.bm_cur = st.bm[it & 1];
    pc.cnt_next = st.cnt + ((it + 1) % 3) * NCLS;
    pc.q_next = st.queue[(it + 1) & 1];
    expansion<BATCH, UNIT, PUSH, true>(st, pc, st.queue[it & 1], tb);
}

__global__ void k_advance(DevState st, int k) {
    if (threadIdx.x == 0 && blockIdx.x == 0) st.ctrl->it_base += k;
}

__global__ void k_set_cfg(DevState st, int level, int klo, int khi) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st.ctrl->cfg_level = level;
        st.ctrl->cfg_klo = klo;
        st.ctrl->cfg_khi = khi;
        st.ctrl->cfg_pad = 1;
    }
}

// ------------------------------------------------------------------------------------
// the reference's other schedules (SSSP_GPU solutions 0-5, BFS_REC_GPU solutions 0 and 2)
// ------------------------------------------------------------------------------------
// They run on the host-loop and graph drives after the scan compaction and k_decide of the
// consolidated path, over the same class-segmented queue; they differ from it, and from one
// another, only in how queued vertices are mapped to workers (sssp_wrapper.cu:151-473,
// sssp_kernel.cu:34-327; bfs_rec_kernel.cu:41-140).  A vertex is heavy when its out-degree is
// at least 2^cfg_klo (THRESHOLD 64 = 2^6, sssp_kernel.cu:7), i.e. its class is >= cfg_klo + 1.
// Every relaxation is the reference's: alt = dist[u] + w; if (dist[v] > alt) atomicMin.

// One lane walks all edges of one vertex (the thread loop of unorder_threadQueue_kernel,
// sssp_kernel.cu:34-55), UNR edges' loads in flight at a time.
template <bool UNIT>
__device__ __forceinline__ void relax_serial(const DevState &st, const PushCtx &pc, int slot) {
    const int e0 = st.row[slot], e1 = st.row[slot + 1];
    const int du = ld_coherent(st.dist + slot);
    for (int base = e0; base < e1; base += UNR) {
        int tv[UNR], alt[UNR], cur[UNR];
#pragma unroll
        for (int k = 0; k < UNR; ++k) {
            const int e = base + k < e1 ? base + k : e0;
            tv[k] = st.col[e];
            alt[k] = du + (UNIT ? 1 : st.w[e]);
        }
#pragma unroll
        for (int k = 0; k < UNR; ++k) cur[k] = base + k < e1 ? ld_coherent(st.dist + tv[k]) : 0;
#pragma unroll
        for (int k = 0; k < UNR; ++k) {
            int c = 0;
            if (base + k < e1 && alt[k] < cur[k]) relax_tail<false>(st, pc, tv[k], alt[k], c);
        }
    }
}

// All lanes of `stride` workers (t = this one's index) walk the edges of one vertex.
template <bool UNIT>
__device__ __forceinline__ void relax_vertex(const DevState &st, const PushCtx &pc, int slot, int t, int stride) {
    const int e0 = st.row[slot], e1 = st.row[slot + 1];
    relax_range<UNIT, false, true>(st, pc, 0, e0, e1, ld_coherent(st.dist + slot), t, stride);
}

// Class prefix of the queued vertices of classes [c0, c1) into pre[0..NCLS] (thread 0).
__device__ __forceinline__ void sol_prefix(const int *cnt, int c0, int c1, long long *pre) {
    long long acc = 0;
    for (int c = 0; c < NCLS; ++c) {
        pre[c] = acc;
        if (c >= c0 && c < c1) acc += cnt[c];
    }
    pre[NCLS] = acc;
}

struct SolBufs {
    int *buf;  // global delayed buffer (solution 3): heavy vertices of one expansion
    int *bcnt; // [3] its fill counters, rotating by iteration
};

enum SolMode { SOL_THREAD = 0, SOL_DUAL = 1, SOL_LDS = 2, SOL_GLOBAL = 3, SOL_WAVE = 4, SOL_WGLIST = 5 };

// The parent kernel of solutions 0-5: one lane per queued vertex.
//   SOL_THREAD  every vertex walked by its lane (solution 0, unorder_threadQueue_kernel)
//   SOL_DUAL    light classes only; the heavy ones are the second queue (solution 1)
//   SOL_LDS     heavy vertices of the workgroup's 256 into an LDS buffer, then the whole
//               workgroup walks each one's edges (solution 2,
//               unorder_threadQueue_shared_delaybuffer_kernel, sssp_kernel.cu:112-172)
//   SOL_GLOBAL  heavy vertices appended to a global buffer (one atomic per wavefront),
//               drained by k_sol_drain (solution 3, sssp_kernel.cu:179-209)
//   SOL_WAVE    each wavefront then walks its own heavy vertices one by one with 64 lanes
//               (solution 4: a 64-thread nested launch per heavy vertex, sssp_kernel.cu:237-286)
//   SOL_WGLIST  heavy vertices of the workgroup into an LDS list; its four wavefronts take
//               them in turn, 64 lanes per vertex (solution 5: one nested launch per block
//               over the block's buffer, 64 threads per vertex, sssp_kernel.cu:289-327)
template <int MODE, bool UNIT>
__global__ __launch_bounds__(WG) void k_sol_parent(DevState st, SolBufs sb, int it_off) {
    __shared__ long long s_pre[NCLS + 1];
    __shared__ int s_seg[NCLS];
    __shared__ int s_list[WG];
    __shared__ int s_nlist;
    Ctrl *ctrl = st.ctrl;
    if (ld_coherent(&ctrl->done) || ld_coherent(&ctrl->error) || ld_coherent(&ctrl->skip)) return;
    const int it = ld_coherent(&ctrl->it_base) + it_off;
    const int hc = ld_coherent(&ctrl->cfg_klo) + 1;  // first heavy class
    const int *cnt = st.cnt + (it % 3) * NCLS;
    if (MODE == SOL_GLOBAL && blockIdx.x == 0 && threadIdx.x == 0) st_coherent(sb.bcnt + (it + 1) % 3, 0);
    if (threadIdx.x < NCLS) s_seg[threadIdx.x] = st.seg_off[threadIdx.x];
    if (threadIdx.x == 0) {
        int lc[NCLS];
        for (int c = 0; c < NCLS; ++c) lc[c] = ld_coherent(cnt + c);
        sol_prefix(lc, 1, MODE == SOL_DUAL ? hc : NCLS, s_pre);
        s_nlist = 0;
    }
    __syncthreads();
    const long long total = s_pre[NCLS];
    const int *q = st.queue[it & 1];
    PushCtx pc;
    pc.bm_next = st.bm[0];
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    for (long long base = (long long)blockIdx.x * WG; base < total; base += (long long)gridDim.x * WG) {
        const long long i = base + threadIdx.x;
        int slot = 0, c = 0;
        const bool valid = i < total;
        if (valid) {
            c = find_class(s_pre, i);
            slot = ld_coherent(q + s_seg[c] + (i - s_pre[c]));
        }
        const bool heavy = valid && MODE != SOL_THREAD && MODE != SOL_DUAL && c >= hc;
        if (valid && !heavy) relax_serial<UNIT>(st, pc, slot);
        if (MODE == SOL_LDS || MODE == SOL_WGLIST) {
            if (heavy) s_list[atomicAdd(&s_nlist, 1)] = slot;
            __syncthreads();
            const int nl = s_nlist;
            if (MODE == SOL_LDS) {
                for (int k = 0; k < nl; ++k) relax_vertex<UNIT>(st, pc, s_list[k], threadIdx.x, WG);
            } else {
                for (int k = wave; k < nl; k += WAVES) relax_vertex<UNIT>(st, pc, s_list[k], lane, 64);
            }
            __syncthreads();
            if (threadIdx.x == 0) s_nlist = 0;
            __syncthreads();
        } else if (MODE == SOL_WAVE) {
            unsigned long long m = __ballot(heavy);
            while (m) {
                const int l = __ffsll((long long)m) - 1;
                m &= m - 1;
                relax_vertex<UNIT>(st, pc, __shfl(slot, l), lane, 64);
            }
        } else if (MODE == SOL_GLOBAL) {
            const unsigned long long m = __ballot(heavy);
            if (m) {
                int at = 0;
                if (lane == 0) at = atomicAdd(sb.bcnt + it % 3, __popcll(m));
                at = __shfl(at, 0);
                const long long bi = (long long)at + __popcll(m & __lanemask_lt());
                if (heavy && seg_ok(ctrl, bi, 0, st.seg_off[NCLS])) st_coherent(sb.buf + bi, slot);
            }
        }
    }
}

// Second launch of solution 1: one workgroup per heavy vertex (unorder_blockQueue_kernel,
// sssp_kernel.cu:57-78, over the heavy queue).
template <bool UNIT>
__global__ __launch_bounds__(WG) void k_sol_heavy(DevState st, int it_off) {
    __shared__ long long s_pre[NCLS + 1];
    __shared__ int s_seg[NCLS];
    Ctrl *ctrl = st.ctrl;
    if (ld_coherent(&ctrl->done) || ld_coherent(&ctrl->error) || ld_coherent(&ctrl->skip)) return;
    const int it = ld_coherent(&ctrl->it_base) + it_off;
    const int hc = ld_coherent(&ctrl->cfg_klo) + 1;
    const int *cnt = st.cnt + (it % 3) * NCLS;
    if (threadIdx.x < NCLS) s_seg[threadIdx.x] = st.seg_off[threadIdx.x];
    if (threadIdx.x == 0) {
        int lc[NCLS];
        for (int c = 0; c < NCLS; ++c) lc[c] = ld_coherent(cnt + c);
        sol_prefix(lc, hc, NCLS, s_pre);
    }
    __syncthreads();
    const long long total = s_pre[NCLS];
    const int *q = st.queue[it & 1];
    PushCtx pc;
    pc.bm_next = st.bm[0];
    for (long long k = blockIdx.x; k < total; k += gridDim.x) {
        const int c = find_class(s_pre, k);
        relax_vertex<UNIT>(st, pc, ld_coherent(q + s_seg[c] + (k - s_pre[c])), threadIdx.x, WG);
    }
}

// Second launch of solution 3: one wavefront per buffered vertex (NESTED_BLOCK_SIZE 64 threads
// of sssp_process_buffer / unorder_blockQueue_kernel).
template <bool UNIT>
__global__ __launch_bounds__(WG) void k_sol_drain(DevState st, SolBufs sb, int it_off) {
    Ctrl *ctrl = st.ctrl;
    if (ld_coherent(&ctrl->done) || ld_coherent(&ctrl->error) || ld_coherent(&ctrl->skip)) return;
    const int it = ld_coherent(&ctrl->it_base) + it_off;
    const int nb = ld_coherent(sb.bcnt + it % 3);
    PushCtx pc;
    pc.bm_next = st.bm[0];
    const long long nw = (long long)gridDim.x * WAVES;
    for (long long k = (long long)blockIdx.x * WAVES + (threadIdx.x >> 6); k < nb; k += nw)
        relax_vertex<UNIT>(st, pc, ld_coherent(sb.buf + k), lane_id(), 64);
}

// BFS_REC_GPU solution 0 (bfs_kernel_flat, bfs_rec_kernel.cu:41-58): level-synchronous and
// topology-driven -- every vertex is inspected at every level; those at `level` set their
// unreached or deeper children to level + 1 (plain stores: every writer of a level writes the
// same value) and raise *changed.
__global__ __launch_bounds__(WG) void k_bfs_flat(const int *row, const int *col, int *lvl, int n, int level,
                                                 int *changed) {
    int any = 0;
    for (long long v = (long long)blockIdx.x * WG + threadIdx.x; v < n; v += (long long)gridDim.x * WG) {
        if (lvl[v] != level) continue;
        for (int e = row[v], e1 = row[v + 1]; e < e1; ++e) {
            const int t = col[e];
            if (lvl[t] > level + 1) { lvl[t] = level + 1; any = 1; }
        }
    }
    if (__ballot(any) && lane_id() == 0) st_coherent(changed, 1);
}

// BFS_REC_GPU solution 2 (bfs_kernel_dp_hier, bfs_rec_kernel.cu:95-140): two levels per
// expansion.  A queued vertex u's children are taken 64 at a time by one wavefront, a lane
// per child (the reference's block per child); each child whose level this lane lowered to
// level(u) + 1 is expanded at once by the whole wavefront (the reference's threads over the
// grandchildren), and the grandchildren it lowers to level(u) + 2 are flagged for the next
// expansion (the reference's nested launch per grandchild).  Levels are exact: a vertex is
// expanded after every lowering of its level, as in the consolidated path.
__device__ __forceinline__ int hier_chunks(int c) { return c <= 6 ? 1 : 1 << (c - 6); }

__global__ __launch_bounds__(WG) void k_bfs_hier(DevState st, int it_off) {
    __shared__ long long s_pre[NCLS + 1];
    __shared__ int s_seg[NCLS];
    Ctrl *ctrl = st.ctrl;
    if (ld_coherent(&ctrl->done) || ld_coherent(&ctrl->error) || ld_coherent(&ctrl->skip)) return;
    const int it = ld_coherent(&ctrl->it_base) + it_off;
    const int *cnt = st.cnt + (it % 3) * NCLS;
    if (threadIdx.x < NCLS) s_seg[threadIdx.x] = st.seg_off[threadIdx.x];
    if (threadIdx.x == 0) {
        long long acc = 0;
        for (int c = 0; c < NCLS; ++c) {
            s_pre[c] = acc;
            if (c >= 1) acc += (long long)ld_coherent(cnt + c) * hier_chunks(c);
        }
        s_pre[NCLS] = acc;
    }
    __syncthreads();
    const long long total = s_pre[NCLS];
    const int *q = st.queue[it & 1];
    PushCtx pc;
    pc.bm_next = st.bm[0];
    const int lane = lane_id();
    const long long nw = (long long)gridDim.x * WAVES;
    for (long long k = (long long)blockIdx.x * WAVES + (threadIdx.x >> 6); k < total; k += nw) {
        const int c = find_class(s_pre, k);
        const long long r = k - s_pre[c];
        const int sh = c <= 6 ? 0 : c - 6;
        const int u = ld_coherent(q + s_seg[c] + (r >> sh));
        const int j = (int)(r & ((1ll << sh) - 1));
        const long long e = (long long)st.row[u] + (long long)j * 64 + lane;
        const int e1 = st.row[u + 1];
        const int du = ld_coherent(st.dist + u);
        bool won = false;
        int child = 0;
        if (e < e1) {
            child = st.col[e];
            const int alt = du + 1;
            if (alt < ld_coherent(st.dist + child)) won = alt < atomicMin(&st.dist[child], alt);
        }
        unsigned long long m = __ballot(won);
        while (m) {
            const int l = __ffsll((long long)m) - 1;
            m &= m - 1;
            const int ch = __shfl(child, l);
            relax_range<true, false, true>(st, pc, 0, st.row[ch], st.row[ch + 1], du + 1, lane, 64);
        }
    }
}

// ------------------------------------------------------------------------------------
// source-packed batch: up to 32 sources of one graph solved together
// ------------------------------------------------------------------------------------
// Distances are vertex-major, 2^kshift sources per vertex row: dist[(v << kshift) + i].
// The distances of every source at one vertex share a cache line (64 B for 16 sources),
// so one read of an edge's col/w serves every source pending at its tail, and the check
// of dist[v][*] is one coalesced line instead of one random sector per source.
// mflag[v] collects the sources whose distance at v improved since v was last queued;
// bm[0] has one bit per vertex with a nonzero mflag (set by the relaxation that made it
// nonzero), so the compaction scans n/8 bytes and reads mflag only where work is pending.
// A queue entry carries the source mask it was queued with (qmask).  The work of a
// queued vertex is out-degree x pending sources; its items are numbered edge-major,
// item t = (edge t / P, t % P-th pending source), so the P lanes of one edge are adjacent.
// The relaxation per (edge, source) is the reference's: alt = dist[u][i] + w;
// if (dist[v][i] > alt) atomicMin -- int32, as in the single-source kernels.

// Position of the k-th (0-based) set bit of m (k < popcount(m)).
__device__ __forceinline__ int nth_set_bit(unsigned m, int k) {
    int pos = 0, c;
    c = __popc(m & 0xFFFFu); if (k >= c) { k -= c; m >>= 16; pos += 16; }
    c = __popc(m & 0xFFu);   if (k >= c) { k -= c; m >>= 8;  pos += 8; }
    c = __popc(m & 0xFu);    if (k >= c) { k -= c; m >>= 4;  pos += 4; }
    c = __popc(m & 0x3u);    if (k >= c) { k -= c; m >>= 2;  pos += 2; }
    c = (int)(m & 1u);       if (k >= c) { pos += 1; }
    return pos;
}

// Lane-bin vertices of a packed expansion have out-degree < 2^(PACK_LANE_CLS - 1): each lane
// of a lane tile walks the edges of one vertex in sequence (pk_lane_tile), so the bin is kept to
// short rows; longer ones go to the wavefront and wider bins whatever the schedule says.
constexpr int PACK_LANE_CLS = 10;

// UNR packed relaxation steps of a wavefront: in step k every lane holds one (edge, source)
// item -- target tv[k], pending source s, candidate alt[k] = dist[u][s] + w, loaded dist[v][s]
// cur[k] -- or none (ok[k] false).  The lanes of one edge form a run of consecutive lanes (rs =
// the run's first lane inside this wavefront, last = this lane ends its run), so they share v.
// The reference's relaxation per item (sssp_kernel.cu:46-52): if (dist[v][s] > alt) atomicMin,
// issued without a return value; each source that passed the check sets its bit in mflag[v],
// OR-ed across the run first so that one lane per run updates it, and the first bit of mflag[v]
// sets v's frontier bit.  Flagging on the check alone flags exactly the (v, s) the atomic's old
// value would: when alt < cur but a concurrent atomicMin has already stored less, that writer
// passed its own check in this same launch and flags (v, s) itself (every earlier store is
// visible to the check).  So no lane waits for an atomic's return before the next step.
// Vertices without out-edges are flagged too and dropped by the compaction.  Every lane of the
// wavefront calls it together.
__device__ __forceinline__ void pk_relax_steps(const DevState &st, const bool *ok, const int *tv, int s,
                                               const int *alt, const int *cur, int rs, bool last) {
#pragma unroll
    for (int k = 0; k < UNR; ++k) {
        if (ok[k] && alt[k] < cur[k])
            atomicMin(st.dist + (((unsigned)tv[k] << st.kshift) + (unsigned)s), alt[k]);
    }
    // every step's mflag OR is issued before any old value is tested: one wait for all UNR
    const int lane = lane_id();
    unsigned prev[UNR];
#pragma unroll
    for (int k = 0; k < UNR; ++k) {
        prev[k] = 1u;
        const unsigned bit = (ok[k] && alt[k] < cur[k]) ? 1u << s : 0u;
        if (__ballot(bit != 0) == 0) continue;
        unsigned acc = bit;
#pragma unroll
        for (int d = 1; d < 32; d <<= 1) {  // runs are at most 32 lanes long (one per source)
            const unsigned y = __shfl_up(acc, d);
            if (lane - d >= rs) acc |= y;
        }
        if (last && ok[k] && acc) prev[k] = atomicOr(&st.mflag[tv[k]], acc);
    }
#pragma unroll
    for (int k = 0; k < UNR; ++k)
        if (prev[k] == 0) atomicOr(&st.bm[0][tv[k] >> 5], 1u << (tv[k] & 31));
}

// Items of one queued vertex u -- edges [e0, e1) x the P pending sources of M -- on `stride`
// workers (t = this worker's index; M, e0, e1 are the same for all of them).  Worker t takes
// source j = t mod P, whose distance at u it reads once, and the edges e0 + t / P + k q,
// q = stride / P: the P workers of an edge are adjacent and share its col/w read and the cache
// line of dist[v][*]; the last stride mod P workers idle.
template <bool UNIT>
__device__ __forceinline__ void pk_relax_vertex(const DevState &st, int u, unsigned M, int e0, int e1, int t,
                                                int stride) {
    M = (unsigned)__builtin_amdgcn_readfirstlane((int)M);
    const int P = __popc(M);
    if (P == 0 || e0 >= e1) return;
    const int ks = st.kshift;
    const int q = stride / P;
    const int eo = t / P, j = t - eo * P;
    const bool on = eo < q;
    const int s = on ? nth_set_bit(M, j) : 0;
    const int du = on ? ld_coherent(st.dist + (((unsigned)u << ks) + (unsigned)s)) : 0;
    const int lane = lane_id();
    const int rs = lane - j > 0 ? lane - j : 0;
    const bool last = j == P - 1 || lane == 63;
    for (int e = e0 + eo; __ballot(on && e < e1) != 0; e += UNR * q) {
        int tv[UNR], alt[UNR], cur[UNR];
        bool ok[UNR];
#pragma unroll
        for (int k = 0; k < UNR; ++k) {
            const int ek = e + k * q;
            ok[k] = on && ek < e1;
            const unsigned ec = ok[k] ? ek : e0;
            tv[k] = st.col[ec];
            alt[k] = du + (UNIT ? 1 : st.w[ec]);
        }
#pragma unroll
        for (int k = 0; k < UNR; ++k)
            cur[k] = ok[k] ? ld_coherent(st.dist + (((unsigned)tv[k] << ks) + (unsigned)s)) : 0;
        pk_relax_steps(st, ok, tv, s, alt, cur, rs, last);
    }
}

// Wavefront-scope form of pk_relax_vertex (stride 64) with the neighbour tile staged in LDS: the
// wavefront copies up to PK_TILE consecutive edges' col (and w) into its own LDS slice with
// coalesced loads, one edge per lane per load, then takes the edges of its relaxation steps from
// there.  Without staging every step re-reads col/w from global memory, q = 64 / P distinct edges
// per load, and waits for them before its distance gathers.  tcol / tw: the wavefront's slice.
constexpr int PK_TILE = 256;
constexpr bool PK_STAGE = true;
// a group of UNR steps spans up to UNR * 64 edges (one pending source): a tile must hold one,
// or the rounded tile below would be empty and its loop would not advance
static_assert(PK_TILE >= UNR * 64 && PK_TILE % 64 == 0, "PK_TILE must hold a step group of one source");
template <bool UNIT>
__device__ __forceinline__ void pk_relax_vertex_staged(const DevState &st, int u, unsigned M, int e0, int e1,
                                                       int *tcol, int *tw) {
    M = (unsigned)__builtin_amdgcn_readfirstlane((int)M);
    const int P = __popc(M);
    if (P == 0 || e0 >= e1) return;
    const int ks = st.kshift;
    const int lane = lane_id();
    const int q = 64 / P;
    const int eo = lane / P, j = lane - eo * P;
    const bool on = eo < q;
    const int s = on ? nth_set_bit(M, j) : 0;
    const int du = on ? ld_coherent(st.dist + (((unsigned)u << ks) + (unsigned)s)) : 0;
    const int rs = lane - j > 0 ? lane - j : 0;
    const bool last = j == P - 1 || lane == 63;
    const int span = UNR * q;                  // edges of one group of UNR steps (<= PK_TILE)
    const int tile = (PK_TILE / span) * span;  // whole groups per tile: no group straddles two tiles
    for (int tb = e0; tb < e1; tb += tile) {
        const int nt = e1 - tb < tile ? e1 - tb : tile;
        __builtin_amdgcn_wave_barrier();       // the previous tile's reads come first
        int c[PK_TILE / 64], wv[PK_TILE / 64];
#pragma unroll
        for (int r = 0; r < PK_TILE / 64; ++r) {
            const int i = lane + 64 * r;
            const int ic = tb + (i < nt ? i : nt - 1);  // lanes past the tile re-read its last edge
            c[r] = st.col[ic];
            if (!UNIT) wv[r] = st.w[ic];
        }
#pragma unroll
        for (int r = 0; r < PK_TILE / 64; ++r) {
            tcol[lane + 64 * r] = c[r];
            if (!UNIT) tw[lane + 64 * r] = wv[r];
        }
        __builtin_amdgcn_wave_barrier();
        for (int e = eo; __ballot(on && e < nt) != 0; e += span) {
            int tv[UNR], alt[UNR], cur[UNR];
            bool ok[UNR];
#pragma unroll
            for (int k = 0; k < UNR; ++k) {
                const int ek = e + k * q;
                ok[k] = on && ek < nt;
                const int ec = ok[k] ? ek : 0;
                tv[k] = tcol[ec];
                alt[k] = du + (UNIT ? 1 : tw[ec]);
            }
#pragma unroll
            for (int k = 0; k < UNR; ++k)
                cur[k] = ok[k] ? ld_coherent(st.dist + (((unsigned)tv[k] << ks) + (unsigned)s)) : 0;
            pk_relax_steps(st, ok, tv, s, alt, cur, rs, last);
        }
    }
}

// Lane bin: 64 queued vertices (queue[qidx .. qidx+nvalid)) of one out-degree class.  Their
// (vertex, pending source) pairs go to the lanes, 64 at a time, the pairs of one vertex adjacent;
// each lane walks its vertex's edges in order for its source, so the lanes of one vertex read the
// same edge (one col/w address, one dist[v][*] line) at every step.
template <bool UNIT>
__device__ __forceinline__ void pk_lane_tile(const DevState &st, const int *q, const unsigned *qm,
                                             long long qidx, int nvalid) {
    const int lane = lane_id();
    const int ks = st.kshift;
    int u = 0, s0 = 0, d = 0, P = 0;
    unsigned M = 0;
    if (lane < nvalid) {
        u = ld_coherent(q + qidx + lane);
        M = ld_coherent_u(qm + qidx + lane);
        s0 = st.row[u];
        d = st.row[u + 1] - s0;
        P = d > 0 ? __popc(M) : 0;
    }
    const int incl = wave_incl_scan(P);
    const int total = __shfl(incl, 63);
    const int excl = incl - P;
    if (total == 0) return;
    // lanes past the last pair read the first edge of the first vertex with pairs (a valid edge)
    const int e_safe = __shfl(s0, __ffsll((long long)__ballot(P > 0)) - 1);
    for (int base = 0; base < total; base += 64) {
        const int k = base + lane;
        const bool on = k < total;
        int o = 0;  // owner: the last lane whose exclusive offset is <= k
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
            const int x = __shfl(excl, o + step);
            if (x <= k) o += step;
        }
        const int j = k - __shfl(excl, o);
        const int Po = __shfl(P, o);
        const unsigned Mo = __shfl(M, o);
        const int uo = __shfl(u, o);
        // every lane takes part in each shuffle: a lane left out would return 0 to its readers
        const int so = __shfl(s0, o);
        const int dso = __shfl(d, o);
        const int e0 = on ? so : e_safe;
        const int deg = on ? dso : 0;
        const int s = on ? nth_set_bit(Mo, j) : 0;
        const int du = on ? ld_coherent(st.dist + (((unsigned)uo << ks) + (unsigned)s)) : 0;
        const int rs = lane - j > 0 ? lane - j : 0;
        const bool last = j == Po - 1 || lane == 63;
        for (int i = 0; __ballot(i < deg) != 0; i += UNR) {
            int tv[UNR], alt[UNR], cur[UNR];
            bool ok[UNR];
#pragma unroll
            for (int kk = 0; kk < UNR; ++kk) {
                ok[kk] = i + kk < deg;
                const unsigned ec = e0 + (ok[kk] ? i + kk : 0);  // a lane with deg 0 reads a valid edge of its tile
                tv[kk] = st.col[ec];
                alt[kk] = du + (UNIT ? 1 : st.w[ec]);
            }
#pragma unroll
            for (int kk = 0; kk < UNR; ++kk)
                cur[kk] = ok[kk] ? ld_coherent(st.dist + (((unsigned)tv[kk] << ks) + (unsigned)s)) : 0;
            pk_relax_steps(st, ok, tv, s, alt, cur, rs, last);
        }
    }
}

template <bool UNIT>
__device__ void pk_expansion(const DevState &st, const int *q, const unsigned *qm, ItemTables &tb) {
    const long long G = tb.pre[0][NCLS], K = tb.pre[1][NCLS], W = tb.pre[2][NCLS],
                    T = tb.pre[3][NCLS];
    for (long long k = blockIdx.x; k < K; k += gridDim.x) {
        int c = find_class(tb.pre[1], k);
        long long qi = tb.seg[c] + (k - tb.pre[1][c]);
        int u = ld_coherent(q + qi);
        unsigned M = ld_coherent_u(qm + qi);
        pk_relax_vertex<UNIT>(st, u, M, st.row[u], st.row[u + 1], threadIdx.x, WG);
    }
    const long long NW = (long long)gridDim.x * WAVES;
    const long long gw = (long long)blockIdx.x * WAVES + (threadIdx.x >> 6);
    const long long U = G + W + T;
    __shared__ int s_tcol[WAVES][PK_TILE], s_tw[WAVES][PK_TILE];
    int *tcol = s_tcol[threadIdx.x >> 6], *tw = s_tw[threadIdx.x >> 6];
    for (long long x = gw; x < U; x += NW) {
        if (x < G) {
            int c = find_class(tb.pre[0], x);
            long long r = x - tb.pre[0][c];
            int sh = c > GRID_CHUNK_SHIFT ? c - GRID_CHUNK_SHIFT : 0;
            long long qi = tb.seg[c] + (r >> sh);
            int j = (int)(r & ((1ll << sh) - 1));
            int u = ld_coherent(q + qi);
            unsigned M = ld_coherent_u(qm + qi);
            int s = st.row[u], e = st.row[u + 1];
            int e0 = s + (j << GRID_CHUNK_SHIFT);
            int e1 = e0 + (1 << GRID_CHUNK_SHIFT);
            if (e1 > e) e1 = e;
            if (PK_STAGE) pk_relax_vertex_staged<UNIT>(st, u, M, e0, e1, tcol, tw);
            else pk_relax_vertex<UNIT>(st, u, M, e0, e1, lane_id(), 64);
        } else if (x < G + W) {
            long long i = x - G;
            int c = find_class(tb.pre[2], i);
            long long qi = tb.seg[c] + (i - tb.pre[2][c]);
            int u = ld_coherent(q + qi);
            unsigned M = ld_coherent_u(qm + qi);
            if (PK_STAGE) pk_relax_vertex_staged<UNIT>(st, u, M, st.row[u], st.row[u + 1], tcol, tw);
            else pk_relax_vertex<UNIT>(st, u, M, st.row[u], st.row[u + 1], lane_id(), 64);
        } else {
            long long t = x - G - W;
            int c = find_class(tb.pre[3], t);
            long long ti = t - tb.pre[3][c];
            long long first = ti * 64;
            int nvalid = (int)((long long)tb.cnt[c] - first < 64 ? (long long)tb.cnt[c] - first : 64);
            pk_lane_tile<UNIT>(st, q, qm, tb.seg[c] + first, nvalid);
        }
    }
}

// Number of the pack's sources that have reached v (finite entries of its distance row; the
// row's unused entries stay at inf).
__device__ __forceinline__ int pk_reached(const DevState &st, long long v) {
    const int *row = st.dist + (v << st.kshift);
    int r = 0;
    if (st.kshift < 2) {
        for (int i = 0; i < (1 << st.kshift); ++i) r += ld_coherent(row + i) < st.inf;
        return r;
    }
    const int4 *row4 = (const int4 *)row;
    for (int q = 0; q < (1 << (st.kshift - 2)); ++q) {
        const int4 x = row4[q];
        r += (x.x < st.inf) + (x.y < st.inf) + (x.z < st.inf) + (x.w < st.inf);
    }
    return r;
}

// The pending sources of v (mask m) whose distance at v is below the window T; *dmin
// receives the least distance among the others.  The row is read with 16-byte loads
// through the caches: a value older than the last atomicMin is only larger, so it can
// only keep a pair pending one pass longer (every pending pair is expanded eventually).
__device__ __forceinline__ unsigned pk_take(const DevState &st, long long v, unsigned m, int T, int *dmin) {
    *dmin = NO_WINDOW;
    if (T == NO_WINDOW) return m;
    const int *row = st.dist + (v << st.kshift);
    unsigned take = 0;
    int dm = NO_WINDOW;
    if (st.kshift < 2) {  // rows of 1 or 2 distances
        for (unsigned b = m; b; b &= b - 1) {
            const int i = __ffs(b) - 1, d = row[i];
            if (d < T) take |= 1u << i;
            else dm = d < dm ? d : dm;
        }
    } else {
        const int4 *row4 = (const int4 *)row;
        const int nq = 1 << (st.kshift - 2);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (q >= nq || !((m >> (4 * q)) & 0xFu)) continue;
            const int4 x = row4[q];
            const int d[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (!((m >> (4 * q + k)) & 1u)) continue;
                if (d[k] < T) take |= 1u << (4 * q + k);
                else dm = d[k] < dm ? d[k] : dm;
            }
        }
    }
    *dmin = dm;
    return take;
}

// Vertices with pending sources -> class-segmented queue with their source masks.  With a
// priority window (T < NO_WINDOW) only the (vertex, source) pairs with dist < T are queued;
// the others stay pending in mflag[v] (and v in the bitmap), and the pass adds their count and
// least distance to defer_cnt / defer_min.
__device__ __forceinline__ void pk_compaction(const DevState &st, int *cnt_out, int *q_out, unsigned *qm_out,
                                              int T, int *defer_cnt, int *defer_min) {
    __shared__ int s_lc[NCLS], s_lb[NCLS];
    __shared__ int s_dc, s_dm;
    unsigned *bm = st.bm[0];
    if (threadIdx.x == 0) { s_dc = 0; s_dm = NO_WINDOW; }
    if (threadIdx.x < NCLS) s_lc[threadIdx.x] = 0;
    __syncthreads();
    int pairs = 0, my_dc = 0, my_dm = NO_WINDOW;
    const long long nv = st.nwords * 32;
    // Without a window every pending vertex is taken whole and nothing stays flagged but the hubs
    // held for one pass (hold_cls): the frontier is as sparse as a single source's, so a thread
    // scans one bitmap word as compaction_impl does.
    for (long long wb = (long long)blockIdx.x * WG; T == NO_WINDOW && wb < st.nwords;
         wb += (long long)gridDim.x * WG) {
        const long long wi = wb + threadIdx.x;
        const unsigned word = wi < st.nwords ? ld_coherent_u(bm + wi) : 0u;
        unsigned queued = 0, held = 0, hw = 0;
        bool hw_read = false;
        for (unsigned w = word; w; w &= w - 1) {
            const int b = __ffs(w) - 1;
            const int c = st.cls[wi * 32 + b];
            if (!c) continue;  // no out-edges: dropped (its mflag stays set, so it is not flagged again)
            if (st.hold_cls && c >= st.hold_cls) {
                if (!hw_read) { hw = ld_coherent_u(st.bm[1] + wi); hw_read = true; }
                if (!((hw >> b) & 1u) && pk_reached(st, wi * 32 + b) < st.hold_min) {
                    held |= 1u << b;  // more sources are on their way: wait one pass, keep the flag
                    my_dc++;
                    continue;
                }
            }
            queued |= 1u << b;
            atomicAdd(&s_lc[c], 1);
        }
        reserve_classes(s_lc, s_lb, cnt_out);
        for (unsigned w = queued; w; w &= w - 1) {
            const long long v = wi * 32 + (__ffs(w) - 1);
            const int c = st.cls[v];
            const long long qi = st.seg_off[c] + s_lb[c] + atomicAdd(&s_lc[c], 1);
            const unsigned t = ld_coherent_u(st.mflag + v);
            st_coherent_u(st.mflag + v, 0u);
            if (seg_ok(st.ctrl, qi, st.seg_off[c], st.seg_off[c + 1])) {
                st_coherent(q_out + qi, (int)v);
                st_coherent_u(qm_out + qi, t);
            }
            pairs += __popc(t);
        }
        if (word) st_coherent_u(bm + wi, held);
        // a held vertex stays flagged, so the held bits are always a subset of the flags: they
        // become exactly this pass's held vertices (the ones held before are queued now)
        if (hw_read && hw != held) st_coherent_u(st.bm[1] + wi, held);
        __syncthreads();
        if (threadIdx.x < NCLS) s_lc[threadIdx.x] = 0;  // the second loop counted in s_lc
        __syncthreads();
    }
    for (long long vb = (long long)blockIdx.x * (WG * CPT); T != NO_WINDOW && vb < nv;
         vb += (long long)gridDim.x * (WG * CPT)) {
        int qc[CPT], qp[CPT];
        unsigned qt[CPT];
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const long long v = vb + k * WG + threadIdx.x;
            const unsigned word = v < nv ? ld_coherent_u(bm + (v >> 5)) : 0u;
            bool keep = false;
            qc[k] = 0;
            qp[k] = 0;
            qt[k] = 0;
            if ((word >> (v & 31)) & 1u) {
                const int c = st.cls[v];
                if (c) {  // no out-edges: dropped (its mflag stays set, so it is not flagged again)
                    const unsigned m = ld_coherent_u(st.mflag + v);
                    int dmin;
                    const unsigned t = pk_take(st, v, m, T, &dmin);
                    if (t) {
                        qc[k] = c;
                        qt[k] = t;
                        qp[k] = atomicAdd(&s_lc[c], 1);
                        st_coherent_u(st.mflag + v, m & ~t);
                    }
                    if (t != m) {
                        keep = true;
                        my_dc += __popc(m & ~t);
                        my_dm = dmin < my_dm ? dmin : my_dm;
                    }
                }
            }
            rewrite_word(bm, v, nv, word, keep);
        }
        reserve_classes(s_lc, s_lb, cnt_out);
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            if (!qc[k]) continue;
            const long long qi = st.seg_off[qc[k]] + s_lb[qc[k]] + qp[k];
            const long long v = vb + k * WG + threadIdx.x;
            if (seg_ok(st.ctrl, qi, st.seg_off[qc[k]], st.seg_off[qc[k] + 1])) {
                st_coherent(q_out + qi, (int)v);
                st_coherent_u(qm_out + qi, qt[k]);
            }
            pairs += __popc(qt[k]);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) pairs += __shfl_xor(pairs, o);
    if (lane_id() == 0 && pairs) atomicAdd((unsigned long long *)&st.ctrl->pairs_total, (unsigned long long)pairs);

    if (T != NO_WINDOW || st.hold_cls) flush_deferred(my_dc, my_dm, &s_dc, &s_dm, defer_cnt, defer_min);
}

// Lane-bin bound of a packed expansion (see PACK_LANE_CLS); the other bins keep theirs.
__device__ __forceinline__ Cfg pk_cfg(Cfg c) {
    if (c.b1 > PACK_LANE_CLS) c.b1 = PACK_LANE_CLS;
    return c;
}

template <bool UNIT>
__global__ __launch_bounds__(WG) void k_pk_persistent(DevState st, SchedDev sp) {
    __shared__ ItemTables tb;
    __shared__ int s_total;
    Ctrl *ctrl = st.ctrl;
    unsigned long long gen = 0;
    const unsigned nb = gridDim.x;
    long long t_prev = 0;
    // priority window: every workgroup computes the same sequence of thresholds
    int T = sp.delta > 0 ? sp.delta : NO_WINDOW;
    int pass = 0;
    for (int it = 0;; ++it) {
        int *cnt_cur = st.cnt + (it % 3) * NCLS;
        int *cnt_zero = st.cnt + ((it + 2) % 3) * NCLS;
        int *q_cur = st.queue[it & 1];
        unsigned *qm_cur = st.qmask[it & 1];
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            long long now = (long long)wall_clock64();
            if (it > 0 && sp.trace && it - 1 < sp.trace_cap) sp.trace[it - 1].t_end = now;
            t_prev = now;
        }
        if (blockIdx.x == 0 && threadIdx.x < NCLS) st_coherent(cnt_zero + threadIdx.x, 0);
        for (;;) {
            int *dc = ctrl->defer_cnt + pass % 3, *dm = ctrl->defer_min + pass % 3;
            // the statistics slot of the next pass was last read before the previous barrier
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                st_coherent(ctrl->defer_cnt + (pass + 1) % 3, 0);
                st_coherent(ctrl->defer_min + (pass + 1) % 3, NO_WINDOW);
            }
            pk_compaction(st, cnt_cur, q_cur, qm_cur, T, dc, dm);
            if (!grid_sync(ctrl, nb, gen)) return;
            pass++;
            if (threadIdx.x < NCLS) {
                tb.cnt[threadIdx.x] = ld_coherent(cnt_cur + threadIdx.x);
                tb.seg[threadIdx.x] = st.seg_off[threadIdx.x];
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                long long tot = 0;
                for (int c = 1; c < NCLS; ++c) tot += tb.cnt[c];
                s_total = tot > 0x7FFFFFFF ? 0x7FFFFFFF : (int)tot;
            }
            __syncthreads();
            if (s_total > 0) break;
            // nothing queued: done unless pairs were kept back (below the window, or held
            // vertices, which the next pass takes); a window rises to the least pending
            // distance + delta
            if (ld_coherent(dc) == 0) break;
            if (T != NO_WINDOW) {
                const int mn = ld_coherent(dm);
                T = mn > NO_WINDOW - sp.delta ? NO_WINDOW : mn + sp.delta;
            }
        }
        const int total = s_total;
        if (total == 0 || (sp.max_iters > 0 && it >= sp.max_iters)) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                ctrl->iters = it;
                ctrl->done = total == 0 ? 1 : 0;
                if (total != 0) ctrl->error = NP_ERR_NOT_CONVERGED;
            }
            return;
        }
        Cfg cfg = decide(sp, tb.cnt, it);
        if (threadIdx.x == 0) build_tables(tb, pk_cfg(cfg));
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            ctrl->frontier_total += total;
            record_trace(sp, it, cfg, tb.cnt, t_prev, ctrl);
        }
        __syncthreads();
        pk_expansion<UNIT>(st, q_cur, qm_cur, tb);
        if (!grid_sync(ctrl, nb, gen)) return;
    }
}

__global__ __launch_bounds__(WG) void k_pk_compact(DevState st, int it_off) {
    Ctrl *ctrl = st.ctrl;
    if (ld_coherent(&ctrl->done) || ld_coherent(&ctrl->error)) return;
    const int it = ld_coherent(&ctrl->it_base) + it_off;
    const int pass = ld_coherent(&ctrl->pass);
    const int T = ld_coherent(&ctrl->thresh);
    if (blockIdx.x == 0 && threadIdx.x < NCLS) st.cnt[((it + 2) % 3) * NCLS + threadIdx.x] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctrl->defer_cnt[(pass + 1) % 3] = 0;
        ctrl->defer_min[(pass + 1) % 3] = NO_WINDOW;
    }
    pk_compaction(st, st.cnt + (it % 3) * NCLS, st.queue[it & 1], st.qmask[it & 1], T,
                  ctrl->defer_cnt + pass % 3, ctrl->defer_min + pass % 3);
}

// The packed expansion is a latency-bound gather: it is held to 64 VGPRs so that eight waves share
// each SIMD (C3 packed, graph drive: 41.7 -> 31.3 ms per solve, a few registers spilled).  The
// persistent kernel keeps the compiler's allocation: held to 64 it spills its compaction and
// window code, which made C3 faster (42.2 -> 37.2 ms) but the grid with a window slower
// (125 -> 155 ms), and the graph drive is the faster drive for both.
// Work of the queue the last compaction built (traced solves on the host-loop and graph drives):
// out-degrees of the queued vertices into Ctrl::edges_total, out-degree x queued sources into
// Ctrl::items_total.  Kept out of the compaction, where even the untaken counting branch cost
// the windowed grid 20 %.
__global__ __launch_bounds__(WG) void k_pk_count(DevState st, int it_off) {
    __shared__ long long s_pre[NCLS + 1];
    __shared__ int s_seg[NCLS];
    Ctrl *ctrl = st.ctrl;
    if (ld_coherent(&ctrl->done) || ld_coherent(&ctrl->error)) return;
    const int it = ld_coherent(&ctrl->it_base) + it_off;
    if (threadIdx.x < NCLS) s_seg[threadIdx.x] = st.seg_off[threadIdx.x];
    if (threadIdx.x == 0) {
        int lc[NCLS];
        for (int c = 0; c < NCLS; ++c) lc[c] = ld_coherent(st.cnt + (it % 3) * NCLS + c);
        sol_prefix(lc, 1, NCLS, s_pre);
    }
    __syncthreads();
    const long long total = s_pre[NCLS];
    const int *q = st.queue[it & 1];
    const unsigned *qm = st.qmask[it & 1];
    unsigned long long edges = 0, items = 0;
    for (long long i = (long long)blockIdx.x * WG + threadIdx.x; i < total; i += (long long)gridDim.x * WG) {
        const int c = find_class(s_pre, i);
        const long long qi = s_seg[c] + (i - s_pre[c]);
        const int u = ld_coherent(q + qi);
        const unsigned d = (unsigned)(st.row[u + 1] - st.row[u]);
        edges += d;
        items += (unsigned long long)d * __popc(ld_coherent_u(qm + qi));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        edges += __shfl_xor(edges, o);
        items += __shfl_xor(items, o);
    }
    if (lane_id() == 0 && edges) {
        atomicAdd((unsigned long long *)&ctrl->edges_total, edges);
        atomicAdd((unsigned long long *)&ctrl->items_total, items);
    }
}

template <bool UNIT>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(8))) void k_pk_expand(DevState st, int it_off) {
    __shared__ ItemTables tb;
    Ctrl *ctrl = st.ctrl;
    if (ld_coherent(&ctrl->done) || ld_coherent(&ctrl->error) || ld_coherent(&ctrl->skip)) return;
    const int it = ld_coherent(&ctrl->it_base) + it_off;
    if (threadIdx.x < NCLS) {
        tb.cnt[threadIdx.x] = ld_coherent(st.cnt + (it % 3) * NCLS + threadIdx.x);
        tb.seg[threadIdx.x] = st.seg_off[threadIdx.x];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Cfg cfg = make_cfg(ld_coherent(&ctrl->cfg_level), ld_coherent(&ctrl->cfg_klo),
                           ld_coherent(&ctrl->cfg_khi));
        build_tables(tb, pk_cfg(cfg));
    }
    __syncthreads();
    pk_expansion<UNIT>(st, st.queue[it & 1], st.qmask[it & 1], tb);
}

__global__ void k_pk_init_sources(DevState st, const int *srcs, int nsrc, int T0) {
    if (blockIdx.x != 0) return;
    if (threadIdx.x == 0) {
        st.ctrl->thresh = T0;
        for (int k = 0; k < 3; ++k) st.ctrl->defer_min[k] = NO_WINDOW;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nsrc; i += blockDim.x) {
        const int v = srcs[i];
        st.dist[((long long)v << st.kshift) + i] = 0;
        if (st.cls[v] == 0) continue;
        atomicOr(&st.mflag[v], 1u << i);
        atomicOr(&st.bm[0][v >> 5], 1u << (v & 31));
    }
}

__global__ void k_unpack(const int *dist, int kshift, int lane, const int *old2new, int n, int *out) {
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (long long)gridDim.x * blockDim.x) {
        const long long r = old2new ? old2new[v] : v;
        out[v] = dist[(r << kshift) + lane];
    }
}

// ------------------------------------------------------------------------------------
// setup kernels
// ------------------------------------------------------------------------------------
__global__ void k_fill_i32(int *p, long long n, int v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u32(unsigned *p, long long n, unsigned v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ __launch_bounds__(WG) void k_degree_class(const int *row, int n, unsigned char *cls,
                                                     int *hist) {
    __shared__ int h[NCLS];
    if (threadIdx.x < NCLS) h[threadIdx.x] = 0;
    __syncthreads();
    for (long long v = (long long)blockIdx.x * WG + threadIdx.x; v < n; v += (long long)gridDim.x * WG) {
        int c = deg_class(row[v + 1] - row[v]);
        cls[v] = (unsigned char)c;
        atomicAdd(&h[c], 1);
    }
    __syncthreads();
    if (threadIdx.x < NCLS && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// CSR well-formedness, so that no kernel can index outside the arrays:
// row[0] == 0, row non-decreasing, row[n] == m, 0 <= col < n.  out[0] counts violations;
// with weights (w non-null) out[1] counts negative weights and out[2] receives the largest
// weight (out[] starts at 0).
__global__ void k_validate_csr(const int *row, const int *col, const int *w, int n, long long m, int *out) {
    int local = 0, neg = 0, wmax = 0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += stride) {
        int r = row[i];
        if (i == 0 && r != 0) local++;
        if (i == n && (long long)r != m) local++;
        if (i < n && row[i + 1] < r) local++;
    }
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride) {
        int c = col[e];
        if (c < 0 || c >= n) local++;
        if (w) {
            int x = w[e];
            if (x < 0) neg++;
            wmax = max(wmax, x);
        }
    }
    if (local) atomicAdd(&out[0], local);
    if (neg) atomicAdd(&out[1], neg);
    if (wmax) atomicMax(&out[2], wmax);
}

// Largest finite value of a solve's distance buffer (count entries): atomicMax into *out,
// which starts at 0.
__global__ __launch_bounds__(WG) void k_max_finite(const int *dist, long long count, int inf, int *out) {
    int mx = 0;
    for (long long i = (long long)blockIdx.x * WG + threadIdx.x; i < count; i += (long long)gridDim.x * WG) {
        int d = dist[i];
        if (d != inf) mx = max(mx, d);
    }
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o));
    if (lane_id() == 0 && mx) atomicMax(out, mx);
}

// Relaxations that int32 distances below inf cannot hold: an edge u -> v whose tail is
// reached, whose target is still at inf and with dist[u] + w >= inf (the relaxation
// alt < dist[v] is false there, so the target stays unreached although a path exists).
// Counts them into *bad.  Every layout: one distance per slot (BATCH: slots of
// concatenated instances) or rows of 2^kshift sources per vertex.
template <bool BATCH>
__global__ __launch_bounds__(WG) void k_dist_escape(DevState st, int *bad) {
    const int ks = st.kshift < 0 ? 0 : st.kshift;
    const long long total = st.nslots << ks;
    int local = 0;
    for (long long i = (long long)blockIdx.x * WG + threadIdx.x; i < total; i += (long long)gridDim.x * WG) {
        const int du = st.dist[i];
        if (du == st.inf) continue;
        const int slot = (int)(i >> ks);
        const int lane = (int)(i & ((1 << ks) - 1));
        int B = 0, VB = 0;
        decode<BATCH>(st, slot, B, VB);
        const int v = slot - B;
        for (int e = st.row[VB + v], e1 = st.row[VB + v + 1]; e < e1; ++e) {
            const long long t = (long long)(B + st.col[e]);
            if (st.dist[(t << ks) + lane] == st.inf && (long long)du + st.w[e] >= (long long)st.inf) local++;
        }
    }
    if (local) atomicAdd(bad, local);
}

template <bool PUSH>
__global__ void k_init_sources(DevState st, const int *srcs, int nsrc, int T0) {
    if (blockIdx.x != 0) return;
    if (threadIdx.x == 0) {
        st.ctrl->thresh = T0;
        for (int k = 0; k < 3; ++k) st.ctrl->defer_min[k] = NO_WINDOW;
    }
    for (int i = threadIdx.x; i < nsrc; i += blockDim.x) {
        int slot = srcs[i];
        st.dist[slot] = 0;
        int c = st.cls[slot];
        if (c == 0) continue;
        atomicOr(&st.bm[0][slot >> 5], 1u << (slot & 31));
        if (PUSH) {
            int p = atomicAdd(&st.cnt[c], 1);  // counters of iteration 0
            st.queue[0][st.seg_off[c] + p] = slot;
        }
    }
}

__global__ __launch_bounds__(WG) void k_reached_work(const int *row, const int *dist, int n, int inf,
                                                     unsigned long long *out2) {
    __shared__ unsigned long long sa[WG], sb[WG];
    unsigned long long a = 0, b = 0;
    for (long long v = (long long)blockIdx.x * WG + threadIdx.x; v < n; v += (long long)gridDim.x * WG) {
        if (dist[v] != inf) { a++; b += (unsigned long long)(row[v + 1] - row[v]); }
    }
    sa[threadIdx.x] = a; sb[threadIdx.x] = b;
    __syncthreads();
    for (int o = WG / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) { sa[threadIdx.x] += sa[threadIdx.x + o]; sb[threadIdx.x] += sb[threadIdx.x + o]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { atomicAdd(&out2[0], sa[0]); atomicAdd(&out2[1], sb[0]); }
}

__global__ void k_rebase_row(const int *row, int n, int add, int *out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i <= n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = row[i] + add;
}
__global__ void k_fill_lut(int *lut, long long first, long long count, int inst) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (long long)gridDim.x * blockDim.x)
        lut[first + i] = inst;
}
__global__ void k_copy_u8(const unsigned char *src, unsigned char *dst, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ------------------------------------------------------------------------------------
// scheduler MLP (fp32): forward, per-sample backward, per-parameter reduction, Adam
// ------------------------------------------------------------------------------------
constexpr int H = 32, NO = 5, NI = NP_MLP_NIN;
// per-sample scratch: x copy not needed; h1, h2 (post-ReLU), g1, g2 (pre-activation grads),
// go (output grads), loss
constexpr int PER_SAMPLE = H + H + H + H + NO + 1;

__device__ __forceinline__ void mlp_fwd_sample(const float *p, const float *x, float *a1, float *h1,
                                               float *a2, float *h2, float *o) {
    const float *W1 = p, *b1 = W1 + H * NI, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H,
                *b3 = W3 + NO * H;
    for (int j = 0; j < H; ++j) {
        float a = b1[j];
        for (int i = 0; i < NI; ++i) a += W1[j * NI + i] * x[i];
        a1[j] = a; h1[j] = a > 0.f ? a : 0.f;
    }
    for (int j = 0; j < H; ++j) {
        float a = b2[j];
        for (int i = 0; i < H; ++i) a += W2[j * H + i] * h1[i];
        a2[j] = a; h2[j] = a > 0.f ? a : 0.f;
    }
    for (int j = 0; j < NO; ++j) {
        float a = b3[j];
        for (int i = 0; i < H; ++i) a += W3[j * H + i] * h2[i];
        o[j] = a;
    }
}

__global__ void k_mlp_forward(const float *p, const float *x, int batch, float *out) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= batch) return;
    float a1[H], h1[H], a2[H], h2[H], o[NO];
    mlp_fwd_sample(p, x + (long long)s * NI, a1, h1, a2, h2, o);
    for (int j = 0; j < NO; ++j) out[(long long)s * NO + j] = o[j];
}

// One thread per sample: forward, softmax-CE + lam * squared error, and the gradients of
// the pre-activations (stored for the reduction).
__global__ void k_mlp_sample_grad(const float *p, const float *x, const int *label,
                                  const float *target, int batch, float lam, float *work) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= batch) return;
    const float *W2 = p + H * NI + H, *W3 = W2 + H * H + H;
    float a1[H], h1[H], a2[H], h2[H], o[NO], go[NO];
    mlp_fwd_sample(p, x + (long long)s * NI, a1, h1, a2, h2, o);
    const float inv = 1.0f / (float)batch;
    float mx = fmaxf(o[0], fmaxf(o[1], o[2]));
    float e0 = expf(o[0] - mx), e1 = expf(o[1] - mx), e2 = expf(o[2] - mx);
    float z = e0 + e1 + e2;
    float pr[3] = {e0 / z, e1 / z, e2 / z};
    int lb = label[s];
    float loss = -logf(fmaxf(pr[lb], 1e-30f));
    for (int j = 0; j < 3; ++j) go[j] = (pr[j] - (j == lb ? 1.f : 0.f)) * inv;
    for (int j = 0; j < 2; ++j) {
        float d = o[3 + j] - target[(long long)s * 2 + j];
        loss += lam * d * d;
        go[3 + j] = 2.f * lam * d * inv;
    }
    float *wk = work + (long long)s * PER_SAMPLE;
    float *wh1 = wk, *wh2 = wk + H, *wg1 = wk + 2 * H, *wg2 = wk + 3 * H, *wgo = wk + 4 * H;
    float g2[H];
    for (int i = 0; i < H; ++i) {
        float g = 0.f;
        for (int j = 0; j < NO; ++j) g += W3[j * H + i] * go[j];
        g2[i] = a2[i] > 0.f ? g : 0.f;
    }
    for (int i = 0; i < H; ++i) {
        float g = 0.f;
        for (int j = 0; j < H; ++j) g += W2[j * H + i] * g2[j];
        wg1[i] = a1[i] > 0.f ? g : 0.f;
        wh1[i] = h1[i];
        wh2[i] = h2[i];
        wg2[i] = g2[i];
    }
    for (int j = 0; j < NO; ++j) wgo[j] = go[j];
    wk[4 * H + NO] = loss * inv;
}

// One thread per parameter: sum of the per-sample contributions (deterministic order).
__global__ void k_mlp_reduce(const float *x, int batch, const float *work, float *grad,
                             float *loss) {
    int pidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (pidx > NP_MLP_NPARAM) return;
    if (pidx == NP_MLP_NPARAM) {
        float l = 0.f;
        for (int s = 0; s < batch; ++s) l += work[(long long)s * PER_SAMPLE + 4 * H + NO];
        loss[0] = l;
        return;
    }
    // locate the parameter
    int q = pidx;
    float acc = 0.f;
    if (q < H * NI) {  // W1[j][i] += g1[j] * x[i]
        int j = q / NI, i = q % NI;
        for (int s = 0; s < batch; ++s) acc += work[(long long)s * PER_SAMPLE + 2 * H + j] * x[(long long)s * NI + i];
    } else if ((q -= H * NI) < H) {  // b1
        for (int s = 0; s < batch; ++s) acc += work[(long long)s * PER_SAMPLE + 2 * H + q];
    } else if ((q -= H) < H * H) {  // W2[j][i] += g2[j] * h1[i]
        int j = q / H, i = q % H;
        for (int s = 0; s < batch; ++s) acc += work[(long long)s * PER_SAMPLE + 3 * H + j] * work[(long long)s * PER_SAMPLE + i];
    } else if ((q -= H * H) < H) {  // b2
        for (int s = 0; s < batch; ++s) acc += work[(long long)s * PER_SAMPLE + 3 * H + q];
    } else if ((q -= H) < NO * H) {  // W3[j][i] += go[j] * h2[i]
        int j = q / H, i = q % H;
        for (int s = 0; s < batch; ++s) acc += work[(long long)s * PER_SAMPLE + 4 * H + j] * work[(long long)s * PER_SAMPLE + H + i];
    } else {  // b3
        q -= NO * H;
        for (int s = 0; s < batch; ++s) acc += work[(long long)s * PER_SAMPLE + 4 * H + q];
    }
    grad[pidx] = acc;
}

__global__ void k_adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi; v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------
static int grid_for(long long n, int block) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (int)g;
}

hipError_t launch_fill_i32(int *p, long long n, int v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}
hipError_t launch_fill_u32(unsigned *p, long long n, unsigned v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}
hipError_t launch_validate_csr(const int *row, const int *col, const int *w, int n, long long m, int *out3,
                               hipStream_t s) {
    long long work = (long long)n + 1 > m ? (long long)n + 1 : m;
    hipLaunchKernelGGL(k_validate_csr, dim3(grid_for(work, 256) > 4096 ? 4096 : grid_for(work, 256)),
                       dim3(256), 0, s, row, col, w, n, m, out3);
    return hipGetLastError();
}

hipError_t launch_max_finite(const DevState &st, int *out, hipStream_t s) {
    const long long count = st.nslots << (st.kshift < 0 ? 0 : st.kshift);
    hipError_t e = launch_fill_i32(out, 1, 0, s);
    if (e != hipSuccess || count <= 0) return e;
    hipLaunchKernelGGL(k_max_finite, dim3(grid_for(count, WG) > 4096 ? 4096 : grid_for(count, WG)), dim3(WG),
                       0, s, (const int *)st.dist, count, st.inf, out);
    return hipGetLastError();
}

hipError_t launch_dist_escape(const DevState &st, bool batch, int *bad, hipStream_t s) {
    const long long count = st.nslots << (st.kshift < 0 ? 0 : st.kshift);
    hipError_t e = launch_fill_i32(bad, 1, 0, s);
    if (e != hipSuccess || count <= 0) return e;
    const dim3 grid(grid_for(count, WG) > 4096 ? 4096 : grid_for(count, WG));
    if (batch) hipLaunchKernelGGL(k_dist_escape<true>, grid, dim3(WG), 0, s, st, bad);
    else hipLaunchKernelGGL(k_dist_escape<false>, grid, dim3(WG), 0, s, st, bad);
    return hipGetLastError();
}
hipError_t launch_degree_class(const int *row, int n, unsigned char *cls, int *hist, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_degree_class, dim3(grid_for(n, WG)), dim3(WG), 0, s, row, n, cls, hist);
    return hipGetLastError();
}
hipError_t launch_init_solve(const DevState &st, const int *d_src_slots, int nsrc, int push, int delta,
                             hipStream_t s) {
    const int T0 = (!push && delta > 0) ? delta : NO_WINDOW;
    hipError_t e;
    if ((e = launch_fill_i32(st.dist, st.nslots, st.inf, s)) != hipSuccess) return e;
    if ((e = launch_fill_u32(st.bm[0], st.nwords, 0u, s)) != hipSuccess) return e;
    if ((e = launch_fill_u32(st.bm[1], st.nwords, 0u, s)) != hipSuccess) return e;
    if ((e = launch_fill_i32(st.cnt, 3 * NCLS, 0, s)) != hipSuccess) return e;
    if (st.sbcnt && (e = launch_fill_i32(st.sbcnt, 3, 0, s)) != hipSuccess) return e;
    if ((e = launch_fill_i32((int *)st.ctrl, (long long)(sizeof(Ctrl) / sizeof(int)), 0, s)) != hipSuccess)
        return e;
    if (push) hipLaunchKernelGGL(k_init_sources<true>, dim3(1), dim3(256), 0, s, st, d_src_slots, nsrc, T0);
    else hipLaunchKernelGGL(k_init_sources<false>, dim3(1), dim3(256), 0, s, st, d_src_slots, nsrc, T0);
    return hipGetLastError();
}
hipError_t launch_reached_work(const int *row, const int *dist, int n, int inf,
                               unsigned long long *out2, hipStream_t s) {
    hipError_t e = launch_fill_u32((unsigned *)out2, 4, 0u, s);
    if (e != hipSuccess) return e;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reached_work, dim3(grid_for(n, WG) > 2048 ? 2048 : grid_for(n, WG)), dim3(WG),
                       0, s, row, dist, n, inf, out2);
    return hipGetLastError();
}
hipError_t launch_rebase_row(const int *row, int n, int add, int *out, hipStream_t s) {
    hipLaunchKernelGGL(k_rebase_row, dim3(grid_for((long long)n + 1, 256)), dim3(256), 0, s, row, n, add, out);
    return hipGetLastError();
}
hipError_t launch_fill_lut(int *lut, long long first, long long count, int inst, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_lut, dim3(grid_for(count, 256)), dim3(256), 0, s, lut, first, count, inst);
    return hipGetLastError();
}
hipError_t launch_copy_cls(const unsigned char *src, unsigned char *dst, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_u8, dim3(grid_for(n, 256)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

template <bool B, bool U, bool P>
static const void *persistent_fn() { return (const void *)&k_persistent<B, U, P>; }

static const void *pick_persistent(bool batch, bool unit, bool push) {
    if (batch) {
        if (unit) return push ? persistent_fn<true, true, true>() : persistent_fn<true, true, false>();
        return push ? persistent_fn<true, false, true>() : persistent_fn<true, false, false>();
    }
    if (unit) return push ? persistent_fn<false, true, true>() : persistent_fn<false, true, false>();
    return push ? persistent_fn<false, false, true>() : persistent_fn<false, false, false>();
}

static const void *pick_packed(bool unit) {
    return unit ? (const void *)&k_pk_persistent<true> : (const void *)&k_pk_persistent<false>;
}

int persistent_max_wgs_per_cu(bool batch, bool unit, bool push, bool packed) {
    int nb = 0;
    const void *fn = packed ? pick_packed(unit) : pick_persistent(batch, unit, push);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, WG, 0) != hipSuccess) return 0;
    return nb;
}

hipError_t launch_persistent(const DevState &st, const SchedDev &sp, bool batch, bool unit, int wgs,
                             hipStream_t s) {
    DevState a0 = st;
    SchedDev a1 = sp;
    void *args[] = {&a0, &a1};
    const void *fn = st.kshift >= 0 ? pick_packed(unit) : pick_persistent(batch, unit, sp.push != 0);
    return hipLaunchCooperativeKernel(fn, dim3(wgs), dim3(WG), args, 0, s);
}

hipError_t launch_compact(const DevState &st, int it_off, bool batch, int wgs, hipStream_t s) {
    (void)batch;
    long long need = (st.nwords + WG - 1) / WG;
    int g = (int)(need < wgs ? (need < 1 ? 1 : need) : wgs);
    if (st.kshift >= 0) {
        hipLaunchKernelGGL(k_pk_compact, dim3(g), dim3(WG), 0, s, st, it_off);
        if (st.count_work) hipLaunchKernelGGL(k_pk_count, dim3(wgs), dim3(WG), 0, s, st, it_off);
    } else {
        hipLaunchKernelGGL(k_compact, dim3(g), dim3(WG), 0, s, st, it_off);
    }
    return hipGetLastError();
}
hipError_t launch_init_packed(const DevState &st, const int *d_srcs, int nsrc, int delta, hipStream_t s) {
    if (st.kshift < 0 || nsrc < 1 || nsrc > (1 << st.kshift) || nsrc > PACK_MAX) return hipErrorInvalidValue;
    hipError_t e;
    if ((e = launch_fill_i32(st.dist, st.nslots << st.kshift, st.inf, s)) != hipSuccess) return e;
    if ((e = launch_fill_u32(st.bm[0], st.nwords, 0u, s)) != hipSuccess) return e;
    if (st.hold_cls && (e = launch_fill_u32(st.bm[1], st.nwords, 0u, s)) != hipSuccess) return e;
    if ((e = launch_fill_u32(st.mflag, st.nslots, 0u, s)) != hipSuccess) return e;
    if ((e = launch_fill_i32(st.cnt, 3 * NCLS, 0, s)) != hipSuccess) return e;
    if (st.sbcnt && (e = launch_fill_i32(st.sbcnt, 3, 0, s)) != hipSuccess) return e;
    if ((e = launch_fill_i32((int *)st.ctrl, (long long)(sizeof(Ctrl) / sizeof(int)), 0, s)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_pk_init_sources, dim3(1), dim3(64), 0, s, st, d_srcs, nsrc, delta > 0 ? delta : NO_WINDOW);
    return hipGetLastError();
}
hipError_t launch_unpack(const int *dist, int kshift, int lane, const int *old2new, int n, int *out,
                         hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_unpack, dim3(grid_for(n, 256)), dim3(256), 0, s, dist, kshift, lane, old2new, n, out);
    return hipGetLastError();
}
hipError_t launch_decide(const DevState &st, const SchedDev &sp, int it_off, hipStream_t s) {
    hipLaunchKernelGGL(k_decide, dim3(1), dim3(WG), 0, s, st, sp, it_off);
    return hipGetLastError();
}
// The expansion of the reference's other schedules (single-source, scan frontier).
static hipError_t launch_scheme(const DevState &st, const SchedDev &sp, int it_off, bool unit, int wgs,
                                hipStream_t s) {
    const SolBufs sb{st.sbuf, st.sbcnt};
#define NP_SOL(M) do { if (unit) hipLaunchKernelGGL((k_sol_parent<M, true>), dim3(wgs), dim3(WG), 0, s, st, sb, it_off); \
                       else hipLaunchKernelGGL((k_sol_parent<M, false>), dim3(wgs), dim3(WG), 0, s, st, sb, it_off); } while (0)
    switch (sp.scheme) {
    case NP_SCHEME_THREAD: NP_SOL(SOL_THREAD); break;
    case NP_SCHEME_DUAL_QUEUE:
        NP_SOL(SOL_DUAL);
        if (unit) hipLaunchKernelGGL(k_sol_heavy<true>, dim3(wgs), dim3(WG), 0, s, st, it_off);
        else hipLaunchKernelGGL(k_sol_heavy<false>, dim3(wgs), dim3(WG), 0, s, st, it_off);
        break;
    case NP_SCHEME_LDS_BUFFER: NP_SOL(SOL_LDS); break;
    case NP_SCHEME_GLOBAL_BUFFER:
        NP_SOL(SOL_GLOBAL);
        if (unit) hipLaunchKernelGGL(k_sol_drain<true>, dim3(wgs), dim3(WG), 0, s, st, sb, it_off);
        else hipLaunchKernelGGL(k_sol_drain<false>, dim3(wgs), dim3(WG), 0, s, st, sb, it_off);
        break;
    case NP_SCHEME_WAVE_NESTED: NP_SOL(SOL_WAVE); break;
    case NP_SCHEME_WG_NESTED: NP_SOL(SOL_WGLIST); break;
    case NP_SCHEME_BFS_HIER: hipLaunchKernelGGL(k_bfs_hier, dim3(wgs), dim3(WG), 0, s, st, it_off); break;
    default: return hipErrorInvalidValue;
    }
#undef NP_SOL
    return hipGetLastError();
}

hipError_t launch_bfs_flat(const DevState &st, int level, int *changed, int wgs, hipStream_t s) {
    hipLaunchKernelGGL(k_bfs_flat, dim3(wgs), dim3(WG), 0, s, st.row, st.col, st.dist, (int)st.nslots, level,
                       changed);
    return hipGetLastError();
}

hipError_t launch_expand(const DevState &st, const SchedDev &sp, int it_off, bool batch, bool unit,
                         int wgs, hipStream_t s) {
    if (sp.scheme != NP_SCHEME_CONSOLIDATE) return launch_scheme(st, sp, it_off, unit, wgs, s);
    const bool push = sp.push != 0;
    if (st.kshift >= 0) {
        if (unit) hipLaunchKernelGGL(k_pk_expand<true>, dim3(wgs), dim3(WG), 0, s, st, it_off);
        else hipLaunchKernelGGL(k_pk_expand<false>, dim3(wgs), dim3(WG), 0, s, st, it_off);
        return hipGetLastError();
    }
#define NP_EXP(B, U, P) hipLaunchKernelGGL((k_expand<B, U, P>), dim3(wgs), dim3(WG), 0, s, st, it_off)
    if (batch) {
        if (unit) { if (push) NP_EXP(true, true, true); else NP_EXP(true, true, false); }
        else { if (push) NP_EXP(true, false, true); else NP_EXP(true, false, false); }
    } else {
        if (unit) { if (push) NP_EXP(false, true, true); else NP_EXP(false, true, false); }
        else { if (push) NP_EXP(false, false, true); else NP_EXP(false, false, false); }
    }
#undef NP_EXP
    return hipGetLastError();
}
hipError_t launch_advance(const DevState &st, int k, hipStream_t s) {
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s, st, k);
    return hipGetLastError();
}
hipError_t launch_set_cfg(const DevState &st, int level, int klo, int khi, hipStream_t s) {
    hipLaunchKernelGGL(k_set_cfg, dim3(1), dim3(64), 0, s, st, level, klo, khi);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(const float *p, const float *x, int batch, float *out, hipStream_t s) {
    if (batch <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_forward, dim3((batch + 127) / 128), dim3(128), 0, s, p, x, batch, out);
    return hipGetLastError();
}
size_t mlp_work_floats(int batch) { return (size_t)(batch > 0 ? batch : 1) * PER_SAMPLE; }
hipError_t launch_mlp_loss_grad(const float *p, const float *x, const int *label, const float *target,
                                int batch, float lam, float *work, float *grad, float *loss,
                                hipStream_t s) {
    if (batch <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mlp_sample_grad, dim3((batch + 127) / 128), dim3(128), 0, s, p, x, label,
                       target, batch, lam, work);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NP_MLP_NPARAM + 1 + 127) / 128), dim3(128), 0, s, x, batch,
                       work, grad, loss);
    return hipGetLastError();
}
hipError_t launch_adam(float *p, const float *grad, float *m, float *v, int n, float lr, float b1,
                       float b2, float eps, int step, hipStream_t s) {
    float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, p, grad, m, v, n, lr, b1, b2, eps,
                       c1, c2);
    return hipGetLastError();
}

}  // namespace np
