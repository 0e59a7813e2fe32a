// Betweenness centrality (Brandes) on the frontier machinery, up to 32 sources per batch.
//
// Reference: apps/bc/src/gpu/bc_wrapper.cu:678-761 (bc_np_consolidate_gpu: a level-synchronous
// forward BFS that counts shortest paths, then a backward sweep over the levels) with the
// kernels of bc_fp_kernel.cu:429-489 (forward: a vertex reached at level d+1 adds the path count
// of every level-d predecessor, sigma[w] += sigma[u]) and bc_bp_kernel.cu:4-36 (backward:
// delta[v] += sigma[v] / sigma[w] * (1 + delta[w]) over the edges v -> w into the next level;
// bc[v] += delta[v] for v != source).  Here the factor (1 + delta[w]) / sigma[w] is formed once
// per slot when w's level is final and read as one value by the level above.  The reference
// keeps an n x n predecessor byte matrix; here an edge u -> w is a predecessor edge exactly
// when level[w] = level[u] + 1, which is what that matrix records, so no n^2 state exists.
//
// Layout: the sources of a batch share rows of 2^ks slots, slot (v << ks) + j for source j, as
// the source-packed SSSP does: level (int), sigma and delta (double: path counts stay exact
// integers up to 2^53).  A forward expansion of level d turns the entries (v, mask of the
// sources for which v is at level d) into level d + 1 flags (nmask); the compaction writes the
// next level's entries after the previous ones, so the backward sweep replays every level's
// entries in reverse order.
//
// Work split: entry o has deg(v) x P' items, P' = popc(mask) rounded up to a power of two, item
// r = (edge r / P', source slot r % P').  Items are cut into 64-item chunks (one wavefront each);
// each entry's chunks are reserved as one contiguous range of the level's chunk counter (one
// atomic per wavefront of entries, no prefix sum) and an owner array maps every chunk to its
// entry.  Chunk starts are multiples of 64, so a lane's
// source slot is lane % P' and the lanes of one source are lane, lane ^ P', ...: the backward
// sweep sums them with xor-shuffles before one atomicAdd per source.
#include <hip/hip_runtime.h>
#include "np_internal.h"

namespace np {

__device__ __forceinline__ int bc_lane() { return threadIdx.x & 63; }

// chunks of the entry (v, m): ceil(deg(v) x P' / 64)
__device__ __forceinline__ long long bc_nchunks(const BcDev &b, int v, unsigned m) {
    const int p = __popc(m);
    const int lp = p <= 1 ? 0 : 32 - __clz(p - 1);
    return (((long long)(b.row[v + 1] - b.row[v]) << lp) + 63) >> 6;
}

// Every lane of the wave calls it: k chunks for this lane's entry (0 for none) are reserved on
// *ctot; returns the first chunk of the lane's range.
__device__ __forceinline__ long long bc_reserve(long long k, unsigned long long *ctot) {
    const int lane = bc_lane();
    long long incl = k;
    for (int o = 1; o < 64; o <<= 1) {
        const long long t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    const long long tot = __shfl(incl, 63);
    long long base = 0;
    if (lane == 0 && tot) base = (long long)atomicAdd(ctot, (unsigned long long)tot);
    base = __shfl(base, 0);
    return base + incl - k;
}

// k-th set bit of m (k < popc(m))
__device__ __forceinline__ int bc_nth_bit(unsigned m, int k) {
    for (int i = 0; i < k; ++i) m &= m - 1;
    return __ffs(m) - 1;
}

__global__ void k_bc_reset(BcDev b) {
    const long long slots = (long long)b.n << b.ks;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < slots;
         i += (long long)gridDim.x * blockDim.x) {
        b.lvl[i] = BC_UNSEEN;
        b.sig[i] = 0.0;
        b.del[i] = 0.0;
        if (i < b.n) b.nmask[i] = 0u;
    }
}

__global__ void k_bc_sources(BcDev b, const int *srcs, int nsrc) {
    const int j = threadIdx.x;
    if (blockIdx.x != 0 || j >= nsrc) return;
    const int s = srcs[j];
    const long long k = ((long long)s << b.ks) + j;
    b.lvl[k] = 0;
    b.sig[k] = 1.0;
    atomicOr(&b.nmask[s], 1u << j);
}

// Flagged vertices -> entries (v, mask) at ent_v/ent_m + *eoff + (position counted in *count),
// and the first chunk of each at cst[position] (chunks reserved on *ctot).  Entries of all
// levels fit in n x sources: every (vertex, source) pair is in one level's entry and every
// entry holds at least one.
__global__ void k_bc_compact(BcDev b, int *ent_v, unsigned *ent_m, const int *eoff, long long *cst, int *count,
                             unsigned long long *ctot) {
    const int E = *eoff;
    for (long long base = (long long)blockIdx.x * blockDim.x; base < b.n; base += (long long)gridDim.x * blockDim.x) {
        const long long v = base + threadIdx.x;
        const unsigned m = v < b.n ? b.nmask[v] : 0u;
        const unsigned long long want = __ballot(m != 0);
        if (want == 0) continue;
        const int lane = bc_lane();
        const int leader = __ffsll((long long)want) - 1;
        int pos0 = 0;
        if (lane == leader) pos0 = atomicAdd(count, __popcll(want));
        pos0 = __shfl(pos0, leader);
        const long long first = bc_reserve(m ? bc_nchunks(b, (int)v, m) : 0, ctot);
        if (m) {
            const int pos = pos0 + __popcll(want & ((1ull << lane) - 1));
            b.nmask[v] = 0u;
            ent_v[E + pos] = (int)v;
            ent_m[E + pos] = m;
            cst[pos] = first;
        }
    }
}

// Chunk ranges of ne entries replayed by the backward sweep: first chunk of each at cst[o].
__global__ void k_bc_chunks(BcDev b, const int *ent_v, const unsigned *ent_m, int ne, long long *cst,
                            unsigned long long *ctot) {
    for (long long base = (long long)blockIdx.x * blockDim.x; base < ne; base += (long long)gridDim.x * blockDim.x) {
        const long long o = base + threadIdx.x;  // the trip count is uniform over the workgroup
        const long long first = bc_reserve(o < ne ? bc_nchunks(b, ent_v[o], ent_m[o]) : 0, ctot);
        if (o < ne) cst[o] = first;
    }
}

// owner[c] = o for every chunk c of entry o (entries at ent_v/ent_m + *eoff, eoff null: at
// ent_v; first chunk cst[o]; ne from *ne_dev when given), one wavefront per entry.  A level has
// at most m/2 + ne chunks (deg x 32 / 64 + 1 per entry); a range past cap sets *err and
// nothing is written past it.
__global__ void k_bc_owner(BcDev b, const int *ent_v, const unsigned *ent_m, const int *eoff, const long long *cst,
                           int ne, const int *ne_dev, int *owner, long long cap, int *err) {
    if (ne_dev) ne = *ne_dev;
    if (eoff) {
        ent_v += *eoff;
        ent_m += *eoff;
    }
    const int lane = bc_lane();
    const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
    for (long long o = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); o < ne; o += nw) {
        const long long c0 = cst[o], c1 = c0 + bc_nchunks(b, ent_v[o], ent_m[o]);
        if (c1 > cap && lane == 0) *err = 1;
        const long long ce = c1 < cap ? c1 : cap;
        for (long long c = c0 + lane; c < ce; c += 64) owner[c] = (int)o;
    }
}

// One 64-item chunk per wavefront.  FWD: level d -> d + 1 with path counts; else the
// dependencies of level d from level d + 1.  The level's entries start at ent_v + *eoff (eoff
// null: at ent_v); entry o's chunks start at cst[o]; *ctot is the chunk total.
template <bool FWD>
__global__ __launch_bounds__(256) void k_bc_expand(BcDev b, const int *ent_v, const unsigned *ent_m,
                                                   const int *eoff, const long long *cst, const int *owner,
                                                   const unsigned long long *ctot, long long cap, int d) {
    if (eoff) {
        ent_v += *eoff;
        ent_m += *eoff;
    }
    const int lane = bc_lane();
    const long long tot = (long long)*ctot;
    const long long nchunks = tot < cap ? tot : cap;
    const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
    for (long long c = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); c < nchunks; c += nw) {
        const int o = owner[c];
        const int v = ent_v[o];
        const unsigned m = ent_m[o];
        const int p = __popc(m);
        const int lp = p <= 1 ? 0 : 32 - __clz(p - 1);
        const int e0 = b.row[v];
        const long long items = (long long)(b.row[v + 1] - e0) << lp;
        const long long r = ((c - cst[o]) << 6) + lane;
        const int j = lane & ((1 << lp) - 1);
        const bool live = r < items && j < p;
        const int src = live ? bc_nth_bit(m, j) : 0;
        const long long sv = ((long long)v << b.ks) + src;
        if (FWD) {
            if (live) {
                const int w = b.col[e0 + (int)(r >> lp)];
                const long long sw = ((long long)w << b.ks) + src;
                // Levels up to d were set by earlier launches and are final, so a plain read
                // settles most edges without an atomic: <= d, w was reached before (nothing to
                // do); d + 1, w is on the next level already (only the path count).  Only an
                // unseen w needs the atomicMin; a read racing with this launch's writes sees
                // unseen or d + 1, never a wrong level.
                const int cur = b.lvl[sw];
                if (cur == d + 1) {
                    atomicAdd(&b.sig[sw], b.sig[sv]);
                } else if (cur > d + 1) {
                    const int old = atomicMin(&b.lvl[sw], d + 1);
                    if (old >= d + 1) {  // w is at level d + 1 for this source: v is a predecessor
                        atomicAdd(&b.sig[sw], b.sig[sv]);
                        if (old > d + 1) atomicOr(&b.nmask[w], 1u << src);  // first reached
                    }
                }
            }
        } else {
            double x = 0.0;
            if (live) {
                const int w = b.col[e0 + (int)(r >> lp)];
                const long long sw = ((long long)w << b.ks) + src;
                // level d + 1 is final: its slots hold (1 + delta) / sigma (k_bc_accum)
                if (b.lvl[sw] == d + 1) x = b.sig[sv] * b.sig[sw];
            }
            for (int off = 32; off >= (1 << lp) && off > 0; off >>= 1) x += __shfl_xor(x, off);
            if (lane < p && lane == j && x != 0.0) atomicAdd(&b.del[((long long)v << b.ks) + bc_nth_bit(m, lane)], x);
        }
    }
}

// Once a level's dependencies are final: bc[v] += delta of v for every source of its entry
// (level >= 1: v != source), and the slot's sigma is replaced by (1 + delta) / sigma, the
// factor the level above needs from it (one gather there instead of two).  Each (vertex,
// source) slot is in exactly one level's entries, so it is transformed once, after its own
// level's sweep; the deepest level (delta 0) is passed through here too.
// Half a wavefront per entry, lane j on the j-th source of its mask: the transforms run in
// parallel and the entry's deltas are summed with shuffles.
__global__ void k_bc_accum(BcDev b, const int *ent_v, const unsigned *ent_m, int ne, double *bc) {
    const int j = threadIdx.x & 31;
    const long long per = (long long)gridDim.x * (blockDim.x / 32);
    const long long nn = ((long long)ne + 1) / 2 * 2;  // whole waves: both halves make the same trips
    for (long long o = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / 32; o < nn; o += per) {
        double dl = 0.0;
        int v = 0;
        if (o < ne) {
            v = ent_v[o];
            const unsigned m = ent_m[o];
            if (j < __popc(m)) {
                const long long k = ((long long)v << b.ks) + bc_nth_bit(m, j);
                dl = b.del[k];
                b.sig[k] = (1.0 + dl) / b.sig[k];
            }
        }
        for (int off = 16; off > 0; off >>= 1) dl += __shfl_xor(dl, off);
        if (o < ne && j == 0) bc[v] += dl;  // one entry per vertex per level
    }
}

// After level d's forward expansion (one thread): lv[2d], lv[2d + 1] = the level's first entry
// and entry count; the next level's entries start after them.
__global__ void k_bc_advance(int *eoff, int *count, int *lv, int d) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int E = *eoff, ne = *count;
    lv[2 * d] = E;
    lv[2 * d + 1] = ne;
    *eoff = E + ne;
    *count = 0;
}

__global__ void k_bc_to_float(const double *bc, int n, float *out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = (float)bc[i];
}

static int bc_grid(long long work, int cap) {
    long long g = (work + 255) / 256;
    return (int)(g < 1 ? 1 : (g > cap ? cap : g));
}

hipError_t launch_bc_reset(const BcDev &b, hipStream_t s) {
    hipLaunchKernelGGL(k_bc_reset, dim3(bc_grid((long long)b.n << b.ks, 8192)), dim3(256), 0, s, b);
    return hipGetLastError();
}
hipError_t launch_bc_sources(const BcDev &b, const int *srcs, int nsrc, hipStream_t s) {
    hipLaunchKernelGGL(k_bc_sources, dim3(1), dim3(64), 0, s, b, srcs, nsrc);
    return hipGetLastError();
}
hipError_t launch_bc_compact(const BcDev &b, int *ent_v, unsigned *ent_m, const int *eoff, long long *cst,
                             int *count, unsigned long long *ctot, hipStream_t s) {
    hipLaunchKernelGGL(k_bc_compact, dim3(bc_grid(b.n, 4096)), dim3(256), 0, s, b, ent_v, ent_m, eoff, cst, count,
                       ctot);
    return hipGetLastError();
}
hipError_t launch_bc_advance(int *eoff, int *count, int *lv, int d, hipStream_t s) {
    hipLaunchKernelGGL(k_bc_advance, dim3(1), dim3(64), 0, s, eoff, count, lv, d);
    return hipGetLastError();
}
hipError_t launch_bc_chunks(const BcDev &b, const int *ent_v, const unsigned *ent_m, int ne, long long *cst,
                            unsigned long long *ctot, hipStream_t s) {
    hipLaunchKernelGGL(k_bc_chunks, dim3(bc_grid(ne, 4096)), dim3(256), 0, s, b, ent_v, ent_m, ne, cst, ctot);
    return hipGetLastError();
}
hipError_t launch_bc_owner(const BcDev &b, const int *ent_v, const unsigned *ent_m, const int *eoff,
                           const long long *cst, int ne, const int *ne_dev, int *owner, long long cap, int *err,
                           hipStream_t s) {
    // 4 wavefronts per workgroup, one entry each, striding; with the count on the device ne is
    // only a bound, so the grid is capped lower
    const long long g = ((long long)ne + 3) / 4;
    const long long gmax = ne_dev ? 2048 : 16384;
    const int grid = (int)(g < 1 ? 1 : (g > gmax ? gmax : g));
    hipLaunchKernelGGL(k_bc_owner, dim3(grid), dim3(256), 0, s, b, ent_v, ent_m, eoff, cst, ne, ne_dev, owner, cap,
                       err);
    return hipGetLastError();
}
hipError_t launch_bc_expand(const BcDev &b, bool fwd, const int *ent_v, const unsigned *ent_m, const int *eoff,
                            const long long *cst, const int *owner, const unsigned long long *ctot, long long cap,
                            int d, int num_cus, hipStream_t s) {
    const int grid = num_cus * 16;  // grid-stride over the chunks (their count is on the device)
    if (fwd)
        hipLaunchKernelGGL(k_bc_expand<true>, dim3(grid), dim3(256), 0, s, b, ent_v, ent_m, eoff, cst, owner, ctot,
                           cap, d);
    else
        hipLaunchKernelGGL(k_bc_expand<false>, dim3(grid), dim3(256), 0, s, b, ent_v, ent_m, eoff, cst, owner, ctot,
                           cap, d);
    return hipGetLastError();
}
hipError_t launch_bc_accum(const BcDev &b, const int *ent_v, const unsigned *ent_m, int ne, double *bc,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_bc_accum, dim3(bc_grid((long long)ne * 32, 4096)), dim3(256), 0, s, b, ent_v, ent_m, ne, bc);
    return hipGetLastError();
}
hipError_t launch_bc_to_float(const double *bc, int n, float *out, hipStream_t s) {
    hipLaunchKernelGGL(k_bc_to_float, dim3(bc_grid(n, 4096)), dim3(256), 0, s, bc, n, out);
    return hipGetLastError();
}

}  // namespace np
