// PageRank of the reference's PAGERANK_GPU (apps/pagerank/src/gpu/pagerank_wrapper.cu:147-244,
// 565-661; kernels pagerank_kernel.cu:60-107; constants pagerank_wrapper.cu:5-6 and pagerank.h:15).
//
// The reference pulls over the inverted graph (invertGraph, common/global.cpp:237-276): per
// iteration new[v] = damping * sum over in-edges p -> v of rank[p] / outdeg[p]
//                    + (1 - damping) * TOTAL / n + damping * (rank held by dangling vertices) / n,
// from rank = TOTAL / n, until no vertex changes by more than EPSILON (damping 0.5, TOTAL 1000,
// EPSILON 1e-13).  Parallel edges count once per edge in both the sum and outdeg.
//
// MI355X mapping: the in-edge CSR is built once on the device (a rocPRIM radix sort of the
// (target, source) pairs, offsets from the sorted targets); each iteration computes the
// contributions rank/outdeg with the dangling sum (workgroup reduction, one float64 atomic per
// workgroup), then the pull with the
// largest change (the same reduction, an integer atomicMax on the non-negative float64 bit
// pattern).  The pull splits vertices by in-degree (bins built once): fewer than 32 in-edges,
// 8 lanes per vertex (consecutive in-edges, coalesced); up to PR_HUGE, one wavefront per vertex;
// longer (R-MAT hubs reach 10^5 in-edges, a single wavefront on one would be the whole pass's
// tail), PR_CHUNK-entry chunks on one wavefront each, whose partial sums a last kernel adds in
// chunk order, so results do not depend on scheduling.  All in
// float64, so the reference's EPSILON is reachable (float ranks near 1000/n cannot move by
// less than an ulp ~1e-9); ranks are returned as float32, the reference's FLOAT_T.  The file also
// holds SpMV (apps/SpMV), the same row-split product without PageRank's terms.
#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include "np_internal.h"

namespace np {

__device__ __forceinline__ int pr_lane() { return threadIdx.x & 63; }

static int pr_grid(long long work, int cap) {
    long long g = (work + 255) / 256;
    return (int)(g < 1 ? 1 : (g > cap ? cap : g));
}

// Sum over entries [e0, e1) of term(e), lanes striding by 64 with four loads in flight per lane,
// reduced over the wavefront (every lane returns the total).
template <class IDX, class F>
__device__ __forceinline__ double pr_wave_sum(IDX e0, IDX e1, F term) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    IDX e = e0 + pr_lane();
    for (; e + 192 < e1; e += 256) {
        s0 += term(e);
        s1 += term(e + 64);
        s2 += term(e + 128);
        s3 += term(e + 192);
    }
    for (; e < e1; e += 64) s0 += term(e);
    double s = (s0 + s1) + (s2 + s3);
    for (int k = 32; k > 0; k >>= 1) s += __shfl_xor(s, k);
    return s;
}

// The in-edge CSR is the transpose: a stable radix sort of the (target, source) pairs puts every
// vertex's in-edges together, sources ascending (the order the reference's invertGraph lists
// them, common/global.cpp:237-276), with no per-edge atomics (hub targets would serialise them).
// src[e] = the source of edge e: each vertex with out-edges writes its id at its first edge
// (heads, zero elsewhere), and an inclusive max-scan over the edges carries it to the rest of
// its edges (ids grow with the position).  One wavefront per vertex left most lanes idle on
// R-MAT's short rows (0.2 ms at scale 20).
__global__ void k_pr_heads(const int *row, int n, int *heads) {
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x)
        if (row[v + 1] > row[v]) heads[row[v]] = (int)v;
}

// irow[u] = the first sorted position with target >= u, for u = 0..n (item i = m closes the
// tail: irow[key[m-1]+1 .. n] = m).
__global__ void k_pr_offsets(const unsigned *key, long long m, int n, long long *irow) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i <= m; i += (long long)gridDim.x * blockDim.x) {
        const long long prev = i > 0 ? (long long)key[i - 1] : -1;
        const long long k = i < m ? (long long)key[i] : n;
        for (long long u = prev + 1; u <= k; ++u) irow[u] = i;
    }
}

// Row bins (rows: a vertex's in-edges for PageRank, a matrix row for SpMV), built once.  Rows
// of PR_HEAVY..PR_HUGE-1 entries go to med[]; longer rows to huge[], with their chunks (row,
// chunk index) in crow/ck from hc0[h] on: ceil(len / PR_CHUNK) <= 1.5 len / PR_CHUNK chunks per
// huge row, so at most 1.5 m / PR_CHUNK in all (pr_chunk_cap); the cap check never fires.
// Each workgroup takes a contiguous range of rows: it counts its medium rows, reserves their
// slots with one global atomic (one per wave or per row would serialise on the counter), then
// writes them at offsets taken from an LDS counter.  Huge rows are rare: plain atomics.
template <class OFF>
__global__ __launch_bounds__(256) void k_pr_bins(const OFF *off, int n, RowBins rb, long long chunk_cap) {
    __shared__ int cnt, base;
    const long long per = ((long long)n + gridDim.x - 1) / gridDim.x;
    const long long v0 = (long long)blockIdx.x * per;
    const long long v1 = v0 + per < n ? v0 + per : n;
    const int lane = pr_lane();
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int mine = 0;
    for (long long v = v0 + threadIdx.x; v < v1; v += blockDim.x) {
        const long long len = (long long)(off[v + 1] - off[v]);
        mine += len >= PR_HEAVY && len < PR_HUGE;
        if (len >= PR_HUGE) {
            const int k = (int)((len + PR_CHUNK - 1) / PR_CHUNK);
            const int h = atomicAdd(rb.nhuge, 1);
            const int c0 = atomicAdd(rb.nchunk, k);
            rb.huge[h] = (int)v;
            rb.hc0[h] = c0;
            if (c0 + (long long)k <= chunk_cap)
                for (int j = 0; j < k; ++j) { rb.crow[c0 + j] = (int)v; rb.ck[c0 + j] = j; }
        }
    }
    for (int k = 32; k > 0; k >>= 1) mine += __shfl_xor(mine, k);
    if (lane == 0 && mine) atomicAdd(&cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0) {
        base = cnt ? atomicAdd(rb.nmed, cnt) : 0;
        cnt = 0;
    }
    __syncthreads();
    for (long long t = v0; t < v1; t += blockDim.x) {  // trip count uniform over the workgroup
        const long long v = t + threadIdx.x;
        const long long len = v < v1 ? (long long)(off[v + 1] - off[v]) : 0;
        const bool take = len >= PR_HEAVY && len < PR_HUGE;
        const unsigned long long want = __ballot(take);
        if (want == 0) continue;
        const int leader = __ffsll((long long)want) - 1;
        int b = 0;
        if (lane == leader) b = atomicAdd(&cnt, __popcll(want));
        b = __shfl(b, leader);
        if (take) rb.med[base + b + __popcll(want & ((1ull << lane) - 1))] = (int)v;
    }
}

__global__ void k_pr_init(double *rank, int n, double r0) {
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x)
        rank[v] = r0;
}

// workgroup sum of one double per thread -> one atomicAdd (all threads call it)
__device__ __forceinline__ void pr_block_add(double x, double *dst) {
    __shared__ double s[4];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if (pr_lane() == 0) s[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double t = s[0] + s[1] + s[2] + s[3];
        if (t != 0.0) atomicAdd(dst, t);
    }
    __syncthreads();
}

// workgroup max of one non-negative double per thread -> one atomicMax on its bit pattern
__device__ __forceinline__ void pr_block_max(double x, unsigned long long *dst) {
    __shared__ double s[4];
    for (int o = 32; o > 0; o >>= 1) { const double y = __shfl_xor(x, o); x = y > x ? y : x; }
    if (pr_lane() == 0) s[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = s[0];
        for (int k = 1; k < 4; ++k) t = s[k] > t ? s[k] : t;
        if (t > 0.0) atomicMax(dst, (unsigned long long)__double_as_longlong(t));
    }
    __syncthreads();
}

// Start of iteration i (one thread): ctl[0] = done, ctl[1] = iterations run.  Iteration i - 1's
// largest change (*dmax_prev) at most eps ends the run, as the reference's loop does after that
// iteration; otherwise iteration i is counted and its sums (*dsum, *dmax) start at zero.  The
// host launches iterations in batches and reads ctl once per batch; iterations past the end do
// nothing.
__global__ void k_pr_gate(int i, int *ctl, const unsigned long long *dmax_prev, double *dsum,
                          unsigned long long *dmax, double eps) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || ctl[0]) return;
    if (i > 0 && !(__longlong_as_double((long long)*dmax_prev) > eps)) { ctl[0] = 1; return; }
    ctl[1] = i + 1;
    *dsum = 0.0;
    *dmax = 0;
}

// contrib[p] = rank[p] / outdeg[p]; dangling vertices add their rank to *dsum instead
__global__ __launch_bounds__(256) void k_pr_contrib(const int *done, const int *row, const double *rank, int n,
                                                     double *contrib, double *dsum) {
    if (*done) return;  // uniform over the workgroup, before any barrier
    double dang = 0.0;
    for (long long base = (long long)blockIdx.x * blockDim.x; base < n; base += (long long)gridDim.x * blockDim.x) {
        const long long v = base + threadIdx.x;
        if (v < n) {
            const int d = row[v + 1] - row[v];
            if (d > 0) contrib[v] = rank[v] / (double)d;
            else { contrib[v] = 0.0; dang += rank[v]; }
        }
    }
    pr_block_add(dang, dsum);
}

// light vertices (fewer than PR_HEAVY in-edges): PR_GROUP lanes each, reading consecutive
// in-edges (coalesced), summed with shuffles.  base_add = the random-walk and dangling terms of
// this iteration; *dmax receives the largest |new - old|.
__global__ __launch_bounds__(256) void k_pr_pull_light(const int *done, const long long *irow, const int *pcol,
                                                        const double *contrib,
                                                        const double *rank, double *nrank, int n, double damping,
                                                        const double *dsum, double rw, unsigned long long *dmax) {
    if (*done) return;  // uniform over the workgroup, before any barrier
    const double base_add = rw + damping * (*dsum) / (double)n;
    const int sub = threadIdx.x & (PR_GROUP - 1);
    const long long groups = (long long)gridDim.x * (blockDim.x / PR_GROUP);
    double mx = 0.0;
    // every lane of a wave runs the same number of trips (the bound is rounded to whole waves)
    const long long nv = ((long long)n + 63) / 64 * 64;
    for (long long v = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / PR_GROUP; v < nv; v += groups) {
        const bool light = v < n && irow[v + 1] - irow[v] < PR_HEAVY;
        double s = 0.0;
        if (light)
            for (long long i = irow[v] + sub; i < irow[v + 1]; i += PR_GROUP) s += contrib[pcol[i]];
        for (int o = PR_GROUP / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (light && sub == 0) {
            const double r = damping * s + base_add;
            nrank[v] = r;
            const double dd = fabs(r - rank[v]);
            mx = dd > mx ? dd : mx;
        }
    }
    pr_block_max(mx, dmax);
}

// chunks of huge vertices (items 0..nchunk-1: partial sums into part[]) and then medium
// vertices (one wavefront each, new rank written here), one wavefront per item
__global__ __launch_bounds__(256) void k_pr_pull_wave(const int *done, RowBins rb, const long long *irow, const int *pcol,
                                                       const double *contrib, const double *rank, double *nrank,
                                                       int n, double damping, const double *dsum, double rw,
                                                       unsigned long long *dmax) {
    if (*done) return;  // uniform over the workgroup, before any barrier
    const double base_add = rw + damping * (*dsum) / (double)n;
    const int nc = *rb.nchunk, nm = *rb.nmed;
    const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
    auto term = [&](long long i) { return contrib[pcol[i]]; };
    double mx = 0.0;
    for (long long o = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); o < (long long)nc + nm; o += nw) {
        if (o < nc) {
            const int v = rb.crow[o];
            const long long i0 = irow[v] + (long long)rb.ck[o] * PR_CHUNK;
            const long long i1 = i0 + PR_CHUNK < irow[v + 1] ? i0 + PR_CHUNK : irow[v + 1];
            const double s = pr_wave_sum(i0, i1, term);
            if (pr_lane() == 0) rb.part[o] = s;
        } else {
            const int v = rb.med[o - nc];
            const double s = pr_wave_sum(irow[v], irow[v + 1], term);
            if (pr_lane() == 0) {
                const double r = damping * s + base_add;
                nrank[v] = r;
                const double dd = fabs(r - rank[v]);
                mx = dd > mx ? dd : mx;
            }
        }
    }
    pr_block_max(mx, dmax);
}

// huge vertices: their chunks' partial sums added in chunk order, then the new rank
__global__ __launch_bounds__(256) void k_pr_pull_huge(const int *done, RowBins rb, const long long *irow, const double *rank,
                                                       double *nrank, int n, double damping, const double *dsum,
                                                       double rw, unsigned long long *dmax) {
    if (*done) return;  // uniform over the workgroup, before any barrier
    const double base_add = rw + damping * (*dsum) / (double)n;
    const int nh = *rb.nhuge;
    double mx = 0.0;
    for (long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x; h < nh; h += (long long)gridDim.x * blockDim.x) {
        const int v = rb.huge[h];
        const int k = (int)((irow[v + 1] - irow[v] + PR_CHUNK - 1) / PR_CHUNK);
        double s = 0.0;
        for (int j = 0; j < k; ++j) s += rb.part[rb.hc0[h] + j];
        const double r = damping * s + base_add;
        nrank[v] = r;
        const double dd = fabs(r - rank[v]);
        mx = dd > mx ? dd : mx;
    }
    pr_block_max(mx, dmax);
}

// ---- SpMV (apps/SpMV/src/gpu/SpMV_kernel.cu:77-92, csr_spmv_kernel): y[v] = sum over the
// row's entries of data[e] * x[col[e]], with the pull's row bins.
__global__ __launch_bounds__(256) void k_spmv_light(const int *row, const int *col, const double *data,
                                                     const double *x, double *y, int n) {
    const int sub = threadIdx.x & (PR_GROUP - 1);
    const long long groups = (long long)gridDim.x * (blockDim.x / PR_GROUP);
    const long long nv = ((long long)n + 63) / 64 * 64;  // whole waves: same trips in every lane
    for (long long v = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / PR_GROUP; v < nv; v += groups) {
        const bool light = v < n && row[v + 1] - row[v] < PR_HEAVY;
        double s = 0.0;
        if (light)
            for (int e = row[v] + sub; e < row[v + 1]; e += PR_GROUP) s += data[e] * x[col[e]];
        for (int o = PR_GROUP / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (light && sub == 0) y[v] = s;
    }
}
__global__ __launch_bounds__(256) void k_spmv_wave(RowBins rb, const int *row, const int *col, const double *data,
                                                    const double *x, double *y) {
    const int nc = *rb.nchunk, nm = *rb.nmed;
    const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
    auto term = [&](int e) { return data[e] * x[col[e]]; };
    for (long long o = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); o < (long long)nc + nm; o += nw) {
        if (o < nc) {
            const int v = rb.crow[o];
            const int e0 = row[v] + rb.ck[o] * PR_CHUNK;
            const int e1 = e0 + PR_CHUNK < row[v + 1] ? e0 + PR_CHUNK : row[v + 1];
            const double s = pr_wave_sum(e0, e1, term);
            if (pr_lane() == 0) rb.part[o] = s;
        } else {
            const int v = rb.med[o - nc];
            const double s = pr_wave_sum(row[v], row[v + 1], term);
            if (pr_lane() == 0) y[v] = s;
        }
    }
}
__global__ void k_spmv_huge(RowBins rb, const int *row, double *y) {
    const int nh = *rb.nhuge;
    for (long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x; h < nh; h += (long long)gridDim.x * blockDim.x) {
        const int v = rb.huge[h];
        const int k = (row[v + 1] - row[v] + PR_CHUNK - 1) / PR_CHUNK;
        double s = 0.0;
        for (int j = 0; j < k; ++j) s += rb.part[rb.hc0[h] + j];
        y[v] = s;
    }
}

__global__ void k_pr_to_float(const double *rank, int n, float *out) {
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x)
        out[v] = (float)rank[v];
}

hipError_t pr_sources(void *temp, size_t *temp_bytes, const int *row, int n, long long m, int *heads, int *src,
                      hipStream_t s) {
    if (!temp)
        return rocprim::inclusive_scan(nullptr, *temp_bytes, heads, src, (size_t)m, rocprim::maximum<int>(), s);
    hipError_t e = hipMemsetAsync(heads, 0, sizeof(int) * (size_t)m, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pr_heads, dim3(pr_grid(n, 4096)), dim3(256), 0, s, row, n, heads);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return rocprim::inclusive_scan(temp, *temp_bytes, heads, src, (size_t)m, rocprim::maximum<int>(), s);
}
hipError_t launch_pr_offsets(const unsigned *key, long long m, int n, long long *irow, hipStream_t s) {
    hipLaunchKernelGGL(k_pr_offsets, dim3(pr_grid(m + 1, 8192)), dim3(256), 0, s, key, m, n, irow);
    return hipGetLastError();
}
long long pr_chunk_cap(long long m) { return 3 * m / (2 * PR_CHUNK) + 1; }
hipError_t launch_pr_bins(const long long *irow, int n, long long m, const RowBins &rb, hipStream_t s) {
    hipLaunchKernelGGL(k_pr_bins<long long>, dim3(pr_grid(n, 2048)), dim3(256), 0, s, irow, n, rb, pr_chunk_cap(m));
    return hipGetLastError();
}
hipError_t launch_spmv_bins(const int *row, int n, long long m, const RowBins &rb, hipStream_t s) {
    hipLaunchKernelGGL(k_pr_bins<int>, dim3(pr_grid(n, 2048)), dim3(256), 0, s, row, n, rb, pr_chunk_cap(m));
    return hipGetLastError();
}
hipError_t launch_pr_init(double *rank, int n, double r0, hipStream_t s) {
    hipLaunchKernelGGL(k_pr_init, dim3(pr_grid(n, 4096)), dim3(256), 0, s, rank, n, r0);
    return hipGetLastError();
}
hipError_t launch_pr_iteration(int i, int *ctl, const int *row, const long long *irow, const int *pcol,
                               const RowBins &rb, const double *rank, double *nrank, double *contrib, int n,
                               double damping, double rw, double eps, double *dsum2, unsigned long long *dmax2,
                               int num_cus, hipStream_t s) {
    double *dsum = dsum2 + (i & 1);
    unsigned long long *dmax = dmax2 + (i & 1);
    hipLaunchKernelGGL(k_pr_gate, dim3(1), dim3(64), 0, s, i, ctl, dmax2 + ((i + 1) & 1), dsum, dmax, eps);
    hipLaunchKernelGGL(k_pr_contrib, dim3(pr_grid(n, num_cus * 4)), dim3(256), 0, s, ctl, row, rank, n, contrib, dsum);
    hipLaunchKernelGGL(k_pr_pull_light, dim3(pr_grid((long long)n * PR_GROUP, num_cus * 16)), dim3(256), 0, s, ctl,
                       irow, pcol, contrib, rank, nrank, n, damping, dsum, rw, dmax);
    hipLaunchKernelGGL(k_pr_pull_wave, dim3(num_cus * 8), dim3(256), 0, s, ctl, rb, irow, pcol, contrib, rank, nrank,
                       n, damping, dsum, rw, dmax);
    hipLaunchKernelGGL(k_pr_pull_huge, dim3(4), dim3(256), 0, s, ctl, rb, irow, rank, nrank, n, damping, dsum, rw,
                       dmax);
    return hipGetLastError();
}
hipError_t launch_spmv(const int *row, const int *col, const double *data, const double *x, double *y, int n,
                       const RowBins &rb, int num_cus, hipStream_t s) {
    hipLaunchKernelGGL(k_spmv_light, dim3(pr_grid((long long)n * PR_GROUP, num_cus * 16)), dim3(256), 0, s, row, col,
                       data, x, y, n);
    hipLaunchKernelGGL(k_spmv_wave, dim3(num_cus * 8), dim3(256), 0, s, rb, row, col, data, x, y);
    hipLaunchKernelGGL(k_spmv_huge, dim3(4), dim3(256), 0, s, rb, row, y);
    return hipGetLastError();
}
hipError_t launch_pr_to_float(const double *rank, int n, float *out, hipStream_t s) {
    hipLaunchKernelGGL(k_pr_to_float, dim3(pr_grid(n, 4096)), dim3(256), 0, s, rank, n, out);
    return hipGetLastError();
}
// Stable sort of m (key, value) pairs on key bits [0, end_bit).
hipError_t pr_sort(void *temp, size_t *temp_bytes, const unsigned *key_in, unsigned *key_out, const int *val_in,
                   int *val_out, long long m, int end_bit, hipStream_t s) {
    return rocprim::radix_sort_pairs(temp, *temp_bytes, key_in, key_out, val_in, val_out, (size_t)m, 0u,
                                     (unsigned)end_bit, s);
}

}  // namespace np
