// PageRank of the reference's PAGERANK_GPU (apps/pagerank/src/gpu/pagerank_wrapper.cu:147-244,
// 565-661; kernels pagerank_kernel.cu:60-107; constants pagerank_wrapper.cu:5-6 and pagerank.h:15).
//
// The reference pulls over the inverted graph (invertGraph, common/global.cpp:237-276): per
// iteration new[v] = damping * sum over in-edges p -> v of rank[p] / outdeg[p]
//                    + (1 - damping) * TOTAL / n + damping * (rank held by dangling vertices) / n,
// from rank = TOTAL / n, until no vertex changes by more than EPSILON (damping 0.5, TOTAL 1000,
// EPSILON 1e-13).  Parallel edges count once per edge in both the sum and outdeg.
//
// MI355X mapping: the in-edge CSR is built once on the device (count, rocPRIM scan, scatter;
// one wavefront per vertex); each iteration is two kernels: contributions rank/outdeg with the
// dangling sum (workgroup reduction, one float64 atomic per workgroup), then the pull with the
// largest change (the same reduction, an integer atomicMax on the non-negative float64 bit
// pattern).  The pull splits vertices by in-degree: fewer than 32 in-edges, 8 lanes per vertex
// (consecutive in-edges, coalesced); more, one wavefront per vertex (heavy list built once); both
// reduce with shuffles.  All in
// float64, so the reference's EPSILON is reachable (float ranks near 1000/n cannot move by
// less than an ulp ~1e-9); ranks are returned as float32, the reference's FLOAT_T.  The file also
// holds SpMV (apps/SpMV), the same row-split product without PageRank's terms.
#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_scan.hpp>
#include "np_internal.h"

namespace np {

__device__ __forceinline__ int pr_lane() { return threadIdx.x & 63; }

static int pr_grid(long long work, int cap) {
    long long g = (work + 255) / 256;
    return (int)(g < 1 ? 1 : (g > cap ? cap : g));
}

// in-degree counts (one wavefront per vertex over its out-edges)
__global__ void k_pr_count(const int *row, const int *col, int n, long long *icnt) {
    const int lane = pr_lane();
    const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
    for (long long v = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); v < n; v += nw)
        for (int e = row[v] + lane; e < row[v + 1]; e += 64) atomicAdd((unsigned long long *)&icnt[col[e]], 1ull);
}

// in-edge lists: pcol[irow[u] + k] = v for every edge v -> u
__global__ void k_pr_fill(const int *row, const int *col, int n, const long long *irow, int *cursor, int *pcol) {
    const int lane = pr_lane();
    const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
    for (long long v = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); v < n; v += nw)
        for (int e = row[v] + lane; e < row[v + 1]; e += 64) {
            const int u = col[e];
            pcol[irow[u] + atomicAdd(&cursor[u], 1)] = (int)v;
        }
}

// vertices with at least PR_HEAVY in-edges (wave-aggregated append)
constexpr int PR_HEAVY = 32;
__global__ void k_pr_heavy(const long long *irow, int n, int *heavy, int *nheavy) {
    for (long long base = (long long)blockIdx.x * blockDim.x; base < n; base += (long long)gridDim.x * blockDim.x) {
        const long long v = base + threadIdx.x;
        const bool h = v < n && irow[v + 1] - irow[v] >= PR_HEAVY;
        const unsigned long long want = __ballot(h);
        if (want == 0) continue;
        const int lane = pr_lane();
        const int leader = __ffsll((long long)want) - 1;
        int b = 0;
        if (lane == leader) b = atomicAdd(nheavy, __popcll(want));
        b = __shfl(b, leader);
        if (h) heavy[b + __popcll(want & ((1ull << lane) - 1))] = (int)v;
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

// contrib[p] = rank[p] / outdeg[p]; dangling vertices add their rank to *dsum instead
__global__ __launch_bounds__(256) void k_pr_contrib(const int *row, const double *rank, int n, double *contrib,
                                                     double *dsum) {
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
constexpr int PR_GROUP = 8;
__global__ __launch_bounds__(256) void k_pr_pull_light(const long long *irow, const int *pcol, const double *contrib,
                                                        const double *rank, double *nrank, int n, double damping,
                                                        const double *dsum, double rw, unsigned long long *dmax) {
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

// heavy vertices: one wavefront each, lanes over the in-edges
__global__ __launch_bounds__(256) void k_pr_pull_heavy(const int *heavy, const int *nheavy, const long long *irow,
                                                        const int *pcol, const double *contrib, const double *rank,
                                                        double *nrank, int n, double damping, const double *dsum,
                                                        double rw, unsigned long long *dmax) {
    const double base_add = rw + damping * (*dsum) / (double)n;
    const int lane = pr_lane();
    const int nh = *nheavy;
    const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
    double mx = 0.0;
    for (long long o = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); o < nh; o += nw) {
        const int v = heavy[o];
        double s = 0.0;
        for (long long i = irow[v] + lane; i < irow[v + 1]; i += 64) s += contrib[pcol[i]];
        for (int k = 32; k > 0; k >>= 1) s += __shfl_xor(s, k);
        if (lane == 0) {
            const double r = damping * s + base_add;
            nrank[v] = r;
            const double dd = fabs(r - rank[v]);
            mx = dd > mx ? dd : mx;
        }
    }
    pr_block_max(mx, dmax);
}

// ---- SpMV (apps/SpMV/src/gpu/SpMV_kernel.cu:77-92, csr_spmv_kernel): y[v] = sum over the
// row's entries of data[e] * x[col[e]], the same light/heavy split as the pull (rows with fewer
// than PR_HEAVY entries PR_GROUP lanes each, longer rows one wavefront each).
__global__ void k_spmv_heavy_list(const int *row, int n, int *heavy, int *nheavy) {
    for (long long base = (long long)blockIdx.x * blockDim.x; base < n; base += (long long)gridDim.x * blockDim.x) {
        const long long v = base + threadIdx.x;
        const bool h = v < n && row[v + 1] - row[v] >= PR_HEAVY;
        const unsigned long long want = __ballot(h);
        if (want == 0) continue;
        const int lane = pr_lane();
        const int leader = __ffsll((long long)want) - 1;
        int b = 0;
        if (lane == leader) b = atomicAdd(nheavy, __popcll(want));
        b = __shfl(b, leader);
        if (h) heavy[b + __popcll(want & ((1ull << lane) - 1))] = (int)v;
    }
}
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
__global__ __launch_bounds__(256) void k_spmv_heavy(const int *heavy, const int *nheavy, const int *row, const int *col,
                                                     const double *data, const double *x, double *y) {
    const int lane = pr_lane();
    const int nh = *nheavy;
    const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
    for (long long o = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); o < nh; o += nw) {
        const int v = heavy[o];
        double s = 0.0;
        for (int e = row[v] + lane; e < row[v + 1]; e += 64) s += data[e] * x[col[e]];
        for (int k = 32; k > 0; k >>= 1) s += __shfl_xor(s, k);
        if (lane == 0) y[v] = s;
    }
}

__global__ void k_pr_to_float(const double *rank, int n, float *out) {
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x)
        out[v] = (float)rank[v];
}

hipError_t launch_pr_count(const int *row, const int *col, int n, long long *icnt, hipStream_t s) {
    hipLaunchKernelGGL(k_pr_count, dim3(pr_grid((long long)n * 64, 8192)), dim3(256), 0, s, row, col, n, icnt);
    return hipGetLastError();
}
hipError_t launch_pr_fill(const int *row, const int *col, int n, const long long *irow, int *cursor, int *pcol,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_pr_fill, dim3(pr_grid((long long)n * 64, 8192)), dim3(256), 0, s, row, col, n, irow, cursor,
                       pcol);
    return hipGetLastError();
}
hipError_t launch_pr_heavy(const long long *irow, int n, int *heavy, int *nheavy, hipStream_t s) {
    hipLaunchKernelGGL(k_pr_heavy, dim3(pr_grid(n, 4096)), dim3(256), 0, s, irow, n, heavy, nheavy);
    return hipGetLastError();
}
hipError_t launch_pr_init(double *rank, int n, double r0, hipStream_t s) {
    hipLaunchKernelGGL(k_pr_init, dim3(pr_grid(n, 4096)), dim3(256), 0, s, rank, n, r0);
    return hipGetLastError();
}
hipError_t launch_pr_iteration(const int *row, const long long *irow, const int *pcol, const int *heavy,
                               const int *nheavy, const double *rank, double *nrank, double *contrib, int n,
                               double damping, double rw, double *dsum, unsigned long long *dmax, int num_cus,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_pr_contrib, dim3(pr_grid(n, num_cus * 4)), dim3(256), 0, s, row, rank, n, contrib, dsum);
    hipLaunchKernelGGL(k_pr_pull_light, dim3(pr_grid((long long)n * PR_GROUP, num_cus * 16)), dim3(256), 0, s, irow,
                       pcol, contrib, rank,
                       nrank, n, damping, dsum, rw, dmax);
    hipLaunchKernelGGL(k_pr_pull_heavy, dim3(num_cus * 4), dim3(256), 0, s, heavy, nheavy, irow, pcol, contrib, rank,
                       nrank, n, damping, dsum, rw, dmax);
    return hipGetLastError();
}
hipError_t launch_spmv_heavy_list(const int *row, int n, int *heavy, int *nheavy, hipStream_t s) {
    hipLaunchKernelGGL(k_spmv_heavy_list, dim3(pr_grid(n, 4096)), dim3(256), 0, s, row, n, heavy, nheavy);
    return hipGetLastError();
}
hipError_t launch_spmv(const int *row, const int *col, const double *data, const double *x, double *y, int n,
                       const int *heavy, const int *nheavy, int num_cus, hipStream_t s) {
    hipLaunchKernelGGL(k_spmv_light, dim3(pr_grid((long long)n * PR_GROUP, num_cus * 16)), dim3(256), 0, s, row, col,
                       data, x, y, n);
    hipLaunchKernelGGL(k_spmv_heavy, dim3(num_cus * 4), dim3(256), 0, s, heavy, nheavy, row, col, data, x, y);
    return hipGetLastError();
}
hipError_t launch_pr_to_float(const double *rank, int n, float *out, hipStream_t s) {
    hipLaunchKernelGGL(k_pr_to_float, dim3(pr_grid(n, 4096)), dim3(256), 0, s, rank, n, out);
    return hipGetLastError();
}
hipError_t pr_scan(void *temp, size_t *temp_bytes, const long long *in, long long *out, int n, hipStream_t s) {
    return rocprim::exclusive_scan(temp, *temp_bytes, in, out, 0ll, (size_t)n + 1, rocprim::plus<long long>(), s);
}

}  // namespace np
