// np_gen.hip -- device graph generators (R-MAT, Erdos-Renyi, 2-D grid) and CSR
// construction.  The sampling functions are shared with the host generators
// (np_common.h), so the device and host CSR of one (generator, seed) are identical.
//
// CSR order: edges are grouped by source with a stable radix sort of the edge list fed
// in reverse generation order, so each vertex lists its out-edges in reverse generation
// order -- what convertCSR (common/global.cpp:216-221) makes of a file read in that order.
#include <hip/hip_runtime.h>
#include <cstring>
#include <vector>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include "np_internal.h"

namespace np {

static int grid_for(long long n, int block) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (int)g;
}

void rmat_thresholds(double a, double b, double c, uint64_t *ta, uint64_t *tab, uint64_t *tabc) {
    const double S = 4294967296.0;
    double x = a * S, y = (a + b) * S, z = (a + b + c) * S;
    *ta = x >= S ? (uint64_t)S : (uint64_t)x;
    *tab = y >= S ? (uint64_t)S : (uint64_t)y;
    *tabc = z >= S ? (uint64_t)S : (uint64_t)z;
}

__global__ void k_gen_rmat(int scale, long long m, uint64_t key, uint64_t wkey, uint64_t pkey,
                           uint64_t ta, uint64_t tab, uint64_t tabc, int permute, int wmin,
                           uint32_t wrange, uint32_t *src, uint32_t *dst, int *w) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (long long)gridDim.x * blockDim.x) {
        uint32_t s, t;
        rmat_edge(key, (uint64_t)e, scale, ta, tab, tabc, &s, &t);
        if (permute) { s = permute_id(s, scale, pkey); t = permute_id(t, scale, pkey); }
        src[e] = s;
        dst[e] = t;
        w[e] = wmin + (int)uniform_below((uint32_t)draw(wkey, (uint64_t)e), wrange);
    }
}

__global__ void k_gen_er(uint32_t n, long long m, uint64_t key, uint64_t wkey, int wmin,
                         uint32_t wrange, uint32_t *src, uint32_t *dst, int *w) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (long long)gridDim.x * blockDim.x) {
        uint32_t s, t;
        er_edge(key, (uint64_t)e, n, &s, &t);
        src[e] = s;
        dst[e] = t;
        w[e] = wmin + (int)uniform_below((uint32_t)draw(wkey, (uint64_t)e), wrange);
    }
}

// Grid edges of vertex v start at grid_off[v] (exclusive prefix of grid_deg); long-range
// edge j is edge grid_m + j.
__global__ void k_gen_grid(int rows, int cols, long long grid_m, long long n_long,
                           const long long *grid_off, uint64_t lkey, uint64_t wkey, int wmin,
                           uint32_t wrange, uint32_t *src, uint32_t *dst, int *w) {
    const long long n = (long long)rows * cols;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (long long)gridDim.x * blockDim.x) {
        long long o = grid_off[v];
        int d = grid_deg((uint32_t)v, rows, cols);
        for (int k = 0; k < d; ++k) {
            long long e = o + k;
            src[e] = (uint32_t)v;
            dst[e] = grid_nbr((uint32_t)v, rows, cols, k);
            w[e] = wmin + (int)uniform_below((uint32_t)draw(wkey, (uint64_t)e), wrange);
        }
    }
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n_long;
         j += (long long)gridDim.x * blockDim.x) {
        uint32_t s, t;
        er_edge(lkey, (uint64_t)j, (uint32_t)n, &s, &t);
        long long e = grid_m + j;
        src[e] = s;
        dst[e] = t;
        w[e] = wmin + (int)uniform_below((uint32_t)draw(wkey, (uint64_t)e), wrange);
    }
}

__global__ void k_grid_deg(int rows, int cols, long long *deg) {
    const long long n = (long long)rows * cols;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (long long)gridDim.x * blockDim.x)
        deg[v] = grid_deg((uint32_t)v, rows, cols);
}

// keys/vals in reverse generation order: position p holds edge m-1-p.
__global__ void k_rev_keys(const uint32_t *src, long long m, uint32_t *keys, uint32_t *vals) {
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < m;
         p += (long long)gridDim.x * blockDim.x) {
        long long e = m - 1 - p;
        keys[p] = src[e];
        vals[p] = (uint32_t)e;
    }
}
__global__ void k_count_src(const uint32_t *src, long long m, int *deg) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (long long)gridDim.x * blockDim.x)
        atomicAdd(&deg[src[e]], 1);
}
__global__ void k_gather_edges(const uint32_t *vals, long long m, const uint32_t *dst, const int *w,
                               int *col, int *wout) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (long long)gridDim.x * blockDim.x) {
        uint32_t e = vals[i];
        col[i] = (int)dst[e];
        wout[i] = w[e];
    }
}
__global__ void k_zero_i32(int *p, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = 0;
}

#define NP_CHECK(x) do { hipError_t _e = (x); if (_e != hipSuccess) { err = _e; goto done; } } while (0)

hipError_t build_csr_device(const uint32_t *src, const uint32_t *dst, const int *w, int n,
                            long long m, GenOut *out, hipStream_t s) {
    hipError_t err = hipSuccess;
    uint32_t *keys = nullptr, *vals = nullptr, *keys2 = nullptr, *vals2 = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    int *deg = nullptr;
    int bits = 1;
    while (bits < 32 && ((long long)1 << bits) < (long long)n) ++bits;
    out->row = out->col = out->w = nullptr;
    NP_CHECK(hipMalloc(&out->row, sizeof(int) * ((size_t)n + 1)));
    NP_CHECK(hipMalloc(&out->col, sizeof(int) * (size_t)(m > 0 ? m : 1)));
    NP_CHECK(hipMalloc(&out->w, sizeof(int) * (size_t)(m > 0 ? m : 1)));
    NP_CHECK(hipMalloc(&deg, sizeof(int) * ((size_t)n + 1)));
    hipLaunchKernelGGL(k_zero_i32, dim3(grid_for((long long)n + 1, 256)), dim3(256), 0, s, deg, (long long)n + 1);
    if (m > 0) {
        hipLaunchKernelGGL(k_count_src, dim3(grid_for(m, 256)), dim3(256), 0, s, src, m, deg);
        NP_CHECK(hipMalloc(&keys, sizeof(uint32_t) * (size_t)m));
        NP_CHECK(hipMalloc(&vals, sizeof(uint32_t) * (size_t)m));
        NP_CHECK(hipMalloc(&keys2, sizeof(uint32_t) * (size_t)m));
        NP_CHECK(hipMalloc(&vals2, sizeof(uint32_t) * (size_t)m));
        hipLaunchKernelGGL(k_rev_keys, dim3(grid_for(m, 256)), dim3(256), 0, s, src, m, keys, vals);
        NP_CHECK(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, keys2, vals, vals2, (size_t)m, 0,
                                           bits, s));
        NP_CHECK(hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1));
        NP_CHECK(rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, keys2, vals, vals2, (size_t)m, 0, bits,
                                           s));
        hipLaunchKernelGGL(k_gather_edges, dim3(grid_for(m, 256)), dim3(256), 0, s, vals2, m, dst, w,
                           out->col, out->w);
        NP_CHECK(hipGetLastError());
        NP_CHECK(hipFree(tmp)); tmp = nullptr;
    }
    // row = exclusive prefix of deg (deg[n] = 0 -> row[n] = m)
    {
        size_t sb = 0;
        NP_CHECK(rocprim::exclusive_scan(nullptr, sb, deg, out->row, 0, (size_t)n + 1,
                                         rocprim::plus<int>(), s));
        NP_CHECK(hipMalloc(&tmp, sb ? sb : 1));
        NP_CHECK(rocprim::exclusive_scan(tmp, sb, deg, out->row, 0, (size_t)n + 1, rocprim::plus<int>(),
                                         s));
    }
    NP_CHECK(hipStreamSynchronize(s));
done:
    (void)hipFree(tmp); (void)hipFree(keys); (void)hipFree(vals); (void)hipFree(keys2);
    (void)hipFree(vals2); (void)hipFree(deg);
    if (err != hipSuccess) {
        (void)hipFree(out->row); (void)hipFree(out->col); (void)hipFree(out->w);
        out->row = out->col = out->w = nullptr;
    }
    return err;
}

hipError_t gen_rmat_edges(int scale, long long m, uint64_t seed, double a, double b, double c,
                          int permute, int wmin, int wmax, uint32_t *src, uint32_t *dst, int *w,
                          hipStream_t s) {
    uint64_t ta, tab, tabc;
    rmat_thresholds(a, b, c, &ta, &tab, &tabc);
    hipLaunchKernelGGL(k_gen_rmat, dim3(grid_for(m, 256)), dim3(256), 0, s, scale, m,
                       stream_key(seed, ST_RMAT), stream_key(seed, ST_WEIGHT), stream_key(seed, ST_PERM),
                       ta, tab, tabc, permute, wmin, (uint32_t)(wmax - wmin + 1), src, dst, w);
    return hipGetLastError();
}

hipError_t gen_er_edges(int n, long long m, uint64_t seed, int wmin, int wmax, uint32_t *src,
                        uint32_t *dst, int *w, hipStream_t s) {
    hipLaunchKernelGGL(k_gen_er, dim3(grid_for(m, 256)), dim3(256), 0, s, (uint32_t)n, m,
                       stream_key(seed, ST_ER), stream_key(seed, ST_WEIGHT), wmin,
                       (uint32_t)(wmax - wmin + 1), src, dst, w);
    return hipGetLastError();
}

hipError_t gen_grid_edges(int rows, int cols, long long n_long, uint64_t seed, int wmin, int wmax,
                          const long long *grid_off, uint32_t *src, uint32_t *dst, int *w,
                          hipStream_t s) {
    long long n = (long long)rows * cols;
    long long grid_m = 2ll * ((long long)rows * (cols - 1) + (long long)cols * (rows - 1));
    hipLaunchKernelGGL(k_gen_grid, dim3(grid_for(n > n_long ? n : n_long, 256)), dim3(256), 0, s, rows,
                       cols, grid_m, n_long, grid_off, stream_key(seed, ST_LONG),
                       stream_key(seed, ST_WEIGHT), wmin, (uint32_t)(wmax - wmin + 1), src, dst, w);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Solver layout: vertices renumbered by decreasing in-degree (ties keep their order), so
// the targets most relaxations hit -- R-MAT concentrates most in-edges on few vertices --
// share a few megabytes of the distance array and stay in L2.  The renumbering is
// internal: sources are mapped in, distances mapped back out.
// ---------------------------------------------------------------------------------
__global__ void k_count_indeg(const int *col, long long m, unsigned *indeg) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (long long)gridDim.x * blockDim.x)
        atomicAdd(&indeg[col[e]], 1u);
}
__global__ void k_indeg_keys(const unsigned *indeg, int n, unsigned *keys, unsigned *vals) {
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (long long)gridDim.x * blockDim.x) {
        keys[v] = ~indeg[v];
        vals[v] = (unsigned)v;
    }
}
__global__ void k_inverse_perm(const unsigned *new2old, int n, int *old2new, int *deg2, const int *row) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        unsigned o = new2old[i];
        old2new[o] = (int)i;
        deg2[i] = row[o + 1] - row[o];
    }
}
// One 64-lane wavefront per renumbered vertex copies its edge list.
__global__ void k_relabel_edges(const int *row, const int *col, const int *w, const unsigned *new2old,
                                const int *old2new, const int *row2, int n, int *col2, int *w2) {
    const long long nw = (long long)gridDim.x * (blockDim.x / 64);
    const int lane = threadIdx.x & 63;
    for (long long i = (long long)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); i < n; i += nw) {
        const unsigned o = new2old[i];
        const int s = row[o], d = row[o + 1] - s, t = row2[i];
        for (int k = lane; k < d; k += 64) {
            col2[t + k] = old2new[col[s + k]];
            if (w) w2[t + k] = w[s + k];
        }
    }
}
__global__ void k_unpermute(const int *src, const int *old2new, int n, int *out) {
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (long long)gridDim.x * blockDim.x)
        out[v] = src[old2new[v]];
}

hipError_t launch_unpermute(const int *src, const int *old2new, int n, int *out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_unpermute, dim3(grid_for(n, 256)), dim3(256), 0, s, src, old2new, n, out);
    return hipGetLastError();
}

hipError_t relabel_by_indegree(const int *row, const int *col, const int *w, int n, long long m,
                               double min_skew, Relabel *out, hipStream_t s) {
    hipError_t err = hipSuccess;
    unsigned *indeg = nullptr, *keys = nullptr, *vals = nullptr, *keys2 = nullptr, *vals2 = nullptr;
    int *deg2 = nullptr;
    void *tmp = nullptr;
    size_t tb = 0;
    *out = Relabel{};
    if (n < 2 || m < 1) return hipSuccess;
    NP_CHECK(hipMalloc(&indeg, sizeof(unsigned) * (size_t)n));
    NP_CHECK(hipMalloc(&keys, sizeof(unsigned) * (size_t)n));
    NP_CHECK(hipMalloc(&vals, sizeof(unsigned) * (size_t)n));
    NP_CHECK(hipMalloc(&keys2, sizeof(unsigned) * (size_t)n));
    NP_CHECK(hipMalloc(&vals2, sizeof(unsigned) * (size_t)n));
    hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(n, 256)), dim3(256), 0, s, (int *)indeg, (long long)n);
    hipLaunchKernelGGL(k_count_indeg, dim3(grid_for(m, 256)), dim3(256), 0, s, col, m, indeg);
    hipLaunchKernelGGL(k_indeg_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, indeg, n, keys, vals);
    NP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, keys, keys2, vals, vals2, (size_t)n, 0, 32, s));
    NP_CHECK(hipMalloc(&tmp, tb ? tb : 1));
    NP_CHECK(rocprim::radix_sort_pairs(tmp, tb, keys, keys2, vals, vals2, (size_t)n, 0, 32, s));
    {
        // skew: share of the edges received by the top n/16 vertices
        const int top = n / 16 > 0 ? n / 16 : 1;
        std::vector<unsigned> h((size_t)top);
        NP_CHECK(hipMemcpyAsync(h.data(), keys2, sizeof(unsigned) * (size_t)top, hipMemcpyDeviceToHost, s));
        NP_CHECK(hipStreamSynchronize(s));
        double sum = 0;
        for (unsigned k : h) sum += (double)(~k);
        out->skew = sum / (double)m;
    }
    if (out->skew < min_skew) goto done;
    NP_CHECK(hipMalloc(&out->row, sizeof(int) * ((size_t)n + 1)));
    NP_CHECK(hipMalloc(&out->col, sizeof(int) * (size_t)m));
    NP_CHECK(hipMalloc(&out->w, sizeof(int) * (size_t)m));
    NP_CHECK(hipMalloc(&out->old2new, sizeof(int) * (size_t)n));
    NP_CHECK(hipMalloc(&deg2, sizeof(int) * ((size_t)n + 1)));
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, deg2 + n, 1ll);
    hipLaunchKernelGGL(k_inverse_perm, dim3(grid_for(n, 256)), dim3(256), 0, s, vals2, n, out->old2new, deg2, row);
    NP_CHECK(hipFree(tmp)); tmp = nullptr;
    {
        size_t sb = 0;
        NP_CHECK(rocprim::exclusive_scan(nullptr, sb, deg2, out->row, 0, (size_t)n + 1, rocprim::plus<int>(), s));
        NP_CHECK(hipMalloc(&tmp, sb ? sb : 1));
        NP_CHECK(rocprim::exclusive_scan(tmp, sb, deg2, out->row, 0, (size_t)n + 1, rocprim::plus<int>(), s));
    }
    hipLaunchKernelGGL(k_relabel_edges, dim3(grid_for((long long)n * 64, 256) > 16384 ? 16384 : grid_for((long long)n * 64, 256)),
                       dim3(256), 0, s, row, col, w, vals2, out->old2new, out->row, n, out->col, out->w);
    NP_CHECK(hipGetLastError());
    NP_CHECK(hipStreamSynchronize(s));
    out->active = true;
done:
    (void)hipFree(tmp); (void)hipFree(indeg); (void)hipFree(keys); (void)hipFree(vals);
    (void)hipFree(keys2); (void)hipFree(vals2); (void)hipFree(deg2);
    if (err != hipSuccess || !out->active) {
        (void)hipFree(out->row); (void)hipFree(out->col); (void)hipFree(out->w); (void)hipFree(out->old2new);
        double sk = out->skew;
        *out = Relabel{};
        out->skew = sk;
    }
    return err;
}

// Exclusive prefix of grid out-degrees (grid_off[n] = grid edge count).
hipError_t grid_offsets(int rows, int cols, long long *grid_off, hipStream_t s) {
    long long n = (long long)rows * cols;
    long long *deg = nullptr;
    void *tmp = nullptr;
    size_t sb = 0;
    hipError_t err = hipMalloc(&deg, sizeof(long long) * (size_t)(n + 1));
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_grid_deg, dim3(grid_for(n, 256)), dim3(256), 0, s, rows, cols, deg);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, (int *)(deg + n), 2ll);
    err = rocprim::exclusive_scan(nullptr, sb, deg, grid_off, 0ll, (size_t)n + 1, rocprim::plus<long long>(), s);
    if (err == hipSuccess) err = hipMalloc(&tmp, sb ? sb : 1);
    if (err == hipSuccess)
        err = rocprim::exclusive_scan(tmp, sb, deg, grid_off, 0ll, (size_t)n + 1, rocprim::plus<long long>(), s);
    if (err == hipSuccess) err = hipStreamSynchronize(s);
    (void)hipFree(tmp);
    (void)hipFree(deg);
    return err;
}

}  // namespace np
