// Graph coloring of the reference's GRAPH_COLOR_GPU (apps/graph-color/src/gpu/
// graph_color_wrapper.cu:598-635, 665-699; kernels graph_color_kernel.cu:14-35, 779-834).
//
// The reference colors in rounds c = 1, 2, ...: an uncolored vertex v takes color c unless one
// of its out-neighbours u with u < v is uncolored or has taken c in the same round
// (graph_color_kernel.cu:26-27).  A smaller neighbour that is uncolored when the round starts
// blocks v whether or not it is colored during the round, and one colored in an earlier round
// never does, so the outcome does not depend on thread timing:
//     color[v] = 1 + max { color[u] : v -> u, u < v }   (1 when there is no such u),
// the length of the longest path of decreasing ids that starts at v.
//
// MI355X mapping (method 0): the same values in O(m) work instead of O(rounds * m).  dep[v]
// counts the edges v -> u with u < v; the transposed lists notify[u] = { v : v -> u, u < v }
// are built once (count, rocPRIM scan, scatter; one wavefront per vertex).  Round c's frontier is the vertices of color
// c; expanding it decrements dep[v] for every v in their notify lists, and a vertex whose count
// reaches zero in round c gets color c + 1 and joins the next frontier.  There are as many
// rounds as colors (about a thousand on R-MAT s20) and each is small, so no round waits on the
// host: one kernel per round, one wavefront per frontier vertex, the frontier count read on the
// device; rounds are captured 30 to a hipGraph and replayed until the frontier is empty.
// Method 1 runs the reference's rounds as they are (every uncolored vertex re-checks its
// smaller neighbours each round, one host readback per round): the baseline.
#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_scan.hpp>
#include "np_internal.h"

namespace np {

__device__ __forceinline__ int col_lane() { return threadIdx.x & 63; }

static int col_grid(long long work, int cap) {
    long long g = (work + 255) / 256;
    return (int)(g < 1 ? 1 : (g > cap ? cap : g));
}

// dep[v] = #{v -> u : u < v}; ncnt[u] += 1 for each such edge.  One wavefront per vertex, lanes
// over its edges (a hub's edges would serialise in one thread).
__global__ void k_color_count(const int *row, const int *col, int n, int *dep, long long *ncnt) {
    const int lane = col_lane();
    const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
    for (long long v = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); v < n; v += nw) {
        int d = 0;
        for (int k = row[v]; k < row[v + 1]; k += 64) {
            const int e = k + lane;
            const bool small = e < row[v + 1] && col[e] < v;
            if (small) atomicAdd((unsigned long long *)&ncnt[col[e]], 1ull);
            d += __popcll(__ballot(small));
        }
        if (lane == 0) dep[v] = d;
    }
}

// notify lists: ncol[nrow[u] + k] = v for the edges v -> u, u < v (cursor[] starts at 0); one
// wavefront per vertex, so a hub's slot reservations are issued 64 at a time.
__global__ void k_color_fill(const int *row, const int *col, int n, const long long *nrow, int *cursor, int *ncol) {
    const int lane = col_lane();
    const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
    for (long long v = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); v < n; v += nw)
        for (int k = row[v]; k < row[v + 1]; k += 64) {
            const int e = k + lane;
            if (e < row[v + 1]) {
                const int u = col[e];
                if (u < v) ncol[nrow[u] + atomicAdd(&cursor[u], 1)] = (int)v;
            }
        }
}

// Appends v to q (wave-aggregated) when `take`; every lane of the wave calls it.
__device__ __forceinline__ void col_append(bool take, int v, int *q, int *qn) {
    const unsigned long long want = __ballot(take);
    if (want == 0) return;
    const int lane = col_lane();
    const int leader = __ffsll((long long)want) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(qn, __popcll(want));
    base = __shfl(base, leader);
    if (take) q[base + __popcll(want & ((1ull << lane) - 1))] = v;
}

// Round 1: the vertices with no smaller out-neighbour.
__global__ void k_color_init(const int *dep, int n, int *color, int *q, int *qn) {
    for (long long base = (long long)blockIdx.x * blockDim.x; base < n; base += (long long)gridDim.x * blockDim.x) {
        const long long v = base + threadIdx.x;
        const bool take = v < n && dep[v] == 0;
        if (v < n) color[v] = take ? 1 : 0;
        col_append(take, (int)v, q, qn);
    }
}

// Round r of the dependency-counting drive (every kernel of a captured chunk is one round):
// the frontier q_in (count *qn_in, all of one color c) decrements the counters of its notify
// lists; a vertex whose counter reaches zero gets c + 1 and joins q_out.  *qn_zero is the
// output count of the next round: zeroed here (nothing reads it during this round).  A round
// with an empty frontier does nothing else, so a chunk may run past the last color.
__global__ __launch_bounds__(256) void k_color_step(const int *q_in, const int *qn_in, int *q_out, int *qn_out,
                                                    int *qn_zero, const long long *nrow, const int *ncol, int *dep,
                                                    int *color, int *rounds) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *qn_zero = 0;
    const int ne = *qn_in;
    if (ne == 0) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) *rounds += 1;
    const int c = color[q_in[0]];
    const int lane = col_lane();
    const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
    for (long long o = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); o < ne; o += nw) {
        const int u = q_in[o];
        const long long e0 = nrow[u], e1 = nrow[u + 1];
        for (long long k = e0; k < e1; k += 64) {  // one wavefront per frontier vertex
            const long long i = k + lane;
            bool take = false;
            int v = 0;
            if (i < e1) {
                v = ncol[i];
                if (atomicSub(&dep[v], 1) == 1) {  // its last smaller neighbour was colored now
                    color[v] = c + 1;
                    take = true;
                }
            }
            col_append(take, v, q_out, qn_out);
        }
    }
}

// The reference's round as it is (method 1): every uncolored vertex checks its smaller
// out-neighbours (graph_color_kernel.cu:14-35).
__global__ void k_color_round(const int *row, const int *col, int n, int *color, int c, int *nonstop) {
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x) {
        if (color[v] != 0) continue;
        bool ok = true;
        for (int e = row[v]; e < row[v + 1]; ++e) {
            const int u = col[e];
            if (u < v) {
                const int cu = __hip_atomic_load(color + u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (cu == 0 || cu == c) { ok = false; break; }
            }
        }
        if (ok) {
            __hip_atomic_store(color + v, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            *nonstop = 1;
        }
    }
}

hipError_t launch_color_count(const int *row, const int *col, int n, int *dep, long long *ncnt, hipStream_t s) {
    hipLaunchKernelGGL(k_color_count, dim3(col_grid((long long)n * 64, 8192)), dim3(256), 0, s, row, col, n, dep, ncnt);
    return hipGetLastError();
}
hipError_t launch_color_fill(const int *row, const int *col, int n, const long long *nrow, int *cursor, int *ncol,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_color_fill, dim3(col_grid((long long)n * 64, 8192)), dim3(256), 0, s, row, col, n, nrow,
                       cursor, ncol);
    return hipGetLastError();
}
hipError_t launch_color_init(const int *dep, int n, int *color, int *q, int *qn, hipStream_t s) {
    hipLaunchKernelGGL(k_color_init, dim3(col_grid(n, 8192)), dim3(256), 0, s, dep, n, color, q, qn);
    return hipGetLastError();
}
hipError_t launch_color_step(const int *q_in, const int *qn_in, int *q_out, int *qn_out, int *qn_zero,
                             const long long *nrow, const int *ncol, int *dep, int *color, int *rounds, int num_cus,
                             hipStream_t s) {
    // frontiers average ~n / colors vertices: a small grid that strides, since a round that ends at
    // once still pays for dispatching every workgroup
    hipLaunchKernelGGL(k_color_step, dim3(num_cus * 2), dim3(256), 0, s, q_in, qn_in, q_out, qn_out, qn_zero, nrow, ncol,
                       dep, color, rounds);
    return hipGetLastError();
}
hipError_t launch_color_round(const int *row, const int *col, int n, int *color, int c, int *nonstop, hipStream_t s) {
    hipLaunchKernelGGL(k_color_round, dim3(col_grid(n, 8192)), dim3(256), 0, s, row, col, n, color, c, nonstop);
    return hipGetLastError();
}
// Exclusive prefix of in[0..n] into out[0..n] (in[n] = 0 gives out[n] = total).
hipError_t color_scan(void *temp, size_t *temp_bytes, const long long *in, long long *out, int n, hipStream_t s) {
    return rocprim::exclusive_scan(temp, *temp_bytes, in, out, 0ll, (size_t)n + 1, rocprim::plus<long long>(), s);
}

}  // namespace np
