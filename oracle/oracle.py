"""ctypes front-end of the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.

Wraps liboracle.so (the C restatement in np_oracle.c) and, when it was built, the
reference's own CPU code in _ref/ (see oracle/Makefile).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
INF = 1073741824            # common/global.h:19
UNDEFINED = 0x7FFFFFFF      # apps/bfs-rec/src/gpu/bfs_rec.h:13

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_lib = None
_ref_sssp = None
_ref_bfs = None


def build():
    """Compile liboracle.so and, where the reference sources exist, oracle/_ref."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        ll = ctypes.c_longlong
        L.orc_dijkstra_lazy.restype = ll
        L.orc_dijkstra_lazy.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, ctypes.c_int, _i32p]
        L.orc_dijkstra_lazy_multi.restype = ll
        L.orc_dijkstra_lazy_multi.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, _i32p, ctypes.c_int, _i32p]
        L.orc_dijkstra_lazy_many.restype = ctypes.c_int
        L.orc_dijkstra_lazy_many.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, _i32p, ctypes.c_int,
                                             _i32p, ctypes.c_int]
        L.orc_dijkstra_dheap.restype = ctypes.c_int
        L.orc_dijkstra_dheap.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, ctypes.c_int, ctypes.c_int, _i32p]
        L.orc_bellman_ford.restype = ctypes.c_int
        L.orc_bellman_ford.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, _i32p, ctypes.c_int, _i32p]
        L.orc_bfs_levels.restype = ctypes.c_int
        L.orc_bfs_levels.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.c_int, _i32p]
        L.orc_check_sssp.restype = ctypes.c_int
        L.orc_check_sssp.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, _i32p, ctypes.c_int, _i32p,
                                     ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.orc_reached_work.restype = None
        L.orc_reached_work.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.c_int,
                                       ctypes.POINTER(ll), ctypes.POINTER(ll)]
        L.orc_mlp_nparam.restype = ctypes.c_int
        L.orc_mlp_forward.restype = None
        L.orc_pagerank.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.orc_pagerank.restype = ctypes.c_int
        L.orc_spmv.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.orc_spmv.restype = None
        L.orc_color.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p]
        L.orc_color.restype = None
        L.orc_bc.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.orc_bc.restype = ctypes.c_int
        L.orc_mlp_forward.argtypes = [_f32p, _f32p, ctypes.c_int, _f32p]
        L.orc_mlp_loss_grad.restype = ctypes.c_float
        L.orc_mlp_loss_grad.argtypes = [_f32p, _f32p, _i32p, _f32p, ctypes.c_int, ctypes.c_float, _f32p]
        _lib = L
    return _lib


def _arr(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def dijkstra(row, col, w, src):
    """Distances from src (restated SSSP_CPU).  Returns (dist, iteration_count)."""
    row, col, w = _arr(row), _arr(col), _arr(w)
    n = len(row) - 1
    dist = np.empty(n, np.int32)
    it = lib().orc_dijkstra_lazy(n, row, col, w if len(w) else np.zeros(1, np.int32), int(src), dist)
    if it < 0:
        raise MemoryError("oracle heap allocation failed")
    return dist, int(it)


def dijkstra_multi(row, col, w, srcs):
    row, col, w, srcs = _arr(row), _arr(col), _arr(w), _arr(srcs)
    n = len(row) - 1
    dist = np.empty(n, np.int32)
    it = lib().orc_dijkstra_lazy_multi(n, row, col, w if len(w) else np.zeros(1, np.int32), srcs,
                                       len(srcs), dist)
    if it < 0:
        raise MemoryError("oracle heap allocation failed")
    return dist


def dijkstra_many(row, col, w, srcs, threads=0):
    """One independent Dijkstra per source, sources spread over OpenMP threads."""
    row, col, w, srcs = _arr(row), _arr(col), _arr(w), _arr(srcs)
    n = len(row) - 1
    dist = np.empty((len(srcs), n), np.int32)
    if lib().orc_dijkstra_lazy_many(n, row, col, w if len(w) else np.zeros(1, np.int32), srcs,
                                    len(srcs), dist, int(threads)) != 0:
        raise MemoryError("oracle heap allocation failed")
    return dist


def dijkstra_dheap(row, col, w, src, arity=2):
    row, col, w = _arr(row), _arr(col), _arr(w)
    n = len(row) - 1
    dist = np.empty(n, np.int32)
    if lib().orc_dijkstra_dheap(n, row, col, w if len(w) else np.zeros(1, np.int32), int(src),
                                int(arity), dist) != 0:
        raise MemoryError
    return dist


def bellman_ford(row, col, w, srcs):
    row, col, w = _arr(row), _arr(col), _arr(w)
    srcs = _arr(np.atleast_1d(srcs))
    n = len(row) - 1
    dist = np.empty(n, np.int32)
    if lib().orc_bellman_ford(n, row, col, w if len(w) else np.zeros(1, np.int32), srcs, len(srcs),
                              dist) < 0:
        raise ValueError("negative cycle")
    return dist


def bfs_levels(row, col, src):
    row, col = _arr(row), _arr(col)
    n = len(row) - 1
    lev = np.empty(n, np.int32)
    if lib().orc_bfs_levels(n, row, col if len(col) else np.zeros(1, np.int32), int(src), lev) != 0:
        raise MemoryError
    return lev


CHECK_CODES = {0: "ok", 1: "source not at 0", 2: "value out of range", 3: "edge not relaxed",
               4: "reached vertex without a tight in-edge", 5: "non-positive weight",
               6: "out of memory"}


def check_sssp(row, col, w, srcs, dist, inf=INF, unit=False):
    """Shortest-path certificate.  Returns (code, bad_vertex); code 0 = exact."""
    row, col, dist = _arr(row), _arr(col), _arr(dist)
    w = _arr(w) if (w is not None and len(w)) else np.zeros(1, np.int32)
    srcs = _arr(np.atleast_1d(srcs))
    n = len(row) - 1
    bad = ctypes.c_int(-1)
    code = lib().orc_check_sssp(n, row, col if len(col) else np.zeros(1, np.int32), w, srcs,
                                len(srcs), dist, int(inf), 1 if unit else 0, ctypes.byref(bad))
    return code, bad.value


def reached_work(row, dist, inf=INF):
    row, dist = _arr(row), _arr(dist)
    a, b = ctypes.c_longlong(), ctypes.c_longlong()
    lib().orc_reached_work(len(row) - 1, row, dist, int(inf), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


MLP_NPARAM = 10 * 32 + 32 + 32 * 32 + 32 + 5 * 32 + 5


def pagerank(row, col, damping=0.5, total=1000.0, eps=1e-13, max_iters=100000):
    """PageRank as PAGERANK_GPU iterates it (float64[n], iterations)."""
    row, col = _arr(row), _arr(col)
    n = len(row) - 1
    out = np.zeros(max(n, 1), np.float64)
    it = lib().orc_pagerank(n, row, col if len(col) else np.zeros(1, np.int32), damping, total, eps, max_iters,
                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if it < 0:
        raise ValueError("orc_pagerank failed")
    return out[:n], it


def spmv(row, col, data, x):
    """y = A x over the CSR (float64)."""
    row, col = _arr(row), _arr(col)
    n = len(row) - 1
    d = np.ascontiguousarray(data, np.float64)
    xx = np.ascontiguousarray(x, np.float64)
    y = np.zeros(max(n, 1), np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    lib().orc_spmv(n, row, col if len(col) else np.zeros(1, np.int32),
                   (d if len(d) else np.zeros(1)).ctypes.data_as(dp), xx.ctypes.data_as(dp), y.ctypes.data_as(dp))
    return y[:n]


def spmv_reference_values(n, m):
    """The reference's inputs (SpMV_wrapper.cu:104-113): data[i] = i % 13 + 0.5, x[i] = i % 17 + 0.5."""
    return np.arange(m) % 13 + 0.5, np.arange(n) % 17 + 0.5


def color(row, col):
    """Colors of the reference's graph-coloring rounds (int32[n])."""
    row, col = _arr(row), _arr(col)
    n = len(row) - 1
    out = np.zeros(max(n, 1), np.int32)
    lib().orc_color(n, row, col if len(col) else np.zeros(1, np.int32), out)
    return out[:n]


def bc(row, col, srcs=None):
    """Brandes betweenness (float64[n]) summed over srcs (None: every vertex)."""
    row, col = _arr(row), _arr(col)
    n = len(row) - 1
    s = np.zeros(1, np.int32) if srcs is None else _arr(srcs)
    ns = 0 if srcs is None else len(s)
    out = np.zeros(max(n, 1), np.float64)
    rc = lib().orc_bc(n, row, col if len(col) else np.zeros(1, np.int32), s if len(s) else np.zeros(1, np.int32),
                      ns, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if rc != 0:
        raise ValueError("orc_bc failed: %d" % rc)
    return out[:n]


def mlp_forward(params, x):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty((x.shape[0], 5), np.float32)
    lib().orc_mlp_forward(params, x, x.shape[0], out)
    return out


def mlp_loss_grad(params, x, label, target, lam):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    label = np.ascontiguousarray(label, np.int32)
    target = np.ascontiguousarray(target, np.float32)
    grad = np.empty(MLP_NPARAM, np.float32)
    loss = lib().orc_mlp_loss_grad(params, x, label, target, x.shape[0], float(lam), grad)
    return float(loss), grad


# --------------------------------------------------------------------------------------
# The reference's own CPU code (oracle/_ref), when built.
# --------------------------------------------------------------------------------------
def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libref_sssp.so"))


def _ref_sssp_lib():
    global _ref_sssp
    if _ref_sssp is None:
        L = ctypes.CDLL(os.path.join(HERE, "_ref", "libref_sssp.so"))
        L.ref_sssp_cpu.restype = ctypes.c_int
        L.ref_sssp_cpu.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _i32p, ctypes.c_int, _i32p,
                                   ctypes.POINTER(ctypes.c_longlong)]
        L.ref_dheap_dijkstra.restype = ctypes.c_int
        L.ref_dheap_dijkstra.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, ctypes.c_int, ctypes.c_int, _i32p]
        L.ref_read_graph.restype = ctypes.c_int
        L.ref_read_graph.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.ref_copy_csr.restype = None
        L.ref_copy_csr.argtypes = [_i32p, _i32p, _i32p]
        L.ref_output_cost.restype = ctypes.c_int
        L.ref_output_cost.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _i32p]
        _ref_sssp = L
    return _ref_sssp


def ref_sssp_cpu(row, col, w, src):
    """The reference SSSP_CPU() itself.  Returns (dist, printed iteration number)."""
    row, col, w = _arr(row), _arr(col), _arr(w)
    n, m = len(row) - 1, len(col)
    dist = np.empty(n, np.int32)
    it = ctypes.c_longlong(-1)
    _ref_sssp_lib().ref_sssp_cpu(n, m, row, col if m else np.zeros(1, np.int32),
                                 w if m else np.zeros(1, np.int32), int(src), dist, ctypes.byref(it))
    return dist, it.value


def ref_dheap_dijkstra(row, col, w, src, arity=2):
    row, col, w = _arr(row), _arr(col), _arr(w)
    n = len(row) - 1
    dist = np.empty(n, np.int32)
    _ref_sssp_lib().ref_dheap_dijkstra(n, row, col if len(col) else np.zeros(1, np.int32),
                                       w if len(w) else np.zeros(1, np.int32), int(src), int(arity), dist)
    return dist


def ref_read_graph(fmt, path):
    """Reference reader + convertCSR.  Returns (row, col, w, source)."""
    L = _ref_sssp_lib()
    n, m, s = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rc = L.ref_read_graph(int(fmt), path.encode(), ctypes.byref(n), ctypes.byref(m), ctypes.byref(s))
    if rc:
        raise IOError("reference reader failed: %d" % rc)
    row = np.empty(n.value + 1, np.int32)
    col = np.empty(max(m.value, 1), np.int32)
    w = np.empty(max(m.value, 1), np.int32)
    L.ref_copy_csr(row, col, w)
    return row, col[:m.value], w[:m.value], s.value


def ref_output_cost(path, src, dist):
    dist = _arr(dist)
    return _ref_sssp_lib().ref_output_cost(path.encode(), len(dist), int(src), dist)


_ref_bc = None


def ref_bc_cpu(row, col):
    """The reference's bc_cpu (every vertex a source, float32) from oracle/_ref."""
    global _ref_bc
    if _ref_bc is None:
        L = ctypes.CDLL(os.path.join(HERE, "_ref", "libref_bc.so"))
        L.ref_bc_cpu.restype = ctypes.c_int
        L.ref_bc_cpu.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, ctypes.POINTER(ctypes.c_float)]
        _ref_bc = L
    row, col = _arr(row), _arr(col)
    n = len(row) - 1
    out = np.zeros(max(n, 1), np.float32)
    _ref_bc.ref_bc_cpu(n, len(col), row, col if len(col) else np.zeros(1, np.int32),
                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return out[:n]


def ref_bc_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libref_bc.so"))


def ref_bfs_cpu(row, col, src):
    global _ref_bfs
    if _ref_bfs is None:
        L = ctypes.CDLL(os.path.join(HERE, "_ref", "libref_bfs.so"))
        L.ref_bfs_cpu.restype = ctypes.c_int
        L.ref_bfs_cpu.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, ctypes.c_int, _i32p]
        _ref_bfs = L
    row, col = _arr(row), _arr(col)
    n = len(row) - 1
    lev = np.empty(n, np.int32)
    _ref_bfs.ref_bfs_cpu(n, len(col), row, col if len(col) else np.zeros(1, np.int32), int(src), lev)
    return lev
