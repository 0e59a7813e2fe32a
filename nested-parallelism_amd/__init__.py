"""MI355X (gfx950) SSSP / recursive-BFS nested-parallelism path -- Python host side.

A thin ctypes mirror of the C ABI in include/np_sssp.h (libnpsssp.so, built in-tree by
build.py).  The reference's operator interface is SSSP_GPU() / BFS_REC_GPU() over the
global CSR (apps/sssp/src/gpu/sssp_wrapper.cu:548-583, apps/bfs-rec/src/gpu/
bfs_rec_wrapper.cu:231-269); here the same solves take the CSR explicitly:

    ctx = Context(0)
    g = Graph.from_csr(ctx, row, col, w)        # prepare_gpu's upload
    dist, stats = ctx.sssp(g, source, Sched(level=1))   # solution 6, CONSOLIDATE_LEVEL=1
    levels, stats = ctx.bfs(g, source)          # bfs-rec solution 3

There is no CPU fallback: every solve runs the HIP kernels, and a missing library or
device raises.
"""
import atexit
import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# NP_SSSP_LIB=bounds selects the bounds-checked build (libnpsssp_bounds.so: every frontier-queue
# and delayed-buffer write checked against its segment, NP_ERR_BOUNDS on a violation)
LIB_PATH = os.path.join(HERE, "libnpsssp_bounds.so" if os.environ.get("NP_SSSP_LIB") == "bounds" else "libnpsssp.so")

INF = 1073741824          # common/global.h:19
UNDEFINED = 0x7FFFFFFF    # apps/bfs-rec/src/gpu/bfs_rec.h:13
NCLASS = 32
MLP_NIN = 10
CTX_RELABEL = 1           # np_set_option key: 0 caller ids, 1 always renumber, 2 when skewed
CTX_PACK = 2              # np_set_option key: 1 batch instances sharing a graph are source-packed
CTX_PACK_MAX = 3          # np_set_option key: most sources per pack (2..32)
MLP_NPARAM = 32 * 10 + 32 + 32 * 32 + 32 + 5 * 32 + 5
NOALLOC_MAX_VERTICES = 2097152  # np_sssp_device_noalloc's largest graph

LEVEL_WAVE, LEVEL_WORKGROUP, LEVEL_GRID, LEVEL_LEARNED = 0, 1, 2, -1
DRIVE_PERSISTENT, DRIVE_HOSTLOOP, DRIVE_GRAPH = 0, 1, 2
FRONTIER_SCAN, FRONTIER_PUSH = 0, 1
# np_sched.scheme (the reference's config.solution schedules)
(SCHEME_CONSOLIDATE, SCHEME_THREAD, SCHEME_DUAL_QUEUE, SCHEME_LDS_BUFFER, SCHEME_GLOBAL_BUFFER,
 SCHEME_WAVE_NESTED, SCHEME_WG_NESTED, SCHEME_BFS_FLAT, SCHEME_BFS_HIER) = range(9)

ERRORS = {0: "NP_OK", -1: "NP_ERR_ARG", -2: "NP_ERR_HIP", -3: "NP_ERR_NOMEM", -4: "NP_ERR_CAPACITY",
          -5: "NP_ERR_NOT_CONVERGED", -6: "NP_ERR_TIMEOUT", -7: "NP_ERR_UNSUPPORTED", -8: "NP_ERR_IO",
          -9: "NP_ERR_BOUNDS"}


class NpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class Sched(ctypes.Structure):
    """np_sched: consolidation level (CONSOLIDATE_LEVEL), log2 bin thresholds, drive."""
    _fields_ = [("level", ctypes.c_int32), ("log2_t_lo", ctypes.c_int32), ("log2_t_hi", ctypes.c_int32),
                ("drive", ctypes.c_int32), ("frontier", ctypes.c_int32), ("wgs_per_cu", ctypes.c_int32),
                ("graph_chunk", ctypes.c_int32), ("max_iters", ctypes.c_int32), ("trace", ctypes.c_int32),
                ("scheme", ctypes.c_int32), ("delta", ctypes.c_int32), ("pack_hold", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 4)]

    def __init__(self, level=1, log2_t_lo=6, log2_t_hi=6, drive=DRIVE_PERSISTENT, frontier=FRONTIER_SCAN,
                 wgs_per_cu=0, graph_chunk=0, max_iters=0, trace=0, scheme=0, delta=0, pack_hold=0):
        super().__init__(level, log2_t_lo, log2_t_hi, drive, frontier, wgs_per_cu, graph_chunk, max_iters,
                         trace, scheme, delta, pack_hold)

    def copy(self, **kw):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}
        d.update(kw)
        return Sched(**d)

    def __repr__(self):
        return "Sched(%s)" % ", ".join("%s=%d" % (f, getattr(self, f)) for f, _ in self._fields_ if f != "reserved")


class Stats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("status", ctypes.c_int32), ("kernel_ms", ctypes.c_double),
                ("frontier_total", ctypes.c_int64), ("drive", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("pairs_total", ctypes.c_int64), ("edges_total", ctypes.c_int64),
                ("items_total", ctypes.c_int64), ("reserved", ctypes.c_int64 * 2)]

    def __repr__(self):
        return "Stats(iterations=%d, kernel_ms=%.4f, frontier_total=%d)" % (
            self.iterations, self.kernel_ms, self.frontier_total)


class TraceRec(ctypes.Structure):
    _fields_ = [("iteration", ctypes.c_int32), ("level", ctypes.c_int32), ("log2_t_lo", ctypes.c_int32),
                ("log2_t_hi", ctypes.c_int32), ("count", ctypes.c_int32 * NCLASS), ("t_begin", ctypes.c_int64),
                ("t_end", ctypes.c_int64), ("pairs", ctypes.c_int64), ("edges", ctypes.c_int64)]


_lib = None
_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p

# (name, restype, argtypes) of every exported entry point of include/np_sssp.h
_SIGS = [
    ("np_last_error", ctypes.c_char_p, []),
    ("np_version", ctypes.c_char_p, []),
    ("np_sched_default", None, [ctypes.c_int32, ctypes.POINTER(Sched)]),
    ("np_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    ("np_set_stream", ctypes.c_int, [_vp, _vp]),
    ("np_get_stream", _vp, [_vp]),
    ("np_destroy", ctypes.c_int, [_vp]),
    ("np_device_info", ctypes.c_int, [_vp, _i32p, _i32p]),
    ("np_set_option", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32]),
    ("np_graph_relabeled", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double)]),
    ("np_dev_alloc", ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    ("np_dev_free", ctypes.c_int, [_vp, _vp]),
    ("np_dev_copy", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint64, ctypes.c_int32]),
    ("np_sync", ctypes.c_int, [_vp]),
    ("np_csr_validate", ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, _vp, _vp, _vp, _i32p]),
    ("np_graph_upload", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int64, _vp, _vp, _vp, ctypes.POINTER(_vp)]),
    ("np_graph_wrap_device", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int64, _vp, _vp, _vp,
                                            ctypes.POINTER(_vp)]),
    ("np_graph_info", ctypes.c_int, [_vp, _i32p, ctypes.POINTER(ctypes.c_int64)]),
    ("np_graph_download", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("np_graph_device_ptrs", ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    ("np_graph_free", ctypes.c_int, [_vp]),
    ("np_sssp", ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.POINTER(Sched), _vp, ctypes.POINTER(Stats)]),
    ("np_sssp_device", ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.POINTER(Sched), _vp,
                                      ctypes.POINTER(Stats)]),
    ("np_sssp_device_noalloc", ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.POINTER(Sched), _vp,
                                              ctypes.POINTER(Stats)]),
    ("np_bfs", ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.POINTER(Sched), _vp, ctypes.POINTER(Stats)]),
    ("np_bfs_device", ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.POINTER(Sched), _vp, ctypes.POINTER(Stats)]),
    ("np_batch_create", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(_vp), _vp, ctypes.c_int32,
                                       ctypes.POINTER(_vp)]),
    ("np_batch_solve", ctypes.c_int, [_vp, _vp, ctypes.POINTER(Sched), ctypes.POINTER(Stats)]),
    ("np_batch_get", ctypes.c_int, [_vp, _vp, ctypes.c_int32, _vp]),
    ("np_batch_device_dist", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(_vp)]),
    ("np_batch_layout", ctypes.c_int, [_vp, _i32p, _i32p]),
    ("np_batch_free", ctypes.c_int, [_vp]),
    ("np_reached_work", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(ctypes.c_int64)]),
    ("np_trace_get", ctypes.c_int, [_vp, ctypes.POINTER(TraceRec), ctypes.c_int32, _i32p]),
    ("np_model_set", ctypes.c_int, [_vp, _vp]),
    ("np_model_get", ctypes.c_int, [_vp, _vp]),
    ("np_features", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, _vp]),
    ("np_heuristic", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(Sched)]),
    ("np_model_train", ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                      ctypes.c_float, _vp]),
    ("np_model_loss_grad", ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_float, _vp, _vp]),
    ("np_model_forward", ctypes.c_int, [_vp, _vp, ctypes.c_int32, _vp]),
    ("np_step_begin", ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.POINTER(Stats)]),
    ("np_step_frontier", ctypes.c_int, [_vp, _vp, _i32p]),
    ("np_step_try", ctypes.c_int, [_vp, ctypes.POINTER(Sched), ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]),
    ("np_step_advance", ctypes.c_int, [_vp, ctypes.POINTER(Sched), ctypes.POINTER(ctypes.c_double)]),
    ("np_step_end", ctypes.c_int, [_vp, _vp]),
    ("np_gen_rmat", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64,
                                   ctypes.c_int32, ctypes.POINTER(_vp)]),
    ("np_gen_er", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                 ctypes.c_uint64, ctypes.POINTER(_vp)]),
    ("np_gen_grid", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                   ctypes.c_int32, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    ("np_gen_rmat_host", ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64,
                                        ctypes.c_int32, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                        ctypes.POINTER(_vp), _i32p, ctypes.POINTER(ctypes.c_int64)]),
    ("np_gen_er_host", ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_uint64, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    ("np_gen_grid_host", ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                        ctypes.c_int32, ctypes.c_uint64, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                        ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_int64)]),
    ("np_pick_sources", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, _vp]),
    ("np_host_free", None, [_vp]),
    ("np_read_graph", ctypes.c_int, [ctypes.c_int32, ctypes.c_char_p, _i32p, ctypes.POINTER(ctypes.c_int64), _i32p,
                                     ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    ("np_read_graph_stream", ctypes.c_int, [ctypes.c_int32, _vp, _i32p, ctypes.POINTER(ctypes.c_int64), _i32p,
                                            ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    ("np_bc", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32, _vp, ctypes.POINTER(Stats)]),
    ("np_color", ctypes.c_int, [_vp, _vp, ctypes.c_int32, _vp, ctypes.POINTER(Stats)]),
    ("np_pagerank", ctypes.c_int, [_vp, _vp, ctypes.c_int32, _vp, ctypes.POINTER(Stats)]),
    ("np_spmv", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.POINTER(Stats)]),
    ("np_csr_save", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, _i32p, _i32p,
                                   _i32p]),
    ("np_csr_load", ctypes.c_int, [ctypes.c_char_p, _i32p, ctypes.POINTER(ctypes.c_int64), _i32p,
                                   ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    ("np_autotune", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32, ctypes.POINTER(Sched), ctypes.POINTER(Sched),
                                   ctypes.POINTER(ctypes.c_double)]),
    ("np_batch_autotune", ctypes.c_int, [_vp, _vp, ctypes.POINTER(Sched), ctypes.POINTER(Sched),
                                         ctypes.POINTER(ctypes.c_double)]),
]
EXPORTS = [s[0] for s in _SIGS]


def load(path=LIB_PATH):
    """Loads libnpsssp.so (raises if it was not built: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libnpsssp.so not found at %s: build it with nested-parallelism_amd/build.py" % path)
    L = ctypes.CDLL(path)
    for name, res, args in _SIGS:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def last_error():
    return load().np_last_error().decode(errors="replace")


def _check(rc):
    if rc != 0:
        raise NpError(rc, last_error())


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None and a.size else None


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _host_array(p, count, dtype=np.int32):
    """Copies a malloc'd host array into numpy and frees it."""
    L = load()
    if count == 0:
        out = np.zeros(0, dtype)
    else:
        buf = (ctypes.c_int32 * count).from_address(p.value)
        out = np.frombuffer(buf, dtype=dtype, count=count).copy()
    L.np_host_free(p)
    return out


def default_sched(level=1):
    s = Sched()
    load().np_sched_default(level, ctypes.byref(s))
    return s


# ---------------------------------------------------------------------------------------
# host-only helpers (no device needed)
# ---------------------------------------------------------------------------------------
def csr_validate(row, col, w=None):
    """np_csr_validate: raises NpError (NP_ERR_ARG for a malformed CSR or a negative weight,
    NP_ERR_CAPACITY for a weight >= INF); returns the largest weight."""
    row, col = _i32(row), _i32(col)
    w = None if w is None else _i32(w)
    mx = ctypes.c_int32()
    _check(load().np_csr_validate(len(row) - 1, len(col), _ptr(row), _ptr(col), _ptr(w), ctypes.byref(mx)))
    return mx.value


def read_graph(fmt, path):
    """readInputDIMACS9/10/SLNDC + convertCSR (fmt 0/1/2; 3 = the binary CSR cache).
    Returns (row, col, w, source)."""
    L = load()
    n, m, s = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32()
    r, c, w = _vp(), _vp(), _vp()
    _check(L.np_read_graph(fmt, os.fsencode(path), ctypes.byref(n), ctypes.byref(m), ctypes.byref(s),
                           ctypes.byref(r), ctypes.byref(c), ctypes.byref(w)))
    return (_host_array(r, n.value + 1), _host_array(c, m.value), _host_array(w, m.value), s.value)


def save_csr(path, row, col, w=None, source=0):
    """Binary CSR cache (np_csr_save): row[n+1], col[m] and optional w[m] as int32."""
    L = load()
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    wp = None
    if w is not None:
        w = np.ascontiguousarray(w, np.int32)
        wp = w.ctypes.data_as(_i32p)
    _check(L.np_csr_save(os.fsencode(path), len(row) - 1, len(col), source, row.ctypes.data_as(_i32p),
                         col.ctypes.data_as(_i32p), wp))


def load_csr(path):
    """np_csr_load: returns (row, col, w, source); unit weights when the file has none."""
    L = load()
    n, m, s = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32()
    r, c, w = _vp(), _vp(), _vp()
    _check(L.np_csr_load(os.fsencode(path), ctypes.byref(n), ctypes.byref(m), ctypes.byref(s),
                         ctypes.byref(r), ctypes.byref(c), ctypes.byref(w)))
    return (_host_array(r, n.value + 1), _host_array(c, m.value), _host_array(w, m.value), s.value)


def gen_rmat_host(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, wmin=1, wmax=255, seed=1, permute=True):
    L = load()
    r, cc, w = _vp(), _vp(), _vp()
    n, m = ctypes.c_int32(), ctypes.c_int64()
    _check(L.np_gen_rmat_host(scale, edge_factor, a, b, c, wmin, wmax, seed, 1 if permute else 0,
                              ctypes.byref(r), ctypes.byref(cc), ctypes.byref(w), ctypes.byref(n), ctypes.byref(m)))
    return _host_array(r, n.value + 1), _host_array(cc, m.value), _host_array(w, m.value)


def gen_er_host(n, m, wmin=1, wmax=100, seed=1):
    L = load()
    r, c, w = _vp(), _vp(), _vp()
    _check(L.np_gen_er_host(n, m, wmin, wmax, seed, ctypes.byref(r), ctypes.byref(c), ctypes.byref(w)))
    return _host_array(r, n + 1), _host_array(c, m), _host_array(w, m)


def grid_long_range(rows, cols):
    """1% of the grid's directed edges (the long-range edge count used for the grid graphs)."""
    return 2 * (rows * (cols - 1) + cols * (rows - 1)) // 100


def gen_grid_host(rows, cols, n_long=None, wmin=1, wmax=1000, seed=5):
    L = load()
    if n_long is None:
        n_long = grid_long_range(rows, cols)
    r, c, w = _vp(), _vp(), _vp()
    m = ctypes.c_int64()
    _check(L.np_gen_grid_host(rows, cols, n_long, wmin, wmax, seed, ctypes.byref(r), ctypes.byref(c),
                              ctypes.byref(w), ctypes.byref(m)))
    n = rows * cols
    return _host_array(r, n + 1), _host_array(c, m.value), _host_array(w, m.value)


def pick_sources(row, k, seed):
    row = _i32(row)
    out = np.zeros(k, np.int32)
    _check(load().np_pick_sources(_ptr(row), len(row) - 1, k, seed, _ptr(out)))
    return out


def features(count, it, n):
    count = _i32(count)
    x = np.zeros(MLP_NIN, np.float32)
    _check(load().np_features(_ptr(count), it, n, _ptr(x)))
    return x


# ---------------------------------------------------------------------------------------
# device objects
# ---------------------------------------------------------------------------------------
_LIVE = weakref.WeakSet()


@atexit.register
def _close_live_contexts():
    """Contexts still open when the interpreter exits are destroyed here, while the HIP
    runtime (and any tool attached to it, such as a profiler) is still fully up: no HIP
    call of this library is left for the C runtime's exit handlers or for object
    finalisation during interpreter teardown."""
    for c in list(_LIVE):
        try:
            c.close()
        except Exception:
            pass


class Context:
    """np_ctx: one HIP device and one stream (sssp_wrapper.cu:58-140 prepare_gpu's device)."""

    def __init__(self, device=0):
        L = load()
        h = _vp()
        _check(L.np_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device
        _LIVE.add(self)

    def close(self):
        if self.h:
            load().np_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr):
        _check(load().np_set_stream(self.h, ctypes.c_void_p(stream_ptr)))

    def sync(self):
        _check(load().np_sync(self.h))

    @property
    def stream(self):
        return load().np_get_stream(self.h)

    def set_option(self, key, value):
        _check(load().np_set_option(self.h, int(key), int(value)))

    def device_info(self):
        a, b = ctypes.c_int32(), ctypes.c_int32()
        _check(load().np_device_info(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def sssp(self, g, src, sched=None):
        """Distances from src (SSSP_GPU, solution 6).  Returns (dist int32[n], Stats)."""
        dist = np.empty(g.n, np.int32)
        st = Stats()
        _check(load().np_sssp(self.h, g.h, int(src), ctypes.byref(sched or Sched()), _ptr(dist), ctypes.byref(st)))
        return dist, st

    def sssp_device(self, g, src, d_dist_ptr, sched=None):
        st = Stats()
        _check(load().np_sssp_device(self.h, g.h, int(src), ctypes.byref(sched or Sched()),
                                     ctypes.c_void_p(d_dist_ptr), ctypes.byref(st)))
        return st

    def sssp_device_noalloc(self, g, src, d_dist_ptr, sched=None):
        """sssp_device with no device allocation (np_sssp_device_noalloc): graphs of at most
        NOALLOC_MAX_VERTICES vertices; calls overlapping on one GPU race."""
        st = Stats()
        _check(load().np_sssp_device_noalloc(self.h, g.h, int(src), ctypes.byref(sched or Sched()),
                                             ctypes.c_void_p(d_dist_ptr), ctypes.byref(st)))
        return st

    def bfs(self, g, src, sched=None):
        """BFS levels from src (BFS_REC_GPU, solution 3).  Returns (level int32[n], Stats)."""
        lev = np.empty(g.n, np.int32)
        st = Stats()
        _check(load().np_bfs(self.h, g.h, int(src), ctypes.byref(sched or Sched(level=2)), _ptr(lev),
                             ctypes.byref(st)))
        return lev, st

    def bfs_device(self, g, src, d_level_ptr, sched=None):
        """BFS levels into device memory (np_bfs_device).  Returns Stats."""
        st = Stats()
        _check(load().np_bfs_device(self.h, g.h, int(src), ctypes.byref(sched or Sched(level=2)),
                                    ctypes.c_void_p(d_level_ptr), ctypes.byref(st)))
        return st

    def bc(self, g, srcs=None):
        """Betweenness centrality (BC_GPU, solution 7) summed over `srcs` (None: every vertex).
        Returns (bc float32[n], Stats)."""
        out = np.empty(g.n, np.float32)
        st = Stats()
        s = None if srcs is None else np.ascontiguousarray(srcs, np.int32)
        ns = 0 if s is None else len(s)
        if s is not None and ns == 0:
            raise ValueError("empty source list (None means every vertex)")
        _check(load().np_bc(self.h, g.h, _ptr(s), ns, _ptr(out), ctypes.byref(st)))
        return out, st

    def color(self, g, method=0):
        """Graph coloring of GRAPH_COLOR_GPU: method 0 dependency counting, 1 the reference's
        rounds.  Returns (color int32[n], Stats)."""
        out = np.empty(g.n, np.int32)
        st = Stats()
        _check(load().np_color(self.h, g.h, int(method), _ptr(out), ctypes.byref(st)))
        return out, st

    def pagerank(self, g, max_iters=0):
        """PageRank of PAGERANK_GPU (damping 0.5, total 1000, until no change > 1e-13).
        Returns (rank float32[n], Stats)."""
        out = np.empty(g.n, np.float32)
        st = Stats()
        _check(load().np_pagerank(self.h, g.h, int(max_iters), _ptr(out), ctypes.byref(st)))
        return out, st

    def spmv(self, g, data, x):
        """SpMV_GPU: y = A x over g's CSR with edge values data (float64).  Returns (y, Stats)."""
        data = np.ascontiguousarray(data, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        if len(data) != g.m or len(x) != g.n:
            raise ValueError("data must have m entries and x n")
        y = np.empty(g.n, np.float64)
        st = Stats()
        _check(load().np_spmv(self.h, g.h, _ptr(data), _ptr(x), _ptr(y), ctypes.byref(st)))
        return y, st

    def reached_work(self, g, d_dist_ptr, inf=INF):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        _check(load().np_reached_work(self.h, g.h, ctypes.c_void_p(d_dist_ptr), inf, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def trace(self):
        k = ctypes.c_int32()
        _check(load().np_trace_get(self.h, None, 0, ctypes.byref(k)))
        recs = (TraceRec * max(k.value, 1))()
        _check(load().np_trace_get(self.h, recs, k.value, ctypes.byref(k)))
        return [recs[i] for i in range(k.value)]

    # learned scheduler
    def model_set(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == MLP_NPARAM
        _check(load().np_model_set(self.h, _ptr(p)))

    def model_get(self):
        p = np.empty(MLP_NPARAM, np.float32)
        _check(load().np_model_get(self.h, _ptr(p)))
        return p

    def model_train(self, x, label, target, epochs, lr=1e-2, lam=0.1):
        x = np.ascontiguousarray(x, np.float32)
        label = _i32(label)
        target = np.ascontiguousarray(target, np.float32)
        loss = np.empty(epochs, np.float32)
        _check(load().np_model_train(self.h, _ptr(x), _ptr(label), _ptr(target), x.shape[0], epochs, lr, lam,
                                     _ptr(loss)))
        return loss

    def model_loss_grad(self, x, label, target, lam=0.1):
        x = np.ascontiguousarray(x, np.float32)
        label = _i32(label)
        target = np.ascontiguousarray(target, np.float32)
        loss = np.empty(1, np.float32)
        grad = np.empty(MLP_NPARAM, np.float32)
        _check(load().np_model_loss_grad(self.h, _ptr(x), _ptr(label), _ptr(target), x.shape[0], lam, _ptr(loss),
                                         _ptr(grad)))
        return float(loss[0]), grad

    def model_forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty((x.shape[0], 5), np.float32)
        _check(load().np_model_forward(self.h, _ptr(x), x.shape[0], _ptr(out)))
        return out

    # stepping (timing labels)
    def step_begin(self, g, src):
        _check(load().np_step_begin(self.h, g.h, int(src), None))

    def step_frontier(self):
        cnt = np.zeros(NCLASS, np.int32)
        done = ctypes.c_int32()
        _check(load().np_step_frontier(self.h, _ptr(cnt), ctypes.byref(done)))
        return cnt, bool(done.value)

    def step_try(self, sched, reps=1):
        ms = ctypes.c_double()
        _check(load().np_step_try(self.h, ctypes.byref(sched), reps, ctypes.byref(ms)))
        return ms.value

    def step_advance(self, sched):
        ms = ctypes.c_double()
        _check(load().np_step_advance(self.h, ctypes.byref(sched), ctypes.byref(ms)))
        return ms.value

    def step_end(self, n):
        d = np.empty(n, np.int32)
        _check(load().np_step_end(self.h, _ptr(d)))
        return d

    def autotune(self, g, srcs, base=None):
        srcs = _i32(srcs)
        best = Sched()
        ms = ctypes.c_double()
        _check(load().np_autotune(self.h, g.h, _ptr(srcs), len(srcs), ctypes.byref(base or Sched()),
                                  ctypes.byref(best), ctypes.byref(ms)))
        return best, ms.value


class DeviceBuffer:
    """Device memory owned through the library (np_dev_alloc); int32 by default."""

    def __init__(self, ctx, count, dtype=np.int32):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        p = _vp()
        _check(load().np_dev_alloc(ctx.h, self.count * self.dtype.itemsize, ctypes.byref(p)))
        self.ptr = p.value

    def to_host(self):
        out = np.empty(self.count, self.dtype)
        _check(load().np_dev_copy(self.ctx.h, _ptr(out), ctypes.c_void_p(self.ptr), out.nbytes, 1))
        return out

    def from_host(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size == self.count
        _check(load().np_dev_copy(self.ctx.h, ctypes.c_void_p(self.ptr), _ptr(a), a.nbytes, 0))

    def free(self):
        if self.ptr:
            load().np_dev_free(self.ctx.h, ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Graph:
    """np_graph: a device-resident CSR (vertexArray/edgeArray/weightArray of common/global.h)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle
        n, m = ctypes.c_int32(), ctypes.c_int64()
        _check(load().np_graph_info(handle, ctypes.byref(n), ctypes.byref(m)))
        self.n, self.m = n.value, m.value

    @property
    def relabeled(self):
        """(renumbered?, in-degree skew) of the solver layout."""
        sk = ctypes.c_double()
        r = load().np_graph_relabeled(self.h, ctypes.byref(sk))
        if r < 0:
            _check(r)
        return bool(r), sk.value

    @classmethod
    def from_csr(cls, ctx, row, col, w):
        row, col, w = _i32(row), _i32(col), _i32(w)
        h = _vp()
        _check(load().np_graph_upload(ctx.h, len(row) - 1, len(col), _ptr(row), _ptr(col), _ptr(w), ctypes.byref(h)))
        return cls(ctx, h)

    @classmethod
    def rmat(cls, ctx, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, wmin=1, wmax=255, seed=1, permute=True):
        h = _vp()
        _check(load().np_gen_rmat(ctx.h, scale, edge_factor, a, b, c, wmin, wmax, seed, 1 if permute else 0,
                                  ctypes.byref(h)))
        return cls(ctx, h)

    @classmethod
    def er(cls, ctx, n, m, wmin=1, wmax=100, seed=1):
        h = _vp()
        _check(load().np_gen_er(ctx.h, n, m, wmin, wmax, seed, ctypes.byref(h)))
        return cls(ctx, h)

    @classmethod
    def grid(cls, ctx, rows, cols, n_long=None, wmin=1, wmax=1000, seed=5):
        if n_long is None:
            n_long = grid_long_range(rows, cols)
        h = _vp()
        _check(load().np_gen_grid(ctx.h, rows, cols, n_long, wmin, wmax, seed, ctypes.byref(h)))
        return cls(ctx, h)

    def download(self):
        row = np.empty(self.n + 1, np.int32)
        col = np.empty(self.m, np.int32)
        w = np.empty(self.m, np.int32)
        _check(load().np_graph_download(self.ctx.h, self.h, _ptr(row), _ptr(col), _ptr(w)))
        return row, col, w

    def free(self):
        if self.h:
            load().np_graph_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    """np_batch: many (graph, source) instances solved in one launch."""

    def __init__(self, ctx, graphs, srcs, unit=False):
        self.ctx = ctx
        self.graphs = list(graphs)
        arr = (_vp * len(self.graphs))(*[g.h for g in self.graphs])
        srcs = _i32(srcs)
        h = _vp()
        _check(load().np_batch_create(ctx.h, len(self.graphs), arr, _ptr(srcs), 1 if unit else 0, ctypes.byref(h)))
        self.h = h
        self.count = len(self.graphs)

    def solve(self, sched=None):
        st = Stats()
        _check(load().np_batch_solve(self.ctx.h, self.h, ctypes.byref(sched or Sched()), ctypes.byref(st)))
        return st

    def get(self, i):
        d = np.empty(self.graphs[i].n, np.int32)
        _check(load().np_batch_get(self.ctx.h, self.h, i, _ptr(d)))
        return d

    def device_dist(self, i):
        p = _vp()
        _check(load().np_batch_device_dist(self.h, i, ctypes.byref(p)))
        return p.value

    def autotune(self, base=None):
        """Fastest schedule for this batch (np_batch_autotune): (Sched, ms per solve)."""
        best, ms = Sched(), ctypes.c_double()
        _check(load().np_batch_autotune(self.ctx.h, self.h, ctypes.byref(base or Sched()), ctypes.byref(best),
                                        ctypes.byref(ms)))
        return best, ms.value

    def layout(self):
        """(number of source-packed parts, instances in the concatenated part)."""
        a, b = ctypes.c_int32(), ctypes.c_int32()
        _check(load().np_batch_layout(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def free(self):
        if self.h:
            load().np_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
