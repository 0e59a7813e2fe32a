"""np_sssp_device_noalloc (solver state in the library's static device array, no device
allocation) against np_sssp_device, through the C ABI: bit-exact on consecutive calls, the
second on a smaller graph (the array keeps the first call's contents), over levels, drives
and a scheme with a delayed buffer; graphs above the array's limit are rejected."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(npa, ctx, g, src, d_ptr, sched):
    st = npa.Stats()
    rc = npa.load().np_sssp_device_noalloc(ctx.h, g.h, int(src), ctypes.byref(sched), ctypes.c_void_p(d_ptr),
                                           ctypes.byref(st))
    return rc, st


def test_noalloc_matches_sssp_device(npa, ctx, orc):
    big = npa.Graph.rmat(ctx, 16, 16, wmin=1, wmax=255, seed=71)      # renumbered (skewed)
    small = npa.Graph.er(ctx, 3000, 24000, 1, 100, seed=72)             # caller's ids
    assert big.relabeled[0]
    n = big.n
    ref_buf = npa.DeviceBuffer(ctx, n)
    na_buf = npa.DeviceBuffer(ctx, n)
    scheds = [npa.Sched(level=1), npa.Sched(level=2, log2_t_lo=5, log2_t_hi=10, drive=npa.DRIVE_HOSTLOOP),
              npa.Sched(level=0, drive=npa.DRIVE_GRAPH, delta=64),
              npa.Sched(level=1, drive=npa.DRIVE_HOSTLOOP, scheme=npa.SCHEME_GLOBAL_BUFFER)]
    hosts = {id(big): big.download(), id(small): small.download()}
    for sch in scheds:
        for g in (big, small):          # two calls in a row, the second on a smaller graph
            row = hosts[id(g)][0]
            src = int(npa.pick_sources(row, 1, 73)[0])
            na_buf.from_host(np.full(n, -7, np.int32))
            ctx.sssp_device(g, src, ref_buf.ptr, sch)
            rc, st = _noalloc(npa, ctx, g, src, na_buf.ptr, sch)
            assert rc == 0, npa.last_error()
            assert st.status == 0 and st.iterations > 0
            if sch.drive == npa.DRIVE_GRAPH:
                assert st.drive == npa.DRIVE_HOSTLOOP
            got = na_buf.to_host()
            assert np.array_equal(got[:g.n], ref_buf.to_host()[:g.n]), (sch, g.n)
            assert np.all(got[g.n:] == -7)              # nothing written past the graph
            if sch.level == 1 and sch.scheme == 0:
                assert np.array_equal(got[:g.n], orc.dijkstra(*hosts[id(g)], src)[0])
    ref_buf.free()
    na_buf.free()
    big.free()
    small.free()


def test_noalloc_limits(npa, ctx):
    nmax = npa.NOALLOC_MAX_VERTICES
    row = np.zeros(nmax + 2, np.int32)                 # nmax + 1 vertices, no edges
    G = npa.Graph.from_csr(ctx, row, np.zeros(0, np.int32), np.zeros(0, np.int32))
    buf = npa.DeviceBuffer(ctx, nmax + 1)
    rc, _ = _noalloc(npa, ctx, G, 0, buf.ptr, npa.Sched(level=1))
    assert rc == -4 and "exceeds" in npa.last_error()
    G.free()
    row = np.zeros(nmax + 1, np.int32)                 # exactly nmax vertices: accepted
    G = npa.Graph.from_csr(ctx, row, np.zeros(0, np.int32), np.zeros(0, np.int32))
    rc, st = _noalloc(npa, ctx, G, nmax - 1, buf.ptr, npa.Sched(level=1))
    assert rc == 0, npa.last_error()
    d = buf.to_host()[:nmax]
    assert d[nmax - 1] == 0 and np.all(d[:nmax - 1] == npa.INF)
    rc, _ = _noalloc(npa, ctx, G, nmax, buf.ptr, npa.Sched(level=1))   # source out of range
    assert rc == -1
    rc, _ = _noalloc(npa, ctx, G, 0, 0, npa.Sched(level=1))            # no output buffer
    assert rc == -1
    rc, _ = _noalloc(npa, ctx, G, 0, buf.ptr, npa.Sched(level=1, scheme=npa.SCHEME_BFS_FLAT))
    assert rc == -1
    G.free()
    buf.free()
