"""The benchmark configurations at full size, on the GPU: the same inputs bench.py builds
(its build_workload), every instance of every mode checked with the shortest-path
certificate (exact for positive weights) and one instance per configuration re-solved by the
oracle's Dijkstra, bit-exact.

  C2  R-MAT s22 (4,194,304 vertices, 67,108,864 edges), 8 sources (seed 2): block, warp and
      grid consolidation per source, and the source-packed batch
  C3  R-MAT s24 (16,777,216 / 268,435,456), 16 sources (seed 3): the plain fixed-threshold
      schedule, the learned scheduler, the source-packed batch (with and without the hub hold)
      and the learned batch; BFS levels of the same sources, alone and source-packed
  C4  R-MAT s20 x 64 sources (seed 4) + 256 R-MAT graphs of scale U{12..18} (seeds
      1000-1255): the concatenated launch and the packed + concatenated batch
  C5  4096 x 4096 grid + 1% long-range edges (seed 5), 32 sources: the batch with a priority
      window, learned and fixed
"""
import importlib.util
import os
import types

import numpy as np
import pytest

from conftest import PKG_DIR, ROOT

pytestmark = pytest.mark.gpu


def _bench():
    spec = importlib.util.spec_from_file_location("np_bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _workload(npa, ctx, cfg):
    args = types.SimpleNamespace(scale=24, sources=16)
    return _bench().build_workload(npa, ctx, cfg, args)


def _certify(orc, host, src, dist, what):
    code, bad = orc.check_sssp(*host, src, dist)
    assert code == 0, "%s: %s at vertex %d" % (what, orc.CHECK_CODES[code], bad)


def _load_model(npa, ctx):
    ctx.model_set(np.fromfile(os.path.join(PKG_DIR, "sched_mlp.f32"), dtype=np.float32))


# --------------------------------------------------------------------------- C2
@pytest.mark.timeout(400)
def test_c2_rmat_s22_levels_and_packed(npa, ctx, orc):
    W = _workload(npa, ctx, "c2")
    G = W.graphs[0]
    assert (G.n, G.m) == (4194304, 67108864)
    host = G.download()
    srcs = [s for _, s in W.inst]
    assert len(srcs) == 8
    plain = []
    for s in srcs:
        d1, _ = ctx.sssp(G, s, npa.Sched(level=1))
        _certify(orc, host, s, d1, "C2 block level source %d" % s)
        for lv in (0, 2):
            d, _ = ctx.sssp(G, s, npa.Sched(level=lv))
            assert np.array_equal(d, d1), ("C2 level", lv, s)
        plain.append(d1)
    ref, _ = orc.dijkstra(*host, srcs[0])
    assert np.array_equal(plain[0], ref)
    B = npa.Batch(ctx, [G] * len(srcs), srcs)
    assert B.layout() == (1, 0)
    for sch in (npa.Sched(level=2, log2_t_lo=5, log2_t_hi=12, drive=npa.DRIVE_GRAPH),
                npa.Sched(level=1, drive=npa.DRIVE_PERSISTENT),
                npa.Sched(level=2, log2_t_lo=4, log2_t_hi=8, drive=npa.DRIVE_GRAPH, pack_hold=10)):
        B.solve(sch)
        for i in range(len(srcs)):
            assert np.array_equal(B.get(i), plain[i]), ("C2 packed", sch, i)
    B.free()
    G.free()


# --------------------------------------------------------------------------- C3
@pytest.fixture(scope="module")
def c3(npa, ctx, orc):
    W = _workload(npa, ctx, "c3")
    G = W.graphs[0]
    assert (G.n, G.m) == (16777216, 268435456)
    host = G.download()
    srcs = [s for _, s in W.inst]
    assert len(srcs) == 16
    plain = []
    for s in srcs:
        d, _ = ctx.sssp(G, s, npa.Sched(level=1))
        _certify(orc, host, s, d, "C3 plain source %d" % s)
        plain.append(d)
    yield types.SimpleNamespace(G=G, host=host, srcs=srcs, plain=plain)
    G.free()


@pytest.mark.timeout(400)
def test_c3_rmat_s24_plain_vs_dijkstra(npa, orc, c3):
    ref, _ = orc.dijkstra(*c3.host, c3.srcs[0])
    assert np.array_equal(c3.plain[0], ref)


@pytest.mark.timeout(300)
def test_c3_rmat_s24_learned(npa, ctx, c3):
    _load_model(npa, ctx)
    for drive in (npa.DRIVE_PERSISTENT, npa.DRIVE_GRAPH):
        sch = npa.Sched(level=npa.LEVEL_LEARNED, drive=drive)
        for s, ref in zip(c3.srcs, c3.plain):
            d, _ = ctx.sssp(c3.G, s, sch)
            assert np.array_equal(d, ref), ("C3 learned", drive, s)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("learned", [False, True])
def test_c3_rmat_s24_batched(npa, ctx, c3, learned):
    if learned:
        _load_model(npa, ctx)
    B = npa.Batch(ctx, [c3.G] * len(c3.srcs), c3.srcs)
    assert B.layout() == (1, 0)
    lv = npa.LEVEL_LEARNED if learned else 2
    for sch in (npa.Sched(level=lv, log2_t_lo=6, log2_t_hi=12, drive=npa.DRIVE_GRAPH),
                npa.Sched(level=lv, log2_t_lo=5, log2_t_hi=10, drive=npa.DRIVE_PERSISTENT),
                npa.Sched(level=lv, log2_t_lo=4, log2_t_hi=10, drive=npa.DRIVE_GRAPH, pack_hold=11),
                npa.Sched(level=lv, log2_t_lo=4, log2_t_hi=10, drive=npa.DRIVE_PERSISTENT, pack_hold=10)):
        st = B.solve(sch)
        assert st.iterations > 0
        for i, ref in enumerate(c3.plain):
            assert np.array_equal(B.get(i), ref), ("C3 batched", learned, sch, i)
    B.free()


@pytest.mark.timeout(300)
def test_c3_rmat_s24_bfs_levels(npa, ctx, orc, c3):
    """BFS levels (BFS_REC_GPU's result) on the C3 graph against the oracle's queue BFS: one
    source alone at block level, and the 16 sources source-packed with and without the hub hold."""
    from concurrent.futures import ThreadPoolExecutor
    row, col, _ = c3.host
    with ThreadPoolExecutor(max(1, min(16, os.cpu_count() or 1))) as ex:
        refs = list(ex.map(lambda s: orc.bfs_levels(row, col, s), c3.srcs))
    lev, _ = ctx.bfs(c3.G, c3.srcs[0], npa.Sched(level=1))
    assert np.array_equal(lev, refs[0])
    Bu = npa.Batch(ctx, [c3.G] * len(c3.srcs), c3.srcs, unit=True)
    assert Bu.layout() == (1, 0)
    for sch in (npa.Sched(level=2, log2_t_lo=4, log2_t_hi=12, drive=npa.DRIVE_GRAPH, pack_hold=7),
                npa.Sched(level=2, log2_t_lo=4, log2_t_hi=12, drive=npa.DRIVE_PERSISTENT)):
        st = Bu.solve(sch)
        assert st.iterations > 0
        for i, ref in enumerate(refs):
            assert np.array_equal(Bu.get(i), ref), ("C3 BFS packed", sch, i)
    Bu.free()


# --------------------------------------------------------------------------- C4
@pytest.mark.timeout(500)
def test_c4_many_sources_and_uneven_graphs(npa, ctx, orc):
    W = _workload(npa, ctx, "c4")
    assert len(W.graphs) == 257 and len(W.inst) == 320
    assert (W.graphs[0].n, W.graphs[0].m) == (1048576, 16777216)
    scales = [int(np.log2(g.n)) for g in W.graphs[1:]]
    assert min(scales) >= 12 and max(scales) <= 18 and len(set(scales)) == 7
    hosts = [g.download() for g in W.graphs]
    graphs = [W.graphs[gi] for gi, _ in W.inst]
    srcs = [s for _, s in W.inst]
    ctx.set_option(npa.CTX_PACK, 0)
    try:
        Bc = npa.Batch(ctx, graphs, srcs)
    finally:
        ctx.set_option(npa.CTX_PACK, 1)
    assert Bc.layout() == (0, 320)
    Bc.solve(npa.Sched(level=2, log2_t_lo=5, log2_t_hi=12, drive=npa.DRIVE_GRAPH))
    concat = []
    for i, (gi, s) in enumerate(W.inst):
        d = Bc.get(i)
        _certify(orc, hosts[gi], s, d, "C4 concatenated instance %d" % i)
        concat.append(d)
    Bc.free()
    for i in (0, 64, 319):  # the s20 graph, the first and last uneven graphs
        gi, s = W.inst[i]
        ref, _ = orc.dijkstra(*hosts[gi], s)
        assert np.array_equal(concat[i], ref), i
    B = npa.Batch(ctx, graphs, srcs)
    assert B.layout() == (2, 256)
    for sch in (npa.Sched(level=2, log2_t_lo=5, log2_t_hi=10, drive=npa.DRIVE_GRAPH),
                npa.Sched(level=1, drive=npa.DRIVE_PERSISTENT)):
        B.solve(sch)
        for i in range(len(srcs)):
            assert np.array_equal(B.get(i), concat[i]), ("C4 batched", sch, i)
    B.free()
    for g in W.graphs:
        g.free()


# --------------------------------------------------------------------------- C5
@pytest.mark.timeout(400)
def test_c5_grid_batched_learned_and_window(npa, ctx, orc):
    W = _workload(npa, ctx, "c5")
    G = W.graphs[0]
    assert G.n == 4096 * 4096 and G.m > 67092480
    host = G.download()
    srcs = [s for _, s in W.inst]
    assert len(srcs) == 32
    B = npa.Batch(ctx, [G] * len(srcs), srcs)
    assert B.layout() == (1, 0)
    B.solve(npa.Sched(level=1, log2_t_lo=6, log2_t_hi=12, drive=npa.DRIVE_GRAPH, delta=4004))
    first = []
    for i, s in enumerate(srcs):
        d = B.get(i)
        _certify(orc, host, s, d, "C5 batched window source %d" % s)
        first.append(d)
    ref, _ = orc.dijkstra(*host, srcs[0])
    assert np.array_equal(first[0], ref)
    _load_model(npa, ctx)
    # a window turns the hub hold off: the pack_hold of the last schedule changes nothing
    for sch in (npa.Sched(level=npa.LEVEL_LEARNED, log2_t_lo=6, log2_t_hi=12, drive=npa.DRIVE_GRAPH, delta=4004),
                npa.Sched(level=npa.LEVEL_LEARNED, drive=npa.DRIVE_PERSISTENT),
                npa.Sched(level=2, log2_t_lo=4, log2_t_hi=14, drive=npa.DRIVE_GRAPH, delta=4004, pack_hold=9)):
        B.solve(sch)
        for i in range(len(srcs)):
            assert np.array_equal(B.get(i), first[i]), ("C5 learned batch", sch, i)
    B.free()
    d, _ = ctx.sssp(G, srcs[5], npa.Sched(level=1))
    assert np.array_equal(d, first[5])
    G.free()
