"""GPU parity of the reference's other schedules (np_sched.scheme: SSSP_GPU solutions 0-5,
BFS_REC_GPU solutions 0-2), the input rules (negative weights, weights and distances at
INF = 2^30) and the lifetime rules of contexts and their objects.  Distances and levels are
bit-exact against the oracle and the reference fixtures."""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, hub_csr, load_golden, random_csr

pytestmark = pytest.mark.gpu

FIXTURES = sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN, "*.npz")))
SSSP_SCHEMES = [1, 2, 3, 4, 5, 6]   # NP_SCHEME_THREAD .. NP_SCHEME_WG_NESTED (solutions 0-5)
BFS_SCHEMES = [7, 8]                # NP_SCHEME_BFS_FLAT, NP_SCHEME_BFS_HIER (bfs-rec 0 and 2)
INF = 1 << 30


def _scheme(npa, scheme, drive=1, klo=6):
    return npa.Sched(level=1, log2_t_lo=klo, log2_t_hi=klo, drive=drive, scheme=scheme)


@pytest.mark.parametrize("name", FIXTURES)
def test_schemes_on_fixtures(npa, ctx, name):
    g = load_golden(name)
    G = npa.Graph.from_csr(ctx, g["row"], g["col"], g["w"])
    for k, s in enumerate(g["sources"]):
        for sc in SSSP_SCHEMES:
            for drive in (1, 2):
                d, st = ctx.sssp(G, s, _scheme(npa, sc, drive))
                assert np.array_equal(d, g["dist"][k]), (name, s, sc, drive)
                assert st.drive == drive
            d, st = ctx.sssp(G, s, _scheme(npa, sc, 0))     # persistent request: host loop
            assert np.array_equal(d, g["dist"][k]) and st.drive == 1
        for sc in SSSP_SCHEMES + BFS_SCHEMES:
            lev, _ = ctx.bfs(G, s, _scheme(npa, sc))
            assert np.array_equal(lev, g["level"][k]), (name, s, sc)
    G.free()


@pytest.mark.parametrize("klo", [0, 3, 6, 30])
def test_schemes_rmat_and_hubs(npa, ctx, orc, klo):
    """Power-law graph (heavy vertices in every scheme's hand-off) and rows of 4095..20000
    edges; the heavy threshold swept from every vertex heavy (2^0) to none (2^30)."""
    G = npa.Graph.rmat(ctx, 15, 16, wmin=1, wmax=255, seed=61)
    row, col, w = G.download()
    rng = np.random.default_rng(61)
    hrow, hcol = hub_csr(rng, 30000, 120000, [4095, 4096, 4097, 9000, 20000, 64, 63])
    hw = rng.integers(1, 100, len(hcol)).astype(np.int32)
    H = npa.Graph.from_csr(ctx, hrow, hcol, hw)
    for gr, host, srcs in ((G, (row, col, w), [int(s) for s in npa.pick_sources(row, 2, 61)]),
                           (H, (hrow, hcol, hw), [0, 4, 29999])):
        for s in srcs:
            ref, _ = orc.dijkstra(*host, s)
            lref = orc.bfs_levels(host[0], host[1], s)
            for sc in SSSP_SCHEMES:
                d, _ = ctx.sssp(gr, s, _scheme(npa, sc, 1, klo))
                assert np.array_equal(d, ref), (s, sc, klo)
                lev, _ = ctx.bfs(gr, s, _scheme(npa, sc, 2, klo))
                assert np.array_equal(lev, lref), (s, sc, klo)
            for sc in BFS_SCHEMES:
                lev, _ = ctx.bfs(gr, s, _scheme(npa, sc, 1, klo))
                assert np.array_equal(lev, lref), (s, sc, klo)
    G.free()
    H.free()


def test_schemes_high_diameter_and_edge_cases(npa, ctx, orc):
    G = npa.Graph.grid(ctx, 64, 48, None, 1, 1000, seed=62)
    row, col, w = G.download()
    ref, _ = orc.dijkstra(row, col, w, 100)
    # every reached vertex with out-edges is queued at least once (the source included)
    queued_min = int(np.count_nonzero((ref < INF) & (np.diff(row) > 0)))
    for sc in SSSP_SCHEMES:
        d, st = ctx.sssp(G, 100, _scheme(npa, sc))
        assert np.array_equal(d, ref), sc
        assert st.iterations > 1 and st.frontier_total >= queued_min, (sc, st)
    for sc in BFS_SCHEMES:
        lev, st = ctx.bfs(G, 100, _scheme(npa, sc))
        assert np.array_equal(lev, orc.bfs_levels(row, col, 100)), sc
    G.free()
    rng = np.random.default_rng(63)
    cases = [(np.array([0, 0], np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32), 0),
             (np.array([0, 0, 1], np.int32), np.array([0], np.int32), np.array([5], np.int32), 0)]
    r, c, ww = random_csr(rng, 3000, 9000, 0, 3)          # zero weights, self loops, duplicates
    cases.append((r, c, ww, 1))
    for r, c, ww, s in cases:
        E = npa.Graph.from_csr(ctx, r, c, ww)
        ref, _ = orc.dijkstra(r, c, ww, s)
        for sc in SSSP_SCHEMES:
            assert np.array_equal(ctx.sssp(E, s, _scheme(npa, sc))[0], ref), (len(r), sc)
        for sc in SSSP_SCHEMES + BFS_SCHEMES:
            assert np.array_equal(ctx.bfs(E, s, _scheme(npa, sc))[0], orc.bfs_levels(r, c, s)), (len(r), sc)
        E.free()


def test_scheme_rules(npa, ctx):
    G = npa.Graph.rmat(ctx, 8, 8, wmin=1, wmax=9, seed=64)
    for sc in BFS_SCHEMES:                       # BFS-only schemes on a weighted solve
        with pytest.raises(npa.NpError) as e:
            ctx.sssp(G, 0, _scheme(npa, sc))
        assert e.value.code == -1
    with pytest.raises(npa.NpError):
        ctx.sssp(G, 0, _scheme(npa, 9))
    B = npa.Batch(ctx, [G, G], [0, 1])
    with pytest.raises(npa.NpError) as e:
        B.solve(_scheme(npa, 1))
    assert e.value.code == -1
    B.free()
    G.free()


def test_dropin_solutions_select_distinct_schedules(npa, ctx, orc, tmp_path):
    """gpu-sssp -s 0..6 and gpu-bfs-rec -s 0..3 (the reference's drivers over SSSP_GPU /
    BFS_REC_GPU) produce the reference fixture's output for every solution."""
    import subprocess
    from conftest import PKG_DIR
    g = load_golden("d9_rand.npz")
    src = int(g["source"])
    k = list(g["sources"]).index(src)
    expect = ["From Node %d -> %d is %d" % (src + 1, i + 1, d) for i, d in enumerate(g["dist"][k])]
    for sol in range(7):
        out = tmp_path / ("c%d.txt" % sol)
        r = subprocess.run([os.path.join(PKG_DIR, "gpu-sssp"), "-f", "0", "-i", os.path.join(GOLDEN, str(g["path"])),
                            "-s", str(sol), "-o", str(out)], cwd=tmp_path, stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert out.read_text().splitlines() == expect, sol
    for sol in range(4):
        r = subprocess.run([os.path.join(PKG_DIR, "gpu-bfs-rec"), "-f", "0", "-i", os.path.join(GOLDEN, str(g["path"])),
                            "-s", str(sol)], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        assert "PASS: GPU bfs rec !" in r.stdout, sol


# ------------------------------------------------------------------ input rules
def _path_csr(weights):
    n = len(weights) + 1
    row = np.array(list(range(n - 1)) + [n - 1, n - 1], np.int32)[: n + 1]
    return row, np.arange(1, n, dtype=np.int32), np.asarray(weights, np.int32)


def test_negative_and_huge_weights_rejected(npa, ctx):
    row, col, w = _path_csr([4, -2, 7])
    with pytest.raises(npa.NpError) as e:
        npa.Graph.from_csr(ctx, row, col, w)
    assert e.value.code == -1
    row, col, w = _path_csr([4, INF])
    with pytest.raises(npa.NpError) as e:
        npa.Graph.from_csr(ctx, row, col, w)
    assert e.value.code == -4
    with pytest.raises(npa.NpError) as e:          # the device-side check of generated graphs
        npa.Graph.rmat(ctx, 8, 8, wmin=-3, wmax=5, seed=1)
    assert e.value.code == -1
    # device-resident arrays wrapped in place get the same device-side checks
    row, col, w = _path_csr([1, -5, 2])
    bufs = [npa.DeviceBuffer(ctx, len(a)) for a in (row, col, w)]
    for b, a in zip(bufs, (row, col, w)):
        b.from_host(a)
    import ctypes
    h = ctypes.c_void_p()
    rc = npa.load().np_graph_wrap_device(ctx.h, len(row) - 1, len(col), ctypes.c_void_p(bufs[0].ptr),
                                         ctypes.c_void_p(bufs[1].ptr), ctypes.c_void_p(bufs[2].ptr), ctypes.byref(h))
    assert rc == -1
    for b in bufs:
        b.free()


def test_distance_reaching_inf_is_reported(npa, ctx, orc):
    """A vertex whose shortest distance would be >= INF = 2^30: the reference's relaxation never
    stores such a value and reports the vertex unreached; here the solve fails with
    NP_ERR_CAPACITY.  One below INF is exact."""
    row, col, w = _path_csr([1 << 29, (1 << 29) - 1])
    G = npa.Graph.from_csr(ctx, row, col, w)
    for lv in (0, 1, 2):
        d, _ = ctx.sssp(G, 0, npa.Sched(level=lv))
        assert list(d) == [0, 1 << 29, INF - 1]
    G.free()
    row, col, w = _path_csr([1 << 29, 1 << 29, 5])
    G = npa.Graph.from_csr(ctx, row, col, w)
    for sch in (npa.Sched(level=1), npa.Sched(level=2, drive=2), _scheme(npa, 1)):
        with pytest.raises(npa.NpError) as e:
            ctx.sssp(G, 0, sch)
        assert e.value.code == -4
    d, _ = ctx.sssp(G, 2, npa.Sched(level=1))      # from vertex 2 every distance is small
    assert list(d) == [INF, INF, 0, 5]
    B = npa.Batch(ctx, [G, G], [0, 2])              # source-packed
    with pytest.raises(npa.NpError) as e:
        B.solve(npa.Sched(level=1))
    assert e.value.code == -4
    B.free()
    E = npa.Graph.rmat(ctx, 6, 4, wmin=1, wmax=9, seed=3)
    ctx.set_option(npa.CTX_PACK, 0)
    try:
        B = npa.Batch(ctx, [E, G], [0, 0])          # concatenated
        for drive in (0, 2):
            with pytest.raises(npa.NpError) as e:
                B.solve(npa.Sched(level=1, drive=drive))
            assert e.value.code == -4
        B.free()
    finally:
        ctx.set_option(npa.CTX_PACK, 1)
    E.free()
    G.free()


# ------------------------------------------------------------------ lifetimes
def test_context_closed_before_its_objects(npa):
    """Closing a context first releases what its graphs, batches and buffers hold on the
    device; their later frees only delete host objects and calls on them fail cleanly."""
    c = npa.Context(0)
    G = npa.Graph.rmat(c, 10, 8, wmin=1, wmax=9, seed=5)
    B = npa.Batch(c, [G, G], [0, 1])
    B.solve(npa.Sched(level=1))
    buf = npa.DeviceBuffer(c, 1024)
    c.close()
    c2 = npa.Context(0)
    with pytest.raises(npa.NpError):                 # the graph is detached from any context
        npa._check(npa.load().np_graph_download(c2.h, G.h, None, None, None))
    with pytest.raises(npa.NpError):
        c2.sssp(G, 0)
    B.free()
    G.free()
    buf.free()
    G2 = npa.Graph.rmat(c2, 10, 8, wmin=1, wmax=9, seed=5)  # the device is usable afterwards
    d, _ = c2.sssp(G2, 0)
    assert d[0] == 0
    G2.free()
    c2.close()
