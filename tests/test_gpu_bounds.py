"""The bounds-checked build (libnpsssp_bounds.so, NP_SSSP_LIB=bounds): every frontier-queue,
source-mask and delayed-buffer write is checked against its segment; a violation is skipped and
fails the solve with NP_ERR_BOUNDS.  One child process runs the solves that fill those buffers
-- every drive and frontier mode, the consolidation levels, the schemes of the reference's
solutions (delayed buffers included), the source-packed batch with and without a priority
window and with the hub hold, the concatenated batch and BFS -- and checks every result against
the oracle, so a passing run means no write left its segment and the checks change no result.
A second child does the same at full size: the R-MAT s24 graph and 16 sources of the headline
benchmark, source-packed with the hub hold on the graph and persistent drives, SSSP and BFS."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

CHILD = r'''
import os, sys
import numpy as np
ROOT = sys.argv[1]
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from conftest import load_pkg, random_csr
import oracle as orc
npa = load_pkg()
L = npa.load()
assert b"bounds-checked" in L.np_version(), L.np_version()
ctx = npa.Context(0)
G = npa.Graph.rmat(ctx, 14, 16, wmin=1, wmax=255, seed=61)
row, col, w = G.download()
srcs = [int(s) for s in npa.pick_sources(row, 3, 61)]
refs = [orc.dijkstra(row, col, w, s)[0] for s in srcs]
n_solves = 0
for drive in (0, 1, 2):
    for fr in (0, 1):
        for lv, lo, hi in ((0, 6, 6), (1, 6, 6), (2, 3, 8), (1, 4, 12)):
            d, _ = ctx.sssp(G, srcs[0], npa.Sched(level=lv, log2_t_lo=lo, log2_t_hi=hi, drive=drive, frontier=fr))
            assert np.array_equal(d, refs[0]), (drive, fr, lv, lo, hi)
            n_solves += 1
    d, _ = ctx.sssp(G, srcs[1], npa.Sched(level=1, drive=drive, delta=300))
    assert np.array_equal(d, refs[1]), ("window", drive)
    lev, _ = ctx.bfs(G, srcs[1], npa.Sched(level=2, drive=drive))
    assert np.array_equal(lev, orc.bfs_levels(row, col, srcs[1])), ("bfs", drive)
    n_solves += 2
for scheme in range(1, 7):
    d, _ = ctx.sssp(G, srcs[2], npa.Sched(level=1, log2_t_lo=4, log2_t_hi=4, drive=1, scheme=scheme))
    assert np.array_equal(d, refs[2]), ("scheme", scheme)
    n_solves += 1
rng = np.random.default_rng(61)
E = npa.Graph.from_csr(ctx, *random_csr(rng, 3000, 24000, 1, 60)[:3])
he = E.download()
grid = npa.Graph.grid(ctx, 90, 70, None, 1, 1000, seed=61)
hg = grid.download()
bsrc = [int(s) for s in npa.pick_sources(row, 20, 62)]
gsrc = [int(s) for s in npa.pick_sources(hg[0], 9, 63)]
graphs = [G] * len(bsrc) + [grid] * len(gsrc) + [E, E]
allsrc = bsrc + gsrc + [0, 2999]
hosts = [(row, col, w)] * len(bsrc) + [hg] * len(gsrc) + [he, he]
brefs = [orc.dijkstra(*h, s)[0] for h, s in zip(hosts, allsrc)]
for pack in (1, 0):
    ctx.set_option(npa.CTX_PACK, pack)
    B = npa.Batch(ctx, graphs, allsrc)
    ctx.set_option(npa.CTX_PACK, 1)
    for drive in (0, 1, 2):
        for delta in (0, 500):
            B.solve(npa.Sched(level=2, log2_t_lo=4, log2_t_hi=8, drive=drive, delta=delta))
            for i, ref in enumerate(brefs):
                assert np.array_equal(B.get(i), ref), (pack, drive, delta, i)
            n_solves += 1
        for hold in ((1, 5) if pack else ()):
            B.solve(npa.Sched(level=2, log2_t_lo=4, log2_t_hi=8, drive=drive, pack_hold=hold))
            for i, ref in enumerate(brefs):
                assert np.array_equal(B.get(i), ref), ("hold", hold, drive, i)
            n_solves += 1
    B.free()
E.free(); grid.free(); G.free(); ctx.close()
print("bounds-checked solves exact:", n_solves, flush=True)
'''


@pytest.mark.timeout(300)
def test_bounds_checked_build_runs_clean():
    env = dict(os.environ, NP_SSSP_LIB="bounds")
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=280)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "bounds-checked solves exact:" in r.stdout, r.stdout[-3000:]


CHILD_FULL = r'''
import importlib.util, os, sys, types
from concurrent.futures import ThreadPoolExecutor
import numpy as np
ROOT = sys.argv[1]
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from conftest import load_pkg
import oracle as orc
npa = load_pkg()
assert b"bounds-checked" in npa.load().np_version()
spec = importlib.util.spec_from_file_location("np_bench", os.path.join(ROOT, "bench.py"))
bench = importlib.util.module_from_spec(spec)
spec.loader.exec_module(bench)
ctx = npa.Context(0)
W = bench.build_workload(npa, ctx, "c3", types.SimpleNamespace(scale=24, sources=16))
G = W.graphs[0]
srcs = [s for _, s in W.inst]
host = G.download()
B = npa.Batch(ctx, [G] * len(srcs), srcs)
first = None
for sch in (npa.Sched(level=2, log2_t_lo=4, log2_t_hi=10, drive=npa.DRIVE_GRAPH, pack_hold=11),
            npa.Sched(level=2, log2_t_lo=4, log2_t_hi=10, drive=npa.DRIVE_PERSISTENT, pack_hold=10)):
    B.solve(sch)
    out = [B.get(i) for i in range(len(srcs))]
    if first is None:
        for s, d in zip(srcs, out):
            code, bad = orc.check_sssp(*host, s, d)
            assert code == 0, (s, orc.CHECK_CODES[code], bad)
        first = out
    else:
        assert all(np.array_equal(a, b) for a, b in zip(first, out)), sch
B.free()
Bu = npa.Batch(ctx, [G] * len(srcs), srcs, unit=True)
Bu.solve(npa.Sched(level=2, log2_t_lo=4, log2_t_hi=12, drive=npa.DRIVE_GRAPH, pack_hold=7))
with ThreadPoolExecutor(4) as ex:
    refs = list(ex.map(lambda s: orc.bfs_levels(host[0], host[1], s), srcs[:4]))
for i, ref in enumerate(refs):
    assert np.array_equal(Bu.get(i), ref), ("bfs", i)
Bu.free()
G.free()
ctx.close()
print("bounds-checked full-size solves exact", flush=True)
'''


@pytest.mark.timeout(300)
def test_bounds_checked_build_full_size_packed_hold():
    env = dict(os.environ, NP_SSSP_LIB="bounds")
    r = subprocess.run([sys.executable, "-c", CHILD_FULL, ROOT], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=280)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "bounds-checked full-size solves exact" in r.stdout, r.stdout[-3000:]
