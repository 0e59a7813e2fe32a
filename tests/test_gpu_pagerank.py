"""np_pagerank (PAGERANK_GPU) against the oracle: float64 iterations on both sides, the GPU's
ranks returned as float32; bar |gpu - oracle| <= 1e-5 * |oracle| + 1e-9.  Both runs stop past
convergence: the oracle's ranks after min(gpu, oracle) iterations equal its final ranks to 1e-12
relative.  The counts agree within one where rounding cannot reach the 1e-13 stop (16 ulp of
the largest rank below it); with hub ranks ~140 (1e-13 is 3.5 ulp) the last changes are
summation-order noise and the counts may differ by a few."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG_DIR, hub_csr, load_golden, random_csr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["rmat", "er", "grid", "multigraph", "hubs"])
def test_pagerank_matches_oracle(npa, ctx, orc, kind):
    if kind == "hubs":                                           # in-degrees 4095..20000: chunked pulls
        row, col = hub_csr(np.random.default_rng(76), 6000, 30000, [4095, 4096, 4097, 9000, 20000], inward=True)
        G = npa.Graph.from_csr(ctx, row, col, np.ones(len(col), np.int32))
    elif kind == "rmat":
        G = npa.Graph.rmat(ctx, 13, 16, wmin=1, wmax=1, seed=71)
    elif kind == "er":
        G = npa.Graph.er(ctx, 5000, 40000, 1, 1, seed=72)
    elif kind == "grid":
        G = npa.Graph.grid(ctx, 50, 60, None, 1, 1, seed=73)
    else:
        rng = np.random.default_rng(74)
        row, col, _ = random_csr(rng, 3000, 9000, 1, 1)         # dangling vertices, loops
        col[::4] = col[::4] % 3                                  # heavy in-degree at 0..2
        G = npa.Graph.from_csr(ctx, row, col, np.ones(len(col), np.int32))
    row, col, _ = G.download()
    ref, it = orc.pagerank(row, col)
    got, st = ctx.pagerank(G)
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-9), float(np.max(np.abs(got - ref) / ref))
    early, _ = orc.pagerank(row, col, max_iters=min(st.iterations, it))
    assert np.allclose(early, ref, rtol=1e-12, atol=0), (st.iterations, it)
    if float(ref.max()) * 2.0 ** -52 * 16 < 1e-13:
        assert abs(st.iterations - it) <= 1, (st.iterations, it)
    assert abs(float(got.astype(np.float64).sum()) - 1000.0) < 1e-2
    G.free()


def test_pagerank_without_edges(npa, ctx, orc):
    """Every vertex dangles: the dangling term alone keeps every rank at 1000 / n, so the first
    iteration changes nothing and the run stops after it."""
    n = 9
    G = npa.Graph.from_csr(ctx, np.zeros(n + 1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32))
    got, st = ctx.pagerank(G)
    assert np.allclose(got, 1000.0 / n, rtol=1e-6)
    ref, it = orc.pagerank(np.zeros(n + 1, np.int32), np.zeros(0, np.int32))
    assert st.iterations == it == 1
    G.free()


def test_pagerank_iteration_cap(npa, ctx):
    G = npa.Graph.rmat(ctx, 10, 8, wmin=1, wmax=1, seed=75)
    with pytest.raises(npa.NpError) as e:
        ctx.pagerank(G, max_iters=2)
    assert e.value.code == -5
    G.free()


def test_gpu_pagerank_cli_matches_oracle(tmp_path, orc):
    """gpu-pagerank (solution 6): invertGraph, PAGERANK_GPU, outputRank's "%0.10lf" lines."""
    g = load_golden("d9_rand.npz")
    out = tmp_path / "ranks.txt"
    r = subprocess.run([os.path.join(PKG_DIR, "gpu-pagerank"), "-f", "0", "-i", os.path.join(GOLDEN, str(g["path"])),
                        "-s", "6", "-o", str(out)], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = out.read_text().splitlines()
    assert [l.split(":")[0] for l in lines] == ["Node %d" % (i + 1) for i in range(len(g["row"]) - 1)]
    got = np.array([float(l.split(":")[1]) for l in lines])
    ref, _ = orc.pagerank(g["row"], g["col"])
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-9)
