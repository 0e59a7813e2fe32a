"""np_bc (betweenness centrality, BC_GPU solution 7) against the oracle.

The GPU accumulates path counts and dependencies in double (parallel atomics: the summation
order varies) and returns float32, as the reference's float *bc; the bar is
|gpu - oracle| <= 1e-5 * |oracle| + 1e-5."""
import numpy as np
import pytest

from conftest import random_csr

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-5, 1e-5


def _graph(npa, ctx, row, col):
    return npa.Graph.from_csr(ctx, row, col, np.ones(len(col), np.int32))


@pytest.mark.parametrize("n", [40, 300])   # 300: depth 299, forward levels in many growing runs
def test_bc_directed_path_closed_form(npa, ctx, n):
    row = np.minimum(np.arange(n + 1), n - 1).astype(np.int32)
    col = np.arange(1, n, dtype=np.int32)
    G = _graph(npa, ctx, row, col)
    bc, st = ctx.bc(G)
    assert np.allclose(bc, [v * (n - 1 - v) for v in range(n)], rtol=RTOL, atol=ATOL)
    assert st.iterations > 0
    G.free()


@pytest.mark.parametrize("kind", ["rmat", "er", "grid", "multigraph"])
def test_bc_every_vertex_matches_oracle(npa, ctx, orc, kind):
    """All sources (32 per batch): R-MAT with hubs, ER, a grid, and a multigraph with parallel
    edges, self loops and isolated vertices."""
    if kind == "rmat":
        G = npa.Graph.rmat(ctx, 10, 16, wmin=1, wmax=1, seed=51)
    elif kind == "er":
        G = npa.Graph.er(ctx, 700, 5000, 1, 1, seed=52)
    elif kind == "grid":
        G = npa.Graph.grid(ctx, 24, 30, None, 1, 1, seed=53)
    else:
        rng = np.random.default_rng(54)
        row, col, _ = random_csr(rng, 500, 3000, 1, 1)
        col[::7] = col[::7] % 20              # many parallel edges into a few vertices
        G = _graph(npa, ctx, row, col)
    row, col, _ = G.download()
    bc, _ = ctx.bc(G)
    ref = orc.bc(row, col)
    assert np.allclose(bc, ref, rtol=RTOL, atol=ATOL), float(np.max(np.abs(bc - ref)))
    G.free()


def test_bc_source_subsets(npa, ctx, orc):
    """Batch sizes 1, 5 (rows of 8 slots), 32 and 33 (a second batch of one), duplicate sources,
    a degree-0 source."""
    G = npa.Graph.rmat(ctx, 12, 8, wmin=1, wmax=1, seed=55)
    row, col, _ = G.download()
    deg = np.diff(row)
    zero = int(np.flatnonzero(deg == 0)[0])
    picks = [int(s) for s in npa.pick_sources(row, 40, 55)]
    for srcs in (picks[:1], picks[:5], picks[:32], picks[:33], [picks[0], picks[0], picks[1]], [zero] + picks[:3]):
        bc, st = ctx.bc(G, srcs)
        ref = orc.bc(row, col, srcs)
        assert np.allclose(bc, ref, rtol=RTOL, atol=ATOL), (len(srcs), float(np.max(np.abs(bc - ref))))
    G.free()


def test_bc_rejects_bad_sources(npa, ctx):
    G = npa.Graph.rmat(ctx, 8, 4, wmin=1, wmax=1, seed=56)
    with pytest.raises(npa.NpError):
        ctx.bc(G, [0, G.n])
    with pytest.raises(npa.NpError):
        ctx.bc(G, [-1])
    G.free()
