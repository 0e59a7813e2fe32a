"""The betweenness-centrality oracle (orc_bc) pinned on the CPU.

What pins it: the reference's own bc_cpu (apps/bc/src/cpu, built in oracle/_ref when the
reference sources are present), closed forms, and the definition
bc(v) = sum over s != v != t of sigma_st(v) / sigma_st, evaluated by brute force from all-pairs
BFS path counts (no dependency recursion), on graphs with parallel edges and self loops."""
import numpy as np
import pytest

from conftest import random_csr


def csr_from_edges(n, edges):
    edges = sorted(edges, key=lambda e: e[0])
    row = np.zeros(n + 1, np.int32)
    for u, _ in edges:
        row[u + 1] += 1
    row = np.cumsum(row).astype(np.int32)
    col = np.array([v for _, v in edges], np.int32)
    return row, col


def brute_force_bc(row, col, srcs=None):
    """sigma_st(v) = sigma_sv * sigma_vt when d(s,v) + d(v,t) = d(s,t): the definition."""
    n = len(row) - 1
    dist = np.full((n, n), -1, np.int64)
    sig = np.zeros((n, n), np.float64)
    for s in range(n):
        dist[s, s] = 0
        sig[s, s] = 1.0
        frontier = [s]
        while frontier:
            nxt = []
            for u in frontier:
                for e in range(row[u], row[u + 1]):
                    w = col[e]
                    if dist[s, w] < 0:
                        dist[s, w] = dist[s, u] + 1
                        nxt.append(w)
                    if dist[s, w] == dist[s, u] + 1:
                        sig[s, w] += sig[s, u]
            frontier = nxt
    bc = np.zeros(n)
    for s in (range(n) if srcs is None else srcs):
        for t in range(n):
            if t == s or dist[s, t] < 0:
                continue
            for v in range(n):
                if v in (s, t) or dist[s, v] < 0 or dist[v, t] < 0:
                    continue
                if dist[s, v] + dist[v, t] == dist[s, t]:
                    bc[v] += sig[s, v] * sig[v, t] / sig[s, t]
    return bc


def test_closed_forms(orc):
    n = 12
    row, col = csr_from_edges(n, [(i, i + 1) for i in range(n - 1)])          # directed path
    assert np.allclose(orc.bc(row, col), [v * (n - 1 - v) for v in range(n)])
    k = 9                                                                       # star, both ways
    row, col = csr_from_edges(k + 1, [(0, i) for i in range(1, k + 1)] + [(i, 0) for i in range(1, k + 1)])
    assert np.allclose(orc.bc(row, col), [k * (k - 1)] + [0] * k)
    n = 7                                                                       # complete digraph
    row, col = csr_from_edges(n, [(u, v) for u in range(n) for v in range(n) if u != v])
    assert np.allclose(orc.bc(row, col), 0)
    # two disjoint equal routes 0 -> {1, 2} -> 3: each middle vertex carries half of (0, 3)
    row, col = csr_from_edges(4, [(0, 1), (0, 2), (1, 3), (2, 3)])
    assert np.allclose(orc.bc(row, col), [0, 0.5, 0.5, 0])
    # a doubled edge doubles the paths through it: 0 => 1 -> 3 twice, 0 -> 2 -> 3 once
    row, col = csr_from_edges(4, [(0, 1), (0, 1), (0, 2), (1, 3), (2, 3)])
    assert np.allclose(orc.bc(row, col), [0, 2 / 3, 1 / 3, 0])


@pytest.mark.parametrize("seed", range(6))
def test_matches_definition_by_brute_force(orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 14))
    row, col, _ = random_csr(rng, n, int(rng.integers(0, 4 * n)), 1, 1)      # parallel edges, self loops
    assert np.allclose(orc.bc(row, col), brute_force_bc(row, col), rtol=1e-12, atol=1e-12)
    srcs = sorted(set(int(x) for x in rng.integers(0, n, size=3)))
    assert np.allclose(orc.bc(row, col, srcs), brute_force_bc(row, col, srcs), rtol=1e-12, atol=1e-12)


def test_sources_add_up(orc):
    rng = np.random.default_rng(7)
    row, col, _ = random_csr(rng, 300, 1500, 1, 1)
    parts = [orc.bc(row, col, [s]) for s in range(300)]
    assert np.allclose(orc.bc(row, col), np.sum(parts, axis=0), rtol=1e-12)
    assert np.allclose(orc.bc(row, col, [5, 5]), 2 * parts[5])


@pytest.mark.parametrize("seed", range(3))
def test_matches_reference_bc_cpu(orc, seed):
    if not orc.ref_bc_available():
        pytest.skip("oracle/_ref/libref_bc.so not built (reference sources absent)")
    rng = np.random.default_rng(100 + seed)
    n = [60, 200, 400][seed]
    row, col, _ = random_csr(rng, n, 5 * n, 1, 1)
    ref = orc.ref_bc_cpu(row, col).astype(np.float64)
    # the reference accumulates in float (delta, bc) from int path counts
    assert np.allclose(orc.bc(row, col), ref, rtol=2e-5, atol=1e-3)
