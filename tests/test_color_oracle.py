"""The graph-coloring oracle (orc_color) pinned on the CPU.

The reference (apps/graph-color/src/gpu/graph_color_kernel.cu:14-35) colors in rounds with
racing threads: in round c an uncolored v takes c unless an out-neighbour u < v is uncolored
or already holds c.  The simulation below runs those rounds literally, visiting the uncolored
vertices in a random order per round and writing colors immediately (the effect of any thread
interleaving); every order must give orc_color's one-pass values."""
import numpy as np
import pytest

from conftest import random_csr


def rounds_in_random_order(row, col, rng):
    n = len(row) - 1
    color = np.zeros(n, np.int64)
    c = 1
    while (color == 0).any():
        for v in rng.permutation(np.flatnonzero(color == 0)):
            nb = col[row[v]:row[v + 1]]
            nb = nb[nb < v]
            if not np.any((color[nb] == 0) | (color[nb] == c)):
                color[v] = c
        c += 1
    return color


@pytest.mark.parametrize("seed", range(8))
def test_matches_the_reference_rounds_in_any_order(orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 120))
    row, col, _ = random_csr(rng, n, int(rng.integers(0, 6 * n)), 1, 1)   # parallel edges, self loops
    want = orc.color(row, col)
    for _ in range(3):
        assert np.array_equal(rounds_in_random_order(row, col, rng), want)


def test_closed_forms(orc):
    n = 30
    row = np.array([0] + [max(0, v) for v in range(n)], np.int32)          # v -> v - 1 for v >= 1
    col = np.arange(0, n - 1, dtype=np.int32)
    assert np.array_equal(orc.color(row, col), np.arange(1, n + 1))        # longest decreasing path
    row_up = np.array([min(v, n - 1) for v in range(n + 1)], np.int32)
    col_up = np.arange(1, n, dtype=np.int32)                                # v -> v + 1: never blocked
    assert np.array_equal(orc.color(row_up, col_up), np.ones(n))
    k = 9                                                                   # complete digraph
    row_k = np.arange(0, k * (k - 1) + 1, k - 1, dtype=np.int32)
    col_k = np.array([u for v in range(k) for u in range(k) if u != v], np.int32)
    assert np.array_equal(orc.color(row_k, col_k), np.arange(1, k + 1))
    row_s = np.array([0, 1, 2, 3], np.int32)                                # self loops only
    assert np.array_equal(orc.color(row_s, np.array([0, 1, 2], np.int32)), [1, 1, 1])


def test_proper_coloring_on_symmetric_graphs(orc):
    rng = np.random.default_rng(11)
    n = 400
    u = rng.integers(0, n, 3000)
    v = rng.integers(0, n, 3000)
    keep = u != v
    src = np.concatenate([u[keep], v[keep]])
    dst = np.concatenate([v[keep], u[keep]])
    order = np.argsort(src, kind="stable")
    row = np.zeros(n + 1, np.int32)
    np.add.at(row, src + 1, 1)
    row = np.cumsum(row).astype(np.int32)
    col = dst[order].astype(np.int32)
    color = orc.color(row, col)
    assert np.all(color[src] != color[dst])
    assert color.min() == 1
