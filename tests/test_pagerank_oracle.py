"""The PageRank oracle (orc_pagerank) pinned on the CPU.

The reference iteration (apps/pagerank/src/gpu/pagerank_wrapper.cu:199-244) converges to the
fixed point r = d * P r + (1 - d) * T / n + d * (dangling mass) / n with P[v, p] = (edges p -> v) /
outdeg[p]; that linear system, solved directly with scipy, pins the oracle independently of the
iteration.  Closed forms: a directed cycle is uniform; a two-way star is solved by hand."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import random_csr

D, T = 0.5, 1000.0


def fixed_point(row, col):
    n = len(row) - 1
    src = np.repeat(np.arange(n), np.diff(row))
    outdeg = np.diff(row).astype(np.float64)
    P = sp.coo_matrix((1.0 / outdeg[src], (col, src)), shape=(n, n)).tocsr()   # duplicates add up
    dang = (outdeg == 0).astype(np.float64)
    M = sp.identity(n) - D * P - D * sp.csr_matrix(np.outer(np.ones(n), dang) / n)
    return spla.spsolve(M.tocsc(), np.full(n, (1 - D) * T / n))


@pytest.mark.parametrize("seed", range(5))
def test_matches_the_fixed_point_system(orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 300))
    row, col, _ = random_csr(rng, n, int(rng.integers(0, 5 * n)), 1, 1)   # dangling vertices, parallel edges, loops
    r, it = orc.pagerank(row, col)
    assert it > 1
    assert np.allclose(r, fixed_point(row, col), rtol=1e-10, atol=1e-12)
    assert abs(r.sum() - T) < 1e-8


def test_closed_forms(orc):
    n = 17
    row = np.arange(n + 1, dtype=np.int32)
    col = (np.arange(n) + 1) % n                                           # directed cycle
    r, it = orc.pagerank(row, col.astype(np.int32))
    assert np.allclose(r, T / n) and it == 1
    k = 6                                                                  # two-way star, center 0
    row = np.array([0, k] + list(range(k + 1, 2 * k + 1)), np.int32)
    col = np.array(list(range(1, k + 1)) + [0] * k, np.int32)
    r, _ = orc.pagerank(row, col)
    # center c = D * k * leaf + (1-D) T/(k+1); leaf l = D * c / k + (1-D) T/(k+1)
    base = (1 - D) * T / (k + 1)
    c = (base + D * k * base) / (1 - D * D)
    l = D * c / k + base
    assert np.allclose(r, [c] + [l] * k, rtol=1e-12)
