"""Input rules of the C ABI, on the host (np_csr_validate: no device needed).  The reference
checks nothing; the rules here are what its int32 arithmetic needs (weights >= 0 for shortest
paths, weights < INF = 2^30 so that dist[u] + w with dist[u] < INF stays below 2^31)."""
import numpy as np
import pytest

INF = 1 << 30


def _path(weights):
    n = len(weights) + 1
    row = np.array(list(range(n - 1)) + [n - 1, n - 1], np.int32)[: n + 1]
    col = np.arange(1, n, dtype=np.int32)
    return row, col, np.asarray(weights, np.int32)


def test_good_csr_returns_largest_weight(npa):
    row, col, w = _path([3, 0, 17, 5])
    assert npa.csr_validate(row, col, w) == 17
    assert npa.csr_validate(row, col) == 0          # no weights (BFS)
    assert npa.csr_validate([0], [], []) == 0       # one vertex, no edge


@pytest.mark.parametrize("weights,code", [([1, -1, 2], -1), ([INF], -4), ([INF - 1, 5], 0),
                                          ([0x7FFFFFFF], -4), ([-0x80000000], -1)])
def test_weight_rules(npa, weights, code):
    row, col, w = _path(weights)
    if code == 0:
        assert npa.csr_validate(row, col, w) == max(weights)
    else:
        with pytest.raises(npa.NpError) as e:
            npa.csr_validate(row, col, w)
        assert e.value.code == code


@pytest.mark.parametrize("row,col", [([1, 1], [0]),          # row[0] != 0
                                     ([0, 2, 1], [1]),       # decreasing row
                                     ([0, 1, 1], [2]),       # target out of range
                                     ([0, 1, 1], [-1]),      # negative target
                                     ([0, 1, 2], [1])])      # row[n] != m
def test_structure_rules(npa, row, col):
    with pytest.raises(npa.NpError) as e:
        npa.csr_validate(row, col, np.ones(len(col), np.int32))
    assert e.value.code == -1


def test_reference_fixtures_pass(npa):
    from conftest import load_golden
    for name in ("d9_rand.npz", "d10_small.npz", "rmat_s10.npz", "grid_16x16.npz"):
        g = load_golden(name)
        assert npa.csr_validate(g["row"], g["col"], g["w"]) == int(g["w"].max())
