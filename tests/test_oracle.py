"""Pins the CPU oracle (oracle/np_oracle.c) before anything is checked against it.

Sources of truth: the golden fixtures written from the reference's own compiled code
(tests/golden/make_golden.py), the reference code itself when oracle/_ref is built,
brute force (Bellman-Ford), closed forms, and planted mistakes that must be caught.
"""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden, random_csr

FIXTURES = sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


@pytest.mark.parametrize("name", FIXTURES)
def test_dijkstra_matches_reference_fixture(orc, name):
    g = load_golden(name)
    for k, s in enumerate(g["sources"]):
        d, it = orc.dijkstra(g["row"], g["col"], g["w"], s)
        assert np.array_equal(d, g["dist"][k])
        # same pop count as the reference's printed "Iteration number": the heap is
        # restated operation for operation, not just any Dijkstra
        assert it == g["iterations"][k]


@pytest.mark.parametrize("name", FIXTURES)
def test_dheap_bellman_ford_bfs_match_fixture(orc, name):
    g = load_golden(name)
    for k, s in enumerate(g["sources"]):
        for arity in (2, 3, 4):
            assert np.array_equal(orc.dijkstra_dheap(g["row"], g["col"], g["w"], s, arity), g["dist"][k])
        assert np.array_equal(orc.bellman_ford(g["row"], g["col"], g["w"], s), g["dist"][k])
        assert np.array_equal(orc.bfs_levels(g["row"], g["col"], s), g["level"][k])


@pytest.mark.parametrize("name", FIXTURES)
def test_certificate_accepts_fixture(orc, name):
    g = load_golden(name)
    positive = bool((g["w"] > 0).all())
    for k, s in enumerate(g["sources"]):
        code, _ = orc.check_sssp(g["row"], g["col"], g["w"], s, g["dist"][k])
        assert code == (0 if positive else 5) or (not positive and code == 0)
        code, _ = orc.check_sssp(g["row"], g["col"], None, s, g["level"][k], inf=orc.UNDEFINED, unit=True)
        assert code == 0


def test_closed_form_path(orc):
    n = 200
    w = (np.arange(n - 1) % 13 + 1).astype(np.int32)
    row = np.concatenate([np.arange(n, dtype=np.int32), [n - 1]]).astype(np.int32)
    col = np.arange(1, n, dtype=np.int32)
    d, _ = orc.dijkstra(row, col, w, 0)
    assert np.array_equal(d, np.concatenate([[0], np.cumsum(w)]).astype(np.int32))
    d, _ = orc.dijkstra(row, col, w, 100)
    assert (d[:100] == orc.INF).all() and d[100] == 0


def test_closed_form_unit_grid_is_manhattan(npa, orc):
    rows, cols = 23, 31
    row, col, _ = npa.gen_grid_host(rows, cols, 0, 1, 1, seed=5)
    w = np.ones(len(col), np.int32)
    s = 7 * cols + 11
    d, _ = orc.dijkstra(row, col, w, s)
    r, c = np.divmod(np.arange(rows * cols), cols)
    assert np.array_equal(d, np.abs(r - 7) + np.abs(c - 11))
    assert np.array_equal(orc.bfs_levels(row, col, s), d)


def test_closed_form_star(orc):
    n = 1000
    w = (np.arange(n - 1) * 7 % 50 + 1).astype(np.int32)
    row = np.zeros(n + 1, np.int32)
    row[1:] = n - 1
    col = np.arange(1, n, dtype=np.int32)
    d, _ = orc.dijkstra(row, col, w, 0)
    assert np.array_equal(d[1:], w) and d[0] == 0


def test_unit_weights_equal_bfs(orc):
    rng = np.random.default_rng(3)
    row, col, _ = random_csr(rng, 3000, 12000)
    w = np.ones(len(col), np.int32)
    d, _ = orc.dijkstra(row, col, w, 5)
    lev = orc.bfs_levels(row, col, 5)
    d2 = np.where(d == orc.INF, orc.UNDEFINED, d)
    assert np.array_equal(d2, lev)


def test_brute_force_random(orc):
    rng = np.random.default_rng(11)
    for n, m in [(1, 0), (2, 1), (17, 40), (120, 700), (400, 1200)]:
        row, col, w = random_csr(rng, n, m, 0, 9)
        for s in {0, n // 2, n - 1}:
            d, _ = orc.dijkstra(row, col, w, s)
            assert np.array_equal(d, orc.bellman_ford(row, col, w, s))
            assert np.array_equal(d, orc.dijkstra_dheap(row, col, w, s, 3))


def test_multi_source_equals_min_over_sources(orc):
    rng = np.random.default_rng(5)
    row, col, w = random_csr(rng, 500, 3000)
    srcs = [3, 77, 400]
    dm = orc.dijkstra_multi(row, col, w, srcs)
    single = np.min(np.stack([orc.dijkstra(row, col, w, s)[0] for s in srcs]), axis=0)
    assert np.array_equal(dm, single)
    assert orc.check_sssp(row, col, w, srcs, dm)[0] == 0
    many = orc.dijkstra_many(row, col, w, srcs, threads=2)
    for k, s in enumerate(srcs):
        assert np.array_equal(many[k], orc.dijkstra(row, col, w, s)[0])


def test_certificate_catches_planted_mistakes(orc):
    rng = np.random.default_rng(9)
    row, col, w = random_csr(rng, 400, 3000, 1, 50)
    d, _ = orc.dijkstra(row, col, w, 0)
    reached = np.flatnonzero((d != orc.INF) & (np.arange(400) != 0))
    v = int(reached[len(reached) // 2])
    bad = d.copy(); bad[v] += 1          # too large: some in-edge is not relaxed
    assert orc.check_sssp(row, col, w, 0, bad)[0] == 3
    bad = d.copy(); bad[v] -= 1          # too small: no tight in-edge
    assert orc.check_sssp(row, col, w, 0, bad)[0] in (3, 4)
    bad = d.copy(); bad[0] = 1           # source not at zero
    assert orc.check_sssp(row, col, w, 0, bad)[0] == 1
    unreached = np.flatnonzero(d == orc.INF)
    if len(unreached):
        bad = d.copy(); bad[unreached[0]] = 12345   # unreachable vertex given a value
        assert orc.check_sssp(row, col, w, 0, bad)[0] == 4
    bad = d.copy(); bad[v] = orc.INF      # reachable vertex left at INF
    assert orc.check_sssp(row, col, w, 0, bad)[0] == 3


def test_reached_work(orc):
    g = load_golden("d9_small.npz")
    n_r, m_r = orc.reached_work(g["row"], g["dist"][0])
    deg = np.diff(g["row"])
    fin = g["dist"][0] != orc.INF
    assert n_r == fin.sum() and m_r == deg[fin].sum()


@pytest.mark.skipif(not __import__("oracle").ref_available(), reason="oracle/_ref not built here")
def test_against_live_reference_code(orc):
    rng = np.random.default_rng(21)
    for n, m in [(50, 200), (2000, 20000), (5000, 10000)]:
        row, col, w = random_csr(rng, n, m, 1, 1000)
        for s in (0, n - 1):
            d, it = orc.dijkstra(row, col, w, s)
            dr, itr = orc.ref_sssp_cpu(row, col, w, s)
            assert np.array_equal(d, dr) and it == itr
            assert np.array_equal(orc.dijkstra_dheap(row, col, w, s, 2), orc.ref_dheap_dijkstra(row, col, w, s, 2))
            assert np.array_equal(orc.bfs_levels(row, col, s), orc.ref_bfs_cpu(row, col, s))


# ---- scheduler MLP restatement: pinned by torch autograd (float64) ----------------
def _torch_loss(params, x, label, target, lam):
    import torch
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    xs = torch.tensor(x, dtype=torch.float64)
    o = 0
    W1 = p[o:o + 320].view(32, 10); o += 320
    b1 = p[o:o + 32]; o += 32
    W2 = p[o:o + 1024].view(32, 32); o += 1024
    b2 = p[o:o + 32]; o += 32
    W3 = p[o:o + 160].view(5, 32); o += 160
    b3 = p[o:o + 5]
    h = torch.relu(torch.nn.functional.linear(xs, W1, b1))
    h = torch.relu(torch.nn.functional.linear(h, W2, b2))
    out = torch.nn.functional.linear(h, W3, b3)
    ce = torch.nn.functional.cross_entropy(out[:, :3], torch.tensor(label, dtype=torch.long), reduction="sum")
    reg = lam * ((out[:, 3:] - torch.tensor(target, dtype=torch.float64)) ** 2).sum()
    loss = (ce + reg) / x.shape[0]
    loss.backward()
    return float(loss), p.grad.numpy(), out.detach().numpy()


def test_mlp_forward_and_gradient_vs_torch(orc):
    rng = np.random.default_rng(0)
    params = (rng.standard_normal(orc.MLP_NPARAM) * 0.3).astype(np.float32)
    x = rng.random((64, 10)).astype(np.float32)
    label = rng.integers(0, 3, 64).astype(np.int32)
    target = (rng.random((64, 2)) * 10).astype(np.float32)
    loss_t, grad_t, out_t = _torch_loss(params.astype(np.float64), x, label, target, 0.1)
    out = orc.mlp_forward(params, x)
    np.testing.assert_allclose(out, out_t, rtol=1e-4, atol=1e-4)
    loss, grad = orc.mlp_loss_grad(params, x, label, target, 0.1)
    assert abs(loss - loss_t) <= 1e-4 * max(1.0, abs(loss_t))
    np.testing.assert_allclose(grad, grad_t, rtol=2e-3, atol=2e-5)
    # a planted mistake in the restatement (one bias gradient off) is visible at this tolerance
    g2 = grad.copy(); g2[320] += 1e-3
    assert not np.allclose(g2, grad_t, rtol=2e-3, atol=2e-5)
