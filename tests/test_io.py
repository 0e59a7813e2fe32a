"""Graph files and host generators (CPU only).

The readers are pinned by the reference's own readers + convertCSR: the fixtures in
tests/golden hold the CSR that common/global.cpp produced for each file, and when
oracle/_ref is built the same comparison runs live on freshly written files."""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden

READER_FIXTURES = ["d9_small", "d9_rand", "d9_path", "d10_small", "slndc_small"]


@pytest.mark.parametrize("name", READER_FIXTURES)
def test_readers_match_reference_csr(npa, name):
    g = load_golden(name + ".npz")
    path = os.path.join(GOLDEN, str(g["path"]))
    row, col, w, src = npa.read_graph(int(g["fmt"]), path)
    assert np.array_equal(row, g["row"])
    assert np.array_equal(col, g["col"])
    assert np.array_equal(w, g["w"])
    assert src == int(g["source"])


def test_convertcsr_order_is_reversed_file_order(npa, tmp_path):
    p = tmp_path / "order.gr"
    p.write_text("p sp 3 4\na 1 2 10\na 1 3 20\na 2 3 5\na 1 1 7\n")
    row, col, w, src = npa.read_graph(0, str(p))
    assert list(row) == [0, 3, 4, 4]
    assert list(col[:3]) == [0, 2, 1] and list(w[:3]) == [7, 20, 10]
    assert src == 0


def _rand_dimacs9(path, rng, n, m, src):
    with open(path, "w") as f:
        f.write("c x\np sp %d %d\ns %d\n" % (n, m, src + 1))
        for _ in range(m):
            f.write("a %d %d %d\n" % (rng.integers(1, n + 1), rng.integers(1, n + 1), rng.integers(0, 5000)))


@pytest.mark.skipif(not __import__("oracle").ref_available(), reason="oracle/_ref not built here")
def test_readers_live_against_reference(npa, orc, tmp_path):
    rng = np.random.default_rng(8)
    p = str(tmp_path / "r.gr")
    _rand_dimacs9(p, rng, 500, 4000, 42)
    a = npa.read_graph(0, p)
    b = orc.ref_read_graph(0, p)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    assert a[3] == b[3] == 42
    # DIMACS10 with rand() weights
    n = 80
    adj = [[] for _ in range(n)]
    for _ in range(200):
        u, v = rng.integers(0, n, 2)
        if u != v:
            adj[u].append(v); adj[v].append(u)
    p = str(tmp_path / "r.graph")
    with open(p, "w") as f:
        f.write("%%c\n%d %d\n" % (n, sum(len(x) for x in adj) // 2))
        for a_ in adj:
            f.write(" ".join(str(v + 1) for v in a_) + "\n")
    a = npa.read_graph(1, p)
    b = orc.ref_read_graph(1, p)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)


def test_reader_errors(npa, tmp_path):
    with pytest.raises(npa.NpError) as e:
        npa.read_graph(0, str(tmp_path / "missing.gr"))
    assert e.value.code == -8
    p = tmp_path / "bad.gr"
    p.write_text("p sp 3 2\na 1 2 1\n")           # header says 2 arcs, file has 1
    with pytest.raises(npa.NpError):
        npa.read_graph(0, str(p))
    p.write_text("p sp 3 1\na 1 9 1\n")           # vertex id out of range
    with pytest.raises(npa.NpError):
        npa.read_graph(0, str(p))
    with pytest.raises(npa.NpError):
        npa.read_graph(7, str(p))


def test_binary_csr_cache_round_trip(npa, tmp_path):
    """np_csr_save / np_csr_load: a graph parsed from a reference format comes back identical,
    through load_csr and through np_read_graph fmt 3; without weights it reads as unit weights;
    the empty graph and a graph with isolated vertices round-trip too."""
    g = load_golden("d9_rand.npz")
    row, col, w, src = npa.read_graph(int(g["fmt"]), os.path.join(GOLDEN, str(g["path"])))
    p = str(tmp_path / "g.npcsr")
    npa.save_csr(p, row, col, w, source=src)
    for got in (npa.load_csr(p), npa.read_graph(3, p)):
        assert np.array_equal(got[0], row) and np.array_equal(got[1], col)
        assert np.array_equal(got[2], w) and got[3] == src
    assert os.path.getsize(p) == 32 + 4 * (len(row) + 2 * len(col))
    npa.save_csr(p, row, col)                        # unweighted
    r2, c2, w2, s2 = npa.load_csr(p)
    assert np.array_equal(c2, col) and np.all(w2 == 1) and s2 == 0
    for r, c in ((np.zeros(1, np.int32), np.zeros(0, np.int32)),           # n = 0, m = 0
                 (np.array([0, 2, 2, 2, 3], np.int32), np.array([3, 1, 0], np.int32))):
        npa.save_csr(p, r, c, np.arange(1, len(c) + 1, dtype=np.int32))
        r3, c3, w3, _ = npa.load_csr(p)
        assert np.array_equal(r3, r) and np.array_equal(c3, c) and np.array_equal(w3, np.arange(1, len(c) + 1))


def test_binary_csr_cache_rejects_damaged_files(npa, tmp_path):
    """Each damage is caught by the header, the file length or the CSR check: NP_ERR_IO."""
    row = np.array([0, 2, 3, 3], np.int32)
    col = np.array([1, 2, 0], np.int32)
    w = np.array([5, 6, 7], np.int32)
    good = str(tmp_path / "good.npcsr")
    npa.save_csr(good, row, col, w)
    data = open(good, "rb").read()
    head = 32

    def variant(name, blob):
        q = str(tmp_path / name)
        open(q, "wb").write(blob)
        with pytest.raises(npa.NpError) as e:
            npa.load_csr(q)
        assert e.value.code == -8, name

    variant("magic", b"X" + data[1:])
    variant("truncated", data[:-4])
    variant("longer", data + b"\0\0\0\0")
    variant("header_only", data[:head])
    bad = bytearray(data)
    bad[head + 4:head + 8] = np.array([3], np.int32).tobytes()     # row = [0, 3, 2, 3]: decreasing
    bad[head + 8:head + 12] = np.array([2], np.int32).tobytes()
    variant("row_order", bytes(bad))
    bad = bytearray(data)
    bad[head + 16:head + 20] = np.array([7], np.int32).tobytes()   # col[0] = 7 >= n
    variant("col_range", bytes(bad))
    bad = bytearray(data)
    bad[head + 12:head + 16] = np.array([2], np.int32).tobytes()   # row[n] != m
    variant("row_end", bytes(bad))
    with pytest.raises(npa.NpError) as e:
        npa.load_csr(str(tmp_path / "missing.npcsr"))
    assert e.value.code == -8


def test_host_generators_deterministic_and_shaped(npa):
    a = npa.gen_rmat_host(11, 16, seed=5)
    b = npa.gen_rmat_host(11, 16, seed=5)
    c = npa.gen_rmat_host(11, 16, seed=6)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[1], c[1])
    row, col, w = a
    assert len(row) == 2049 and len(col) == 16 * 2048 and row[-1] == len(col)
    assert col.min() >= 0 and col.max() < 2048 and w.min() >= 1 and w.max() <= 255
    deg = np.diff(row)
    assert deg.max() > 20 * deg.mean()            # skewed (R-MAT a=0.57)
    row, col, w = npa.gen_er_host(4096, 65536, 1, 100, seed=1)
    src = np.repeat(np.arange(4096), np.diff(row))
    assert (src != col).all() and len(col) == 65536 and w.min() >= 1 and w.max() <= 100
    row, col, w = npa.gen_grid_host(64, 64, None, 1, 1000, seed=5)
    assert len(col) == 2 * (64 * 63 * 2) + npa.grid_long_range(64, 64)


def test_rmat_permutation_is_a_bijection(npa):
    # permuted and unpermuted graphs have the same degree multiset
    r1, _, _ = npa.gen_rmat_host(10, 8, seed=3, permute=True)
    r0, _, _ = npa.gen_rmat_host(10, 8, seed=3, permute=False)
    assert sorted(np.diff(r1)) == sorted(np.diff(r0))


def test_pick_sources(npa):
    row, _, _ = npa.gen_rmat_host(10, 16, seed=2)
    s = npa.pick_sources(row, 16, 3)
    assert len(set(s.tolist())) == 16
    assert (np.diff(row)[s] >= 1).all()
    assert np.array_equal(s, npa.pick_sources(row, 16, 3))


def test_features(npa):
    cnt = np.zeros(32, np.int32)
    cnt[1] = 3; cnt[7] = 1    # three degree-1 vertices, one of degree 64..127
    x = npa.features(cnt, 5, 1 << 16)
    assert abs(x[0] - np.log2(5) / 24) < 1e-6
    assert abs(x[1] - np.log2(1 + 3 * 0.75 * 2 + 0.75 * 128) / 32) < 1e-6
    assert np.allclose(x[2:8], [0.75, 0, 0, 0.25, 0, 0])
    assert abs(x[8] - 5 / 64) < 1e-7 and abs(x[9] - 0.5) < 1e-7
    assert np.allclose(npa.features(np.zeros(32, np.int32), 0, 1), np.zeros(10))
