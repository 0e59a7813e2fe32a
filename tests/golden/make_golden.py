"""Writes the golden fixtures of the SSSP / BFS path.

Run in the build container, where the reference sources exist and oracle/_ref was built
from them (make -C oracle).  For each input it records what the REFERENCE'S OWN CODE
produces: its readers + convertCSR (common/global.cpp), its SSSP_CPU Dijkstra with its
Heap (apps/sssp/src/cpu/sssp.cpp, common/heap.cpp, including the printed iteration
count), its bfs_cpu (apps/bfs-rec/src/gpu/bfs_rec.cpp) and its outputCost format.

Inputs are synthetic graphs written here (the reference ships no graph fixtures; its
datasets/ link is dangling).  Outputs: tests/golden/graphs/* and tests/golden/*.npz.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle  # noqa: E402
from conftest import load_pkg  # noqa: E402

GRAPHS = os.path.join(HERE, "graphs")


def write_dimacs9(path, n, arcs, src=None, comments=("generated fixture",)):
    with open(path, "w") as f:
        for c in comments:
            f.write("c %s\n" % c)
        f.write("p sp %d %d\n" % (n, len(arcs)))
        if src is not None:
            f.write("s %d\n" % (src + 1))
        for (u, v, w) in arcs:
            f.write("a %d %d %d\n" % (u + 1, v + 1, w))


def write_dimacs10(path, n, adj):
    m2 = sum(len(a) for a in adj)
    assert m2 % 2 == 0
    with open(path, "w") as f:
        f.write("%% generated fixture\n%d %d\n" % (n, m2 // 2))
        for a in adj:
            f.write(" ".join(str(v + 1) for v in a) + "\n")


def write_slndc(path, n_hdr, pairs):
    with open(path, "w") as f:
        f.write("# Directed graph (each unordered pair of nodes is saved once): fixture.txt\n")
        f.write("# generated fixture\n")
        f.write("# Nodes: %d Edges: %d\n" % (n_hdr, len(pairs)))
        f.write("# FromNodeId\tToNodeId\n")
        for (u, v) in pairs:
            f.write("%d\t%d\n" % (u, v))


def record(name, fmt, path, sources=None):
    row, col, w, src = oracle.ref_read_graph(fmt, path)
    srcs = [src] if sources is None else list(sources)
    dists, iters, levels = [], [], []
    for s in srcs:
        d, it = oracle.ref_sssp_cpu(row, col, w, s)
        dists.append(d)
        iters.append(it)
        levels.append(oracle.ref_bfs_cpu(row, col, s))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), fmt=np.int32(fmt),
                        path=np.array(os.path.relpath(path, HERE)), row=row, col=col, w=w,
                        source=np.int32(src), sources=np.array(srcs, np.int32),
                        dist=np.array(dists, np.int32), iterations=np.array(iters, np.int64),
                        level=np.array(levels, np.int32))
    print("%-22s n=%-6d m=%-7d sources=%s iterations=%s" % (name, len(row) - 1, len(col), srcs, iters))
    return row, col, w, srcs


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref is not built (make -C oracle, with the reference sources present)")
    os.makedirs(GRAPHS, exist_ok=True)
    rng = np.random.default_rng(2024)

    # DIMACS9, hand-made: duplicates, a self-loop, a zero weight, an unreachable vertex
    arcs = [(0, 1, 4), (0, 2, 1), (2, 1, 2), (1, 3, 1), (2, 3, 5), (3, 4, 3), (4, 0, 7), (1, 3, 1),
            (3, 3, 2), (4, 5, 0), (5, 6, 9), (2, 6, 20)]
    p = os.path.join(GRAPHS, "d9_small.gr")
    write_dimacs9(p, 8, arcs, src=0)
    record("d9_small", 0, p, sources=[0, 2, 7])

    # DIMACS9, random with a source line
    n, m = 300, 2400
    arcs = [(int(u), int(v), int(x)) for u, v, x in zip(rng.integers(0, n, m), rng.integers(0, n, m),
                                                         rng.integers(1, 1001, m))]
    p = os.path.join(GRAPHS, "d9_rand.gr")
    write_dimacs9(p, n, arcs, src=17)
    record("d9_rand", 0, p, sources=[17, 0, 123])

    # DIMACS9 path and star (closed forms)
    n = 64
    p = os.path.join(GRAPHS, "d9_path.gr")
    write_dimacs9(p, n, [(i, i + 1, 1 + (i % 7)) for i in range(n - 1)], src=0)
    record("d9_path", 0, p)

    # DIMACS10: undirected adjacency (weights drawn by the reader with rand() after srand(7))
    n = 60
    und = set()
    while len(und) < 150:
        u, v = rng.integers(0, n, 2)
        if u != v:
            und.add((min(u, v), max(u, v)))
    adj = [[] for _ in range(n)]
    for u, v in sorted(und):
        adj[u].append(int(v)); adj[v].append(int(u))
    for a in adj:
        rng.shuffle(a)
    p = os.path.join(GRAPHS, "d10_small.graph")
    write_dimacs10(p, n, adj)
    record("d10_small", 1, p, sources=[0, 5, 59])

    # SLNDC: sparse ids with gaps, ids compacted by the reader, weights rand()%1000+1
    ids = np.sort(rng.choice(np.arange(0, 160), 70, replace=False))
    pairs = [(int(ids[rng.integers(0, 70)]), int(ids[rng.integers(0, 70)])) for _ in range(260)]
    p = os.path.join(GRAPHS, "slndc_small.txt")
    write_slndc(p, 80, pairs)
    record("slndc_small", 2, p, sources=[0, 10, 33])

    # outputCost format of the reference, for the command-line driver
    row, col, w, src = oracle.ref_read_graph(0, os.path.join(GRAPHS, "d9_small.gr"))
    d, _ = oracle.ref_sssp_cpu(row, col, w, src)
    oracle.ref_output_cost(os.path.join(HERE, "d9_small.costs.txt"), src, d)

    # generated graphs of the benchmark families, small, with the reference Dijkstra
    npa = load_pkg()
    for name, (row, col, w), srcs in [
        ("er_1k", npa.gen_er_host(1024, 16384, 1, 100, seed=1), [0, 1, 500]),
        ("rmat_s10", npa.gen_rmat_host(10, 16, wmin=1, wmax=255, seed=2), None),
        ("grid_16x16", npa.gen_grid_host(16, 16, None, 1, 1000, seed=5), [0, 100, 255]),
    ]:
        if srcs is None:
            srcs = list(npa.pick_sources(row, 4, 2))
        dists, iters, levels = [], [], []
        for s in srcs:
            dd, it = oracle.ref_sssp_cpu(row, col, w, s)
            dists.append(dd); iters.append(it); levels.append(oracle.ref_bfs_cpu(row, col, s))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), row=row, col=col, w=w,
                            sources=np.array(srcs, np.int32), dist=np.array(dists, np.int32),
                            iterations=np.array(iters, np.int64), level=np.array(levels, np.int32))
        print("%-22s n=%-6d m=%-7d sources=%s iterations=%s" % (name, len(row) - 1, len(col), srcs, iters))


if __name__ == "__main__":
    main()
