"""Benchmark of the SSSP nested-parallelism path on one MI355X.

Default workload (the one the headline metric is quoted on, --config c3): SSSP on R-MAT
scale 24 (16,777,216 vertices, 268,435,456 directed edges, a=.57 b=.19 c=.19, vertex ids
permuted, duplicates kept), int32 weights U[1,255], 16 sources drawn with seed 3 among
vertices of out-degree >= 1.  A step solves every instance once.  Modes (same inputs,
distances verified identical):
  plain    the reference's fixed schedule: block-level consolidation, THRESHOLD 64,
           one persistent launch per source
  tuned    bins chosen by the autotuner, one launch per source
  learned  the trained MLP picks level and bins per iteration, one launch per source
  concat   all instances in one persistent launch over a concatenated slot space
  batched  np_batch with source packing: the sources of one graph share distance rows
           (one persistent launch per 32 sources), bins from the batch autotuner
  learned_batched  the same batch with the trained MLP picking level and bins per iteration,
           on the batch autotuner's drive and window (its fixed bins are the fallback)
The learned modes run on the drive and priority window their tuned counterparts chose, so
learned vs tuned compares the per-iteration choice of bins only.
The JSON value is the fastest mode; every mode is listed under "modes".
The other configurations (--config c1|c2|c4|c5) are the parity/evaluation cases of the
same path (ER 64K, R-MAT s22 per level, many sources + many uneven graphs, grid road
graph); they print the same JSON schema.  --config bc measures betweenness centrality
(BC_GPU) on the same frontier machinery (R-MAT s16, 1024 sources); --config color the graph
coloring of GRAPH_COLOR_GPU (R-MAT s20); --config pagerank PAGERANK_GPU (R-MAT s20); --config
spmv SpMV_GPU (R-MAT s22); --config bfs the BFS levels of BFS_REC_GPU on the R-MAT s24 graph and
16 sources of c3 (the reference's bfs-rec solutions, the consolidation levels and a packed batch).

The CPU leg (skipped with --no-cpu) is the checker and the CPU baseline: every distance of
every instance and mode is checked with the shortest-path certificate, a bounded sample is
re-solved with the CPU Dijkstra (bit-exact) and timed.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--config c3]
Prints one JSON line.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "nested-parallelism_amd")
METRIC = "SSSP GTEPS + %HBM peak @R-MAT s24; learned+batched vs plain speedup, exact"
HBM_PEAK_GBS = 8000.0


def load_pkg():
    name = "nested_parallelism_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Workload:
    """Instances (graph, source) of one configuration."""

    def __init__(self, name, desc, graphs, inst, cpu_sample):
        self.name, self.desc = name, desc
        self.graphs = graphs          # distinct npa.Graph objects
        self.inst = inst              # list of (graph index, source)
        self.cpu_sample = cpu_sample  # instances re-solved by the CPU Dijkstra


def build_workload(npa, ctx, cfg, args):
    if cfg == "c1":
        g = npa.Graph.er(ctx, 65536, 1048576, 1, 100, seed=1)
        return Workload(cfg, "SSSP, directed Erdos-Renyi n=65,536 m=1,048,576 (self-loops redrawn), weights "
                             "int32 U[1,100], source 0", [g], [(0, 0)], 1)
    if cfg in ("c2", "c3"):
        scale, k, seed = (22, 8, 2) if cfg == "c2" else (args.scale, args.sources, 3)
        g = npa.Graph.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, 1, 255, seed=seed, permute=True)
        row = g.download()[0]
        srcs = npa.pick_sources(row, k, seed)
        return Workload(cfg, "SSSP, R-MAT scale %d, edge factor 16, a=.57 b=.19 c=.19, ids permuted, weights int32 "
                             "U[1,255], %d sources (seed %d)" % (scale, k, seed), [g],
                        [(0, int(s)) for s in srcs], 1)
    if cfg == "c4":
        big = npa.Graph.rmat(ctx, 20, 16, 0.57, 0.19, 0.19, 1, 255, seed=4, permute=True)
        row = big.download()[0]
        srcs = npa.pick_sources(row, 64, 4)
        graphs, inst = [big], [(0, int(s)) for s in srcs]
        for i in range(256):
            seed = 1000 + i
            scale = 12 + int(npa_scale_draw(seed) % 7)
            g = npa.Graph.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, 1, 255, seed=seed, permute=True)
            s = int(npa.pick_sources(g.download()[0], 1, seed)[0])
            graphs.append(g)
            inst.append((len(graphs) - 1, s))
        return Workload(cfg, "SSSP, 64 sources (seed 4) on R-MAT s20 + 256 R-MAT graphs of scale U{12..18} (seeds "
                             "1000-1255, one source each), edge factor 16, weights int32 U[1,255]", graphs, inst, 2)
    if cfg == "c5":
        g = npa.Graph.grid(ctx, 4096, 4096, None, 1, 1000, seed=5)
        row = g.download()[0]
        srcs = npa.pick_sources(row, 32, 5)
        return Workload(cfg, "SSSP, 4096x4096 grid, 4-neighbour bidirectional + 1% random long-range edges (seed 5), "
                             "weights int32 U[1,1000], 32 sources (seed 5)", [g], [(0, int(s)) for s in srcs], 1)
    raise SystemExit("unknown config %s" % cfg)


def npa_scale_draw(seed):
    """Scale draw of the uneven-graph batch: SplitMix64 finaliser of the graph seed."""
    z = (seed * 0x9E3779B97F4A7C15 + 7) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


def time_mode(ctx, step, steps, warmup):
    for _ in range(warmup):
        step()
    ctx.sync()
    t = time.perf_counter()
    launch_ms = []
    for _ in range(steps):
        launch_ms += step()
    ctx.sync()
    return (time.perf_counter() - t) * 1e3 / steps, launch_ms


def host_cpu():
    """(CPU model string, logical CPUs of the host, CPUs this process may run on)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def cpu_leg(W, hosts, finals, modes, m_r, k):
    """Checker and CPU baseline (the only place the oracle is used).

    1. every instance of every mode: the shortest-path certificate, and modes equal;
    2. per-source single-core time: the CPU Dijkstra (restated SSSP_CPU) on k instances, one
       thread, one after another;
    3. aggregate wall time: every instance of the workload, one Dijkstra per thread on
       min(#instances, usable CPUs, 16) threads (16 = one GPU's share of the host); every
       distance compared with the GPU's, bit-exact."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    exact = True
    t = time.time()
    for i, (gi, s) in enumerate(W.inst):
        row, col, w = hosts[gi]
        code, bad = oracle.check_sssp(row, col, w, s, finals[modes[0]][i])
        if code != 0:
            exact = False
            log("CERTIFICATE FAILED instance %d source %d: %s at vertex %d" % (i, s, oracle.CHECK_CODES[code], bad))
        for mode in modes[1:]:
            if not np.array_equal(finals[mode][i], finals[modes[0]][i]):
                exact = False
                log("MODE MISMATCH %s vs %s instance %d" % (mode, modes[0], i))
    log("certificates: %d instances x %d modes in %.1fs: %s" % (len(W.inst), len(modes), time.time() - t, exact))
    k = min(k, len(W.inst))
    t = time.perf_counter()
    for i in range(k):
        gi, s = W.inst[i]
        if not np.array_equal(oracle.dijkstra(*hosts[gi], s)[0], finals[modes[0]][i]):
            exact = False
            log("DIJKSTRA MISMATCH instance %d" % i)
    t1 = time.perf_counter() - t
    per_source = t1 / max(1, k)
    model, ncpu, usable = host_cpu()
    threads = max(1, min(len(W.inst), usable, 16))
    t = time.perf_counter()
    for gi in sorted(set(g for g, _ in W.inst)):
        ids = [i for i, (g, _) in enumerate(W.inst) if g == gi]
        dist = oracle.dijkstra_many(*hosts[gi], [W.inst[i][1] for i in ids], threads=threads)
        for j, i in enumerate(ids):
            if not np.array_equal(dist[j], finals[modes[0]][i]):
                exact = False
                log("DIJKSTRA MISMATCH instance %d (aggregate)" % i)
        del dist
    agg = time.perf_counter() - t
    total_m = float(sum(m_r))
    cpu = {"value": round(total_m / agg / 1e9, 5), "unit": "GTEPS", "cores": threads, "kind": "port",
           "nproc": ncpu, "usable_cpus": usable, "cpu_model": model,
           "per_source_s": round(per_source, 3), "aggregate_s": round(agg, 3),
           "per_source_gteps_1core": round(float(sum(m_r[:k])) / t1 / 1e9, 5),
           "sample": "CPU Dijkstra (restated SSSP_CPU + binary heap, common/heap.cpp): all %d instances of the "
                     "same workload, one Dijkstra per thread on %d threads, %.1f s wall (value = reached edges / "
                     "that wall time); per source on one core: %.1f s (%d instance(s) in sequence); every distance "
                     "bit-exact vs the GPU" % (len(W.inst), threads, agg, per_source, k)}
    log("cpu baseline: %.5f GTEPS aggregate (%d threads, %.1f s), %.1f s per source on one core; %s, nproc %d"
        % (cpu["value"], threads, agg, per_source, model, ncpu))
    return cpu, exact


BC_METRIC = "BC GTEPS (Brandes betweenness, 1024 sources) + %HBM peak @R-MAT s16, within 1e-5 of the oracle"


def bc_bench(npa, ctx, args):
    """--config bc: betweenness centrality (BC_GPU, solution 7) on R-MAT scale 16 (65,536
    vertices, 1,048,576 edges, ids permuted, unit lengths), 1024 sources drawn with seed 6 among
    vertices of out-degree >= 1; a step is one np_bc call over all of them (32 batches of 32).
    GTEPS = sum over sources of the out-edges of reached vertices / step time.  Algorithmic
    bytes per source = 40 per reached edge (forward: col 4, level 4, path count 8; backward:
    col 4, level 4, path count 8, dependency 8) + 36 per reached vertex (row bounds twice 16,
    own path count and dependency 16, level 4)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    t0 = time.time()
    G = npa.Graph.rmat(ctx, 16, 16, wmin=1, wmax=1, seed=6)
    row, col, _ = G.download()
    srcs = [int(s) for s in npa.pick_sources(row, 1024, 6)]
    deg = np.diff(row).astype(np.int64)
    n_r = np.zeros(len(srcs), np.int64)
    m_r = np.zeros(len(srcs), np.int64)
    for i, s in enumerate(srcs):  # reached work per source, from BFS levels (not timed)
        lev, _ = ctx.bfs(G, s)
        reached = lev != npa.UNDEFINED
        n_r[i], m_r[i] = int(reached.sum()), int(deg[reached].sum())
    log("bc workload built in %.1fs" % (time.time() - t0))
    for _ in range(args.warmup):
        ctx.bc(G, srcs)
    wall, dev = [], []
    for _ in range(args.steps):
        t = time.perf_counter()
        bc, st = ctx.bc(G, srcs)
        wall.append((time.perf_counter() - t) * 1e3)
        dev.append(st.kernel_ms)
    ms = float(np.mean(wall))
    dev_ms = float(np.mean(dev))
    total_m = int(m_r.sum())
    b_alg = float((40 * m_r + 36 * n_r).sum())
    achieved = b_alg / (dev_ms * 1e-3) / 1e9
    exact = cpu = None
    if not args.no_cpu:
        ref = oracle.bc(row, col, srcs)
        err = np.abs(bc.astype(np.float64) - ref)
        exact = bool(np.all(err <= 1e-5 * np.abs(ref) + 1e-5))
        k = 64
        t = time.perf_counter()
        oracle.bc(row, col, srcs[:k])
        tc = time.perf_counter() - t
        cpu = {"value": round(float(m_r[:k].sum()) / tc / 1e9, 5), "unit": "GTEPS", "cores": 1, "kind": "port",
               "sample": "oracle Brandes (restated bc_cpu, double) on %d of the %d sources, one thread, %.1f s; "
                         "the GPU result is within 1e-5 of the oracle over all sources: %s" % (k, len(srcs), tc, exact)}
    line = {
        "metric": BC_METRIC, "value": round(total_m / (ms * 1e-3) / 1e9, 3), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "betweenness centrality, R-MAT scale 16, edge factor 16, a=.57 b=.19 c=.19, ids "
                               "permuted, unit lengths, 1024 sources (seed 6); one step = one np_bc over all sources",
                   "n": int(G.n), "m": int(G.m), "sources": len(srcs), "reached_edges_per_step": total_m,
                   "levels": int(st.iterations), "device_ms": round(dev_ms, 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_bc_compact+k_bc_owner+k_bc_expand+k_bc_accum (host loop)",
                     "bytes_per_launch": b_alg},
        "cpu_baseline": cpu, "exact": exact,
    }
    print(json.dumps(line), flush=True)
    G.free()


COLOR_METRIC = "graph-color GTEPS + %HBM peak @R-MAT s20, dependency counting vs the reference's rounds, exact"


def color_bench(npa, ctx, args):
    """--config color: GRAPH_COLOR_GPU on R-MAT scale 20 (1,048,576 vertices, 16,777,216 edges,
    ids permuted).  Modes: `rounds` = the reference's algorithm (every uncolored vertex re-checks
    its smaller neighbours each round), `depcount` = the dependency-counting frontier; both are
    compared with the oracle (int32, exact).  GTEPS = m / step time.  Algorithmic bytes of
    depcount = 8 per edge (col read by the count and the fill passes) + 12 per smaller-id edge
    (its notify entry written and read, its counter decremented) + 40 per vertex (row bounds
    twice 16, notify bounds twice 16, counter 4, color 4)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    G = npa.Graph.rmat(ctx, 20, 16, wmin=1, wmax=1, seed=7)
    row, col, _ = G.download()
    src = np.repeat(np.arange(G.n, dtype=np.int64), np.diff(row))
    m_small = int((col < src).sum())
    results = {}
    outs = {}
    for name, method in (("rounds", 1), ("depcount", 0)):
        for _ in range(args.warmup):
            ctx.color(G, method)
        wall, dev = [], []
        for _ in range(args.steps):
            t = time.perf_counter()
            outs[name], st = ctx.color(G, method)
            wall.append((time.perf_counter() - t) * 1e3)
            dev.append(st.kernel_ms)
        ms = float(np.mean(wall))
        results[name] = {"gteps": G.m / (ms * 1e-3) / 1e9, "ms_per_step": ms, "device_ms": float(np.mean(dev)),
                         "rounds": int(st.iterations)}
        log("%-9s %8.2f ms/step  %7.3f GTEPS  %d rounds" % (name, ms, results[name]["gteps"], st.iterations))
    exact = cpu = None
    if not args.no_cpu:
        t = time.perf_counter()
        ref = oracle.color(row, col)
        tc = time.perf_counter() - t
        exact = bool(all(np.array_equal(o, ref) for o in outs.values()))
        cpu = {"value": round(G.m / tc / 1e9, 5), "unit": "GTEPS", "cores": 1, "kind": "port",
               "sample": "oracle one-pass coloring (increasing ids) over the whole graph, one thread, %.2f s; "
                         "both GPU methods equal it: %s" % (tc, exact)}
    best = results["depcount"]
    b_alg = 8.0 * G.m + 12.0 * m_small + 40.0 * G.n
    achieved = b_alg / (best["device_ms"] * 1e-3) / 1e9
    line = {
        "metric": COLOR_METRIC, "value": round(best["gteps"], 3), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(best["ms_per_step"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "graph coloring (GRAPH_COLOR_GPU), R-MAT scale 20, edge factor 16, a=.57 b=.19 c=.19, "
                               "ids permuted; one step colors the whole graph", "n": int(G.n), "m": int(G.m),
                   "smaller_id_edges": m_small, "colors": int(outs["depcount"].max())},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_color_count+k_color_fill+k_color_step (rounds captured in hipGraphs)", "bytes_per_launch": b_alg},
        "cpu_baseline": cpu, "exact": exact,
        "modes": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                  for k, v in results.items()},
        "speedup_depcount_vs_rounds": round(results["depcount"]["gteps"] / results["rounds"]["gteps"], 3),
    }
    print(json.dumps(line), flush=True)
    G.free()


PR_METRIC = "PageRank GTEPS (edges x iterations) + %HBM peak @R-MAT s20, within 1e-5 of the oracle"


def pr_bench(npa, ctx, args):
    """--config pagerank: PAGERANK_GPU on R-MAT scale 20 (1,048,576 vertices, 16,777,216 edges, ids
    permuted): float64 pull iterations until no rank moves by more than 1e-13 (damping 0.5, total
    1000).  GTEPS = m x iterations / step time (the step includes building the in-edge CSR).
    Algorithmic bytes = per iteration 12 per edge (in-edge id 4, contribution gather 8) + 48 per
    vertex (in-edge bounds 16, rank read twice 16, contribution and new rank written 16), plus 16
    per edge once for the in-edge CSR (out-edge ids read twice 8, in-edge id written 4, slot
    counter 4)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    G = npa.Graph.rmat(ctx, 20, 16, wmin=1, wmax=1, seed=8)
    row, col, _ = G.download()
    for _ in range(args.warmup):
        ctx.pagerank(G)
    wall, dev = [], []
    for _ in range(args.steps):
        t = time.perf_counter()
        rank, st = ctx.pagerank(G)
        wall.append((time.perf_counter() - t) * 1e3)
        dev.append(st.kernel_ms)
    ms, dev_ms, it = float(np.mean(wall)), float(np.mean(dev)), int(st.iterations)
    b_alg = it * (12.0 * G.m + 48.0 * G.n) + 16.0 * G.m
    achieved = b_alg / (dev_ms * 1e-3) / 1e9
    exact = cpu = None
    if not args.no_cpu:
        t = time.perf_counter()
        ref, rit = oracle.pagerank(row, col)
        tc = time.perf_counter() - t
        exact = bool(np.all(np.abs(rank - ref) <= 1e-5 * np.abs(ref) + 1e-9))
        cpu = {"value": round(G.m * rit / tc / 1e9, 5), "unit": "GTEPS", "cores": 1, "kind": "port",
               "sample": "oracle PageRank (restated iteration, float64) over the whole graph, one thread, %d "
                         "iterations in %.1f s; the GPU ranks are within 1e-5 of it: %s" % (rit, tc, exact)}
    line = {
        "metric": PR_METRIC, "value": round(G.m * it / (ms * 1e-3) / 1e9, 3), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "PageRank (PAGERANK_GPU), R-MAT scale 20, edge factor 16, a=.57 b=.19 c=.19, ids "
                               "permuted; one step = one np_pagerank to convergence", "n": int(G.n), "m": int(G.m),
                   "iterations": it, "device_ms": round(dev_ms, 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic("pagerank"),
                     "kernel": "k_pr_gate+k_pr_contrib+k_pr_pull_light+k_pr_pull_wave+k_pr_pull_huge "
                               "(8 iterations per readback)",
                     "bytes_per_launch": b_alg},
        "cpu_baseline": cpu, "exact": exact,
    }
    print(json.dumps(line), flush=True)
    G.free()


SPMV_METRIC = "SpMV GFLOP/s + %HBM peak @R-MAT s22 (float64), exact"


def copy_calibration(npa, ctx, nbytes):
    """Three device-to-device copies of nbytes (a copy kernel of known size for scaling the
    FETCH_SIZE / WRITE_SIZE counters of a profiled run); nothing when nbytes is 0."""
    if nbytes <= 0:
        return
    n4 = nbytes // 4
    a, b = npa.DeviceBuffer(ctx, n4), npa.DeviceBuffer(ctx, n4)
    for _ in range(3):
        npa._check(npa.load().np_dev_copy(ctx.h, npa.ctypes.c_void_p(b.ptr), npa.ctypes.c_void_p(a.ptr), n4 * 4, 2))
    log("copy calibration: 3 device-to-device copies of %d bytes" % (n4 * 4))
    a.free()
    b.free()


def stored_traffic(config, head):
    """(bytes, note) of the mode `head` from profiles/traffic_<config>.json, or (None, None)."""
    tpath = os.path.join(ROOT, "profiles", "traffic_%s.json" % config)
    try:
        tj = json.load(open(tpath))
    except (ValueError, OSError):
        return None, None
    rec = tj.get(head) or (tj.get("batched") if head == "learned_batched" else None)
    if not isinstance(rec, dict):
        return None, None
    return rec.get("bytes"), {"source": "stored, not measured in this run: profiles/traffic_%s.json" % config,
                              "unscaled_bytes": rec.get("unscaled_bytes"), "fetch_scale": rec.get("fetch_scale"),
                              "write_scale": rec.get("write_scale"), "profile": rec.get("profile")}


def measured_traffic(config):
    """HBM bytes per launch measured with rocprofv3 counters (profiles/traffic_<config>.json), or None."""
    tpath = os.path.join(ROOT, "profiles", "traffic_%s.json" % config)
    try:
        return json.load(open(tpath)).get(config)
    except (ValueError, OSError):
        return None


def sched_arg(s):
    """The --sched value of schedule s: level,log2_t_lo,log2_t_hi,wgs_per_cu,drive,delta,pack_hold."""
    return ",".join(str(int(x)) for x in (s.level, s.log2_t_lo, s.log2_t_hi, s.wgs_per_cu, s.drive, s.delta,
                                             s.pack_hold))


def traffic_child(npa, ctx, args):
    """--traffic-child: the solve whose HBM traffic the parent measures, run under rocprofv3 --pmc.
    The workload and batch of --config, one warm-up solve of the --sched solve=... schedule, then
    a reached-work count, the measured solve, a second count, and three 1 GiB device-to-device
    copies.  The two k_reached_work dispatches delimit the solve's dispatches in the counter file;
    the copies scale its counters."""
    unit = args.config == "bfs"
    W = build_workload(npa, ctx, "c3" if unit else args.config, args)
    v = [int(x) for x in [x for x in args.sched if x.startswith("solve=")][0].split("=")[1].split(",")]
    sch = npa.Sched(level=v[0], log2_t_lo=v[1], log2_t_hi=v[2], wgs_per_cu=v[3], drive=v[4], delta=v[5],
                    pack_hold=v[6])
    if sch.level == npa.LEVEL_LEARNED:
        ctx.model_set(np.fromfile(os.path.join(PKG_DIR, "sched_mlp.f32"), dtype=np.float32))
    ctx.set_option(npa.CTX_PACK, 1)
    B = npa.Batch(ctx, [W.graphs[gi] for gi, _ in W.inst], [s for _, s in W.inst], unit=unit)
    G = W.graphs[0]
    d = npa.DeviceBuffer(ctx, G.n)
    B.solve(sch)
    ctx.reached_work(G, d.ptr)
    st = B.solve(sch)
    ctx.reached_work(G, d.ptr)
    copy_calibration(npa, ctx, 1 << 30)
    print(json.dumps({"traffic_child": True, "drive": int(st.drive), "requested_drive": int(sch.drive),
                      "kernel_ms": float(st.kernel_ms)}), flush=True)
    d.free()
    B.free()
    for g in W.graphs:
        g.free()
    return 0


def _counter_pass(prof, counter, child_args, tmp):
    """One rocprofv3 --pmc pass over a child bench.py: (solve KiB, copy KiB, dispatches, child line)
    or a string saying why there is no figure."""
    import glob
    import signal
    import subprocess
    d = os.path.join(tmp, counter)
    cmd = [prof, "--pmc", counter, "-d", d, "-o", "run", "--output-format", "csv", "--",
           sys.executable, "-u", os.path.abspath(__file__)] + child_args
    with open(os.path.join(tmp, counter + ".out"), "w") as fo, open(os.path.join(tmp, counter + ".err"), "w") as fe:
        p = subprocess.Popen(cmd, stdout=fo, stderr=fe, start_new_session=True)
        try:
            p.wait(timeout=300)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.wait()
            return "the %s pass exceeded 300 s and was stopped" % counter
    child = None
    for ln in open(os.path.join(tmp, counter + ".out")):
        if ln.startswith("{") and "traffic_child" in ln:
            child = json.loads(ln)
    if child is None:
        return "the %s pass ended (status %d) before its solve finished" % (counter, p.returncode)
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        return "the %s pass (status %d) wrote no counter file" % (counter, p.returncode)
    r = solve_counter(files)
    if isinstance(r, str):
        return "the %s pass: %s" % (counter, r)
    return r + (child,)


def solve_counter(files):
    """From rocprofv3 counter-collection CSV files of a --traffic-child run: (the counter summed
    over the dispatches between the last two k_reached_work dispatches, its median over the last
    three copy kernels, the number of dispatches summed), or a string saying what is missing.
    A dispatch's rows (one per counter instance) are summed."""
    import csv
    import statistics
    disp = {}
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                e = disp.setdefault((f, r["Dispatch_Id"]), {"name": r["Kernel_Name"], "t": int(r["Start_Timestamp"]),
                                                            "v": 0.0})
                e["v"] += float(r["Counter_Value"])
    rows = sorted(disp.values(), key=lambda e: e["t"])
    marks = [i for i, e in enumerate(rows) if "k_reached_work" in e["name"]]
    copies = [i for i, e in enumerate(rows) if "copyBuffer" in e["name"]]
    if len(marks) < 2 or len([i for i in copies if i > marks[-1]]) < 3:
        return "no pair of markers followed by three copies"
    window = [e for e in rows[marks[-2] + 1:marks[-1]] if "copyBuffer" not in e["name"]]
    if not window:
        return "no dispatch between the markers"
    return sum(e["v"] for e in window), statistics.median(rows[i]["v"] for i in copies[-3:]), len(window)


def live_traffic(args, config, sch):
    """HBM bytes of one solve of schedule sch, measured in this run: two child processes build the
    same workload and run the solve under rocprofv3 --pmc, one for FETCH_SIZE and one for
    WRITE_SIZE.  Each counter is scaled by what it reads for a 1 GiB device-to-device copy in the
    same child (a copy reads and writes exactly 1 GiB).  Returns (bytes, note) or (None, reason)."""
    import shutil
    import tempfile
    if any(k.startswith("ROCPROF") for k in os.environ):
        return None, "bench.py itself runs under rocprofv3"
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None, "rocprofv3 not found"
    tmp = tempfile.mkdtemp(prefix="np_traffic_")
    child_args = ["--config", config, "--traffic-child", "--scale", str(args.scale), "--sources", str(args.sources),
                  "--sched", "solve=" + sched_arg(sch)]
    got = {}
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            t = time.time()
            r = _counter_pass(prof, counter, child_args, tmp)
            if isinstance(r, str):
                log("traffic: %s" % r)
                return None, r
            kib, copy_kib, nd, child = r
            if child["drive"] != child["requested_drive"]:
                return None, "the profiled solve ran on drive %d, not %d" % (child["drive"], child["requested_drive"])
            got[counter] = (kib, copy_kib, nd)
            log("traffic: %s %.0f KiB over %d dispatches of one solve; 1 GiB copy %.0f KiB (%.1fs)" % (
                counter, kib, nd, copy_kib, time.time() - t))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    (fk, fc, nd), (wk, wc, _) = got["FETCH_SIZE"], got["WRITE_SIZE"]
    fscale, wscale = 1048576.0 / fc, 1048576.0 / wc
    traffic = 1024.0 * (fk * fscale + wk * wscale)
    return int(traffic), {
        "source": "measured in this run: rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, each over one solve of this "
                  "schedule in a child process, scaled by the counter's reading for a 1 GiB device copy",
        "unscaled_bytes": int(1024.0 * (fk + wk)), "fetch_kib": round(fk), "write_kib": round(wk),
        "fetch_scale": round(fscale, 4), "write_scale": round(wscale, 4), "dispatches": nd}


def roofline_traffic(args, config, head, sch, batched):
    """roofline.traffic of the line: measured in this run (--traffic live, for a batch solve), else
    the stored record of profiles/traffic_<config>.json, marked as such."""
    reason = None
    if args.traffic == "live" and batched:
        traffic, note = live_traffic(args, config, sch)
        if traffic is not None:
            stored, _ = stored_traffic(config, head)
            note["stored_bytes"] = stored
            return traffic, note
        reason = note
    if args.traffic == "off":
        return None, None
    traffic, note = stored_traffic(config, head)
    if note is not None and reason:
        note["live_measurement"] = reason
    return traffic, note


def spmv_bench(npa, ctx, args):
    """--config spmv: SpMV_GPU on R-MAT scale 22 (4,194,304 rows, 67,108,864 entries, ids permuted)
    with the reference's values (data[i] = i % 13 + 0.5, x[i] = i % 17 + 0.5, float64: every sum is
    exact, so the check is equality).  GFLOP/s = 2 m / the product's device time (data, x and the
    CSR resident in HBM; the host copies np_spmv also makes are reported apart).  Algorithmic bytes
    = 20 per entry (column 4, value 8, x gather 8) + 16 per row (row bounds 8, y 8)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    G = npa.Graph.rmat(ctx, 22, 16, wmin=1, wmax=1, seed=9)
    row, col, _ = G.download()
    data, x = oracle.spmv_reference_values(G.n, G.m)
    for _ in range(args.warmup):
        ctx.spmv(G, data, x)
    wall, dev = [], []
    for _ in range(args.steps):
        t = time.perf_counter()
        y, st = ctx.spmv(G, data, x)
        wall.append((time.perf_counter() - t) * 1e3)
        dev.append(st.kernel_ms)
    dev_ms, wall_ms = float(np.mean(dev)), float(np.mean(wall))
    b_alg = 20.0 * G.m + 16.0 * G.n
    achieved = b_alg / (dev_ms * 1e-3) / 1e9
    exact = cpu = None
    if not args.no_cpu:
        t = time.perf_counter()
        ref = oracle.spmv(row, col, data, x)
        tc = time.perf_counter() - t
        exact = bool(np.array_equal(y, ref))
        cpu = {"value": round(2.0 * G.m / tc / 1e9, 4), "unit": "GFLOP/s", "cores": 1, "kind": "port",
               "sample": "oracle CSR product over the whole matrix, one thread, %.2f s; GPU result equal: %s" % (tc, exact)}
    line = {
        "metric": SPMV_METRIC, "value": round(2.0 * G.m / (dev_ms * 1e-3) / 1e9, 2), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dev_ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "SpMV (SpMV_GPU), R-MAT scale 22, edge factor 16, a=.57 b=.19 c=.19, ids permuted, "
                               "the reference's values; one step = one product", "n": int(G.n), "m": int(G.m),
                   "wall_ms_with_host_copies": round(wall_ms, 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic("spmv"),
                     "kernel": "k_pr_bins+k_spmv_light+k_spmv_wave+k_spmv_huge", "bytes_per_launch": b_alg},
        "cpu_baseline": cpu, "exact": exact,
    }
    print(json.dumps(line), flush=True)
    G.free()


BFS_METRIC = "BFS_REC_GPU GTEPS + %HBM peak @R-MAT s24, 16 sources, exact"


def bfs_bench(npa, ctx, args):
    """--config bfs: BFS_REC_GPU levels (the same frontier machinery with unit lengths) on the
    R-MAT scale-24 graph of the SSSP configuration (16,777,216 vertices, 268,435,456 edges, ids
    permuted), from its 16 sources (seed 3, out-degree >= 1); a step solves every source, levels
    left on the device (np_bfs_device).  Modes: the reference's bfs-rec solutions 0 (flat: every
    vertex at every level), 1 (naive recursion: a workgroup per frontier vertex, children pushed
    by the relaxers) and 2 (hierarchical: two levels per expansion) on the host loop as the
    reference runs them; solution 3 (consolidation) at levels 0/1/2 and level 2 on the graph
    drive; and the 16 sources source-packed in one batch on the batch autotuner's schedule.  GTEPS = sum over sources of the
    out-edges of reached vertices / step time.  Algorithmic bytes per source = 8 per reached
    edge (col 4 + the level gather 4; no weights) + 16 per reached vertex (row bounds 8, own
    level 4, final level write 4).  Every level array of every mode is compared with the
    oracle's BFS (restated bfs_cpu), which is also the CPU baseline."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    t0 = time.time()
    W = build_workload(npa, ctx, "c3", args)
    G = W.graphs[0]
    srcs = [s for _, s in W.inst]
    row, col, _ = G.download()
    deg = np.diff(row).astype(np.int64)
    log("bfs: %s, built in %.1fs" % (W.desc, time.time() - t0))
    levels = [ctx.bfs(G, s)[0] for s in srcs]
    reached = [lv != npa.UNDEFINED for lv in levels]
    n_r = np.array([int(r.sum()) for r in reached], np.int64)
    m_r = np.array([int(deg[r].sum()) for r in reached], np.int64)
    total_m = int(m_r.sum())
    b_alg = (8 * m_r + 16 * n_r).astype(np.float64)
    scheds = {
        "sol0_flat": npa.Sched(level=1, scheme=npa.SCHEME_BFS_FLAT, drive=npa.DRIVE_HOSTLOOP),
        "sol1_push": npa.Sched(level=1, log2_t_lo=0, log2_t_hi=0, frontier=npa.FRONTIER_PUSH,
                               drive=npa.DRIVE_HOSTLOOP),
        "sol2_hier": npa.Sched(level=1, scheme=npa.SCHEME_BFS_HIER, drive=npa.DRIVE_HOSTLOOP),
        "level0": npa.Sched(level=0), "level1": npa.Sched(level=1), "level2": npa.Sched(level=2),
        "level2_graph": npa.Sched(level=2, log2_t_lo=4, log2_t_hi=10, drive=npa.DRIVE_GRAPH),
        "batched": npa.Sched(level=2, log2_t_lo=4, log2_t_hi=10, drive=npa.DRIVE_GRAPH),
    }
    modes = [m for m in (args.modes.split(",") if args.modes else scheds) if m in scheds]
    d_lev = npa.DeviceBuffer(ctx, G.n)
    B = npa.Batch(ctx, [G] * len(srcs), srcs, unit=True) if "batched" in modes else None
    if B is not None:
        fixed = [x for x in args.sched if x.startswith("batched=")]
        if fixed:
            v = [int(x) for x in fixed[0].split("=")[1].split(",")]
            scheds["batched"] = npa.Sched(level=v[0], log2_t_lo=v[1], log2_t_hi=v[2], wgs_per_cu=v[3],
                                          drive=v[4] if len(v) > 4 else npa.DRIVE_PERSISTENT,
                                          pack_hold=v[6] if len(v) > 6 else 0)
        else:
            t = time.time()
            bs, bms = B.autotune(npa.Sched(level=1))
            scheds["batched"] = bs.copy(trace=0)
            log("batch autotune: %r (%.3f ms/solve, %.1fs)" % (bs, bms, time.time() - t))
    results, finals = {}, {}
    for mode in modes:
        sch = scheds[mode]
        if mode == "batched":
            def step(sch=sch):
                return [B.solve(sch).kernel_ms]
        else:
            def step(sch=sch):
                return [ctx.bfs_device(G, s, d_lev.ptr, sch).kernel_ms for s in srcs]
        ms, launch = time_mode(ctx, step, args.steps, args.warmup)
        results[mode] = {"gteps": total_m / (ms * 1e-3) / 1e9, "ms_per_step": ms,
                         "launch_ms": float(np.mean(launch)), "launches_per_step": len(launch) // args.steps}
        finals[mode] = [B.get(i) for i in range(len(srcs))] if mode == "batched" else \
            [ctx.bfs(G, s, sch)[0] for s in srcs]
        log("%-12s %9.2f ms/step %8.3f GTEPS  (%d launches/step, %.3f ms each)  %r" % (
            mode, ms, results[mode]["gteps"], results[mode]["launches_per_step"], results[mode]["launch_ms"], sch))
    exact = cpu = None
    if not args.no_cpu:
        from concurrent.futures import ThreadPoolExecutor
        model, nproc, usable = host_cpu()
        threads = max(1, min(len(srcs), usable, 16))
        t = time.perf_counter()
        oracle.bfs_levels(row, col, srcs[0])          # one source alone: the single-core time
        t1 = time.perf_counter() - t
        t = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:       # the oracle releases the GIL in its C code
            ref = list(ex.map(lambda s: oracle.bfs_levels(row, col, s), srcs))
        agg = time.perf_counter() - t
        exact = all(np.array_equal(f[i], ref[i]) for f in finals.values() for i in range(len(srcs))) and \
            all(np.array_equal(levels[i], ref[i]) for i in range(len(srcs)))
        cpu = {"value": round(total_m / agg / 1e9, 5), "unit": "GTEPS", "cores": threads, "kind": "port",
               "nproc": nproc, "usable_cpus": usable, "cpu_model": model, "aggregate_s": round(agg, 3),
               "per_source_s": round(t1, 3), "per_source_gteps_1core": round(float(m_r[0]) / t1 / 1e9, 5),
               "sample": "oracle queue BFS (restated bfs_cpu, bfs_rec.cpp:100-119): all %d sources, one BFS per "
                         "thread on %d threads, %.1f s wall (value = reached edges / that wall time); one source "
                         "alone on one core %.1f s; every level of every GPU mode equal: %s" % (
                             len(srcs), threads, agg, t1, exact)}
        log("cpu baseline: %.5f GTEPS aggregate (%d threads, %.1f s), %.1f s per source on one core; %s, nproc %d; "
            "exact %s" % (cpu["value"], threads, agg, t1, model, nproc, exact))
    head = max(results, key=lambda m: results[m]["gteps"])
    hr = results[head]
    per_launch_bytes = float(b_alg.sum()) if head == "batched" else float(b_alg.mean())
    achieved = per_launch_bytes / (hr["launch_ms"] * 1e-3) / 1e9
    traffic, traffic_note = roofline_traffic(args, "bfs", head, scheds[head], head == "batched")
    drive = scheds[head].drive
    kernel = ("k_pk_compact+k_decide+k_pk_expand" if head == "batched" else
              {"sol0_flat": "k_bfs_flat", "sol2_hier": "k_bfs_hier"}.get(head, "k_persistent" if drive == 0 else
                                                                          "k_compact+k_decide+k_expand"))
    kernel += " (%s drive)" % {0: "persistent", 1: "host-loop", 2: "graph"}[drive]
    line = {
        "metric": BFS_METRIC, "value": round(hr["gteps"], 3), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(hr["ms_per_step"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "BFS levels (BFS_REC_GPU), " + W.desc.replace("SSSP, ", "", 1) +
                               " (weights unused); one step solves every source",
                   "path": head, "schedule": repr(scheds[head]), "n": int(G.n), "m": int(G.m),
                   "sources": len(srcs), "reached_edges_per_step": total_m},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_note": traffic_note,
                     "kernel": kernel, "bytes_per_launch": per_launch_bytes},
        "cpu_baseline": cpu, "exact": exact,
        "modes": {k: {"gteps": round(v["gteps"], 3), "ms_per_step": round(v["ms_per_step"], 3),
                      "launch_ms": round(v["launch_ms"], 4), "schedule": repr(scheds[k])} for k, v in results.items()},
    }
    print(json.dumps(line), flush=True)
    copy_calibration(npa, ctx, args.copy_calibration)
    if B is not None:
        B.free()
    d_lev.free()
    G.free()
    return 1 if exact is False else 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--sources", type=int, default=16)
    ap.add_argument("--cpu-sources", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default=None)
    ap.add_argument("--sched", action="append", default=[],
                    help="MODE=level,log2_t_lo,log2_t_hi,wgs_per_cu[,drive[,delta[,pack_hold]]]: use this schedule for MODE (tuned or "
                         "batched) instead of running the autotuner (profiling runs hold only timed launches)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="after the run, copy this many bytes device-to-device 3 times (a copy kernel of known "
                         "size for scaling FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--traffic", choices=("live", "stored", "off"), default="live",
                    help="roofline.traffic of a batched line: live = rocprofv3 --pmc passes over one solve in child "
                         "processes during this run (about a minute), stored = the profiles/traffic_<config>.json "
                         "record, off = null")
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus %d is not supported "
              "(a batch's instances would be split across GPUs, one process each, with no communication)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    npa = load_pkg()
    ctx = npa.Context(0)
    try:
        rc = traffic_child(npa, ctx, args) if args.traffic_child else run_config(npa, ctx, args)
    finally:
        ctx.close()  # every device object is released before the interpreter exits
    if rc:
        sys.exit(rc)


def run_config(npa, ctx, args):
    ncu, occ = ctx.device_info()
    log("device: %d CUs, persistent occupancy %d workgroups/CU" % (ncu, occ))
    if args.config == "bc":
        bc_bench(npa, ctx, args)
        return 0
    if args.config == "color":
        color_bench(npa, ctx, args)
        return 0
    if args.config == "pagerank":
        pr_bench(npa, ctx, args)
        return 0
    if args.config == "spmv":
        spmv_bench(npa, ctx, args)
        return 0
    if args.config == "bfs":
        return bfs_bench(npa, ctx, args)
    t0 = time.time()
    W = build_workload(npa, ctx, args.config, args)
    log("%s: %d graphs, %d instances, built in %.1fs" % (W.desc, len(W.graphs), len(W.inst), time.time() - t0))

    plain = npa.Sched(level=1)
    nmax = max(g.n for g in W.graphs)
    dbuf = npa.DeviceBuffer(ctx, nmax)

    default_modes = {"c1": "plain,tuned", "c2": "level0,level1,level2,tuned,batched",
                     "c3": "plain,tuned,learned,concat,batched,learned_batched",
                     "c4": "plain,tuned,concat,batched", "c5": "plain,tuned,learned,batched,learned_batched"}[args.config]
    modes = [m for m in (args.modes or default_modes).split(",") if m]
    scheds = {"plain": plain, "level0": npa.Sched(level=0), "level1": npa.Sched(level=1),
              "level2": npa.Sched(level=2)}
    fixed = {}
    for spec in args.sched:
        mode, vals = spec.split("=")
        v = [int(x) for x in vals.split(",")]
        fixed[mode] = npa.Sched(level=v[0], log2_t_lo=v[1], log2_t_hi=v[2], wgs_per_cu=v[3],
                                drive=v[4] if len(v) > 4 else npa.DRIVE_PERSISTENT, delta=v[5] if len(v) > 5 else 0,
                                pack_hold=v[6] if len(v) > 6 else 0)
    if "tuned" in fixed:
        scheds["tuned"] = fixed["tuned"]
    elif "tuned" in modes or "concat" in modes or "learned" in modes:
        t = time.time()
        big = max(range(len(W.graphs)), key=lambda i: W.graphs[i].m)
        srcs = [s for gi, s in W.inst if gi == big][:2]
        tuned, tms = ctx.autotune(W.graphs[big], srcs, npa.Sched(level=1))
        scheds["tuned"] = tuned.copy(trace=0)
        log("autotune: %r (%.3f ms/source, %.1fs)" % (tuned, tms, time.time() - t))
    model_path = os.path.join(PKG_DIR, "sched_mlp.f32")
    if os.path.exists(model_path):
        ctx.model_set(np.fromfile(model_path, dtype=np.float32))
        if "tuned" in scheds:
            scheds["learned"] = scheds["tuned"].copy(level=npa.LEVEL_LEARNED)
    elif "learned" in modes or "learned_batched" in modes:
        log("no trained scheduler model at %s: learned modes skipped" % model_path)
        modes = [m for m in modes if m not in ("learned", "learned_batched")]

    results, finals, batches = {}, {}, {}
    for mode in modes:
        if mode in ("batched", "concat", "learned_batched"):
            # batched: instances sharing a graph source-packed; concat: every instance in the
            # concatenated slot space (packing off)
            ctx.set_option(npa.CTX_PACK, 0 if mode == "concat" else 1)
            B = npa.Batch(ctx, [W.graphs[gi] for gi, _ in W.inst], [s for _, s in W.inst])
            ctx.set_option(npa.CTX_PACK, 1)
            batches[mode] = B
            if mode == "concat":
                scheds[mode] = scheds["tuned"]
            elif "batched" in fixed:
                scheds["batched"] = fixed["batched"]
            elif "batched" not in scheds:
                t = time.time()
                bs, bms = B.autotune(npa.Sched(level=1))
                scheds["batched"] = bs.copy(trace=0)
                log("batch autotune: %r (%.3f ms/solve, %.1fs); layout %s" % (bs, bms, time.time() - t, B.layout()))
            if mode == "learned_batched":
                scheds[mode] = scheds["batched"].copy(level=npa.LEVEL_LEARNED)

            def step(B=B, sch=scheds[mode]):
                st = B.solve(sch)
                if st.drive != sch.drive:
                    log("note: the batch ran on drive %d (no cooperative queue for this process)" % st.drive)
                return [st.kernel_ms]
        else:
            sch = scheds[mode]

            def step(sch=sch):
                out = []
                for gi, s in W.inst:
                    st = ctx.sssp_device(W.graphs[gi], s, dbuf.ptr, sch)
                    if st.drive != sch.drive:
                        log("note: a solve ran on drive %d (no cooperative queue for this process)" % st.drive)
                    out.append(st.kernel_ms)
                return out
        ms_step, launch_ms = time_mode(ctx, step, args.steps, args.warmup)
        results[mode] = {"ms_per_step": ms_step, "launch_ms": float(np.mean(launch_ms)),
                         "launches_per_step": len(launch_ms) // args.steps}
        if mode in batches:
            finals[mode] = [batches[mode].get(i) for i in range(len(W.inst))]
        else:
            out = []
            for gi, s in W.inst:
                ctx.sssp_device(W.graphs[gi], s, dbuf.ptr, scheds[mode])
                out.append(dbuf.to_host()[:W.graphs[gi].n])
            finals[mode] = out
        if mode == modes[0]:
            # work counted for throughput, from the distances (identical in every mode)
            work = []
            for (gi, s), d in zip(W.inst, finals[mode]):
                dbuf.from_host(np.concatenate([d, np.zeros(nmax - len(d), np.int32)]))
                work.append(ctx.reached_work(W.graphs[gi], dbuf.ptr))
            n_r = np.array([a for a, _ in work], np.int64)
            m_r = np.array([b for _, b in work], np.int64)
            total_m = int(m_r.sum())
            b_alg = 12 * m_r + 16 * n_r
        results[mode]["gteps"] = total_m / (ms_step * 1e-3) / 1e9
        log("%-8s %9.2f ms/step %8.3f GTEPS  (%d launches/step, %.3f ms each)  %r" % (
            mode, ms_step, results[mode]["gteps"], results[mode]["launches_per_step"], results[mode]["launch_ms"],
            scheds[mode]))

    cpu, exact = None, None
    if not args.no_cpu:
        hosts = [g.download() for g in W.graphs]
        k = args.cpu_sources if args.cpu_sources is not None else W.cpu_sample
        cpu, exact = cpu_leg(W, hosts, finals, modes, m_r, k)

    head = max(results, key=lambda m: results[m]["gteps"])
    hr = results[head]
    # A "launch" is one solve: one cooperative kernel on the persistent drive, or the
    # compaction / decision / expansion kernels of every iteration on the graph and
    # host-loop drives; its duration is the solve's device time (hipEvents on the
    # library's stream).
    persistent = scheds[head].drive == npa.DRIVE_PERSISTENT
    if head in batches:
        npk, ncat = batches[head].layout()
        names = ([("k_pk_persistent" if persistent else "k_pk_compact+k_decide+k_pk_expand")] if npk else []) + \
                ([("k_persistent" if persistent else "k_compact+k_decide+k_expand")] if ncat else [])
        kernel = " and ".join(names)
        per_launch_bytes = float(b_alg.sum()) / max(1, hr["launches_per_step"])
    else:
        kernel = "k_persistent" if persistent else "k_compact+k_decide+k_expand"
        per_launch_bytes = float(b_alg.mean())
    kernel += " (%s drive)" % {0: "persistent", 1: "host-loop", 2: "graph"}[scheds[head].drive]
    achieved = per_launch_bytes / (hr["launch_ms"] * 1e-3) / 1e9
    traffic, traffic_note = roofline_traffic(args, args.config, head, scheds[head], head in batches)
    # work of one traced solve of the fastest mode (not timed): edge rows read (out-degrees of the queued
    # vertices), (edge, source) items and queued (vertex, source) pairs; for packs the mean number of
    # sources pending per edge row read and the re-expansion of pairs over the reached ones
    work = None
    if head in batches:
        st = batches[head].solve(scheds[head].copy(trace=1))
        n_r_total = int(n_r.sum())
        work = {"queued_pairs": int(st.pairs_total), "queued_pairs_over_reached": round(st.pairs_total / max(1, n_r_total), 3),
                "iterations": int(st.iterations)}
        if st.edges_total:
            work.update({"edge_rows": int(st.edges_total), "items": int(st.items_total),
                         "pending_sources_per_edge_row": round(st.items_total / st.edges_total, 2),
                         "edge_rows_over_m": round(st.edges_total / max(1, sum(g.m for g in W.graphs)), 3)})
    base = results.get("plain", {}).get("gteps")
    line = {
        "metric": METRIC if args.config == "c3" else METRIC + " [config %s]" % args.config,
        "value": round(hr["gteps"], 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(hr["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": W.desc + "; one step solves every instance", "path": head,
                   "schedule": repr(scheds.get(head)),
                   "instances": len(W.inst), "graphs": len(W.graphs),
                   "n": int(sum(g.n for g in W.graphs)), "m": int(sum(g.m for g in W.graphs)),
                   "reached_edges_per_step": total_m},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_note": traffic_note,
                     "kernel": kernel, "bytes_per_launch": per_launch_bytes},
        "cpu_baseline": cpu,
        "exact": exact,
        "work": work,
        "modes": {k: {"gteps": round(v["gteps"], 3), "ms_per_step": round(v["ms_per_step"], 3),
                      "launch_ms": round(v["launch_ms"], 4), "schedule": repr(scheds.get(k))}
                  for k, v in results.items()},
        "speedup_learned_vs_plain": round(results["learned"]["gteps"] / base, 3) if base and "learned" in results else None,
        "speedup_batched_vs_plain": round(results["batched"]["gteps"] / base, 3) if base and "batched" in results else None,
        "speedup_learned_batched_vs_plain": (round(results["learned_batched"]["gteps"] / base, 3)
                                             if base and "learned_batched" in results else None),
        "speedup_learned_vs_tuned": (round(results["learned"]["gteps"] / results["tuned"]["gteps"], 3)
                                     if "learned" in results and "tuned" in results else None),
    }
    print(json.dumps(line), flush=True)
    copy_calibration(npa, ctx, args.copy_calibration)
    for B in batches.values():
        B.free()
    dbuf.free()
    for g in W.graphs:
        g.free()
    return 1 if exact is False else 0


if __name__ == "__main__":
    main()
