"""Trains the learned scheduler of the SSSP path (the fp32 MLP of include/np_sssp.h).

Data: per-iteration timings of the consolidated kernels on three graph families (every
graph generated on the device, renumbered by in-degree when skewed, as the library
does by default):
  rmat  R-MAT, scale drawn from 18..22, edge factor 16, (.57,.19,.19), weights U[1,255],
        graph seeds from 100
  er    Erdos-Renyi, n = 2^17..2^20, m = 16 n, weights U[1,100], seeds from 300
  grid  4-neighbour grids 256x256 .. 1024x512 + 1% long-range edges, weights U[1,1000],
        seeds from 400
For each graph one source is solved with the stepping API: at a sampled iteration the
frontier state is saved, every candidate schedule (consolidation level x log2 bin
thresholds) expands it once from that same state and is timed with hipEvents, then the
run advances with the plain schedule (block level, THRESHOLD 64).  Iterations 0..31 are
all sampled, later ones every 16th (grids run thousands of iterations).  A sample is
(features of the frontier, fastest candidate); the label is that candidate's level, the
regression targets its two log2 thresholds.

Training: full-batch Adam on the GPU through np_model_train (the project's forward and
backward kernels).  Output: sched_mlp.f32 (NP_MLP_NPARAM float32) next to this file and
a JSON summary (dataset size, regret of the plain schedule and of the network on held-out
graphs -- the last 20% of each family -- per family).

    python nested-parallelism_amd/train_scheduler.py [--rmat-runs 200] [--er-runs 30] [--grid-runs 30]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_pkg():
    name = "nested_parallelism_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def candidates(npa):
    out = []
    for lo in (3, 4, 5, 6):
        out.append(npa.Sched(level=0, log2_t_lo=lo, log2_t_hi=lo))
    for lv in (1, 2):
        for lo in (3, 4, 5, 6):
            for hi in (6, 9, 12):
                if hi >= lo:
                    out.append(npa.Sched(level=lv, log2_t_lo=lo, log2_t_hi=hi))
    return out


FAMILIES = ("rmat", "er", "grid")
GRID_SHAPES = ((256, 256), (512, 256), (512, 512), (768, 512), (1024, 512))


def make_graph(npa, ctx, family, r, seed, scales):
    if family == "rmat":
        scale = scales[r % len(scales)]
        return npa.Graph.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, 1, 255, seed=seed, permute=True), "s%d" % scale
    if family == "er":
        n = 1 << (17 + r % 4)
        return npa.Graph.er(ctx, n, 16 * n, 1, 100, seed=seed), "n%d" % n
    rows, cols = GRID_SHAPES[r % len(GRID_SHAPES)]
    return npa.Graph.grid(ctx, rows, cols, None, 1, 1000, seed=seed), "%dx%d" % (rows, cols)


def sampled(it):
    return it < 32 or it % 16 == 0


def collect(npa, ctx, plan, scales, reps, log):
    """plan: list of (family, run index, seed)."""
    cands = candidates(npa)
    plain = npa.Sched(level=1, log2_t_lo=6, log2_t_hi=6)
    plain_idx = [i for i, c in enumerate(cands) if (c.level, c.log2_t_lo, c.log2_t_hi) == (1, 6, 6)][0]
    X, T, meta = [], [], []
    for k, (family, r, seed) in enumerate(plan):
        g, shape = make_graph(npa, ctx, family, r, seed, scales)
        row = g.download()[0]
        src = int(npa.pick_sources(row, 1, seed)[0])
        ctx.step_begin(g, src)
        it = 0
        while True:
            cnt, done = ctx.step_frontier()
            if done:
                break
            if sampled(it):
                times = [ctx.step_try(c, reps) for c in cands]
                X.append(npa.features(cnt, it, g.n))
                T.append(times)
                meta.append((FAMILIES.index(family), seed, it))
            ctx.step_advance(plain)
            it += 1
        ctx.step_end(g.n)
        g.free()
        if (k + 1) % 10 == 0:
            log("  collected %d graphs, %d samples (last: %s %s, %d iterations)" % (k + 1, len(X), family, shape, it))
    return np.array(X, np.float32), np.array(T, np.float64), np.array(meta), cands, plain_idx


def labels(T, cands):
    best = T.argmin(axis=1)
    lab = np.array([cands[b].level for b in best], np.int32)
    tgt = np.array([[cands[b].log2_t_lo, cands[b].log2_t_hi] for b in best], np.float32)
    return best, lab, tgt


def predict_index(out, cands):
    """Network output -> the candidate it selects (level = argmax, thresholds rounded)."""
    res = []
    for o in out:
        lv = int(np.argmax(o[:3]))
        lo = int(np.clip(np.rint(o[3]), 0, 30))
        hi = int(np.clip(np.rint(o[4]), 0, 30))
        # nearest candidate of that level
        best, bd = 0, 1e9
        for i, c in enumerate(cands):
            if c.level != lv:
                continue
            d = abs(c.log2_t_lo - lo) + (abs(c.log2_t_hi - hi) if lv > 0 else 0)
            if d < bd:
                best, bd = i, d
        res.append(best)
    return np.array(res)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rmat-runs", type=int, default=200)
    ap.add_argument("--er-runs", type=int, default=30)
    ap.add_argument("--grid-runs", type=int, default=30)
    ap.add_argument("--scales", default="18,19,20,21,22")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--lam", type=float, default=0.05)
    ap.add_argument("--out", default=os.path.join(HERE, "sched_mlp.f32"))
    ap.add_argument("--summary", default=None)
    ap.add_argument("--dataset", default=None)
    args = ap.parse_args()

    def log(*a):
        print(*a, flush=True)

    npa = load_pkg()
    ctx = npa.Context(0)
    t0 = time.time()
    scales = [int(s) for s in args.scales.split(",")]
    runs = {"rmat": args.rmat_runs, "er": args.er_runs, "grid": args.grid_runs}
    seed0 = {"rmat": 100, "er": 300, "grid": 400}
    plan = [(f, r, seed0[f] + r) for f in FAMILIES for r in range(runs[f])]
    X, T, meta, cands, plain_idx = collect(npa, ctx, plan, scales, args.reps, log)
    log("dataset: %d samples from %d graphs in %.1fs" % (len(X), len(plan), time.time() - t0))
    if args.dataset:
        np.savez_compressed(args.dataset, X=X, T=T, meta=meta)
    best, lab, tgt = labels(T, cands)
    # hold out the last 20% of the graphs of each family for evaluation
    fam, seeds = meta[:, 0], meta[:, 1]
    te = np.zeros(len(X), bool)
    for i, f in enumerate(FAMILIES):
        te |= (fam == i) & (seeds >= seed0[f] + int(runs[f] * 0.8))
    tr = ~te
    p0 = init_params(npa.MLP_NPARAM)
    ctx.model_set(p0)
    t1 = time.time()
    losses = ctx.model_train(X[tr], lab[tr], tgt[tr], args.epochs, lr=args.lr, lam=args.lam)
    log("trained %d epochs on the GPU in %.1fs: loss %.4f -> %.4f" % (args.epochs, time.time() - t1,
                                                                       losses[0], losses[-1]))
    params = ctx.model_get()
    params.astype(np.float32).tofile(args.out)

    def summarize(mask):
        if mask.sum() == 0:
            return {}
        out = ctx.model_forward(X[mask])
        pick = predict_index(out, cands)
        Tm = T[mask]
        t_best = Tm.min(axis=1).sum()
        t_plain = Tm[:, plain_idx].sum()
        t_net = Tm[np.arange(len(pick)), pick].sum()
        return {"samples": int(mask.sum()), "ms_best": float(t_best), "ms_plain": float(t_plain),
                "ms_network": float(t_net), "speedup_network_vs_plain": float(t_plain / t_net),
                "speedup_best_vs_plain": float(t_plain / t_best),
                "level_accuracy": float((np.argmax(out[:, :3], 1) == lab[mask]).mean())}

    summary = {"graphs": runs, "samples": int(len(X)), "candidates": len(cands), "epochs": args.epochs,
               "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
               "train": summarize(tr), "heldout": summarize(te),
               "heldout_by_family": {f: summarize(te & (fam == i)) for i, f in enumerate(FAMILIES)},
               "label_histogram": np.bincount(lab, minlength=3).tolist()}
    log(json.dumps(summary, indent=1))
    if args.summary:
        with open(args.summary, "w") as f:
            json.dump(summary, f, indent=1)
    ctx.close()


def init_params(nparam):
    rng = np.random.default_rng(12345)
    p = np.zeros(nparam, np.float32)
    o = 0
    for fan_in, fan_out in ((10, 32), (32, 32), (32, 5)):
        p[o:o + fan_in * fan_out] = rng.standard_normal(fan_in * fan_out) * np.sqrt(2.0 / fan_in)
        o += fan_in * fan_out + fan_out
    return p


if __name__ == "__main__":
    main()
