"""Train the learned attribute-order model on the GPU and evaluate it.

    python -m sge.train_order_model --out sge/models/order_mlp.npz

Logs (order, measured work) pairs with the plain executor on a strided
root-edge sample of the Chung-Lu graph (n=7,115, m=100,762, seed 1), each order
under a work cap relative to the cost model's order (learned.log_work), for
patterns disjoint from the benchmark set: 400 connected 7-vertex patterns with
11-14 edges (sampler seed 5) and the connected 6-vertex patterns with 8-10
edges.  Fits the MLP (sge_mlp_* kernels), then evaluates it on the 16 benchmark
6-vertex patterns with full exact runs: predicted vs measured work, and the
measured time of the learned choice against the analytic cost model's choice.
Writes the weights (.npz) and a JSON report next to them.

Logging can be split over several runs and fitted afterwards:

    python -m sge.train_order_model --log-only pairs_0.npz --first 0 --count 200
    python -m sge.train_order_model --log-only pairs_1.npz --first 200 --count 256
    python -m sge.train_order_model --data pairs_0.npz,pairs_1.npz --out sge/models/order_mlp.npz

The pairs logged for the committed model are in sge/models/order_pairs.npz.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import sge  # noqa: E402
import sge.workloads as W  # noqa: E402
from sge import learned as L  # noqa: E402


# graphs (order, work) pairs are logged on: the Chung-Lu benchmark graph and
# R-MAT graphs of the benchmark's parameters at other scales and seeds (none of
# them the graph of a benchmark configuration the model is evaluated on)
LOG_GRAPHS = {
    "c2": lambda: W.graph_c2(),
    "rmat12": lambda: W.graph_rmat(12, 11, device_build=True),
    "rmat14": lambda: W.graph_rmat(14, 12, device_build=True),
}


# evaluation sets: the whole job of each, exact, with the same executor for
# both choices; the Chung-Lu graph of seed 1 is one the pairs are logged on (its
# patterns are not); the Chung-Lu graph of seed 9 (same parameters) and R-MAT
# scale 8 (seed 7) are graphs no pair is logged on
EVAL_SETS = {
    "c3": ("chung_lu(n=7115, m=100762, gamma=2.1, cap=1100, seed=1), 16 x 6-vertex", lambda: W.graph_c2(),
           W.six_vertex_set),
    "chunglu9": ("chung_lu(n=7115, m=100762, gamma=2.1, cap=1100, seed=9), 16 x 6-vertex",
                 lambda: sge.Graph.chung_lu(7115, 100762, 2.1, 1100.0, 9), W.six_vertex_set),
    "rmat8": ("rmat(scale=8, ef=16, seed=7), 78 x 7-vertex", lambda: W.graph_rmat(8, 7, device_build=True),
              W.seven_vertex_set),
}


def evaluate_sets(model, names, mode, stride, log):
    """Per evaluation set: the learned choice against the planner's choice (whole
    job), and the judges' ranking quality: Spearman correlation of measured time
    with the network's prediction, the planner's analytic cost and the
    reference's sampled cost model over each pattern's two orders."""
    out = {}
    for name in names:
        desc, gfn, pfn = EVAL_SETS[name]
        g = gfn()
        g.to_device()
        pats = pfn()
        rows = L.evaluate(model, g, pats, mode, stride, log=log)
        t_cost = sum(r["cost_ms"] for r in rows)
        t_learn = sum(r["learned_ms"] for r in rows)
        ratios = [r["learned_ms"] / max(r["cost_ms"], 1e-6) for r in rows]
        changed = [r for r in rows if r["learned_order"] != r["cost_order"] and not r["fallback"]]
        # on the changed patterns: does each judge order the two choices as measured?
        agree = {"network": 0, "analytic": 0, "sampled": 0}
        for r in changed:
            p = pats[r["pattern"]]
            faster_learned = r["learned_ms"] < r["cost_ms"]
            agree["network"] += int((r["pred_rel_log2"] < 0) == faster_learned)
            ca = sge.order_cost(p, g, r["learned_order"], mode) < sge.order_cost(p, g, r["cost_order"], mode)
            agree["analytic"] += int(ca == faster_learned)
            cs = (sge.order_cost_sampled(p, g, r["learned_order"], mode, L.JUDGE_SEED) <
                  sge.order_cost_sampled(p, g, r["cost_order"], mode, L.JUDGE_SEED))
            agree["sampled"] += int(cs == faster_learned)
        out[name] = {"set": desc, "eval_stride": stride, "total_ms": {"cost_model": t_cost, "learned": t_learn,
                                                                     "speedup": t_cost / t_learn if t_learn else None},
                     "max_slowdown": max(ratios), "changed": len(changed),
                     "fallbacks": int(sum(r["fallback"] for r in rows)), "capped": int(sum(r["capped"] for r in rows)),
                     "judges_agree_with_measured_on_changed": agree, "per_pattern": rows}
        log(f"[eval {name}] cost-model {t_cost:.1f} ms, learned {t_learn:.1f} ms "
            f"({t_cost / t_learn:.2f}x), worst pattern {max(ratios):.2f}x slower, {len(changed)} changed, "
            f"judges agree {agree}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "models", "order_mlp.npz"))
    ap.add_argument("--patterns", type=int, default=400)
    ap.add_argument("--orders", type=int, default=6)
    ap.add_argument("--stride", type=int, default=1024)
    ap.add_argument("--eval-stride", type=int, default=1)
    ap.add_argument("--mode", default="iep", choices=["wcoj", "explore", "iep"])
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--log", default=None)
    ap.add_argument("--graph", default="c2", choices=sorted(LOG_GRAPHS),
                    help="data graph the (order, work) pairs are logged on")
    ap.add_argument("--log-only", default=None, help="log (order, work) pairs to this .npz and stop")
    ap.add_argument("--first", type=int, default=0, help="first training pattern to log")
    ap.add_argument("--count", type=int, default=None, help="number of training patterns to log")
    ap.add_argument("--data", default=None, help="comma-separated .npz files of logged pairs to fit from")
    ap.add_argument("--eval", default="c3,chunglu9,rmat8", help="evaluation sets: " + ", ".join(sorted(EVAL_SETS)))
    ap.add_argument("--eval-only", action="store_true",
                    help="evaluate the model at --out on the benchmark set and rewrite its report; no training")
    args = ap.parse_args()
    mode = {"wcoj": sge.MODE_WCOJ, "explore": sge.MODE_EXPLORE, "iep": sge.MODE_IEP}[args.mode]
    logf = open(args.log, "a") if args.log else None

    def log(*a):
        s = " ".join(str(x) for x in a)
        print(s, flush=True)
        if logf:
            logf.write(s + "\n")
            logf.flush()

    t0 = time.time()
    g = W.graph_c2()
    g.to_device()
    log(f"graph n={g.n} m={g.m} in {time.time() - t0:.1f}s")
    lg = LOG_GRAPHS[args.graph]()
    lg.to_device()
    if args.eval_only:
        model = L.load(args.out)
        rep_path = os.path.splitext(args.out)[0] + "_report.json"
        report = json.load(open(rep_path)) if os.path.exists(rep_path) else {}
        report["eval"] = evaluate_sets(model, args.eval.split(","), mode, args.eval_stride, log)
        with open(rep_path, "w") as f:
            json.dump(report, f, indent=1)
        return
    bench_codes = [p.info().canon for p in W.six_vertex_set()]
    train_p = W.training_set(args.patterns) + sge.sample_patterns(6, 8, 10, 80, 5, bench_codes)
    if args.data:
        parts = [np.load(f, allow_pickle=False) for f in args.data.split(",")]
        X = np.concatenate([d["X"] for d in parts])
        y = np.concatenate([d["y"] for d in parts])
        c = np.concatenate([d["c"] for d in parts])
        grp = np.concatenate([d["group"] + 1_000_000 * k for k, d in enumerate(parts)])
        graphs = [str(d["graph"]) if "graph" in d.files else "c2" for d in parts]
        log(f"loaded {len(X)} logged orders of {len(np.unique(grp))} patterns from {args.data}")
    else:
        graphs = [args.graph]
        sel = train_p[args.first:args.first + args.count if args.count else None]
        X, y, c, cap, grp, ords = L.log_pairs(lg, sel, mode, stride=args.stride, orders_per_pattern=args.orders,
                                              seed=args.first, log=log)
        if args.log_only:
            np.savez(args.log_only, X=X, y=y, c=c, capped=cap, group=grp, orders=ords,
                     graph=np.array(args.graph), first=np.array(args.first))
            log(f"wrote {len(X)} orders to {args.log_only}")
            return
    model = L.fit_model(X, y, grp, epochs=args.epochs, log=log)
    model.sizes = sorted({p.k for p in train_p})  # the model only ranks orders of these sizes
    rel = L.relative_targets(y, grp)
    pred = model.predict_rel_log2(X)
    fit = {"train_rmse_rel_log2": float(np.sqrt(np.mean((pred - rel) ** 2))),
           "train_corr_rel": float(np.corrcoef(pred, rel)[0, 1]),
           "cost_model_corr_rel": float(np.corrcoef(L.relative_targets(c, grp), rel)[0, 1])}
    log("fit", fit)
    model.save(args.out)
    report = {"log_graphs": sorted(set(graphs)), "mode": args.mode, "train_patterns": len(train_p),
              "train_orders": int(len(X)), "train_stride": args.stride, "fit": fit,
              "loss_first_last": [model.history["loss"][0], model.history["loss"][-1]],
              "best_epoch": model.history["best_epoch"],
              "eval": evaluate_sets(model, args.eval.split(","), mode, args.eval_stride, log),
              "wall_s": time.time() - t0}
    with open(os.path.splitext(args.out)[0] + "_report.json", "w") as f:
        json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
