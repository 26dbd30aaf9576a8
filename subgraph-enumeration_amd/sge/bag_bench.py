"""Time plans with and without a materialised bag (plain executor, one launch
per plan).

    python -m sge.bag_bench [c2] [rmat12] [--out results.json]

For each pattern the cost model's plan is timed without a bag and with every
valid bag of at most `--max-bag` vertices; a bag's figure is its trie build
(enumerate the bag relation, radix sort, per-depth scans) plus the join.
Counts must be identical.  Workloads: the 16 six-vertex patterns over the
Chung-Lu graph n=7,115 m=100,762 (IEP mode), and every fourth of the 78
seven-vertex patterns over R-MAT scale 12 (seed 4, edge factor 16, IEP mode).
"""
import argparse
import json

import sge
import sge.workloads as W


def best_time(g, plan, reps):
    best = None
    for _ in range(reps):
        c, st = sge.count(g, plan)
        tot = st.kernel_ms + st.trie_ms
        if best is None or tot < best[0]:
            best = (tot, c[0], st.trie_ms)
    return best


def run(label, g, pats, mode, max_bag, reps):
    g.to_device()
    res = []
    for i, p in enumerate(pats):
        plan = sge.build_plan(p, g, mode)
        base = best_time(g, plan, reps)
        rows = []
        for b in sge.valid_bags(plan):
            if bin(b).count("1") > max_bag:
                continue
            try:
                r = best_time(g, sge.set_bag(plan, b), reps)
            except sge.SgeError as e:  # relation above the tuple cap
                rows.append({"bag": b, "error": str(e)})
                continue
            if r[1] != base[1]:
                raise SystemExit(f"{label} pattern {i} bag {b}: count {r[1]} != {base[1]}")
            rows.append({"bag": b, "ms": r[0], "trie_ms": r[2]})
            print(f"[{label}] pattern {i} bag {b}: {r[0]:.3f} ms (trie {r[2]:.3f} ms)", flush=True)
        timed = [r for r in rows if "ms" in r]
        best = min(timed, key=lambda r: r["ms"]) if timed else None
        print(f"[{label}] pattern {i}: no bag {base[0]:.3f} ms | best bag {best} of {len(rows)} | "
              f"decomposition bags {sge.plan_bags(p, plan)}", flush=True)
        res.append({"pattern": i, "no_bag_ms": base[0], "count": base[1], "bags": rows})
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workloads", nargs="*", default=["c2", "rmat12"])
    ap.add_argument("--max-bag", type=int, default=4)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    out = {}
    if "c2" in args.workloads:
        out["c2"] = run("chunglu7115-sixvertex16-iep", W.graph_c2(), W.six_vertex_set(), sge.MODE_IEP,
                        args.max_bag, args.reps)
    if "rmat12" in args.workloads:
        out["rmat12"] = run("rmat12-sevenvertex-every4th-iep", W.graph_rmat(12, 4), W.seven_vertex_set()[::4],
                            sge.MODE_IEP, args.max_bag, args.reps)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
