"""Benchmark: exact matches/s over a pattern set, batched vs plain GPU executor.

Default workload (the headline): the 78 connected 7-vertex patterns with 15-19
edges over an undirected, deduplicated R-MAT graph (scale 20, edge factor 16,
a,b,c,d = .57,.19,.19,.05, seed 4), int32 degree-ordered CSR resident in HBM.

One step = one pass of the batched executor over the whole pattern set (one
persistent launch, replayed from a captured hipGraph).  `value` is total matches
per second of that pass.  The plain executor (one launch per pattern) is timed
on the same plans for the comparison.  The CPU baseline is the oracle (a port of
the reference's join) timed on this host over a bounded sample of root tasks.

    python bench.py [--steps K] [--warmup W] [--workload c4|c2house|c3] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "subgraph-enumeration_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import sge  # noqa: E402
import sge.workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def workload(name):
    if name == "c4":
        return ("rmat20_seven_vertex_78", lambda: W.graph_c4(), W.seven_vertex_set,
                {"graph": "rmat(scale=20, ef=16, a=.57, b=.19, c=.19, seed=4)",
                 "patterns": "78 x 7-vertex, 15-19 edges"})
    if name == "c3":
        return ("chunglu7115_six_vertex_16", lambda: W.graph_c2(), W.six_vertex_set,
                {"graph": "chung_lu(n=7115, m=100762, gamma=2.1, cap=1100, seed=1)", "patterns": "16 x 6-vertex, 11-13 edges"})
    if name == "c2house":
        return ("chunglu7115_house_k5e", lambda: W.graph_c2(), lambda: [W.named("house"), W.named("near5clique")],
                {"graph": "chung_lu(n=7115, m=100762, gamma=2.1, cap=1100, seed=1)", "patterns": "house, K5-e"})
    raise SystemExit(f"unknown workload {name}")


def load_traffic(kernel_workload):
    """HBM bytes per launch from committed rocprofv3 --pmc summaries, if present."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            data = json.load(f)
        return data.get(kernel_workload)
    except Exception:
        return None


def cpu_baseline(graph, plans, stride, budget_s=20.0):
    """Oracle (port of the reference's join) on a bounded sample of the same
    root edges the GPU counts (the strided sample), patterns in set order until
    the time budget is spent, all host threads."""
    from oracle import oracle as O

    off, nbr = graph.csr()
    threads = min(os.cpu_count() or 1, 64)
    total_count, total_t, done = 0, 0.0, 0
    t_start = time.time()
    for pl in plans:
        if time.time() - t_start > budget_s:
            break
        t0 = time.perf_counter()
        total_count += O.count_strided(pl, off, nbr, stride, 0, threads)
        total_t += time.perf_counter() - t0
        done += 1
    return {"value": total_count / total_t if total_t > 0 else None, "unit": "matches/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle on the first {done} of {len(plans)} patterns over the same 1/{stride} root-edge "
                      f"sample, {total_count} matches in {total_t:.2f}s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c4")
    ap.add_argument("--mode", default="iep", choices=["wcoj", "explore", "iep"])
    ap.add_argument("--stride", type=int, default=None,
                    help="root-edge sample: every stride-th directed edge (default 4096 for c4, 1 otherwise)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run the known-size copy kernel (scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--learned", default=os.path.join(ROOT, "subgraph-enumeration_amd", "sge", "models",
                                                      "order_mlp.npz"))
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    mode = {"wcoj": sge.MODE_WCOJ, "explore": sge.MODE_EXPLORE, "iep": sge.MODE_IEP}[args.mode]
    wname, gfn, pfn, cfg = workload(args.workload)
    stride = args.stride or (4096 if args.workload == "c4" else 1)
    t0 = time.time()
    g = gfn()
    log(f"graph n={g.n} m={g.m} maxdeg={g.stats.max_degree} built in {time.time() - t0:.1f}s")
    g.to_device()
    if args.calibrate:
        nbytes = 256 << 20
        ms = sge.calibration_copy(nbytes)
        log(f"calibration copy: {nbytes} bytes read + written in {ms:.3f} ms ({2 * nbytes / ms / 1e6:.0f} GB/s)")
    pats = pfn()
    t0 = time.time()
    plans_cost = [sge.build_plan(p, g, mode) for p in pats]
    plan_s = time.time() - t0
    plans_fast = plans_cost
    learned = False
    try:
        from sge import learned as LM

        if os.path.exists(args.learned):
            model = LM.load(args.learned)
            plans_fast = [LM.choose_plan(model, p, g, mode, fallback=pc)[0] for p, pc in zip(pats, plans_cost)]
            learned = True
    except ImportError:
        pass
    log(f"{len(pats)} plans in {plan_s:.2f}s (learned order: {learned})")

    # plain executor: cost-model order, one launch per pattern
    plain_ms = None
    plain_counts = None
    if not args.no_plain:
        t0 = time.perf_counter()
        plain_counts, pst = sge.count_sampled(g, plans_cost, stride)
        plain_ms = (time.perf_counter() - t0) * 1000
        log(f"plain: {plain_ms:.1f} ms, kernel {sum(s.kernel_ms for s in pst):.1f} ms, total {sum(plain_counts)}")

    batch = sge.Batch(g, plans_fast, stride=stride)
    for i in range(args.warmup):
        counts, st = batch.run()
        log(f"warmup {i}: {st.kernel_ms:.1f} ms")
    times, kms, stats = [], [], None
    for i in range(args.steps):
        t0 = time.perf_counter()
        counts, st = batch.run()
        times.append(time.perf_counter() - t0)
        kms.append(st.kernel_ms)
        stats = st
        log(f"step {i}: {times[-1] * 1000:.1f} ms (kernel {st.kernel_ms:.1f} ms)")
    if plain_counts is not None and plain_counts != counts:
        log("COUNT MISMATCH between plain and batched executors")
        sys.exit(3)
    total = sum(counts)
    step_s = float(np.mean(times))
    kernel_s = float(np.mean(kms)) / 1000.0
    achieved = stats.bytes / kernel_s / 1e9 if kernel_s > 0 else 0.0
    res = {
        "metric": "exact matches/sec (batched executor, whole pattern set)",
        "value": total / step_s,
        "unit": "matches/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1000,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "u32 ids / u64 counts",
        "data": "synthetic",
        "config": {"workload": wname, "mode": args.mode, "learned_order": learned, **cfg,
                   "n": g.n, "m": g.m, "patterns": len(pats), "root_edge_stride": stride,
                   "root_edges": (2 * g.m + stride - 1) // stride},
        "pattern_set_ms": {"batched": step_s * 1000, "plain": plain_ms,
                           "speedup_vs_plain": (plain_ms / (step_s * 1000)) if plain_ms else None},
        "total_matches": total,
        "intersections_per_s": stats.intersections / kernel_s if kernel_s > 0 else None,
        "planning_s": plan_s,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(wname),
                     "kernel": "wcoj_kernel<false>", "algorithmic_bytes_per_launch": stats.bytes,
                     "kernel_ms": kernel_s * 1000},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(g, plans_cost, stride)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
