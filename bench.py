"""Benchmark: exact matches/s over a pattern set, batched vs plain GPU executor.

Default workload (`--workload c4`): the 78 connected 7-vertex patterns with
15-19 edges over R-MAT scale 20 (edge factor 16, a,b,c,d = .57,.19,.19,.05,
seed 4; n = 1,048,576, m = 15,702,114, int32 degree-ordered CSR resident in
HBM), IEP mode.  The whole job holds 10^17 or more matches per pattern, so
each pattern runs over its own stated range of root edges (CSR positions),
recorded with the oracle's count in tests/golden/large_c4.json; `--span S`
multiplies every range's length by S (S = 1: the counts are checked against
the oracle's).

One step = one pass over all 78 ranges.  Executors, each warmed up W times and
then timed over K steps (setup outside every timer):
  plain           one launch per pattern (cost-model orders, library defaults)
  batched_cost    one batched launch (hipGraph replay), the same plans: counts
                  must equal plain's
  batched         one batched launch with the learned orders: `value` = its
                  matches / step time.  On a root subset the learned order is
                  chosen among the orders that keep the cost-model order's
                  root edge and symmetry rules (learned.root_preserving): the
                  root edge maps to the plan's first two pattern vertices and
                  the rules follow the order, so only those orders count the
                  same embeddings over the same root edges.  Its counts must
                  equal plain's (and the oracle's at span 1).
The roofline of the batched run's wcoj_kernel comes from the kernel's own byte
counters and the HIP-event time on the launch stream.  The CPU baseline is the
oracle (a C restatement of the reference's join, all host threads) on the same
ranges, patterns in set order within a time budget, printed beside the GPU
plain executor's rate on the same patterns.

The C3 block (`c3_whole_job`) is the largest configuration that runs whole and
exact: the 16 six-vertex patterns over the Chung-Lu graph n=7,115 m=100,762
(gamma 2.1, expected degree capped at 1,100, seed 1), batched learned-order
against plain, counts equal, and the reference's graph_mining_system (built
from its own sources into oracle/_ref/) on all host cores beside it.

Other whole-job workloads: `--workload sixv|house|rmat8` (the C3 pass alone,
house + K5-e over the Chung-Lu graph, the 78 patterns over R-MAT scale 8).
Measurement sweeps that print their own JSON: `--workload bags` (plans with and
without a materialised bag), `--workload graph_build` (host against GPU CSR
build) and `--workload phases` (wavefront cycles per search phase; needs the
phase-profile library: `make -C subgraph-enumeration_amd libsge_phase.so`, then
SGE_LIB=subgraph-enumeration_amd/libsge_phase.so).

    python bench.py [--steps K] [--warmup W] [--workload c4|sixv|house|rmat8|bags|graph_build|phases] [--span S]
                    [--no-c3] [--no-cpu]
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
C4_SUBSETS = os.path.join(ROOT, "tests", "golden", "large_c4.json")
LEARNED = os.path.join(ROOT, "subgraph-enumeration_amd", "sge", "models", "order_mlp.npz")
METRIC = "exact matches/sec (batched learned-order executor, whole pattern set)"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def load_traffic(kernel_workload):
    """HBM bytes per launch from committed rocprofv3 --pmc summaries, if present."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get(kernel_workload)
    except (OSError, ValueError):
        return None


def learned_plans(g, pats, plans_cost, mode, keep_root=False, root_ranges=None):
    try:
        from sge import learned as LM
    except ImportError:
        return plans_cost, False
    if not os.path.exists(LEARNED):
        return plans_cost, False
    model = LM.load(LEARNED)
    return [plan for plan, _ in LM.choose_plans(model, pats, g, mode, plans_cost, keep_root=keep_root,
                                                root_ranges=root_ranges)], True


def timed_steps(run, warmup, steps, label):
    """run() -> (counts, wall-independent kernel ms); W warmups, then K timed steps."""
    counts = None
    for i in range(warmup):
        counts, kms = run()
        log(f"{label} warmup {i}: kernel {kms:.1f} ms")
    walls, kernels = [], []
    for i in range(steps):
        t0 = time.perf_counter()
        c, kms = run()
        walls.append((time.perf_counter() - t0) * 1000)
        kernels.append(kms)
        if counts is not None and c != counts:
            log(f"{label}: counts changed between steps")
            sys.exit(3)
        counts = c
        log(f"{label} step {i}: {walls[-1]:.1f} ms (kernel {kms:.1f} ms)")
    return counts, float(np.mean(walls)), float(np.mean(kernels))


def plain_runner(g, plans, ranges=None, stride=1):
    def run():
        if ranges is not None:
            c, st, pp = sge.count_ranges(g, plans, ranges[0], ranges[1], batched=False)
        else:
            c, pp = sge.count_sampled(g, plans, stride)
        run.per_plan = pp
        return c, sum(s.kernel_ms for s in pp)
    return run


def batch_runner(batch):
    def run():
        c, st = batch.run()
        run.stats = st
        return c, st.kernel_ms
    return run


def roofline(st, kernel_ms, wname):
    """achieved = the per-pair list-byte model (4(|A'|+|B'|) + 4|out| + 8 per CSR
    row for every pairwise intersection of clamped sorted lists) over the kernel
    time; `executed` = the bytes the kernel's strategy touches (it probes the
    longer list of a last-level pair by search or bitmap instead of reading it)."""
    s = kernel_ms / 1000.0
    achieved = st.model_bytes / s / 1e9 if s > 0 else 0.0
    executed = st.bytes / s / 1e9 if s > 0 else 0.0
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(wname), "kernel": "wcoj_kernel<false,false>",
            "algorithmic_bytes_per_launch": st.model_bytes, "kernel_ms": kernel_ms,
            "executed": {"bytes_per_launch": st.bytes, "achieved": executed, "frac": executed / HBM_PEAK_GBS,
                         "note": "bytes the kernel's strategy touches: the last level's probes at 8 B each "
                                 "(driver element + probed element) instead of both whole lists"}}


def cpu_oracle(g, plans, ranges, budget_s=20.0):
    """The oracle on all host threads over the same per-pattern ranges, patterns in
    set order until the budget is spent (each pattern's range is bounded)."""
    from oracle import oracle as O

    off, nbr = g.csr()
    off, nbr = np.ascontiguousarray(off, np.uint32), np.ascontiguousarray(nbr, np.uint32)
    threads = os.cpu_count() or 1
    tot, t_all, done = 0, 0.0, []
    t_start = time.time()
    for i, pl in enumerate(plans):
        if time.time() - t_start > budget_s:
            break
        t0 = time.perf_counter()
        tot += O.count(pl, off, nbr, threads, int(ranges[0][i]), int(ranges[1][i]))
        t_all += time.perf_counter() - t0
        done.append(i)
    return {"value": tot / t_all if t_all > 0 else None, "unit": "matches/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C restatement of the reference's join, {threads} threads) over the root-edge "
                      f"ranges of patterns 0..{done[-1] if done else -1} of {len(plans)} (set order, "
                      f"{budget_s:.0f}s budget): {tot} matches in {t_all:.2f}s"}, done, tot


def cpu_reference(graph, pats, budget_s=30.0):
    """The reference binary (oracle/_ref/graph_mining_system) with all host cores,
    patterns in set order until the budget is spent; None where it is absent."""
    import subprocess
    import tempfile

    from oracle import oracle as O

    if not os.path.exists(O.REF_BIN):
        return None, []
    cores = os.cpu_count() or 1
    tmp = tempfile.mkdtemp(prefix="sge_bench_")
    gpath = os.path.join(tmp, "g.txt")
    graph.write(gpath)
    total, t_exec, done = 0, 0.0, []
    t_start = time.time()
    for i, p in enumerate(pats):
        left = budget_s - (time.time() - t_start)
        if left <= 1:
            break
        qpath = os.path.join(tmp, f"q{i}.txt")
        with open(qpath, "w") as f:
            e = p.edges()
            f.write(f"{p.k} {len(e)}\n" + "".join(f"{a} {b}\n" for a, b in e))
        try:
            r = subprocess.run([O.REF_BIN, "-q", qpath, "-d", gpath, "-u", "-t", str(cores)],
                               capture_output=True, text=True, timeout=left)
        except subprocess.TimeoutExpired:
            break
        cnt = exe = None
        for line in r.stdout.splitlines():
            if line.startswith("Number of matches:"):
                cnt = int(line.split(":")[1])
            if line.startswith("Execution Time:"):
                exe = float(line.split(":")[1])
        if cnt is None or exe is None:
            break
        total += cnt
        t_exec += exe
        done.append(i)
    if not done:
        return None, []
    return {"value": total / t_exec, "unit": "matches/s", "cores": cores, "kind": "reference",
            "sample": f"reference graph_mining_system -u -t {cores} on patterns {done} of {len(pats)} "
                      f"(set order, {budget_s:.0f}s budget), {total} matches in {t_exec:.2f}s of its "
                      f"'Execution Time'"}, done


def gpu_rate_on(per_plan, counts, done):
    """The plain executor's rate (kernel time) on the patterns a CPU baseline covered."""
    if not done:
        return None
    ms = sum(per_plan[i].kernel_ms for i in done)
    return {"value": sum(counts[i] for i in done) / (ms / 1000.0) if ms > 0 else None, "unit": "matches/s",
            "executor": "plain", "patterns": done, "kernel_ms": ms}


def c4_subsets(args):
    wname = "rmat20_sevenvertex78_subsets"
    with open(C4_SUBSETS) as f:
        doc = json.load(f)
    t0 = time.time()
    g = W.graph_c4(device_build=True)
    log(f"graph n={g.n} m={g.m} maxdeg={g.stats.max_degree} built on the GPU in {time.time() - t0:.1f}s")
    assert (g.n, g.m) == (doc["n"], doc["m"])
    pats = W.seven_vertex_set()
    subs = doc["subsets"]
    t0 = time.time()
    plans_cost = [sge.build_plan(pats[e["pattern"]], g, sge.MODE_IEP, e["plan"]["order"]) for e in subs]
    for pl, e in zip(plans_cost, subs):
        assert list(pl.order)[:pl.k] == e["plan"]["order"]
    plan_s = time.time() - t0
    span = max(1, args.span)
    begins = np.array([e["begin"] for e in subs], dtype=np.uint64)
    ends = np.array([e["begin"] + (e["end"] - e["begin"]) * span for e in subs], dtype=np.uint64)
    t0 = time.time()
    plans_fast, learned = learned_plans(g, [pats[e["pattern"]] for e in subs], plans_cost, sge.MODE_IEP,
                                        keep_root=True, root_ranges=(begins, ends))
    learn_s = time.time() - t0
    changed = sum(list(a.order)[:a.k] != list(b.order)[:b.k] for a, b in zip(plans_fast, plans_cost))
    log(f"78 cost-model plans in {plan_s:.2f}s; learned orders in {learn_s:.2f}s ({changed} changed)")
    ranges = (begins, ends)
    root_edges = int((ends - begins).sum())

    from sge import autotune

    tuning = autotune.tuned(wname)  # batched executors only; plain keeps the library defaults
    tuning_desc = {k: getattr(tuning, k) for k, _ in sge.Tuning._fields_} if tuning else "library defaults"
    plain_counts = plain_ms = plain_kms = bcost_counts = bcost_ms = bcost_kms = None
    plain = plain_runner(g, plans_cost, ranges)
    if not args.no_plain:
        plain_counts, plain_ms, plain_kms = timed_steps(plain, args.warmup, args.steps, "plain")
        if span == 1 and plain_counts != [e["count"] for e in subs]:
            log("COUNT MISMATCH: plain executor against the oracle's counts")
            sys.exit(3)
        bc = sge.Batch(g, plans_cost, tuning=tuning, ranges=ranges)
        bcost_counts, bcost_ms, bcost_kms = timed_steps(batch_runner(bc), args.warmup, args.steps, "batched_cost")
        bc.close()
        if bcost_counts != plain_counts:
            log("COUNT MISMATCH between plain and batched executors (same plans)")
            sys.exit(3)
    b = sge.Batch(g, plans_fast, tuning=tuning, ranges=ranges)
    brun = batch_runner(b)
    counts, step_ms, kernel_ms = timed_steps(brun, args.warmup, args.steps, "batched")
    st = brun.stats
    b.close()
    if plain_counts is not None and counts != plain_counts:
        log("COUNT MISMATCH between the batched learned-order and the plain cost-order executors")
        sys.exit(3)
    total = sum(counts)
    res = {
        "metric": METRIC, "value": total / (step_ms / 1000.0), "unit": "matches/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": step_ms, "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "u32 ids / u64 counts", "data": "synthetic",
        "config": {"workload": f"{wname} (78 x 7-vertex, 15-19 edges, over R-MAT scale 20 seed 4; per-pattern "
                               f"root-edge ranges of tests/golden/large_c4.json x{span}: {root_edges} root edges "
                               f"of {2 * g.m})",
                   "mode": "iep", "graph": doc["graph"], "n": g.n, "m": g.m, "patterns": len(subs),
                   "span": span, "root_edges": root_edges, "learned_order": learned, "learned_changed": changed,
                   "tuning": tuning_desc},
        "pattern_set_ms": {"plain": plain_ms, "plain_kernel": plain_kms, "batched_cost_order": bcost_ms,
                           "batched_cost_order_kernel": bcost_kms, "batched": step_ms, "batched_kernel": kernel_ms,
                           "speedup_vs_plain": plain_ms / step_ms if plain_ms else None,
                           "speedup_vs_plain_kernel": plain_kms / kernel_ms if plain_kms and kernel_ms > 0 else None,
                           "speedup_batching_same_plans": plain_ms / bcost_ms if plain_ms else None,
                           "speedup_learned_same_batching": bcost_ms / step_ms if bcost_ms else None,
                           "note": "all three executors count the same embeddings over the same root edges: "
                                   "the learned orders keep the cost-model orders' root edge and symmetry "
                                   "rules"},
        "total_matches": total,
        "counts_checked": ("not checked (--no-plain)" if args.no_plain else
                           "batched == plain == batched_cost" + (" == oracle fixture" if span == 1 else "")),
        "intersections_per_s": st.intersections / (kernel_ms / 1000.0) if kernel_ms > 0 else None,
        "planning_s": plan_s, "learned_choice_s": learn_s,
        "roofline": roofline(st, kernel_ms, wname),
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        cpu, done, _ = cpu_oracle(g, plans_cost, (begins, np.array([e["end"] for e in subs], np.uint64)))
        res["cpu_baseline"] = cpu
        if span == 1 and plain_counts:
            res["cpu_baseline"]["gpu_same_sample"] = gpu_rate_on(plain.per_plan, plain_counts, done)
    return res


def whole_job(args, name):
    c2 = "chung_lu(n=7115, m=100762, gamma=2.1, cap=1100, seed=1)"
    table = {
        "sixv": ("chunglu7115_sixvertex16", lambda: W.graph_c2(), W.six_vertex_set, c2, "16 x 6-vertex, 11-13 edges"),
        "house": ("chunglu7115_house_k5e", lambda: W.graph_c2(), lambda: [W.named("house"), W.named("near5clique")],
                  c2, "house, K5-e"),
        "rmat8": ("rmat8_sevenvertex78", lambda: W.graph_rmat(8, 7, device_build=True), W.seven_vertex_set,
                  "rmat(scale=8, ef=16, a=.57, b=.19, c=.19, seed=7)", "78 x 7-vertex, 15-19 edges"),
    }
    if name not in table:
        raise SystemExit(f"unknown workload {name}")
    wname, gfn, pfn, gdesc, pdesc = table[name]
    mode = {"wcoj": sge.MODE_WCOJ, "explore": sge.MODE_EXPLORE, "iep": sge.MODE_IEP}[args.mode]
    g = gfn()
    g.to_device()
    pats = pfn()
    t0 = time.time()
    plans_cost = [sge.build_plan(p, g, mode) for p in pats]
    plan_s = time.time() - t0
    plans_fast, learned = learned_plans(g, pats, plans_cost, mode)
    from sge import autotune

    tuning = autotune.tuned(wname)
    tuning_desc = {k: getattr(tuning, k) for k, _ in sge.Tuning._fields_} if tuning else "library defaults"
    plain = plain_runner(g, plans_cost)
    plain_counts = plain_ms = plain_kms = None
    if not args.no_plain:
        plain_counts, plain_ms, plain_kms = timed_steps(plain, args.warmup, args.steps, f"{name} plain")
    b = sge.Batch(g, plans_fast, tuning=tuning)
    brun = batch_runner(b)
    counts, step_ms, kernel_ms = timed_steps(brun, args.warmup, args.steps, f"{name} batched")
    st = brun.stats
    b.close()
    if plain_counts is not None and counts != plain_counts:
        log("COUNT MISMATCH between plain and batched executors")
        sys.exit(3)
    total = sum(counts)
    res = {
        "metric": METRIC, "value": total / (step_ms / 1000.0), "unit": "matches/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": step_ms, "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "u32 ids / u64 counts", "data": "synthetic",
        "config": {"workload": wname, "mode": args.mode, "graph": gdesc, "patterns_desc": pdesc, "n": g.n, "m": g.m,
                   "patterns": len(pats), "learned_order": learned, "tuning": tuning_desc, "whole_job": True},
        "pattern_set_ms": {"plain": plain_ms, "plain_kernel": plain_kms, "batched": step_ms,
                           "batched_kernel": kernel_ms, "speedup_vs_plain": plain_ms / step_ms if plain_ms else None,
                           "speedup_vs_plain_kernel": plain_kms / kernel_ms if plain_kms and kernel_ms > 0 else None},
        "total_matches": total,
        "counts_checked": ("not checked (--no-plain)" if plain_counts is None else
                           "plain == batched (whole job: counts do not depend on the order)"),
        "intersections_per_s": st.intersections / (kernel_ms / 1000.0) if kernel_ms > 0 else None,
        "planning_s": plan_s,
        "roofline": roofline(st, kernel_ms, wname),
        "cpu_baseline": None,
    }
    if not args.no_cpu and mode == sge.MODE_IEP:
        cpu, done = cpu_reference(g, pats)
        if cpu is not None and plain_counts is not None:
            cpu["gpu_same_sample"] = gpu_rate_on(plain.per_plan, plain_counts, done)
        res["cpu_baseline"] = cpu
    return res


def bag_times(args):
    """Plans with and without a materialised bag (plain executor, one launch per
    plan): for each pattern the cost model's plan without a bag and with every
    valid bag of at most 4 vertices; a bag's time is its trie build (enumerate
    the bag relation, radix sort, per-depth scans) plus the join.  Counts must
    be identical.  Sets: the 16 six-vertex patterns over the Chung-Lu graph
    n=7,115 m=100,762, and every fourth of the 78 seven-vertex patterns over
    R-MAT scale 12 (seed 4), both in IEP mode."""
    def best_time(g, plan):
        best = None
        for _ in range(max(1, args.steps)):
            c, st = sge.count(g, plan)
            tot = st.kernel_ms + st.trie_ms
            if best is None or tot < best[0]:
                best = (tot, c[0], st.trie_ms)
        return best

    out = {"workload": "bags", "sets": {}}
    for label, g, pats in (("chunglu7115_sixvertex16", W.graph_c2(), W.six_vertex_set()),
                           ("rmat12_sevenvertex_every4th", W.graph_rmat(12, 4), W.seven_vertex_set()[::4])):
        g.to_device()
        res = []
        for i, p in enumerate(pats):
            plan = sge.build_plan(p, g, sge.MODE_IEP)
            base = best_time(g, plan)
            rows = []
            for b in sge.valid_bags(plan):
                if bin(b).count("1") > 4:
                    continue
                try:
                    r = best_time(g, sge.set_bag(plan, b))
                except sge.SgeError as e:  # relation above the tuple cap
                    rows.append({"bag": b, "error": str(e)})
                    continue
                if r[1] != base[1]:
                    log(f"{label} pattern {i} bag {b}: count {r[1]} != {base[1]}")
                    sys.exit(3)
                rows.append({"bag": b, "ms": r[0], "trie_ms": r[2]})
            timed = [r for r in rows if "ms" in r]
            best = min(timed, key=lambda r: r["ms"]) if timed else None
            log(f"{label} pattern {i}: no bag {base[0]:.3f} ms | best bag {best} of {len(rows)}")
            res.append({"pattern": i, "no_bag_ms": base[0], "count": base[1], "bags": rows})
        out["sets"][label] = res
    return out


PHASES = ["other", "materialize", "explore", "terms", "stage", "probe_setup", "probes", "closed_forms", "queue",
          "root_setup", "dfs_iteration", "donate"]
ITEM_EVENTS = ["count_calls", "sibling_handovers", "dfs_donations", "materialised_sets", "longest_count_cycles",
               "longest_f_formation_cycles", "siblings_counted", "count_wave_steps"]


def phase_profile(args):
    """Wavefront clock cycles per search phase of one batched pass (cost-model
    plans), from a -DSGE_PHASE_PROFILE build of the executor: the Chung-Lu
    whole job (16 six-vertex patterns) and four of its patterns alone, and the
    R-MAT 20 and R-MAT 23 root-edge ranges.  `busy` = the phases' cycles over
    (wavefronts x longest wavefront), all dispatches of the pass together;
    `root_pass_busy` the same for the root dispatch alone (1 - its end-of-pass
    idling)."""
    import ctypes as C

    L = sge.lib()
    if not hasattr(L, "sge_phase_profile_read"):
        raise SystemExit("phases: load the phase-profile build (SGE_LIB=subgraph-enumeration_amd/libsge_phase.so)")
    L.sge_phase_profile_read.restype = C.c_int
    L.sge_phase_profile_read.argtypes = [C.c_void_p, C.c_int]
    L.sge_phase_item_read.restype = C.c_int
    L.sge_phase_item_read.argtypes = [C.c_void_p, C.c_int]
    n = len(PHASES) + 1

    def read_item():
        """The root pass's longest item: its phase cycles and events."""
        a = np.zeros(1 + len(PHASES) + len(ITEM_EVENTS) + 4, np.uint64)
        if L.sge_phase_item_read(a.ctypes.data, 1) != 0:
            raise SystemExit("phases: sge_phase_item_read failed")
        r = [int(x) for x in a[1:]]
        np_, ne = len(PHASES), len(ITEM_EVENTS)
        return {"wave": int(a[0] & 0xffffffff), "cycles": r[np_ + ne], "queue_index": r[np_ + ne + 1],
                "plan": r[np_ + ne + 2] & 0xffffffff, "continuation": bool(r[np_ + ne + 2] >> 32),
                "start_cycles": r[np_ + ne + 3],
                "phase_cycles": {p: r[i] for i, p in enumerate(PHASES) if r[i]},
                "events": {e: r[np_ + i] for i, e in enumerate(ITEM_EVENTS)}}

    def read():
        a = np.zeros(n + 10, np.uint64)
        if L.sge_phase_profile_read(a.ctypes.data, 1) != 0:
            raise SystemExit("phases: sge_phase_profile_read failed")
        return a

    out = {"workload": "phases", "runs": []}

    def one(tag, g, plans, ranges=None, want=None):
        b = sge.Batch(g, plans, ranges=ranges)
        c, _ = b.run()  # warm-up
        read()
        read_item()
        c, st = b.run()
        raw = read()
        item = read_item()
        a = raw.astype(np.float64)
        b.close()
        if want is not None and c != want:
            log(f"phases {tag}: COUNT MISMATCH")
            sys.exit(3)
        tot = a[n - 1]
        row = {"run": tag, "kernel_ms": st.kernel_ms,
               "share": {p: a[i] / tot for i, p in enumerate(PHASES) if a[i] > 0},
               "queue_fetches": a[n], "wavefronts": a[n + 1], "busy": tot / max(a[n + 1] * a[n + 2], 1.0),
               "root_pass_busy": a[n + 2 + 1] / max(a[n + 2 + 3] * a[n + 2 + 2], 1.0),
               "root_pass_longest_item": {"cycles": int(raw[n + 6] >> 32) << 8, "queue_index": int(raw[n + 6] & 0xffffffff),
                                          "start_cycles": int(raw[n + 7] & 0xffffffff) << 8,
                                          "longest_wave_cycles": int(raw[n + 4])},
               "longest_f_formation_cycles": int(raw[n + 8]), "longest_count_cycles": int(raw[n + 9]),
               "root_pass_longest_item_detail": item}
        log(f"phases {tag}: {st.kernel_ms:.1f} ms " + ", ".join(f"{k} {100 * v:.1f}%" for k, v in row['share'].items())
            + f"; root pass busy {100 * row['root_pass_busy']:.1f}%, longest item {row['root_pass_longest_item']}, "
            f"longest F formation {row['longest_f_formation_cycles']}, longest count {row['longest_count_cycles']}; "
            f"longest item detail {item}")
        out["runs"].append(row)

    g2 = W.graph_c2()
    g2.to_device()
    p2 = [sge.build_plan(p, g2, sge.MODE_IEP) for p in W.six_vertex_set()]
    one("chunglu7115_sixvertex16 whole job", g2, p2)
    for i in (1, 3, 7, 11):
        one(f"chunglu7115 six-vertex pattern {i}", g2, [p2[i]])
    for name, gfn, pfn in (("rmat20_sevenvertex78_subsets", lambda: W.graph_c4(device_build=True), W.seven_vertex_set),
                           ("rmat23_sixvertex16_k5e_subsets", lambda: W.graph_c5(device_build=True),
                            lambda: W.six_vertex_set() + [W.named("near5clique")])):
        with open(os.path.join(ROOT, "tests", "golden", "large_" + ("c4" if "rmat20" in name else "c5") + ".json")) as f:
            subs = json.load(f)["subsets"]
        g = gfn()
        pats = pfn()
        plans = [sge.build_plan(pats[e["pattern"]], g, sge.MODE_IEP, e["plan"]["order"]) for e in subs]
        ranges = (np.array([e["begin"] for e in subs], np.uint64), np.array([e["end"] for e in subs], np.uint64))
        one(name, g, plans, ranges, [e["count"] for e in subs])
        g.close()
    return out


def graph_build_times(args):
    """The CSR build on the host and on the GPU for R-MAT scales 18, 20, 21
    (seed 4): both deduplicate, drop self-loops, relabel by ascending degree
    and build ascending rows; the arrays must be identical.  Edge generation
    (host) is not part of either figure."""
    out = {"workload": "graph_build", "scales": []}
    for scale in (18, 20, 21):
        h = sge.Graph.rmat(scale, 16, seed=4)
        d = sge.Graph.rmat(scale, 16, seed=4, device_build=True)
        same = all(np.array_equal(a, b) for a, b in zip(h.csr(), d.csr())) and np.array_equal(h.perm(), d.perm())
        log(f"rmat scale {scale}: n {h.n} m {h.m} | host build {h.stats.build_ms:.1f} ms | device build "
            f"{d.stats.build_ms:.1f} ms | identical arrays: {same}")
        if not same:
            sys.exit(3)
        out["scales"].append({"scale": scale, "n": h.n, "m": h.m, "host_ms": h.stats.build_ms,
                              "device_ms": d.stats.build_ms, "identical": same})
        del h, d
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c4")
    ap.add_argument("--mode", default="iep", choices=["wcoj", "explore", "iep"])
    ap.add_argument("--span", type=int, default=1, help="c4: multiply every root-edge range's length by S")
    ap.add_argument("--no-c3", action="store_true", help="c4: leave out the C3 whole-job block")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="time the batched executor only (profiling runs)")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run the known-size copy kernel (scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.calibrate:
        nbytes = 256 << 20
        ms = sge.calibration_copy(nbytes)
        log(f"calibration copy: {nbytes} bytes read + written in {ms:.3f} ms ({2 * nbytes / ms / 1e6:.0f} GB/s)")
    if args.workload in ("c4", "rmat20"):
        res = c4_subsets(args)
        if not args.no_c3:
            res["c3_whole_job"] = whole_job(args, "sixv")
    elif args.workload == "bags":
        res = bag_times(args)
    elif args.workload == "graph_build":
        res = graph_build_times(args)
    elif args.workload == "phases":
        res = phase_profile(args)
    else:
        res = whole_job(args, args.workload)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
