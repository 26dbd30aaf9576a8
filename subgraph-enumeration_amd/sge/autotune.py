"""Autotuner for the batched executor's launch parameters.

Coordinate descent over workgroup size (waves per workgroup), root-queue chunk,
root-edge split span, work-donation budget and the probe-direction ratio of the
last-level count (which list drives an intersection, by length ratio), each
candidate timed as one batched pass over a coarser root-edge sample of the same
plans.  Counts are checked to be identical for every setting.  Results are cached
in a JSON file keyed by the workload name.
"""
from __future__ import annotations

import json
import os
import sys
import time
from typing import Dict, Optional, Sequence

from . import Batch, Graph, Plan, Tuning, default_tuning

SPACE = {  # first value = the library default, the starting point
    "block_threads": [256, 128, 64],
    "chunk": [4, 2, 8],
    "split_span": [32, 16, 64, 8],
    "donate_budget": [1024, 4096, 256, 16384],
    "donate_split": [64, 16, 256],
    "stage_max": [2, 1, 4, 8],
    "bitmap_ratio": [2, 1, 4, 8],
    "pass_items": [0, 1, 2],
}
CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models", "tuning.json")


def _tuning(d: Dict[str, int]) -> Tuning:
    t = default_tuning()
    for k, v in d.items():
        setattr(t, k, int(v))
    return t


def _time(graph: Graph, plans: Sequence[Plan], stride: int, d: Dict[str, int], reps: int = 2, ranges=None):
    b = Batch(graph, plans, _tuning(d), stride=stride, ranges=ranges)
    try:
        counts, st = b.run()  # warm
        best = float("inf")
        for _ in range(reps):
            c2, st = b.run()
            assert c2 == counts
            best = min(best, st.kernel_ms)
        return best, counts
    finally:
        b.close()


def tune(graph: Graph, plans: Sequence[Plan], stride: int, rounds: int = 1, log=print,
         cache: Optional[str] = None, key: Optional[str] = None, budget_s: float = 120.0,
         ranges=None) -> Dict[str, int]:
    """Coordinate descent from the library defaults; `ranges` = per-plan root
    ranges (begins, ends) instead of a strided sample."""
    if cache and key and os.path.exists(cache):
        with open(cache) as f:
            data = json.load(f)
        if key in data:
            return data[key]["params"]
    cur = {k: v[0] for k, v in SPACE.items()}
    t0 = time.time()
    best_ms, ref = _time(graph, plans, stride, cur, ranges=ranges)
    log(f"[autotune] start {cur}: {best_ms:.2f} ms")
    trace = [{"params": dict(cur), "ms": best_ms}]
    for _ in range(rounds):
        for name, values in SPACE.items():
            for v in values:
                if v == cur[name] or time.time() - t0 > budget_s:
                    continue
                trial = dict(cur, **{name: v})
                ms, counts = _time(graph, plans, stride, trial, ranges=ranges)
                if counts != ref:
                    raise RuntimeError(f"autotune: counts changed with {trial}")
                trace.append({"params": trial, "ms": ms})
                log(f"[autotune] {name}={v}: {ms:.2f} ms")
                if ms < best_ms:
                    best_ms, cur = ms, trial
    log(f"[autotune] best {cur}: {best_ms:.2f} ms in {time.time() - t0:.0f}s")
    if cache and key:
        data = {}
        if os.path.exists(cache):
            with open(cache) as f:
                data = json.load(f)
        data[key] = {"params": cur, "ms": best_ms, "stride": stride, "trace": trace}
        with open(cache, "w") as f:
            json.dump(data, f, indent=1)
    return cur


def tuned(key: str, cache: str = CACHE) -> Optional[Tuning]:
    """The cached tuning for `key`, or None."""
    if not os.path.exists(cache):
        return None
    with open(cache) as f:
        data = json.load(f)
    return _tuning(data[key]["params"]) if key in data else None


def main():
    """Tune the batched executor for a benchmark workload with the orders the
    batch runs (the learned model's when it is present) and cache the result
    under the workload's name:
      sixv  the 16 six-vertex patterns over the Chung-Lu graph (n=7,115,
            m=100,762), IEP mode, whole job
      c4    the 78 seven-vertex patterns over R-MAT scale 20 (seed 4), IEP
            mode, over the per-pattern root ranges of tests/golden/large_c4.json"""
    import argparse

    import numpy as np

    from . import MODE_IEP, build_plan
    from . import workloads as W

    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="sixv", choices=["sixv", "c4"])
    ap.add_argument("--stride", type=int, default=1)
    ap.add_argument("--rounds", type=int, default=1)
    ap.add_argument("--budget-s", type=float, default=240.0)
    args = ap.parse_args()
    ranges = None
    if args.workload == "sixv":
        key = "chunglu7115_sixvertex16"
        g = W.graph_c2()
        g.to_device()
        pats = W.six_vertex_set()
        plans = [build_plan(p, g, MODE_IEP) for p in pats]
    else:
        key = "rmat20_sevenvertex78_subsets"
        root = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
        with open(os.path.join(root, "tests", "golden", "large_c4.json")) as f:
            subs = json.load(f)["subsets"]
        g = W.graph_c4(device_build=True)
        all7 = W.seven_vertex_set()
        pats = [all7[e["pattern"]] for e in subs]
        plans = [build_plan(p, g, MODE_IEP, e["plan"]["order"]) for p, e in zip(pats, subs)]
        ranges = (np.array([e["begin"] for e in subs], np.uint64), np.array([e["end"] for e in subs], np.uint64))
    model_path = os.path.join(os.path.dirname(CACHE), "order_mlp.npz")
    if os.path.exists(model_path):
        from . import learned as L

        model = L.load(model_path)
        # on root subsets only the orders that count the same embeddings there
        plans = [pl for pl, _ in L.choose_plans(model, pats, g, MODE_IEP, plans, keep_root=ranges is not None,
                                                root_ranges=ranges)]
    if os.path.exists(CACHE):
        with open(CACHE) as f:
            data = json.load(f)
        data.pop(key, None)
        with open(CACHE, "w") as f:
            json.dump(data, f, indent=1)
    tune(g, plans, args.stride, args.rounds, log=lambda *a: print(*a, flush=True), cache=CACHE, key=key,
         budget_s=args.budget_s, ranges=ranges)


if __name__ == "__main__":
    sys.exit(main())
