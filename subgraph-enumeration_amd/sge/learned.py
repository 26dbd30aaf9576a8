"""Learned attribute-order cost model.

A 64 -> 128 -> 128 -> 1 MLP (BF16 MFMA hidden layers, FP32 accumulation, Adam;
all on the GPU through sge_mlp_*) predicts log2 of the work the plain GPU
executor measures for a (pattern, order) -- wavefront clock cycles, summed over
the wavefronts that ran it -- relative to the pattern's other orders, from the order's features
(sge_order_features: plan structure, degree moments, analytic cost, Knuth
random-walk estimates over sampled root edges).  It only reorders: counts never
change.  The planner's own choice (its analytic cost model) is the starting
point, and the reference's sampled cost model (computeOrderCost over
cardEstimateUnlabeled cardinalities, plan/optimizer.cpp:319-439 and
plan/estimator.cpp:562-661, restated as sge_plan_cost_sampled) is the fallback
and the judge: when it puts the network's choice more than `judge_bits` above
the planner's order, the planner's order is kept.

Training data are (order, measured work) pairs logged by the plain executor on
a strided root-edge sample (sge_count_sampled), for patterns disjoint from the
ones the model is used on.
"""
from __future__ import annotations

import json
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from concurrent.futures import ThreadPoolExecutor

from . import (MLP, Graph, Pattern, Plan, build_plan, count_sampled, order_cost, order_cost_sampled, order_features,
               range_capped, rank_orders, time_capped, work_steps_bound)

FEATURE_SAMPLES = 512
CAPPED_PENALTY = 2.0  # bits added to the time of an order stopped by the work cap
# measured probe (the guard): a strided sample of at most 1/PROBE_MIN_STRIDE of
# a whole job's root edges (about PROBE_EDGES of them), or the root edges next
# to a run's range (as many, disjoint from it)
PROBE_MIN_STRIDE = 8
PROBE_EDGES = 16384
PROBE_CAP_K = 1 << 16  # work cap of the probe's first plan (x1024 wave steps)
PROBE_BITS = 0.2  # the learned order must measure at least this many bits less work


def automorphisms(pattern: Pattern) -> List[List[int]]:
    """Every automorphism of the pattern (g[u] = image of u), by backtracking."""
    k = pattern.k
    adj = [0] * k
    for a, b in pattern.edges():
        adj[a] |= 1 << b
        adj[b] |= 1 << a
    out, g, used = [], [0] * k, [False] * k

    def extend(u):
        if u == k:
            out.append(list(g))
            return
        for v in range(k):
            if used[v] or bin(adj[v]).count("1") != bin(adj[u]).count("1"):
                continue
            if any(((adj[u] >> w) & 1) != ((adj[v] >> g[w]) & 1) for w in range(u)):
                continue
            g[u], used[v] = v, True
            extend(u + 1)
            used[v] = False

    extend(0)
    return out


def symmetry_prefix(pattern: Pattern, order: Sequence[int], aut=None) -> int:
    """Length of the order's prefix that fixes its symmetry rules and root edge:
    the planner's rules come from the stabiliser chain along the order (a rule
    at level i for every other vertex in the orbit of order[i] under the
    automorphisms fixing order[:i]), so once the stabiliser is trivial no rule
    follows; at least the root edge (2 levels)."""
    G = automorphisms(pattern) if aut is None else aut
    j = 0
    while j < pattern.k and len(G) > 1:
        G = [g for g in G if g[order[j]] == order[j]]
        j += 1
    return max(2, j)


def root_preserving(pattern: Pattern, order: Sequence[int], candidates: Sequence[Sequence[int]]) -> List[List[int]]:
    """The candidates that, up to an automorphism, start with `order`'s symmetry
    prefix, each relabelled so that it starts with exactly that prefix.  Such an
    order maps a root edge to the same two pattern vertices and has the same
    symmetry rules, so over any set of root edges it counts exactly the
    embeddings `order` counts (a count over a root subset is comparable)."""
    aut = automorphisms(pattern)
    P = symmetry_prefix(pattern, order, aut)
    out, seen = [], set()
    for o in candidates:
        for g in aut:
            if all(g[o[q]] == order[q] for q in range(P)):
                t = [g[x] for x in o]
                if tuple(t) not in seen:
                    seen.add(tuple(t))
                    out.append(t)
                break
    return out


def candidate_orders(pattern: Pattern, graph: Optional[Graph], mode: int, top: int = 16, rng=None,
                     keep_root_of: Optional[Sequence[int]] = None) -> List[List[int]]:
    """Connected orders (deduplicated up to automorphism) ranked as the planner
    ranks them for `mode`; the `top` best, and with rng up to `top` others drawn
    from ranks [top, 4 * top) -- the neighbourhood the model has to choose in.
    With keep_root_of, only the orders that count the same embeddings as that
    order over every root subset (root_preserving), in rank order."""
    orders, costs, tails = rank_orders(pattern, graph, mode)
    if keep_root_of is not None:
        orders = root_preserving(pattern, list(keep_root_of), orders)
    pick = list(range(min(top, len(orders))))
    hi = min(len(orders), 4 * top)
    if rng is not None and hi > top:
        extra = rng.choice(np.arange(top, hi), size=min(top, hi - top), replace=False)
        pick += [int(i) for i in extra]
    return [orders[i] for i in pick]


def log_work(graph: Graph, pattern: Pattern, orders: Sequence[Sequence[int]], mode: int, stride: int = 1024,
             fine_stride: int = 256, min_ms: float = 20.0, cap_factor: int = 8, pilot_cap_k: int = 1 << 20):
    """Measured work of each order with the plain executor on a strided root
    sample.  orders[0] (the cost model's choice) is the pilot: it runs under
    the absolute cap `pilot_cap_k` (x1024 wave steps) -- None is returned when
    even it is stopped -- and when it takes under `min_ms` the sample is refined
    to `fine_stride`.  The other orders run under a cap of `cap_factor` times the
    pilot's work.  Returns (projected full-job work = wavefront clock cycles
    spent on the sample x stride, algorithmic bytes, capped flags, stride); a
    capped order's work is a lower bound.  Work, unlike device time, does not
    depend on how evenly a small sample spreads over the GPU."""
    plans = [build_plan(pattern, graph, mode, o) for o in orders]
    (pilot, lim), = time_capped(graph, plans[0], stride, pilot_cap_k)
    if lim:
        return None
    if pilot.kernel_ms < min_ms and fine_stride < stride:
        stride = fine_stride
        (pilot, lim), = time_capped(graph, plans[0], stride, pilot_cap_k)
        if lim:
            return None
    cap_k = max(64, cap_factor * work_steps_bound(pilot) // 1024 + 1)
    res = [(pilot, False)] + time_capped(graph, plans[1:], stride, cap_k)
    ms = np.array([float(st.wave_cycles) * stride for st, _ in res])
    by = np.array([st.bytes for st, _ in res], dtype=np.float64)
    capped = np.array([lim for _, lim in res], dtype=bool)
    return ms, by, capped, stride


class LearnedOrder:
    """The network predicts, for each candidate order of one pattern, log2 of
    its measured work relative to the mean over that pattern's logged orders
    (only the ranking within a pattern matters for the choice; the absolute
    scale across patterns is left out of the target)."""

    def __init__(self, mlp: Optional[MLP] = None, seed: int = 0):
        self.mlp = mlp or MLP(seed)
        self.mu = np.zeros(64, np.float32)
        self.sd = np.ones(64, np.float32)
        self.t_sd = 1.0
        self.threshold = 3.0  # bits of analytic cost the learned order may exceed the cost model's by
        self.judge_bits = 1.0  # bits of the sampled (reference) cost model the learned order may exceed it by
        self.margin = 0.25    # predicted bits of improvement needed to leave the cost model's order
        self.sizes: Optional[List[int]] = None  # pattern sizes trained on (None: no restriction)
        self.history: Dict[str, list] = {}

    # ---- data
    def _norm(self, X):
        return (X - self.mu) / self.sd

    def fit(self, X: np.ndarray, y_log2: np.ndarray, groups: np.ndarray, epochs: int = 300, lr: float = 2e-3,
            batch: int = 256, seed: int = 0, val_frac: float = 0.15, check_every: int = 5):
        """X: features, y_log2: log2 measured work, groups: pattern id per row.
        A `val_frac` share of the patterns is held out; the weights with the
        lowest held-out loss are kept (early stopping)."""
        self.mu = X.mean(0).astype(np.float32)
        self.sd = (X.std(0) + 1e-3).astype(np.float32)
        rel = relative_targets(y_log2, groups)
        self.t_sd = float(rel.std() + 1e-6)
        Xn = self._norm(X).astype(np.float32)
        tn = (rel / self.t_sd).astype(np.float32)
        rng = np.random.default_rng(seed)
        ug = np.unique(groups)
        val_g = rng.choice(ug, size=max(1, int(len(ug) * val_frac)), replace=False) if val_frac > 0 else []
        is_val = np.isin(groups, val_g)
        tr = np.nonzero(~is_val)[0]
        va = np.nonzero(is_val)[0]
        losses, vlosses = [], []
        best = (float("inf"), self.mlp.params(), 0)
        for ep in range(epochs):
            perm = rng.permutation(tr)
            el = []
            for i in range(0, len(perm), batch):
                j = perm[i:i + batch]
                el.append(self.mlp.train_step(Xn[j], tn[j], lr))
            losses.append(float(np.mean(el)))
            if len(va) and (ep + 1) % check_every == 0:
                v = float(np.mean((self.mlp.forward(Xn[va]) - tn[va]) ** 2))
                vlosses.append(v)
                if v < best[0]:
                    best = (v, self.mlp.params(), ep + 1)
        if len(va):
            self.mlp.set_params(best[1])
        self.history.update({"loss": losses, "val_loss": vlosses, "best_epoch": best[2]})
        return losses

    def predict_rel_log2(self, X: np.ndarray) -> np.ndarray:
        """Predicted log2 work of each order relative to its pattern's mean."""
        y = self.mlp.forward(self._norm(np.asarray(X, np.float32)).astype(np.float32))
        return y * self.t_sd

    # ---- persistence
    def save(self, path: str):
        np.savez(path, params=self.mlp.params(), mu=self.mu, sd=self.sd,
                 meta=np.array(json.dumps({"t_sd": self.t_sd, "threshold": self.threshold, "margin": self.margin,
                                           "judge_bits": self.judge_bits, "sizes": self.sizes})))

    @classmethod
    def load(cls, path: str) -> "LearnedOrder":
        d = np.load(path, allow_pickle=False)
        m = cls()
        m.mlp.set_params(d["params"])
        m.mu, m.sd = d["mu"], d["sd"]
        meta = json.loads(str(d["meta"]))
        m.t_sd = meta["t_sd"]
        m.threshold = meta["threshold"]
        m.margin = meta.get("margin", m.margin)
        m.judge_bits = meta.get("judge_bits", m.judge_bits)
        m.sizes = meta.get("sizes")
        return m


def relative_targets(y_log2: np.ndarray, groups: np.ndarray) -> np.ndarray:
    """y minus the mean of y over its group."""
    y = np.asarray(y_log2, np.float64)
    g = np.asarray(groups)
    _, inv = np.unique(g, return_inverse=True)
    mean = np.bincount(inv, weights=y) / np.bincount(inv)
    return y - mean[inv]


def load(path: str) -> LearnedOrder:
    return LearnedOrder.load(path)


JUDGE_SEED = 1  # seed of the sampled cost model's sampling (one seed: one judgement)


def _proposal(model: LearnedOrder, pattern: Pattern, graph: Graph, mode: int, fallback: Plan, top: int,
              keep_root: bool = False, root_range=None):
    """The network's pick among the planner's `top` candidate orders, or None
    (the planner's order) when the pick is not predicted `margin` bits better
    or the pattern size is outside the training sizes."""
    if model.sizes and pattern.k not in model.sizes:
        return _proposal_from(model, pattern, fallback, None, None)
    cands, X = _features(pattern, graph, mode, top, _root_of(fallback, pattern) if keep_root else None, root_range)
    return _proposal_from(model, pattern, fallback, cands, X)


def _judge(model: LearnedOrder, pattern: Pattern, graph: Graph, mode: int, fallback: Plan, order, info):
    """The reference's sampled cost model judges the pick: more than
    `judge_bits` above the planner's order -> the planner's plan."""
    c_new = order_cost_sampled(pattern, graph, order, mode, JUDGE_SEED)
    c_old = order_cost_sampled(pattern, graph, info["cost_order"], mode, JUDGE_SEED)
    info["judge_bits_over"] = float(np.log2(max(c_new, 1.0)) - np.log2(max(c_old, 1.0)))
    info["judge_costs"] = [c_old, c_new]
    if info["judge_bits_over"] > model.judge_bits:
        info["fallback"] = True
        return fallback, info
    return build_plan(pattern, graph, mode, order), info


def probe_work(graph: Graph, plans: Sequence[Plan], root_range=None):
    """Measured work (wavefront clock cycles) of each plan on a probe sample of
    the run's task set that the run's own timing does not see in full: for a
    whole job a strided sample (stride >= PROBE_MIN_STRIDE), for a root-edge
    range the range next to it (disjoint from it, at most PROBE_EDGES long).
    plans[0] runs first, under PROBE_CAP_K, and bounds the others by a work cap
    of 4x its own; returns [(work, capped)], a capped plan's work a lower
    bound."""
    nnz = 2 * int(graph.m)
    out = []
    cap_k = PROBE_CAP_K
    for i, pl in enumerate(plans):
        if root_range is None:
            stride = max(PROBE_MIN_STRIDE, nnz // PROBE_EDGES)
            (st, lim), = time_capped(graph, pl, stride, cap_k)
        else:
            b, e = int(root_range[0]), int(root_range[1])
            n = min(e - b, PROBE_EDGES)
            pb, pe = (e, e + n) if e + n <= nnz else (max(0, b - n), b)
            if pe <= pb:
                pb, pe = b, e
            st, lim = range_capped(graph, pl, pb, pe, cap_k)
        out.append((float(st.wave_cycles), lim))
        if i == 0:
            cap_k = max(64, 4 * work_steps_bound(st) // 1024 + 1)
    return out


def _guard(graph: Graph, fallback: Plan, plan: Plan, info, root_range):
    """Keep the learned plan only when the probe measures less work for it."""
    (w_cost, _), (w_new, lim) = probe_work(graph, [fallback, plan], root_range)
    info["probe_work"] = [w_cost, w_new]
    if lim or np.log2(max(w_new, 1.0)) > np.log2(max(w_cost, 1.0)) - PROBE_BITS:
        info["fallback"] = True
        info["reason"] = "probe: no less measured work"
        return fallback, info
    return plan, info


def choose_plan(model: LearnedOrder, pattern: Pattern, graph: Graph, mode: int, fallback: Optional[Plan] = None,
                top: int = 12, keep_root: bool = False, root_range=None, probe: bool = True):
    """Order with the lowest predicted work among the planner's `top`
    candidates, taken only when the network predicts it at least `margin` bits
    below the planner's own order, the reference's sampled cost model does
    not put it more than `judge_bits` above that order (the fallback), and (with
    probe) the measured probe (probe_work) confirms less work.
    keep_root: choose only among the orders that count the same embeddings as
    the fallback's order over any root subset (root_preserving); root_range =
    (begin, end): the run covers those root edges only, and the features'
    random walks start there."""
    if fallback is None:
        fallback = build_plan(pattern, graph, mode)
    order, info = _proposal(model, pattern, graph, mode, fallback, top, keep_root, root_range)
    if order is None:
        return fallback, info
    plan, info = _judge(model, pattern, graph, mode, fallback, order, info)
    if probe and plan is not fallback:
        plan, info = _guard(graph, fallback, plan, info, root_range)
    return plan, info


def choose_plans(model: LearnedOrder, patterns: Sequence[Pattern], graph: Graph, mode: int,
                 fallbacks: Sequence[Plan], top: int = 12, threads: int = 16, keep_root: bool = False,
                 root_ranges=None, probe: bool = True):
    """choose_plan over a pattern set: feature sampling and the sampled cost
    model (host code, the GIL released in the library) run on `threads` threads;
    the network's forward passes stay on the caller's thread.  root_ranges =
    (begins, ends): pattern i's run covers root edges [begins[i], ends[i])."""
    rr = [None] * len(patterns) if root_ranges is None else list(zip(root_ranges[0], root_ranges[1]))
    with ThreadPoolExecutor(max_workers=threads) as ex:
        feats = list(ex.map(lambda a: ((None, None) if model.sizes and a[0].k not in model.sizes
                                       else _features(a[0], graph, mode, top,
                                                      _root_of(a[1], a[0]) if keep_root else None, a[2])),
                            list(zip(patterns, fallbacks, rr))))
        props = []
        for p, fb, (cands, X) in zip(patterns, fallbacks, feats):
            props.append(_proposal_from(model, p, fb, cands, X))
        judged = list(ex.map(lambda a: (a[1] if a[0] is None else
                                        _judge(model, a[2], graph, mode, a[3], a[0], a[1])),
                             [(o, inf, p, fb) for (o, inf), p, fb in zip(props, patterns, fallbacks)]))
    out = []
    for (o, inf), fb, j, r in zip(props, fallbacks, judged, rr):
        if o is None:
            out.append((fb, inf))
        elif probe and j[0] is not fb:
            out.append(_guard(graph, fb, j[0], j[1], r))  # GPU runs: on the caller's thread
        else:
            out.append(j)
    return out


def _root_of(plan: Plan, pattern: Pattern) -> List[int]:
    return list(plan.order)[: pattern.k]


def _features(pattern: Pattern, graph: Graph, mode: int, top: int, keep_root_of=None, root_range=None):
    cands = candidate_orders(pattern, graph, mode, top, keep_root_of=keep_root_of)
    return cands, np.stack([order_features(pattern, graph, o, mode, FEATURE_SAMPLES, 11, root_range) for o in cands])


def _proposal_from(model: LearnedOrder, pattern: Pattern, fallback: Plan, cands, X):
    cost_order = list(fallback.order)[: pattern.k]
    info = {"pred_rel_log2": 0.0, "cost_order": cost_order, "learned_order": cost_order, "fallback": False,
            "judge_bits_over": 0.0}
    if model.sizes and pattern.k not in model.sizes:
        # outside the pattern sizes the network was trained on its ranking is
        # unfounded (5-vertex house in IEP mode: a 3.1x slower order)
        info.update(fallback=True, reason="pattern size outside the training sizes")
        return None, info
    pred = model.predict_rel_log2(X)
    best = int(np.argmin(pred))
    ci = cands.index(cost_order) if cost_order in cands else None
    info["pred_rel_log2"] = float(pred[best] - (pred[ci] if ci is not None else 0.0))
    if ci is not None and pred[ci] - pred[best] < model.margin:
        return None, info
    info["learned_order"] = list(cands[best])
    return list(cands[best]), info


def evaluate(model: "LearnedOrder", graph: Graph, patterns: Sequence[Pattern], mode: int, stride: int = 1,
          reps: int = 2, log=print):
    """Learned choice against the cost model's choice, per pattern: the same
    executor on the same root sample, best of `reps` device times.  The learned
    plan runs under a work cap of 4x the cost-model plan's work (a runaway
    choice is reported as capped, never run to the end); uncapped whole-job
    counts (stride 1) must agree."""
    rows = []
    for i, p in enumerate(patterns):
        cost_plan = build_plan(p, graph, mode)
        plan, info = choose_plan(model, p, graph, mode, fallback=cost_plan)
        t_cost, t_learn, capped = [], [], False
        same = list(plan.order)[: p.k] == list(cost_plan.order)[: p.k]
        for _ in range(reps):
            cnt, st = count_sampled(graph, [cost_plan], stride)
            t_cost.append(st[0].kernel_ms)
            if same:  # the model kept the cost model's order
                t_learn.append(st[0].kernel_ms)
                continue
            cap_k = max(1024, 4 * work_steps_bound(st[0]) // 1024 + 1)
            (sl, lim), = time_capped(graph, plan, stride, cap_k) if not info["fallback"] else [(st[0], False)]
            capped = capped or lim
            if not lim and not info["fallback"]:  # it finishes: time it without the cap's accounting
                _, sl = count_sampled(graph, [plan], stride)
                sl = sl[0]
            t_learn.append(sl.kernel_ms)
        if stride == 1 and not capped and not same:
            # whole-job counts do not depend on the order (a root-edge sample does:
            # another order roots the search at another pattern edge)
            cl, _ = count_sampled(graph, [plan], stride)
            if cl[0] != cnt[0]:
                raise AssertionError(f"pattern {i}: learned plan counted {cl[0]}, cost-model plan {cnt[0]}")
        row = {"pattern": i, "cost_ms": float(min(t_cost)), "learned_ms": float(min(t_learn)), "capped": capped,
               "pred_rel_log2": info["pred_rel_log2"],
               "measured_rel_log2": float(np.log2(max(min(t_learn), 1e-3) / max(min(t_cost), 1e-3))),
               "fallback": info["fallback"], "cost_order": info["cost_order"],
               "learned_order": info["learned_order"]}
        rows.append(row)
        log(f"[evaluate] pattern {i}: cost-model {row['cost_ms']:.2f} ms, learned {row['learned_ms']:.2f} ms"
            f"{' (capped)' if capped else ''}, fallback {info['fallback']}")
    return rows


def log_pairs(graph: Graph, patterns: Sequence[Pattern], mode: int, stride: int = 1024, orders_per_pattern: int = 8,
              seed: int = 0, log=print):
    """(features, log2 projected full-job work, log2 analytic cost, capped, pattern id) of
    `orders_per_pattern` orders per pattern, measured with log_work.  An order
    stopped by its work cap is logged CAPPED_PENALTY bits above its
    (lower-bound) time; a pattern whose pilot order is stopped is skipped."""
    rng = np.random.default_rng(seed)
    X, y, c, cap, grp, ords = [], [], [], [], [], []
    t0 = time.time()
    for i, p in enumerate(patterns):
        orders = candidate_orders(p, graph, mode, orders_per_pattern // 2, rng)[:orders_per_pattern]
        r = log_work(graph, p, orders, mode, stride)
        if r is None:
            log(f"[learned] pattern {i}: pilot order stopped by the cap, skipped")
            continue
        ms, by, capped, used = r
        for o, t, cp in zip(orders, ms, capped):
            ords.append(list(o) + [255] * (8 - len(o)))
            X.append(order_features(p, graph, o, mode, FEATURE_SAMPLES, 11))
            y.append(np.log2(max(t, 1.0)) + (CAPPED_PENALTY if cp else 0.0))
            c.append(np.log2(max(order_cost(p, graph, o, mode), 1.0)))
            cap.append(bool(cp))
            grp.append(seed + i)
        log(f"[learned] logged {i + 1}/{len(patterns)} patterns, {len(X)} orders, stride {used}, "
            f"projected Gcycles {float(np.min(ms)) / 1e9:.1f}..{float(np.max(ms)) / 1e9:.1f}, capped {int(np.sum(capped))}, "
            f"{time.time() - t0:.0f}s")
    return (np.array(X, np.float32), np.array(y, np.float64), np.array(c, np.float64), np.array(cap, bool),
            np.array(grp, np.int64), np.array(ords, np.uint8).reshape(-1, 8))


def fit_model(X: np.ndarray, y: np.ndarray, groups: np.ndarray, epochs: int = 300, seed: int = 0, log=print):
    model = LearnedOrder(seed=seed)
    losses = model.fit(X, y, groups, epochs=epochs, seed=seed)
    log(f"[learned] trained on {len(X)} orders of {len(np.unique(groups))} patterns, "
        f"loss {losses[0]:.3f} -> {losses[-1]:.3f}, kept epoch {model.history['best_epoch']} "
        f"(held-out loss {min(model.history['val_loss'] or [float('nan')]):.3f})")
    return model


def train(graph: Graph, patterns: Sequence[Pattern], mode: int, stride: int = 1024, orders_per_pattern: int = 8,
          epochs: int = 300, seed: int = 0, log=print):
    """Log (order, measured work) pairs with the plain executor and fit the MLP."""
    X, y, c, cap, grp, _ = log_pairs(graph, patterns, mode, stride, orders_per_pattern, seed, log)
    model = fit_model(X, y, grp, epochs, seed, log)
    model.sizes = sorted({p.k for p in patterns})
    return model, (X, y, c, grp)
