"""GPU parity on the two large R-MAT configurations, on bounded root-edge subsets.

R-MAT scale 20 (seed 4; max degree 64,810) with the 78 seven-vertex patterns,
and R-MAT scale 23 (seed 6; max degree 258,497, a 1 GB CSR in HBM) with the 16
six-vertex patterns plus K5-e, both in IEP mode.  Whole jobs hold 10^17 or
more matches per pattern, so each pattern runs over its own contiguous range
of root edges (the ranges and the oracle's counts are in
tests/golden/large_c*.json, written by tests/golden/make_large.py): the plain
executor (one launch per plan), the batched executor (one launch, hipGraph
replay on scale 20) and the oracle must agree bit for bit.  The searches run
through hub rows, so the large-row paths (scratch slabs of max degree + 64
words, global F bitmaps, level-2 root splits of hub edges, continuation items)
are exercised at their real sizes.  A work cap bounds every capped launch; the
uncapped kernels run the same subsets only after the capped ones finished.
"""
import json
import os
import sys

import numpy as np
import pytest

import sge
import sge.workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
PLAN_FIELDS = ("k", "mode", "order", "parent", "lower", "upper", "check", "reuse", "ready", "tail", "tail_kind",
               "multiplier")
CAP_K = 1 << 20  # work cap: 2^30 wave steps per launch, far above any subset here


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def load(name):
    with open(os.path.join(GOLDEN, f"large_{name}.json")) as f:
        return json.load(f)


def capped():
    t = sge.default_tuning()
    t.node_cap_k = CAP_K
    return t


def plans_of(doc, g, patterns):
    """The recorded plans, rebuilt from their orders (and checked field by field)."""
    plans = []
    for e in doc["subsets"]:
        p = patterns[e["pattern"]]
        assert [list(x) for x in p.edges()] == [list(x) for x in e["edges"]], e["pattern"]
        plan = sge.build_plan(p, g, sge.MODE_IEP, e["plan"]["order"])
        got = plan.as_dict()
        for f in PLAN_FIELDS:
            assert got[f] == e["plan"][f], (e["pattern"], f)
        plans.append(plan)
    return plans


def check_graph(g, doc):
    assert (g.n, g.m, g.stats.max_degree) == (doc["n"], doc["m"], doc["max_degree"])


def oracle_spot_checks(g, plans, doc, picks):
    off, nbr = g.csr()
    for i in picks:
        e = doc["subsets"][i]
        say("oracle recount", e["pattern"], e["begin"], e["end"])
        assert O.count(plans[i], off, nbr, 0, e["begin"], e["end"]) == e["count"], e["pattern"]


@pytest.fixture(scope="module")
def c4():
    say("building R-MAT scale 20 (seed 4) on the GPU")
    g = W.graph_c4(device_build=True)
    doc = load("c4")
    check_graph(g, doc)
    return g, doc, plans_of(doc, g, W.seven_vertex_set())


def test_c4_subsets_plain_batched_oracle(c4):
    g, doc, plans = c4
    want = [e["count"] for e in doc["subsets"]]
    b = [e["begin"] for e in doc["subsets"]]
    e = [e["end"] for e in doc["subsets"]]
    say("c4 plain, capped, 78 plans")
    plain, st, _ = sge.count_ranges(g, plans, b, e, batched=False, tuning=capped())
    assert plain == want
    say("c4 batched launch, capped")
    got, _, _ = sge.count_ranges(g, plans, b, e, batched=True, tuning=capped())
    assert got == want
    say("c4 batched launch, uncapped")
    got, _, _ = sge.count_ranges(g, plans, b, e, batched=True)
    assert got == want
    say("c4 hipGraph batch, two replays")
    batch = sge.Batch(g, plans, ranges=(b, e))
    for _ in range(2):
        got, _ = batch.run()
        assert got == want
    batch.close()
    say("c4 plain, uncapped")
    plain, _, _ = sge.count_ranges(g, plans, b, e, batched=False)
    assert plain == want


def test_c4_subsets_tuning_variants(c4):
    """Donation (budget and split), root-split and probe-ratio knobs at their
    extremes on hub-row searches."""
    g, doc, plans = c4
    want = [e["count"] for e in doc["subsets"]]
    b = [e["begin"] for e in doc["subsets"]]
    e = [e["end"] for e in doc["subsets"]]
    for knobs in ({"donate_budget": 64, "split_span": 8}, {"stage_max": 64, "bitmap_ratio": 64, "chunk": 1},
                  {"global_bitmap_min": 16, "head_items": 1 << 30}, {"donate_split": 64, "donate_budget": 256},
                  {"donate_split": 256, "donate_budget": 1024}):
        t = capped()
        for k, v in knobs.items():
            setattr(t, k, v)
        say("c4 batched, tuning", knobs)
        got, _, _ = sge.count_ranges(g, plans, b, e, batched=True, tuning=t)
        assert got == want, knobs


def test_c4_own_items_and_next_pass_alternating(c4):
    """Passes that take their own continuation items (tickets on the pass's
    output, epoch-published slots) and passes that leave them to the next pass
    alternate in one process, batched and plain, four times: every run returns
    the fixture counts (an earlier in-pass scheme lost work in the launch after
    such runs, in 1 of 8 repeats)."""
    g, doc, plans = c4
    want = [e["count"] for e in doc["subsets"]]
    b = [e["begin"] for e in doc["subsets"]]
    e = [e["end"] for e in doc["subsets"]]
    for rep in range(4):
        for pass_items in (2, 1):
            t = capped()
            t.pass_items = pass_items
            say("c4 alternating", rep, "pass_items", pass_items)
            got, _, _ = sge.count_ranges(g, plans, b, e, batched=True, tuning=t)
            assert got == want, (rep, pass_items)
            batch = sge.Batch(g, plans, tuning=t, ranges=(b, e))
            got, _ = batch.run()
            batch.close()
            assert got == want, (rep, pass_items, "graph")
    t = capped()
    t.pass_items = 2
    got, _, _ = sge.count_ranges(g, plans, b, e, batched=False, tuning=t)
    assert got == want


def test_c4_learned_orders_count_the_same_subsets(c4):
    """The learned model, choosing only among orders that keep each cost-model
    order's root edge and symmetry rules, changes orders and leaves every
    subset count equal to the oracle's (the bench's learned-vs-plain timing
    rests on this).  The measured probe is left out so that the network's own
    proposals (which the probe mostly rejects on these ranges) are run."""
    from sge import learned as L

    g, doc, plans = c4
    want = [e["count"] for e in doc["subsets"]]
    b = [e["begin"] for e in doc["subsets"]]
    e = [e["end"] for e in doc["subsets"]]
    model = L.load(os.path.join(os.path.dirname(sge.__file__), "models", "order_mlp.npz"))
    pats = [W.seven_vertex_set()[x["pattern"]] for x in doc["subsets"]]
    say("c4 learned root-preserving orders")
    learned = [pl for pl, _ in L.choose_plans(model, pats, g, sge.MODE_IEP, plans, keep_root=True,
                                              root_ranges=(b, e), probe=False)]
    changed = sum(list(a.order) != list(c.order) for a, c in zip(learned, plans))
    say(f"c4 learned: {changed} of {len(plans)} orders changed; batched, capped")
    assert changed >= 5
    got, _, _ = sge.count_ranges(g, learned, b, e, batched=True, tuning=capped())
    assert got == want
    say("c4 learned: plain, capped")
    got, _, _ = sge.count_ranges(g, learned, b, e, batched=False, tuning=capped())
    assert got == want


def test_c4_oracle_spot_checks(c4):
    g, doc, plans = c4
    oracle_spot_checks(g, plans, doc, [0, 3, 40, 77])


def test_c4_work_cap_stops_batched_run(c4):
    """A cap below the work stops the batch with SGE_ELIMIT and partial counts;
    every kernel exits normally and the next run is exact."""
    g, doc, plans = c4
    b = [e["begin"] for e in doc["subsets"]]
    e = [e["end"] for e in doc["subsets"]]
    t = sge.default_tuning()
    t.node_cap_k = 4
    batch = sge.Batch(g, plans, tuning=t, ranges=(b, e))
    with pytest.raises(sge.SgeError) as ex:
        batch.run()
    assert ex.value.code == sge.ELIMIT
    batch.close()
    got, _, _ = sge.count_ranges(g, plans, b, e, batched=True, tuning=capped())
    assert got == [e["count"] for e in doc["subsets"]]


def test_c5_subsets_plain_batched_oracle():
    say("building R-MAT scale 23 (seed 6) on the GPU")
    g = W.graph_c5(device_build=True)
    doc = load("c5")
    check_graph(g, doc)
    plans = plans_of(doc, g, W.six_vertex_set() + [W.named("near5clique")])
    want = [e["count"] for e in doc["subsets"]]
    b = [e["begin"] for e in doc["subsets"]]
    e = [e["end"] for e in doc["subsets"]]
    say("c5 plain, capped, 17 plans")
    plain, _, _ = sge.count_ranges(g, plans, b, e, batched=False, tuning=capped())
    assert plain == want
    say("c5 batched launch, capped")
    got, _, _ = sge.count_ranges(g, plans, b, e, batched=True, tuning=capped())
    assert got == want
    say("c5 batched launch, uncapped")
    got, _, _ = sge.count_ranges(g, plans, b, e, batched=True)
    assert got == want
    oracle_spot_checks(g, plans, doc, [0, 16])
