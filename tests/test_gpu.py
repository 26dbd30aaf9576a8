"""GPU parity: the HIP executor (through the C ABI) against the oracle, the
reference's golden counts and brute force.  Bit-exact uint64 counts required."""
import itertools
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
MODES = (sge.MODE_WCOJ, sge.MODE_EXPLORE, sge.MODE_IEP)


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def golden(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        return json.load(f)


def random_graph(n, p, seed):
    rng = np.random.default_rng(seed)
    edges = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < p]
    return sge.Graph.from_edges(n, edges if edges else [(0, 1)])


def small_patterns():
    return [W.named(n) for n in ["triangle", "diamond", "house", "near5clique", "k4", "k5"]] + [
        sge.Pattern.from_edges(3, [(0, 1), (1, 2)]),
        sge.Pattern.from_edges(4, [(0, 1), (1, 2), (2, 3), (3, 0)]),
        sge.Pattern.from_edges(4, [(0, 1), (0, 2), (0, 3)]),
        sge.Pattern.from_edges(5, [(0, 1), (0, 2), (0, 3), (0, 4)]),
        sge.Pattern.from_edges(5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)]),
        sge.Pattern.from_edges(6, [(0, 1), (1, 2), (2, 0), (0, 3), (1, 4), (2, 5)]),
        sge.Pattern.from_edges(2, [(0, 1)]),
    ]


def test_c1_all_patterns_all_modes():
    g = W.graph_c1()
    off, nbr = g.csr()
    gold = golden("c1")["patterns"]
    for p in small_patterns() + W.six_vertex_set()[:6]:
        for mode in MODES:
            say("c1", p.edges(), "mode", mode)
            plan = sge.build_plan(p, g, mode)
            got = sge.count(g, plan)[0][0]
            assert got == O.count(plan, off, nbr), (p, mode)
    for name, rec in gold.items():
        for mode in MODES:
            assert sge.count(g, sge.build_plan(W.named(name), g, mode))[0][0] == rec["wcoj"]["count"]


@pytest.mark.parametrize("seed", range(3))
def test_random_graphs_vs_brute_force(seed):
    g = random_graph(40, 0.3, seed)
    off, nbr = g.csr()
    for p in small_patterns():
        if p.k > 5:
            continue
        info = p.info()
        brute = O.count_brute(p.k, p.c.adj, off, nbr)
        want = brute if info.single_bag else brute // info.aut_size
        for mode in MODES:
            say("random", seed, p.edges(), mode)
            assert sge.count(g, sge.build_plan(p, g, mode))[0][0] == want


def test_every_order_same_count():
    g = W.graph_c1()
    for p in [W.named("house"), W.named("diamond")]:
        seen = set()
        for o in p.orders():
            say("order", o)
            seen.add(sge.count(g, sge.build_plan(p, g, sge.MODE_WCOJ, o))[0][0])
        assert len(seen) == 1


def test_c2_golden_all_modes():
    g = W.graph_c2()
    gold = golden("c2")["patterns"]
    for name in ["near5clique", "house"]:
        for mode, key in ((sge.MODE_WCOJ, "wcoj"), (sge.MODE_EXPLORE, "explore"), (sge.MODE_IEP, "iep")):
            say("c2", name, key)
            got, st = sge.count(g, sge.build_plan(W.named(name), g, mode))
            assert got[0] == gold[name][key]["count"], (name, key)


def test_c2_house_root_ranges_vs_oracle():
    g = W.graph_c2()
    off, nbr = g.csr()
    plan = sge.build_plan(W.named("house"), g)
    nnz = int(off[-1])
    for a in [0, nnz // 3, nnz - 3000]:
        say("c2 house range", a)
        got, _ = sge.count_range(g, plan, a, a + 1500)
        assert got == O.count(plan, off, nbr, begin=a, end=a + 1500)


def test_batched_equals_plain_and_oracle():
    g = W.graph_c1()
    off, nbr = g.csr()
    pats = W.six_vertex_set()
    for mode in MODES:
        plans = [sge.build_plan(p, g, mode) for p in pats]
        say("batched six-vertex mode", mode)
        plain, _ = sge.count(g, plans, batched=False)
        batched, _ = sge.count(g, plans, batched=True)
        assert plain == batched
        assert plain == [O.count(pl, off, nbr) for pl in plans]


def test_hipgraph_batch_replays():
    g = W.graph_c1()
    pats = W.six_vertex_set()[:5]
    plans = [sge.build_plan(p, g, sge.MODE_IEP) for p in pats]
    want, _ = sge.count(g, plans)
    b = sge.Batch(g, plans)
    for rep in range(4):
        say("hipgraph replay", rep)
        got, st = b.run()
        assert got == want
        _ = sge.count(g, plans[:1])  # eager work between replays
    b.close()


def test_tuning_variants_same_counts():
    g = W.graph_c1()
    plans = [sge.build_plan(p, g) for p in W.six_vertex_set()[:4]]
    want, _ = sge.count(g, plans)
    for bt, ch, head in [(64, 1, 0), (128, 7, 1 << 30), (192, 64, 3), (256, 1000, 0), (256, 5, 100000)]:
        t = sge.default_tuning()
        t.block_threads, t.chunk, t.head_items = bt, ch, head
        say("tuning", bt, ch, head)
        assert sge.count(g, plans, batched=True, tuning=t)[0] == want
        assert sge.count(g, plans, batched=False, tuning=t)[0] == want


def test_donation_split_same_counts():
    """Donated ranges split into several items (small budgets make every
    search donate) give the same counts in all three modes, batched and plain."""
    g = W.graph_c2()
    pats = [W.named("house"), W.named("near5clique")] + W.six_vertex_set()[:3]
    for mode in (sge.MODE_WCOJ, sge.MODE_EXPLORE, sge.MODE_IEP):
        plans = [sge.build_plan(p, g, mode) for p in pats]
        want, _ = sge.count(g, plans)
        for split, budget, pass_items in ((2, 16, 2), (16, 64, 2), (256, 16, 2), (64, 4096, 2), (2, 16, 1),
                                          (64, 4096, 1), (16, 64, 0)):
            t = sge.default_tuning()
            t.donate_split, t.donate_budget, t.pass_items = split, budget, pass_items
            say("donation split", mode, split, budget, "pass items", pass_items)
            assert sge.count(g, plans, batched=True, tuning=t)[0] == want, (mode, split, budget, pass_items)
            assert sge.count(g, plans, batched=False, tuning=t)[0] == want, (mode, split, budget, pass_items)
            if pass_items != 2:  # items left to later passes: the hipGraph batch drains them
                b = sge.Batch(g, plans, tuning=t)
                assert b.run()[0] == want
                b.close()
    t = sge.default_tuning()
    t.donate_split = 257
    with pytest.raises(sge.SgeError):
        sge.count(g, plans, tuning=t)
    t = sge.default_tuning()
    t.pass_items = 3
    with pytest.raises(sge.SgeError):
        sge.count(g, plans, tuning=t)


def test_probe_direction_ratios_same_counts():
    """Which side of a last-level term drives the probes (LDS-staged F or the
    sibling's CSR row) must not change a count."""
    g = W.graph_c2()
    plans = [sge.build_plan(p, g, sge.MODE_IEP) for p in W.six_vertex_set()[:6]]
    base = None
    for stage, bm in ((8, 32), (1, 1), (64, 256), (4, 8)):
        t = sge.default_tuning()
        t.stage_max, t.bitmap_ratio = stage, bm
        say("probe ratio", stage, "bitmap ratio", bm)
        got = sge.count_sampled(g, plans, 31, 3, tuning=t)[0]
        if base is None:
            base = got
        assert got == base, (stage, bm)


def test_node_cap_stops_and_reports():
    """A work cap ends the call with SGE_ELIMIT and partial stats; a cap the
    search never reaches changes nothing."""
    g = W.graph_c2()
    plan = sge.build_plan(W.six_vertex_set()[1], g, sge.MODE_WCOJ)
    want, st = sge.count_sampled(g, [plan], 16)
    assert st[0].wave_cycles > 0
    say("node cap 1k")
    (s1, lim1), = sge.time_capped(g, plan, 16, 1)
    assert lim1 and 0 < s1.nodes < st[0].nodes
    with pytest.raises(sge.SgeError) as e:
        t = sge.default_tuning()
        t.node_cap_k = 1
        sge.count_sampled(g, [plan], 16, tuning=t)
    assert e.value.code == sge.ELIMIT
    say("node cap above the work")
    (s2, lim2), = sge.time_capped(g, plan, 16, 4 * sge.work_steps_bound(st[0]) // 1024 + 1024)
    assert not lim2 and s2.nodes == st[0].nodes
    assert sge.count_sampled(g, [plan], 16)[0] == want


def test_root_edge_split_same_counts():
    """Heavy root edges split into level-2 slices must not change any count."""
    g = W.graph_c2()
    plans = [sge.build_plan(p, g, m) for p in W.six_vertex_set()[:3] for m in MODES]
    base = None
    for span, budget in ((1 << 30, 1 << 30), (64, 1 << 30), (4, 1 << 30), (1, 1 << 30), (1 << 30, 16), (8, 2),
                         (1, 1)):
        t = sge.default_tuning()
        t.split_span = span
        t.donate_budget = budget
        say("split span", span, "donate budget", budget)
        got = [sge.count_sampled(g, [pl], 97, 5, tuning=t)[0][0] for pl in plans]
        if base is None:
            base = got
        assert got == base, (span, budget)
    t = sge.default_tuning()
    t.donate_budget = 2
    say("batched with donation")
    b = sge.Batch(g, plans, tuning=t, stride=97, offset=5)
    for _ in range(2):
        assert b.run()[0] == base
    b.close()
    off, nbr = g.csr()
    for pl, c in zip(plans, base):
        assert c == O.count_strided(pl, off, nbr, 97, 5)
    b = sge.Batch(g, plans, stride=97, offset=5)
    assert b.run()[0] == base
    b.close()


def test_enumeration_c1_diamond_matches_brute_force():
    g = W.graph_c1()
    p = W.named("diamond")
    plan = sge.build_plan(p, g)
    cnt = sge.count(g, plan)[0][0]
    say("enumerate diamond")
    rows = sge.enumerate_embeddings(g, plan, p, cnt + 10)
    assert rows.shape[0] == cnt
    # brute force canonical set in original ids
    g0 = W.graph_c1(False)
    off, nbr = g0.csr()
    adj = [set(nbr[off[v]:off[v + 1]].tolist()) for v in range(g0.n)]
    auts = [perm for perm in itertools.permutations(range(4))
            if all(((perm[a], perm[b]) in p.edges() or (perm[b], perm[a]) in p.edges()) for a, b in p.edges())]
    want = set()
    for a in range(g0.n):
        for b in adj[a]:
            for c in adj[a] & adj[b]:
                for d in adj[a] & adj[c]:
                    if len({a, b, c, d}) < 4:
                        continue
                    img = (a, b, c, d)
                    if all(img[y] in adj[img[x]] for x, y in p.edges()):
                        want.add(min(tuple(img[q[i]] for i in range(4)) for q in auts))
    got = set(map(tuple, rows.tolist()))
    assert got == want
    assert rows.tolist() == sorted(rows.tolist())


def test_iep_reproduces_choosec_wrap():
    # star graph: hub with 70000 leaves (degree order gives leaves ids 0..69999).
    # K_{1,5} in IEP mode: root edge (hub, leaf j), then the other four leaves are
    # one symmetric group counted by choosec(#leaves above j, 4), whose numerator
    # wraps mod 2^64 once the argument exceeds ~65536.
    n = 70001
    g = sge.Graph.from_edges(n, [(0, i) for i in range(1, n)])
    star5 = sge.Pattern.from_edges(6, [(0, i) for i in range(1, 6)])
    iep = sge.build_plan(star5, g, sge.MODE_IEP)
    assert iep.tail == 4

    def choosec_wrap(x, k):
        if k > x:
            return 0
        nom = x
        for i in range(1, k):
            nom = (nom * (x - i)) % 2**64
        return nom // 24

    want = sum(choosec_wrap(x, 4) for x in range(n - 1)) % 2**64
    import math
    assert want != math.comb(n - 1, 5)  # the wrap changes the result at this size
    say("iep wrap")
    assert sge.count(g, iep)[0][0] == want


def test_edge_cases():
    say("edge cases")
    empty = sge.Graph.from_edges(10, [(0, 1)])
    assert sge.count(empty, sge.build_plan(W.named("triangle"), empty))[0][0] == 0
    g = W.graph_c1()
    edge = sge.Pattern.from_edges(2, [(0, 1)])
    pl = sge.build_plan(edge, g)
    assert edge.info().single_bag == 1  # the reference prints 2m for the edge pattern
    assert sge.count(g, pl)[0][0] == 2 * g.m
    single = sge.Pattern.from_edges(1, [])
    assert sge.count(g, sge.build_plan(single, g))[0][0] == g.n


def test_device_graph_build_matches_host():
    """The GPU CSR build (radix sorts, unique, degree ranking, scan) gives the
    host builder's arrays exactly: duplicates, both orientations and self-loops
    in the input, isolated vertices, with and without degree ordering."""
    rng = np.random.default_rng(11)
    cases = []
    e = rng.integers(0, 300, size=(5000, 2), dtype=np.uint32)
    e = np.concatenate([e, e[:700, ::-1], np.stack([np.arange(40), np.arange(40)], 1).astype(np.uint32)])
    cases.append(("random+dups+loops", 400, e))
    cases.append(("empty", 10, np.zeros((0, 2), np.uint32)))
    cases.append(("single", 2, np.array([[1, 0]], np.uint32)))
    for name, n, edges in cases:
        for order in (True, False):
            say("device build", name, order)
            h = sge.Graph.from_edges(n, edges, degree_order=order)
            d = sge.Graph.from_edges(n, edges, degree_order=order, device_build=True)
            for a, b in zip(h.csr(), d.csr()):
                assert np.array_equal(a, b), (name, order)
            assert np.array_equal(h.perm(), d.perm()), (name, order)
            assert h.stats.max_degree == d.stats.max_degree
    say("device build rmat 14")
    h = sge.Graph.rmat(14, 16, seed=9)
    d = sge.Graph.rmat(14, 16, seed=9, device_build=True)
    for a, b in zip(h.csr(), d.csr()):
        assert np.array_equal(a, b)
    assert np.array_equal(h.perm(), d.perm())
    # a device-built graph is device-resident and counts like the host-built one
    pl = sge.build_plan(W.named("house"), h, sge.MODE_IEP)
    assert sge.count_sampled(d, [pl], 7)[0] == sge.count_sampled(h, [pl], 7)[0]


def test_vertex_relabeling_invariance():
    """Counts do not depend on the data graph's vertex ids: the degree-ordered
    CSR, the original ids and a random relabelling give the same counts."""
    g0 = W.graph_c1()
    off, nbr = g0.csr()
    src = np.repeat(np.arange(len(off) - 1), np.diff(off))
    edges = np.stack([src, nbr], 1)
    edges = edges[edges[:, 0] < edges[:, 1]].astype(np.uint32)
    perm = np.random.default_rng(3).permutation(g0.n).astype(np.uint32)
    graphs = [g0, sge.Graph.from_edges(g0.n, edges, degree_order=False),
              sge.Graph.from_edges(g0.n, perm[edges], degree_order=False)]
    for p in [W.named("house"), W.named("diamond"), W.named("near5clique")] + W.six_vertex_set()[:4]:
        for mode in MODES:
            say("relabel", p.edges(), mode)
            got = {sge.count(g, sge.build_plan(p, g, mode))[0][0] for g in graphs}
            assert len(got) == 1, (p, mode, got)


def test_cli_drop_in(tmp_path):
    """graph_mining_system with the reference's flags prints the reference's
    count lines (C1 goldens) in every mode, writes -r files and canonical
    embeddings, and loads through the GPU CSR builder on request."""
    import subprocess

    cli = os.path.join(os.path.dirname(sge.LIB_PATH), "graph_mining_system")
    g = W.graph_c1()
    dpath = str(tmp_path / "c1.txt")
    g.write(dpath)
    gold = golden("c1")["patterns"]
    for name in ["diamond", "house"]:
        qpath = tmp_path / f"{name}.txt"
        e = gold[name]["edges"]
        qpath.write_text(f"{gold[name]['k']} {len(e)}\n" + "".join(f"{a} {b}\n" for a, b in e))
        for flags in ([], ["-i", "-t", "4"], ["-u"], ["--device-build"]):
            say("cli", name, flags)
            r = subprocess.run([cli, "-q", str(qpath), "-d", dpath] + flags, capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stderr
            lines = [l for l in r.stdout.splitlines() if l.startswith("Number of matches:")]
            assert lines and int(lines[0].split(":")[1]) == gold[name]["wcoj"]["count"], (name, flags, r.stdout)
    # --bag: the same count through a materialised bag, reported in the plan text
    ph = sge.Pattern.from_edges(gold["house"]["k"], [tuple(e) for e in gold["house"]["edges"]])
    for mode_flags, mode in (([], sge.MODE_WCOJ), (["-u"], sge.MODE_IEP)):
        plan = sge.build_plan(ph, g, mode)
        bags = sge.valid_bags(plan)
        assert bags
        for bag in bags[:2]:
            verts = [v for v in range(ph.k) if (bag >> v) & 1]
            args = mode_flags + ["--order", ",".join(str(v) for v in list(plan.order)[:plan.k]),
                                 "--bag", ",".join(map(str, verts))]
            say("cli bag", args)
            r = subprocess.run([cli, "-q", str(tmp_path / "house.txt"), "-d", dpath] + args, capture_output=True,
                               text=True, timeout=120)
            assert r.returncode == 0, r.stderr
            assert "materialised bag:" in r.stdout
            assert "Number of matches: %d" % gold["house"]["wcoj"]["count"] in r.stdout, (args, r.stdout)
    # two --bag flags: two bags, each through its own trie
    plan = sge.build_plan(ph, g, sge.MODE_WCOJ)
    for a, b in sge.valid_bag_pairs(plan)[:2]:
        args = ["--order", ",".join(str(v) for v in list(plan.order)[:plan.k])]
        for m in (a, b):
            args += ["--bag", ",".join(str(v) for v in range(ph.k) if (m >> v) & 1)]
        say("cli two bags", args)
        r = subprocess.run([cli, "-q", str(tmp_path / "house.txt"), "-d", dpath] + args, capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert r.stdout.count("materialised bag:") == 2
        assert "Number of matches: %d" % gold["house"]["wcoj"]["count"] in r.stdout, (args, r.stdout)
    # --bag auto: the first decomposition bag valid for the cost model's order
    for name in ["diamond", "house"]:
        say("cli bag auto", name)
        r = subprocess.run([cli, "-q", str(tmp_path / f"{name}.txt"), "-d", dpath, "--bag", "auto"],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert "materialised bag:" in r.stdout
        assert "Number of matches: %d" % gold[name]["wcoj"]["count"] in r.stdout, (name, r.stdout)
    rpath, epath = str(tmp_path / "out.txt"), str(tmp_path / "emb.txt")
    r = subprocess.run([cli, "-q", str(tmp_path / "diamond.txt"), "-d", dpath, "-r", rpath, "--embeddings", epath],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Number of matches: %d" % gold["diamond"]["wcoj"]["count"] in open(rpath).read()
    lines = [l for l in open(epath).read().splitlines() if l.strip()]
    rows = [list(map(int, l.split())) for l in lines[1:]]
    assert int(lines[0]) == len(rows) == gold["diamond"]["wcoj"]["count"] and rows == sorted(rows)


def test_seven_vertex_set_rmat8_batched_goldens():
    """All 78 seven-vertex patterns over R-MAT scale 8 (the reference's goldens),
    IEP mode, batched and plain: every count equals the reference's WCOJ count
    and its IEP count, except where the reference's IEP mode overcounts
    (documented in test_oracle)."""
    from test_oracle import REF_IEP_OVERCOUNTS

    wcoj = golden("rmat7")["patterns"]
    iep = golden("rmat7_iep")["patterns"]
    g = W.graph_rmat(8, 7)
    plans = [sge.build_plan(p, g, sge.MODE_IEP) for p in W.seven_vertex_set()]
    say("rmat8 78 batched")
    b = sge.Batch(g, plans)
    got = b.run()[0]
    b.close()
    assert got == [r["wcoj"]["count"] for r in wcoj]
    assert sge.count(g, plans)[0] == got
    for i, (c, r) in enumerate(zip(got, iep)):
        if i not in REF_IEP_OVERCOUNTS:
            assert c == r["iep"]["count"], i


# ---------------------------------------------------------------- materialised bags
def spread(bags, cap):
    return bags[::max(1, len(bags) // cap)][:cap]


def smallest_bags(plan, cap):
    return sorted(sge.valid_bags(plan), key=lambda b: (bin(b).count("1"), b))[:cap]


def test_bag_trie_matches_oracle():
    """Device trie (enumerate, radix sort, per-depth scans) = the oracle's
    recursive grouping of the sorted relation, array for array."""
    for gname, g in (("c1", W.graph_c1()), ("random60", random_graph(60, 0.2, 7))):
        off, nbr = g.csr()
        checked = 0
        for p in small_patterns() + W.six_vertex_set()[::3] + W.seven_vertex_set()[::13]:
            if p.k < 4:
                continue
            for mode in (sge.MODE_WCOJ, sge.MODE_IEP):
                plan = sge.build_plan(p, g, mode)
                for b in spread(sge.valid_bags(plan), 3):
                    q = sge.set_bag(plan, b)
                    say("trie", gname, p.edges(), "mode", mode, "bag", b)
                    t = sge.Trie(g, q)
                    vb, val, child = O.bag_trie(q, off, nbr)
                    values, ch = t.arrays()
                    assert t.vb == vb and t.tuples == vb[-1] - vb[-2]
                    assert np.array_equal(values, val) and np.array_equal(ch, child)
                    t.close()
                    checked += 1
        assert checked >= 20


def test_bag_counts_match_oracle_c1():
    g = W.graph_c1()
    off, nbr = g.csr()
    for p in W.six_vertex_set() + W.seven_vertex_set()[::5] + small_patterns():
        for mode in (sge.MODE_WCOJ, sge.MODE_EXPLORE, sge.MODE_IEP):
            plan = sge.build_plan(p, g, mode)
            want = O.count(plan, off, nbr)
            for b in spread(sge.valid_bags(plan), 4):
                say("bag count c1", p.edges(), "mode", mode, "bag", b)
                got, st = sge.count(g, sge.set_bag(plan, b))
                assert got[0] == want, (p.edges(), mode, b)
                assert st.trie_ms > 0


def test_bag_c2_goldens_plain_and_batched():
    """The 16 six-vertex patterns over the Chung-Lu graph n=7,115 with a
    three-vertex bag each (IEP mode), plain and batched, equal the reference's
    counts; so do house and K5-e in WCOJ and IEP mode with every 3-vertex bag."""
    g = W.graph_c2()
    gold3 = golden("c3")["patterns"]
    pats = W.six_vertex_set()
    plans = []
    for rec in gold3:
        plan = sge.build_plan(pats[rec["index"]], g, sge.MODE_IEP)
        bags = smallest_bags(plan, 1)
        plans.append(sge.set_bag(plan, bags[0]) if bags else plan)
    assert sum(1 for pl in plans if pl.bag) >= 8
    say("c2 six-vertex with bags, plain")
    plain, st = sge.count(g, plans)
    assert plain == [r["iep"]["count"] for r in gold3]
    say("c2 six-vertex with bags, batched")
    b = sge.Batch(g, plans)
    for rep in range(2):
        got, st = b.run()
        assert got == plain
        assert st.trie_ms > 0
    b.close()
    gold2 = golden("c2")["patterns"]
    for name in ["near5clique", "house"]:
        for mode, key in ((sge.MODE_WCOJ, "wcoj"), (sge.MODE_EXPLORE, "explore"), (sge.MODE_IEP, "iep")):
            plan = sge.build_plan(W.named(name), g, mode)
            for bag in [b for b in sge.valid_bags(plan) if bin(b).count("1") == 3]:
                say("c2", name, key, "bag", bag)
                assert sge.count(g, sge.set_bag(plan, bag))[0][0] == gold2[name][key]["count"]


def test_bag_with_donation_and_tuning_variants():
    """Continuation items rebuild their path's trie positions (and, in
    exploration mode, the early parent intersections of the non-bag levels):
    tiny donation budgets and unusual launch shapes give the same counts."""
    g = W.graph_c1()
    plans = []
    for p in W.six_vertex_set()[::2] + W.seven_vertex_set()[::11]:
        for mode in (sge.MODE_WCOJ, sge.MODE_EXPLORE):
            plan = sge.build_plan(p, g, mode)
            bags = spread(sge.valid_bags(plan), 2)
            plans += [sge.set_bag(plan, b) for b in bags]
    want, _ = sge.count(g, [sge.set_bag(pl, 0) for pl in plans])
    for bt, ch, budget, span in [(64, 1, 16, 8), (256, 3, 64, 1 << 30), (128, 8, 1, 4)]:
        t = sge.default_tuning()
        t.block_threads, t.chunk, t.donate_budget, t.split_span = bt, ch, budget, span
        say("bag tuning", bt, ch, budget, span)
        assert sge.count(g, plans, batched=True, tuning=t)[0] == want
        assert sge.count(g, plans, batched=False, tuning=t)[0] == want


def test_bag_rmat8_seven_vertex_goldens():
    gold = golden("rmat7")["patterns"]
    g = W.graph_rmat(8, 7)
    pats = W.seven_vertex_set()
    for rec in gold[::4]:
        for mode in (sge.MODE_IEP, sge.MODE_EXPLORE):
            plan = sge.build_plan(pats[rec["index"]], g, mode)
            for b in spread(smallest_bags(plan, 6), 2):
                say("rmat8 bag", rec["index"], mode, b)
                assert sge.count(g, sge.set_bag(plan, b))[0][0] == rec["wcoj"]["count"], (rec["index"], mode, b)


def test_bag_enumeration_unchanged():
    g = W.graph_c1()
    plan = sge.build_plan(W.named("house"), g, sge.MODE_WCOJ, [2, 3, 4, 0, 1])
    house = W.named("house")
    a = sge.enumerate_embeddings(g, plan, house, 100000)
    b = sge.enumerate_embeddings(g, sge.set_bag(plan, [2, 3, 4]), house, 100000)
    assert len(a) == sge.count(g, plan)[0][0] and np.array_equal(a, b)


def test_bag_empty_relation():
    """A triangle bag over a triangle-free graph (a 4 x 4 grid): the relation and
    the trie are empty, every count is 0 and equals the oracle's."""
    edges = [(r * 4 + c, r * 4 + c + 1) for r in range(4) for c in range(3)] + \
            [(r * 4 + c, (r + 1) * 4 + c) for r in range(3) for c in range(4)]
    g = sge.Graph.from_edges(16, edges)
    off, nbr = g.csr()
    house = W.named("house")
    for mode in (sge.MODE_WCOJ, sge.MODE_EXPLORE, sge.MODE_IEP):
        plan = sge.set_bag(sge.build_plan(house, g, mode, [2, 3, 4, 0, 1]), [2, 3, 4])
        say("empty bag relation, mode", mode)
        t = sge.Trie(g, plan)
        assert t.tuples == 0 and t.nodes == [0, 0, 0]
        values, child = t.arrays()
        assert len(values) == 0 and list(child) == [0, 0]
        t.close()
        assert sge.count(g, plan)[0][0] == 0 == O.count(plan, off, nbr)


def test_bag_rmat8_all_78_batched():
    """One bag per seven-vertex pattern (the smallest valid one), all 78 in one
    batched launch over R-MAT scale 8: the reference's WCOJ counts."""
    gold = golden("rmat7")["patterns"]
    g = W.graph_rmat(8, 7)
    plans = []
    for p in W.seven_vertex_set():
        plan = sge.build_plan(p, g, sge.MODE_IEP)
        bags = smallest_bags(plan, 1)
        plans.append(sge.set_bag(plan, bags[0]) if bags else plan)
    assert sum(1 for pl in plans if pl.bag) >= 60
    say("rmat8 78 with bags, batched")
    b = sge.Batch(g, plans)
    got = b.run()[0]
    b.close()
    assert got == [r["wcoj"]["count"] for r in gold]


def bag_pairs(plan, cap):
    pairs = sge.valid_bag_pairs(plan)
    return pairs[::max(1, len(pairs) // cap)][:cap]


def test_two_bags_match_oracle_and_goldens():
    """Two materialised bags per plan, each joined through its own trie: C1
    counts equal the oracle's in all three modes (plain and batched), C2's
    house and K5-e equal the reference's counts for three pairs each, and the
    R-MAT 8 seven-vertex patterns equal the reference's WCOJ counts."""
    g = W.graph_c1()
    off, nbr = g.csr()
    plans, want = [], []
    for p in W.six_vertex_set()[::2] + W.seven_vertex_set()[::9]:
        for mode in (sge.MODE_WCOJ, sge.MODE_EXPLORE, sge.MODE_IEP):
            plan = sge.build_plan(p, g, mode)
            base = O.count(plan, off, nbr)
            for a, b in bag_pairs(plan, 2):
                plans.append(sge.add_bag(sge.set_bag(plan, a), b))
                want.append(base)
    assert len(plans) >= 20
    say("two bags c1 plain", len(plans))
    got, st = sge.count(g, plans)
    assert got == want and st.trie_ms > 0
    say("two bags c1 batched")
    assert sge.count(g, plans, batched=True)[0] == want
    g2 = W.graph_c2()
    gold2 = golden("c2")["patterns"]
    for name in ["near5clique", "house"]:
        for mode, key in ((sge.MODE_WCOJ, "wcoj"), (sge.MODE_EXPLORE, "explore"), (sge.MODE_IEP, "iep")):
            plan = sge.build_plan(W.named(name), g2, mode)
            for a, b in bag_pairs(plan, 3):
                say("two bags c2", name, key, a, b)
                assert sge.count(g2, sge.add_bag(sge.set_bag(plan, a), b))[0][0] == gold2[name][key]["count"]
    gold = golden("rmat7")["patterns"]
    g8 = W.graph_rmat(8, 7)
    pats = W.seven_vertex_set()
    for rec in gold[::6]:
        plan = sge.build_plan(pats[rec["index"]], g8, sge.MODE_IEP)
        for a, b in bag_pairs(plan, 1):
            say("two bags rmat8", rec["index"], a, b)
            assert sge.count(g8, sge.add_bag(sge.set_bag(plan, a), b))[0][0] == rec["wcoj"]["count"]


def test_two_bags_with_donation_and_tuning_variants():
    """Continuation items and level-2 root splits rebuild the positions in both
    tries: tiny donation budgets and odd launch shapes give the same counts."""
    g = W.graph_c1()
    plans = []
    for p in W.six_vertex_set()[1::3] + W.seven_vertex_set()[::13]:
        for mode in (sge.MODE_WCOJ, sge.MODE_EXPLORE, sge.MODE_IEP):
            plan = sge.build_plan(p, g, mode)
            plans += [sge.add_bag(sge.set_bag(plan, a), b) for a, b in bag_pairs(plan, 2)]
    want, _ = sge.count(g, [sge.set_bag(pl, 0) for pl in plans])
    for bt, ch, budget, span in [(64, 1, 16, 8), (256, 3, 64, 1 << 30), (128, 8, 1, 4)]:
        t = sge.default_tuning()
        t.block_threads, t.chunk, t.donate_budget, t.split_span = bt, ch, budget, span
        say("two bags tuning", bt, ch, budget, span)
        assert sge.count(g, plans, batched=True, tuning=t)[0] == want
        assert sge.count(g, plans, batched=False, tuning=t)[0] == want


def test_global_bitmaps_forced_c2_goldens():
    """F sets too large for the LDS stage go to per-wave global bitmaps when
    their siblings' rows will probe them.  With the size threshold forced low
    the path runs on C2 (by default only hub-sized sets use it): counts equal
    the reference's, plain and batched, and repeated calls find the region
    cleared."""
    g = W.graph_c2()
    gold3 = golden("c3")["patterns"]
    pats = W.six_vertex_set()
    plans = [sge.build_plan(pats[r["index"]], g, sge.MODE_IEP) for r in gold3]
    want = [r["iep"]["count"] for r in gold3]
    t = sge.default_tuning()
    t.global_bitmap_min = 16
    for rep in range(2):
        say("forced global bitmaps, plain, rep", rep)
        assert sge.count(g, plans, tuning=t)[0] == want
    say("forced global bitmaps, batched")
    b = sge.Batch(g, plans, tuning=t)
    for _ in range(2):
        assert b.run()[0] == want
    b.close()
    gold2 = golden("c2")["patterns"]
    for name in ["near5clique", "house"]:
        for mode, key in ((sge.MODE_WCOJ, "wcoj"), (sge.MODE_IEP, "iep")):
            assert sge.count(g, sge.build_plan(W.named(name), g, mode), tuning=t)[0][0] == gold2[name][key]["count"]


def test_query_files_c1_goldens():
    """The reference's 13 query files (dataset/light_query, huge_query) on C1 in
    every mode: the reference's WCOJ counts (its IEP mode overcounts one file,
    documented in test_oracle)."""
    gold = golden("queries")
    g = W.graph_c1()
    for rec in gold["patterns"]:
        p = sge.Pattern.from_edges(rec["k"], [tuple(e) for e in rec["edges"]])
        for mode in MODES:
            say("query file", rec["file"], "mode", mode)
            assert sge.count(g, sge.build_plan(p, g, mode))[0][0] == rec["wcoj"]["count"], (rec["file"], mode)


def test_query_files_c2_goldens():
    """The reference's 13 query files on C2 in every mode, plain and batched:
    the reference's WCOJ counts."""
    gold = golden("queries_c2")
    g = W.graph_c2()
    pats = [sge.Pattern.from_edges(r["k"], [tuple(e) for e in r["edges"]]) for r in gold["patterns"]]
    want = [r["wcoj"]["count"] for r in gold["patterns"]]
    for mode in MODES:
        plans = [sge.build_plan(p, g, mode) for p in pats]
        say("query files c2, mode", mode)
        assert sge.count(g, plans)[0] == want
        assert sge.count(g, plans, batched=True)[0] == want


def test_bag_joins_on_query_files_c1():
    """Every valid bag of every query file's plan (WCOJ and IEP) on C1: the
    reference's WCOJ counts through the device tries."""
    gold = golden("queries")
    g = W.graph_c1()
    for rec in gold["patterns"]:
        p = sge.Pattern.from_edges(rec["k"], [tuple(e) for e in rec["edges"]])
        for mode in (sge.MODE_WCOJ, sge.MODE_EXPLORE, sge.MODE_IEP):
            plan = sge.build_plan(p, g, mode)
            for b in sge.valid_bags(plan):
                say("query file bag", rec["file"], mode, b)
                assert sge.count(g, sge.set_bag(plan, b))[0][0] == rec["wcoj"]["count"], (rec["file"], mode, b)


def test_every_small_class_matches_reference():
    """All 139 connected classes on 4-6 vertices over two graphs in every mode,
    plain and batched: the reference's WCOJ counts, and the oracle's (pinned by
    brute force) on the four classes where the reference binary fails."""
    gold = golden("classes")
    pats = [sge.Pattern.from_edges(r["k"], [tuple(e) for e in r["edges"]]) for r in gold["patterns"]]
    graphs = {"c1": W.graph_c1(), "cl60": sge.Graph.chung_lu(60, 600, 3.0, 0.0, 2, True)}
    for name, g in graphs.items():
        off, nbr = g.csr()
        rows = gold["graphs"][name]["counts"]
        for mode in MODES:
            plans = [sge.build_plan(p, g, mode) for p in pats]
            want = [r["wcoj"] if not r["ref_exit"] else O.count(pl, off, nbr) for r, pl in zip(rows, plans)]
            say("every small class", name, "mode", mode)
            assert sge.count(g, plans)[0] == want
            assert sge.count(g, plans, batched=True)[0] == want


def test_every_reference_pattern_file_on_cliques():
    """The reference's 107 pattern files (94 benchmark patterns, 13 query files)
    on K6, K7 and K8 in every mode, plain and batched: n!/(n-k)! embeddings,
    divided by |Aut| unless the pattern is single-bag."""
    import math

    gold = golden("dataset")
    pats = [sge.Pattern.from_edges(r["k"], [tuple(e) for e in r["edges"]]) for d in gold.values() for r in d]
    assert len(pats) == 107
    for n in (6, 7, 8):
        kn = sge.Graph.from_edges(n, [(i, j) for i in range(n) for j in range(i + 1, n)])
        fit = [p for p in pats if p.k <= n]
        want = []
        for p in fit:
            info = p.info()
            emb = math.perm(n, p.k)
            want.append(emb if info.single_bag else emb // info.aut_size)
        for mode in MODES:
            plans = [sge.build_plan(p, kn, mode) for p in fit]
            say("pattern files on K", n, "mode", mode)
            assert sge.count(kn, plans)[0] == want
            assert sge.count(kn, plans, batched=True)[0] == want


def test_c3_wcoj_and_exploration_match_reference_wcoj():
    """The 16 six-vertex patterns on C2 in WCOJ and exploration mode, plain and
    batched: the reference's WCOJ counts (tests/golden/c3_modes.json)."""
    g = W.graph_c2()
    gold = golden("c3_modes")["patterns"]
    pats = W.six_vertex_set()
    want = [r["wcoj"]["count"] for r in gold]
    for mode in (sge.MODE_WCOJ, sge.MODE_EXPLORE):
        plans = [sge.build_plan(pats[r["index"]], g, mode) for r in gold]
        say("c3 six-vertex reference WCOJ counts, mode", mode)
        assert sge.count(g, plans)[0] == want
        assert sge.count(g, plans, batched=True)[0] == want


def test_rmat8_seven_vertex_wcoj_and_exploration_batched():
    """All 78 seven-vertex patterns over R-MAT scale 8 in WCOJ and exploration
    mode, one batched launch each: the reference's WCOJ counts (its own -i mode
    differs on ten of them, documented in test_oracle)."""
    want = [r["wcoj"]["count"] for r in golden("rmat7")["patterns"]]
    g = W.graph_rmat(8, 7)
    pats = W.seven_vertex_set()
    for mode in (sge.MODE_WCOJ, sge.MODE_EXPLORE):
        plans = [sge.build_plan(p, g, mode) for p in pats]
        say("rmat8 78 batched, mode", mode)
        assert sge.count(g, plans, batched=True)[0] == want


def test_every_seven_vertex_class_matches_reference():
    """All 853 connected 7-vertex classes over C1 (IEP plain and batched, WCOJ
    batched): the reference's WCOJ counts, and the fixture's brute-force counts
    on the 83 classes where the reference binary fails."""
    gold = golden("classes7")
    pats = [sge.Pattern.from_edges(r["k"], [tuple(e) for e in r["edges"]]) for r in gold["patterns"]]
    rows = gold["graphs"]["c1"]["counts"]
    want = [r["brute"] if r["ref_exit"] else r["wcoj"] for r in rows]
    g = W.graph_c1()
    plans = [sge.build_plan(p, g, sge.MODE_IEP) for p in pats]
    say("853 seven-vertex classes, IEP batched")
    assert sge.count(g, plans, batched=True)[0] == want
    say("853 seven-vertex classes, IEP plain")
    assert sge.count(g, plans)[0] == want
    plans = [sge.build_plan(p, g, sge.MODE_WCOJ) for p in pats]
    say("853 seven-vertex classes, WCOJ batched")
    assert sge.count(g, plans, batched=True)[0] == want
