"""Host-side logic: pattern analysis, planner invariants, graph I/O, the C ABI."""
import hashlib
import json
import math
import os
import re

import numpy as np
import pytest

import sge
import sge.workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "sge.h")).read()
    names = set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sge_\w+)\s*\(", hdr, re.M))
    assert len(names) > 30
    L = sge.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(sge.EXPORTED_SYMBOLS) <= names


@pytest.mark.parametrize("k,edges,aut", [
    (3, W.TRIANGLE, 6), (4, W.DIAMOND, 4), (5, W.HOUSE, 2), (5, W.NEAR_5_CLIQUE, 12),
    (5, [(0, 1), (0, 2), (0, 3), (0, 4)], 24), (5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)], 10)])
def test_automorphism_group_sizes(k, edges, aut):
    assert sge.Pattern.from_edges(k, edges).info().aut_size == aut


def test_fractional_widths_closed_form():
    cases = [(3, W.TRIANGLE, 1.5), (4, [(0, 1), (1, 2), (2, 3), (3, 0)], 2.0), (4, W.clique(4), 2.0),
             (5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)], 2.0),  # bags {0,1,4},{1,2,3,4}; rho*(C5)=2.5
             (5, W.clique(5), 2.5)]
    for k, e, w in cases:
        info = sge.Pattern.from_edges(k, e).info()
        assert info.fhw == w, (k, e)


def test_single_bag_rule():
    assert sge.Pattern.from_edges(4, W.clique(4)).info().single_bag == 1
    assert sge.Pattern.from_edges(5, W.clique(5)).info().single_bag == 1
    assert W.named("diamond").info().single_bag == 0
    assert W.named("house").info().single_bag == 0


def test_canonical_code_is_isomorphism_invariant():
    rng = np.random.default_rng(0)
    for p in W.six_vertex_set()[:6]:
        c = p.info().canon
        for _ in range(5):
            perm = rng.permutation(p.k)
            q = sge.Pattern.from_edges(p.k, [(int(perm[a]), int(perm[b])) for a, b in p.edges()])
            assert q.info().canon == c


def test_sampled_sets_are_the_reference_benchmark_classes():
    path = os.path.join(GOLDEN, "dataset.json")
    if not os.path.exists(path):
        pytest.skip("dataset.json not generated")
    data = json.load(open(path))
    six = {p.info().canon for p in W.six_vertex_set()}
    seven = {p.info().canon for p in W.seven_vertex_set()}
    assert len(six) == 16 and len(seven) == 78
    assert six == {d["canon"] for d in data["6vertex"]}
    assert seven == {d["canon"] for d in data["7vertex"]}
    train = {p.info().canon for p in W.training_set(50)}
    assert not (train & seven)


def test_bag_counts_match_reference_dataset_runs():
    path = os.path.join(GOLDEN, "rmat7.json")
    if not os.path.exists(path):
        pytest.skip("rmat7.json not generated")
    pats = W.seven_vertex_set()
    for rec in json.load(open(path))["patterns"]:
        if rec["wcoj"].get("bags") is None:
            continue
        info = pats[rec["index"]].info()
        assert info.single_bag == (rec["wcoj"]["bags"] == 1)


def test_plan_masks_are_consistent():
    g = W.graph_c1()
    for p in W.six_vertex_set():
        for mode in (0, 1, 2):
            pl = sge.build_plan(p, g, mode)
            d = pl.as_dict()
            k = d["k"]
            assert sorted(d["order"]) == list(range(k))
            for i in range(k):
                earlier = (1 << i) - 1
                for key in ("parent", "lower", "upper", "check"):
                    assert d[key][i] & ~earlier == 0
                assert (d["parent"][i] | d["lower"][i] | d["upper"][i] | d["check"][i]) == earlier
                if i:
                    assert d["parent"][i]


def test_iep_tail_on_star():
    star = sge.Pattern.from_edges(5, [(0, 1), (0, 2), (0, 3), (0, 4)])
    pl = sge.build_plan(star, None, sge.MODE_IEP)
    assert pl.tail == 3  # four leaves: root edge fixes one, the other three are a choosec group
    assert sge.build_plan(star, None, sge.MODE_WCOJ).tail == 0


def test_graph_text_and_binary_roundtrip(tmp_path):
    g = W.graph_c1(False)
    t, b = str(tmp_path / "g.txt"), str(tmp_path / "g.bin")
    g.write(t)
    g.write(b, binary=True)
    assert open(b, "rb").read(7) == b"BINGRPH"
    for path in (t, b):
        h = sge.Graph.load(path, degree_order=False)
        o1, n1 = g.csr()
        o2, n2 = h.csr()
        assert np.array_equal(o1, o2) and np.array_equal(n1, n2)


def test_degree_order_is_a_relabelling():
    g0 = W.graph_c1(False)
    g1 = W.graph_c1(True)
    perm = g1.perm()
    o0, n0 = g0.csr()
    o1, n1 = g1.csr()
    deg1 = np.diff(o1)
    assert np.all(np.diff(deg1.astype(np.int64)) >= 0)
    e0 = {(min(a, b), max(a, b)) for v in range(g0.n) for a, b in [(v, u) for u in n0[o0[v]:o0[v + 1]]]}
    e1 = {(min(perm[v], perm[u]), max(perm[v], perm[u])) for v in range(g1.n) for u in n1[o1[v]:o1[v + 1]]}
    assert e0 == e1


def test_generators_are_deterministic():
    a, b = W.graph_c2(), W.graph_c2()
    oa, na = a.csr()
    ob, nb = b.csr()
    assert np.array_equal(na, nb)
    assert (a.n, a.m) == (7115, 100762)
    r = W.graph_rmat(12, 3)
    o, n = r.csr()
    h = hashlib.sha256(o.tobytes() + n.tobytes()).hexdigest()
    r2 = W.graph_rmat(12, 3)
    o2, n2 = r2.csr()
    assert h == hashlib.sha256(o2.tobytes() + n2.tobytes()).hexdigest()


def test_bad_inputs_raise(tmp_path):
    bad = tmp_path / "bad.txt"
    bad.write_text("3 5\n0 1\n")
    with pytest.raises(sge.SgeError):
        sge.Graph.load(str(bad))
    with pytest.raises(sge.SgeError):
        sge.Pattern.load(str(tmp_path / "missing.txt"))
    with pytest.raises(sge.SgeError):
        sge.Pattern.from_edges(9, [(0, 1)])
    disc = sge.Pattern.from_edges(4, [(0, 1), (2, 3)])
    with pytest.raises(sge.SgeError):
        sge.build_plan(disc)
    with pytest.raises(sge.SgeError):
        sge.build_plan(W.named("house"), None, 0, [0, 4, 1, 2, 3])  # level 1 not adjacent to level 0


def test_pattern_arguments_are_checked():
    """Every entry point that takes a caller's pattern rejects k outside 1..8
    (and a null pattern) with SGE_EINVAL before reading it; plan cost rejects an
    order that is not a permutation."""
    import ctypes

    L = sge.lib()
    good = W.named("house").c
    for k in (0, 9, 40):
        bad = sge.CPattern()
        bad.k = k
        n = ctypes.c_uint32()
        cost = ctypes.c_double()
        order = np.arange(8, dtype=np.uint32)
        feat = np.zeros(64, dtype=np.float32)
        plan = sge.Plan()
        rows = np.zeros(8, dtype=np.uint32)
        for pat in (ctypes.byref(bad), None):
            assert L.sge_plan_build(pat, None, 0, None, ctypes.byref(plan)) == -1
            assert L.sge_plan_cost(pat, None, 0, order.ctypes.data_as(ctypes.c_void_p), ctypes.byref(cost)) == -1
            assert L.sge_plan_orders(pat, None, 0, ctypes.byref(n)) == -1
            assert L.sge_plan_rank_orders(pat, None, 0, None, None, None, 0, ctypes.byref(n)) == -1
            assert L.sge_plan_bags(pat, ctypes.byref(plan), None, 0, ctypes.byref(n)) == -1
            assert L.sge_canonicalize_embeddings(pat, rows.ctypes.data_as(ctypes.c_void_p), 1) == -1
            assert L.sge_order_features(pat, None, order.ctypes.data_as(ctypes.c_void_p), 0, 0, 0,
                                        feat.ctypes.data_as(ctypes.c_void_p)) == -1
    cost = ctypes.c_double()
    for o in ([0, 1, 2, 3, 3], [0, 1, 2, 3, 5]):
        oa = np.array(o, dtype=np.uint32)
        assert L.sge_plan_cost(ctypes.byref(good), None, 0, oa.ctypes.data_as(ctypes.c_void_p), ctypes.byref(cost)) == -1
    assert L.sge_plan_cost(ctypes.byref(good), None, 0, None, ctypes.byref(cost)) == -1
    oa = np.array([2, 3, 4, 0, 1], dtype=np.uint32)
    assert L.sge_plan_cost(ctypes.byref(good), None, 0, oa.ctypes.data_as(ctypes.c_void_p), ctypes.byref(cost)) == 0
    assert cost.value > 0
    assert L.sge_pattern_load(None, ctypes.byref(sge.CPattern())) == -1


def test_cli_rejects_bad_arguments(tmp_path):
    """Bag vertex ids outside the pattern and an unwritable result file end the
    CLI with an error before anything is counted."""
    import subprocess

    cli = os.path.join(os.path.dirname(sge.LIB_PATH), "graph_mining_system")
    if not os.path.exists(cli):
        pytest.skip("CLI not built")
    W.graph_c1(False).write(str(tmp_path / "g.txt"))
    (tmp_path / "q.txt").write_text("5 6\n0 1\n0 2\n1 3\n2 3\n2 4\n3 4\n")  # house
    base = [cli, "-q", str(tmp_path / "q.txt"), "-d", str(tmp_path / "g.txt"), "--order", "2,3,4,0,1"]
    for bag in ("2,3,33", "2,3,x", "2,,3", "5"):
        r = subprocess.run(base + ["--bag", bag], capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and "--bag" in r.stderr and "Number of matches" not in r.stdout, bag
    for bags in (["2,3,4", "0,2,4", "0,2,3"], ["auto", "2,3,4"]):  # three bags; auto beside another
        r = subprocess.run(base + sum((["--bag", x] for x in bags), []), capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and "--bag" in r.stderr and "Number of matches" not in r.stdout, bags
    r = subprocess.run(base + ["-r", str(tmp_path / "missing" / "out.txt")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "result file" in r.stderr


def test_host_failures_return_codes():
    """A host allocation failure inside an entry point comes back as SGE_ENOMEM
    (no C++ exception reaches the caller), and a failed write as SGE_EIO."""
    import ctypes

    h = ctypes.c_void_p()
    uv = np.zeros(2, dtype=np.uint32)
    # 2^61 edges: the edge vector cannot be sized, before any edge is read
    rc = sge.lib().sge_graph_from_edges(4, uv.ctypes.data_as(ctypes.c_void_p), 1 << 61, 0, ctypes.byref(h))
    assert rc == -4 and not h.value
    assert b"allocation" in sge.lib().sge_last_error()
    # headers are checked against the file before they size anything
    import struct
    import tempfile

    with tempfile.TemporaryDirectory() as d:
        b = os.path.join(d, "short.bin")
        with open(b, "wb") as f:  # claims 2^32 - 1 edges, holds one
            f.write(b"BINGRPH" + struct.pack("<II", 4, 0xFFFFFFFF) + struct.pack("<II", 0, 1))
        with pytest.raises(sge.SgeError) as e:
            sge.Graph.load(b)
        assert e.value.code == -2 and "truncated" in str(e.value)
        t = os.path.join(d, "big_n.txt")
        with open(t, "w") as f:
            f.write("4294967296 1\n0 1\n")
        with pytest.raises(sge.SgeError) as e:
            sge.Graph.load(t)
        assert e.value.code == -2
    # a Chung-Lu request the draw limit cannot meet is an error, not a smaller
    # graph: the complete graph on 100 vertices under weights ~ (i+1)^-5 needs
    # the pair (98, 99), drawn with probability ~1e-20 per draw
    rc = sge.lib().sge_graph_chung_lu(100, 4950, 1.2, 0.0, 1, 0, ctypes.byref(h))
    assert rc == -5 and not h.value and b"distinct edges" in sge.lib().sge_last_error()
    if os.path.exists("/dev/full"):  # every write fails with ENOSPC
        g = W.graph_c1(False)
        for binary in (False, True):
            with pytest.raises(sge.SgeError) as e:
                g.write("/dev/full", binary=binary)
            assert e.value.code == -2


def test_count_without_gpu_fails_loudly():
    import ctypes
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("GPU present")
    g = W.graph_c1()
    with pytest.raises(sge.SgeError):
        sge.count(g, sge.build_plan(W.named("diamond"), g))
    with pytest.raises(sge.SgeError):
        sge.Graph.from_edges(4, [(0, 1), (1, 2)], device_build=True)
    # the CLI as well: an error and a non-zero exit, never a count
    import subprocess

    cli = os.path.join(os.path.dirname(sge.LIB_PATH), "graph_mining_system")
    if os.path.exists(cli):
        import tempfile

        with tempfile.TemporaryDirectory() as d:
            g.write(os.path.join(d, "g.txt"))
            with open(os.path.join(d, "q.txt"), "w") as f:
                f.write("4 5\n0 1\n0 2\n1 2\n1 3\n2 3\n")
            r = subprocess.run([cli, "-q", os.path.join(d, "q.txt"), "-d", os.path.join(d, "g.txt")],
                               capture_output=True, text=True, timeout=60)
        assert r.returncode != 0 and "error" in r.stderr and "Number of matches" not in r.stdout
    bag_plan = sge.set_bag(sge.build_plan(W.named("house"), g, sge.MODE_WCOJ, [2, 3, 4, 0, 1]), [2, 3, 4])
    with pytest.raises(sge.SgeError):
        sge.Trie(g, bag_plan)
    with pytest.raises(sge.SgeError):
        sge.count(g, bag_plan)


def test_bag_validation():
    g = W.graph_c1()
    house = W.named("house")  # square 0-1-3-2, roof 4 on 2 and 3
    plan = sge.build_plan(house, g, sge.MODE_WCOJ, [2, 3, 4, 0, 1])  # vertex 1 is the last level
    q = sge.set_bag(plan, [2, 3, 4])              # the roof triangle: levels 0, 1, 2
    assert q.bag == 0b111 and plan.bag == 0
    assert sge.set_bag(q, 0).bag == 0            # cleared
    assert sge.set_bag(plan, [2, 4, 0]).bag == 0b1101  # levels 0, 2, 3: 2-4 adjacent, 0 adjacent to 2
    for bad in ([2, 3],          # fewer than 3 vertices
                [2, 3, 1],       # vertex 1 is the closed-form level
                [3, 4, 0]):      # vertex 0 has no neighbour among 3 and 4
        with pytest.raises(sge.SgeError):
            sge.set_bag(plan, bad)
    with pytest.raises(sge.SgeError):  # first two bag levels (vertices 1 and 2) not adjacent
        sge.set_bag(sge.build_plan(house, g, sge.MODE_WCOJ, [0, 1, 2, 3, 4]), [1, 2, 3])
    # exploration mode takes bags too
    assert sge.set_bag(sge.build_plan(house, g, sge.MODE_EXPLORE, [2, 3, 4, 0, 1]), [2, 3, 4]).bag == 0b111
    # a second bag: its own level mask in bits 8-15, checked like the first
    q2 = sge.add_bag(q, [2, 4, 0])
    assert q2.bag == 0b111 | (0b1101 << 8) and q.bag == 0b111
    assert sge.set_bag(q2, [2, 3, 4]).bag == 0b111  # set_bag replaces both
    for first, second in ((plan, [2, 4, 0]),        # no first bag
                          (q, [2, 3, 4]),          # the same bag twice
                          (q2, [2, 3, 0]),         # a third bag
                          (q, [3, 4, 0]),          # an invalid second bag
                          (q, [2, 3, 1])):         # reaches the closed-form level
        with pytest.raises(sge.SgeError):
            sge.add_bag(first, second)
    pairs = sge.valid_bag_pairs(plan)
    assert ((1 << 0) | (1 << 2) | (1 << 4), (1 << 2) | (1 << 3) | (1 << 4)) in pairs
    for a, b in pairs:
        assert sge.add_bag(sge.set_bag(plan, a), b).bag >> 8
    # every bag the decomposition proposes is accepted
    for p in W.six_vertex_set():
        pl = sge.build_plan(p, g, sge.MODE_IEP)
        for b in sge.plan_bags(p, pl):
            assert b in sge.valid_bags(pl)
            assert sge.set_bag(pl, b).bag


def test_ctypes_mirrors_match_the_c_abi(tmp_path):
    """Every ctypes mirror in the Python binding has the size and field offsets
    of its struct in include/sge.h (compiled with the host C compiler)."""
    import ctypes
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    mirrors = {"sge_graph_stats": sge.GraphStats, "sge_pattern": sge.CPattern, "sge_pattern_info": sge.PatternInfo,
               "sge_plan": sge.Plan, "sge_tuning": sge.Tuning, "sge_stats": sge.Stats}
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "sge.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'    printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "abi.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "abi"
    subprocess.run([cc, "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus N > 1 ends with a message and a non-zero
    status before any graph is built."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def _complete_graph(n):
    return sge.Graph.from_edges(n, [(i, j) for i in range(n) for j in range(i + 1, n)])


def _asymmetric(k, lo, hi):
    return [p for p in sge.sample_patterns(k, lo, hi, 400, 7) if p.info().aut_size == 1][:2]


def _ref_sampled_card(N, r):
    """cardEstimateUnlabeled (plan/estimator.cpp:562-661) on K_N for an r-vertex
    prefix without symmetry bounds, in closed form.  Every node at depth d has
    N - d - 1 children, so sampling does not matter, only the scaling: all N
    roots are taken (N < SAMPLE_SIZE); depth d >= 1 keeps k_d = max(1, (N-d)//4)
    of its N - d candidates (SAMPLE_PORTION 4; the SAMPLE_SIZE budget never binds
    at this size); the last level is counted, not sampled.  A finished subtree
    is folded into its parent's estimate scaled by the PARENT level's ratio
    (N-d)/k_d (ratio 1 at the root), so the deepest sampled level's own ratio is
    never applied: est[j] = N * prod_{d=1}^{j-1} (N-d) * k_j for the prefix of
    j+1 vertices below the last, and N * prod_{d=1}^{r-3} (N-d) * k_{r-2} * (N-r+1)
    for the whole prefix; the estimate is the largest of them."""
    k = [0] + [max(1, (N - d) // 4) for d in range(1, r)]
    est = [float(N)]
    for j in range(1, r - 1):
        est.append(N * float(math.prod(N - d for d in range(1, j))) * k[j])
    est.append(N * float(math.prod(N - d for d in range(1, r - 2))) * k[r - 2] * (N - r + 1))
    return max(est)


def test_sampled_cost_model_closed_form_on_complete_graph():
    """The reference's sampled cost model (computeOrderCost over cardEstimateUnlabeled
    cardinalities) in closed form on K_N for asymmetric patterns (no symmetry
    bounds): card(1) = N, card(2) = N(N-1) directed edges, card(r >= 3) as
    _ref_sampled_card derives it (the reference's estimator scales a subtree by
    its parent level's sampling ratio, so the deepest sampled level is left
    unscaled; with exact scaling card(r) would be N!/(N-r)!).  listSize is the
    average degree N-1 at level 1, and at level i >= 2 the average degree times
    the backward neighbours not covered by the widest earlier level (>= 2)
    whose backward-neighbour list is a prefix of level i's (orderedSubset)."""
    N = 14
    assert 40000 // N >= 4 * 3 * 3 * 2 * 2 * 2  # the sample budget never binds below depth 6
    g = _complete_graph(N)
    avg = N - 1
    for p in _asymmetric(6, 6, 9) + _asymmetric(7, 7, 12):
        adj = {(a, b) for a, b in p.edges()} | {(b, a) for a, b in p.edges()}
        for order in p.orders()[:8]:
            k = len(order)
            par = [[j for j in range(i) if (order[i], order[j]) in adj] for i in range(k)]
            want = 0.0
            for i in range(1, k):
                card = float(N) if i == 1 else (float(N * (N - 1)) if i == 2 else _ref_sampled_card(N, i))
                if i == 1:
                    ls = avg
                else:
                    cover = []
                    for j in range(2, i):
                        if len(par[j]) > max(1, len(cover)) and par[i][:len(par[j])] == par[j]:
                            cover = par[j]
                    ls = avg * (len(par[i]) - len(cover))
                want += card * ls
            got = sge.order_cost_sampled(p, g, order, sge.MODE_WCOJ)
            assert got == pytest.approx(want, rel=1e-9), (p.edges(), order)


def test_sampled_cost_model_is_deterministic_and_bounds_halve_the_pair():
    """A seed gives one cost (the reference draws from std::random_device); with a
    symmetry bound between the first two levels the two-vertex card is m, not 2m,
    and level 1's list is half the average degree."""
    g = W.graph_c1()
    p = W.named("house")
    for order in p.orders()[:4]:
        a = sge.order_cost_sampled(p, g, order, sge.MODE_IEP, seed=3)
        assert a == sge.order_cost_sampled(p, g, order, sge.MODE_IEP, seed=3)
    tri = W.named("triangle")  # every pair is bounded: cost = n*avg/2 + m*(2 uncovered parents)*(avg/2)
    n, nnz = g.n, 2 * g.m
    avg = nnz / n
    got = sge.order_cost_sampled(tri, g, [0, 1, 2], sge.MODE_WCOJ)
    assert got == pytest.approx(n * avg / 2 + (nnz / 2) * (2 * avg / 2), rel=1e-12)
