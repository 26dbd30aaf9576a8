"""Adds brute-force counts to tests/golden/classes7.json for every 7-vertex class
where the reference binary gives no trustworthy count: the classes it crashes
on, and those where its -u count differs from its WCOJ count.

    python tests/golden/brute_classes7.py

Brute force = all injective homomorphisms of the pattern into C1 (the oracle's
plain backtracking, oracle_count_brute), divided by |Aut| unless the pattern is
single-bag, with the pattern relabelled so that every vertex after the first is
adjacent to an earlier one.  About 8 minutes on 8 cores.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "subgraph-enumeration_amd"))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import sge  # noqa: E402
import sge.workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402
from test_oracle import bfs_relabelled  # noqa: E402


def main():
    path = os.path.join(HERE, "classes7.json")
    with open(path) as f:
        fx = json.load(f)
    pats = [sge.Pattern.from_edges(r["k"], [tuple(e) for e in r["edges"]]) for r in fx["patterns"]]
    rows = fx["graphs"]["c1"]["counts"]
    g = W.graph_c1()
    assert (g.n, g.m) == (fx["graphs"]["c1"]["n"], fx["graphs"]["c1"]["m"])
    off, nbr = g.csr()
    t0 = time.time()
    for i, (p, r) in enumerate(zip(pats, rows)):
        if not (r["ref_exit"] or r["iep"] != r["wcoj"]):
            continue
        q = bfs_relabelled(p)
        info = p.info()
        b = O.count_brute(q.k, q.c.adj, off, nbr)
        r["brute"] = b if info.single_bag else b // info.aut_size
        print(i, r["brute"], f"{time.time() - t0:.0f}s", flush=True)
    with open(path, "w") as f:
        json.dump(fx, f, indent=1)


if __name__ == "__main__":
    main()
