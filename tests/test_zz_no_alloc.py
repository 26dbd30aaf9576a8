"""The allocation-free entry point (sge_count_noalloc) through the C ABI against
the allocating batched count (sge_count with SGE_COUNT_BATCHED): two calls in a
row, the second on a smaller input, each bit-exact; and its size limits."""
import ctypes as C
import sys

import numpy as np
import pytest

import sge
import sge.workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def test_noalloc_matches_batched_count_twice():
    big = W.graph_c2()
    big.to_device()
    pats = W.six_vertex_set()
    plans = [sge.build_plan(p, big, sge.MODE_IEP) for p in pats]
    say("noalloc: C2, 16 six-vertex plans (IEP)")
    want, _ = sge.count(big, plans, batched=True)
    got, st = sge.count_noalloc(big, plans)
    assert got == want
    assert st.tasks > 0
    small = W.graph_c1()
    small.to_device()
    splans = [sge.build_plan(W.named(n), small, m) for n, m in
              (("diamond", sge.MODE_WCOJ), ("house", sge.MODE_EXPLORE), ("near5clique", sge.MODE_IEP))]
    say("noalloc: C1, 3 plans (second, smaller call)")
    want2, _ = sge.count(small, splans, batched=True)
    got2, _ = sge.count_noalloc(small, splans)
    assert got2 == want2
    say("noalloc: C2 again after the smaller call")
    got3, _ = sge.count_noalloc(big, plans[:5])
    assert got3 == want[:5]


def test_noalloc_limits_and_requirements():
    g = W.graph_c1()
    plan = sge.build_plan(W.named("house"), g, sge.MODE_IEP)
    # not on the device yet: refused (uploading would allocate)
    fresh = W.graph_c1()
    with pytest.raises(sge.SgeError) as ex:
        sge.count_noalloc(fresh, [plan])
    assert ex.value.code == -1 and "resident" in str(ex.value)
    g.to_device()
    with pytest.raises(sge.SgeError) as ex:
        sge.count_noalloc(g, [plan] * 257)
    assert ex.value.code == -5
    bag = sge.set_bag(sge.build_plan(W.named("house"), g, sge.MODE_IEP, [2, 3, 4, 0, 1]), [2, 3, 4])
    with pytest.raises(sge.SgeError):
        sge.count_noalloc(g, [bag])
    # a work cap stops it with partial counts; the next call is exact again
    t = sge.default_tuning()
    t.node_cap_k = 1
    c2 = W.graph_c2()
    c2.to_device()
    hplan = sge.build_plan(W.named("house"), c2, sge.MODE_WCOJ)
    with pytest.raises(sge.SgeError) as ex:
        sge.count_noalloc(c2, [hplan], tuning=t)
    assert ex.value.code == sge.ELIMIT
    assert sge.count_noalloc(c2, [hplan])[0] == sge.count(c2, [hplan])[0]
