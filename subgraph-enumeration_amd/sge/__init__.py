"""Python host for the MI355X subgraph-enumeration executor (libsge.so).

Mirrors the reference's plan-to-join interface (main.cpp:12-112 drives
optCostPlan -> sharedJoin / parSharedJoin): load a pattern and a data graph,
build a plan (cost-model order, symmetry clamps, WCOJ / exploration / IEP mode),
count on the GPU.  All compute goes through the C ABI of include/sge.h; there is
no CPU fallback -- if libsge.so is missing or no GPU is present the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

MAX_K = 8
MODE_WCOJ, MODE_EXPLORE, MODE_IEP = 0, 1, 2
COUNT_BATCHED = 1
ELIMIT = -6
GRAPH_DEGREE_ORDER = 1
GRAPH_BUILD_DEVICE = 2

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SGE_LIB") or os.path.join(os.path.dirname(_HERE), "libsge.so")
CLI_PATH = os.path.join(os.path.dirname(_HERE), "graph_mining_system")


class SgeError(RuntimeError):
    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


class GraphStats(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint64), ("max_degree", C.c_uint32),
                ("degree_ordered", C.c_uint32), ("csr_bytes", C.c_uint64), ("triangles_hint", C.c_uint64),
                ("build_ms", C.c_double)]


class CPattern(C.Structure):
    _fields_ = [("k", C.c_uint32), ("adj", C.c_uint32 * MAX_K)]


class PatternInfo(C.Structure):
    _fields_ = [("k", C.c_uint32), ("m", C.c_uint32), ("aut_size", C.c_uint64), ("canon", C.c_uint64),
                ("orbit", C.c_uint32 * MAX_K), ("core_size", C.c_uint32), ("num_bags", C.c_uint32),
                ("single_bag", C.c_uint32), ("fhw", C.c_double), ("connected", C.c_uint32)]


class Plan(C.Structure):
    _fields_ = [("k", C.c_uint32), ("mode", C.c_uint32), ("order", C.c_uint32 * MAX_K),
                ("parent", C.c_uint32 * MAX_K), ("lower", C.c_uint32 * MAX_K), ("upper", C.c_uint32 * MAX_K),
                ("check", C.c_uint32 * MAX_K), ("reuse", C.c_int32 * MAX_K), ("ready", C.c_uint32 * MAX_K),
                ("tail", C.c_uint32), ("tail_kind", C.c_uint32), ("bag", C.c_uint32), ("num_bags", C.c_uint32),
                ("multiplier", C.c_uint64),
                ("aut_size", C.c_uint64), ("cost", C.c_double)]

    def as_dict(self):
        k = self.k
        return {"k": k, "mode": self.mode, "order": list(self.order)[:k], "parent": list(self.parent)[:k],
                "lower": list(self.lower)[:k], "upper": list(self.upper)[:k], "check": list(self.check)[:k],
                "reuse": list(self.reuse)[:k], "ready": list(self.ready)[:k], "tail": self.tail,
                "tail_kind": self.tail_kind, "bag": self.bag,
                "num_bags": self.num_bags, "multiplier": self.multiplier, "aut_size": self.aut_size,
                "cost": self.cost}


class Tuning(C.Structure):
    _fields_ = [("block_threads", C.c_uint32), ("chunk", C.c_uint32), ("stage_max", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("split_span", C.c_uint32), ("donate_budget", C.c_uint32),
                ("bitmap_ratio", C.c_uint32), ("node_cap_k", C.c_uint32),
                ("head_items", C.c_uint32), ("global_bitmap_min", C.c_uint32),
                ("donate_split", C.c_uint32), ("pass_items", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("intersections", C.c_uint64), ("bytes", C.c_uint64), ("tasks", C.c_uint64),
                ("nodes", C.c_uint64), ("kernel_ms", C.c_double), ("wall_ms", C.c_double),
                ("wave_cycles", C.c_uint64), ("trie_ms", C.c_double), ("model_bytes", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None

_SIGS = {
    "sge_last_error": (C.c_char_p, []),
    "sge_version": (C.c_int, []),
    "sge_graph_load": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(C.c_void_p)]),
    "sge_graph_from_edges": (C.c_int, [C.c_uint32, C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p)]),
    "sge_graph_from_csr": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint32,
                                     C.POINTER(C.c_void_p)]),
    "sge_graph_chung_lu": (C.c_int, [C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_uint64, C.c_uint32,
                                     C.POINTER(C.c_void_p)]),
    "sge_graph_rmat": (C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                 C.c_uint32, C.POINTER(C.c_void_p)]),
    "sge_graph_stats_get": (C.c_int, [C.c_void_p, C.POINTER(GraphStats)]),
    "sge_graph_csr": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "sge_graph_perm": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sge_graph_write": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "sge_graph_to_device": (C.c_int, [C.c_void_p]),
    "sge_graph_destroy": (None, [C.c_void_p]),
    "sge_pattern_load": (C.c_int, [C.c_char_p, C.POINTER(CPattern)]),
    "sge_pattern_from_edges": (C.c_int, [C.c_uint32, C.c_void_p, C.c_uint32, C.POINTER(CPattern)]),
    "sge_pattern_analyze": (C.c_int, [C.POINTER(CPattern), C.POINTER(PatternInfo)]),
    "sge_pattern_sample": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_void_p,
                                     C.c_uint32, C.POINTER(CPattern), C.POINTER(C.c_uint32)]),
    "sge_canonicalize_embeddings": (C.c_int, [C.POINTER(CPattern), C.c_void_p, C.c_uint64]),
    "sge_plan_build": (C.c_int, [C.POINTER(CPattern), C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(Plan)]),
    "sge_plan_cost": (C.c_int, [C.POINTER(CPattern), C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(C.c_double)]),
    "sge_plan_cost_sampled": (C.c_int, [C.POINTER(CPattern), C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64,
                                        C.POINTER(C.c_double)]),
    "sge_plan_orders": (C.c_int, [C.POINTER(CPattern), C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "sge_plan_rank_orders": (C.c_int, [C.POINTER(CPattern), C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "sge_plan_text": (C.c_int, [C.POINTER(Plan), C.c_char_p, C.c_size_t]),
    "sge_plan_set_bag": (C.c_int, [C.POINTER(Plan), C.c_uint32]),
    "sge_plan_add_bag": (C.c_int, [C.POINTER(Plan), C.c_uint32]),
    "sge_plan_bags": (C.c_int, [C.POINTER(CPattern), C.POINTER(Plan), C.c_void_p, C.c_uint32,
                                C.POINTER(C.c_uint32)]),
    "sge_trie_build": (C.c_int, [C.c_void_p, C.POINTER(Plan), C.POINTER(C.c_void_p)]),
    "sge_trie_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.c_void_p,
                                 C.POINTER(C.c_double)]),
    "sge_trie_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "sge_trie_destroy": (None, [C.c_void_p]),
    "sge_tuning_default": (None, [C.POINTER(Tuning)]),
    "sge_count": (C.c_int, [C.c_void_p, C.POINTER(Plan), C.c_uint32, C.c_uint32, C.POINTER(Tuning), C.c_void_p,
                            C.POINTER(Stats), C.POINTER(Stats)]),
    "sge_count_range": (C.c_int, [C.c_void_p, C.POINTER(Plan), C.c_uint64, C.c_uint64, C.POINTER(Tuning),
                                  C.POINTER(C.c_uint64), C.POINTER(Stats)]),
    "sge_count_ranges": (C.c_int, [C.c_void_p, C.POINTER(Plan), C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
                                   C.POINTER(Tuning), C.c_void_p, C.POINTER(Stats), C.POINTER(Stats)]),
    "sge_count_noalloc": (C.c_int, [C.c_void_p, C.POINTER(Plan), C.c_uint32, C.POINTER(Tuning), C.c_void_p,
                                    C.POINTER(Stats)]),
    "sge_num_tasks": (C.c_int, [C.c_void_p, C.POINTER(Plan), C.POINTER(C.c_uint64)]),
    "sge_count_sampled": (C.c_int, [C.c_void_p, C.POINTER(Plan), C.c_uint32, C.c_uint64, C.c_uint64,
                                    C.POINTER(Tuning), C.c_void_p, C.POINTER(Stats)]),
    "sge_order_features": (C.c_int, [C.POINTER(CPattern), C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                     C.c_uint64, C.c_void_p]),
    "sge_order_features_range": (C.c_int, [C.POINTER(CPattern), C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                           C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p]),
    "sge_mlp_num_params": (C.c_int, []),
    "sge_mlp_input_dim": (C.c_int, []),
    "sge_mlp_create": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "sge_mlp_set_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "sge_mlp_get_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "sge_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "sge_mlp_gradients": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                    C.POINTER(C.c_float)]),
    "sge_mlp_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float,
                                     C.POINTER(C.c_float)]),
    "sge_mlp_destroy": (None, [C.c_void_p]),
    "sge_enumerate": (C.c_int, [C.c_void_p, C.POINTER(Plan), C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "sge_batch_create": (C.c_int, [C.c_void_p, C.POINTER(Plan), C.c_uint32, C.POINTER(Tuning),
                                   C.POINTER(C.c_void_p)]),
    "sge_batch_create_sampled": (C.c_int, [C.c_void_p, C.POINTER(Plan), C.c_uint32, C.c_uint64, C.c_uint64,
                                           C.POINTER(Tuning), C.POINTER(C.c_void_p)]),
    "sge_batch_create_ranges": (C.c_int, [C.c_void_p, C.POINTER(Plan), C.c_uint32, C.c_void_p, C.c_void_p,
                                          C.POINTER(Tuning), C.POINTER(C.c_void_p)]),
    "sge_batch_run": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Stats)]),
    "sge_batch_destroy": (None, [C.c_void_p]),
    "sge_device_info": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_char_p, C.c_size_t]),
    "sge_calibration_copy": (C.c_int, [C.c_uint64, C.POINTER(C.c_double)]),
}


def lib():
    """Load libsge.so (built in-tree by `make`); raise if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SgeError(f"libsge.so not built at {LIB_PATH}; run __graft_entry__.build() or make")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        msg = lib().sge_last_error().decode(errors="replace")
        raise SgeError(f"{what} failed ({rc}): {msg}", rc)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _graph_flags(degree_order: bool, device_build: bool) -> int:
    return (GRAPH_DEGREE_ORDER if degree_order else 0) | (GRAPH_BUILD_DEVICE if device_build else 0)


class Graph:
    """Data graph: symmetric CSR (u32), optionally relabelled by ascending degree."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    @classmethod
    def load(cls, path: str, degree_order: bool = True, device_build: bool = False) -> "Graph":
        h = C.c_void_p()
        _check(lib().sge_graph_load(path.encode(), _graph_flags(degree_order, device_build), C.byref(h)),
               "sge_graph_load")
        return cls(h)

    @classmethod
    def from_edges(cls, n: int, edges, degree_order: bool = True, device_build: bool = False) -> "Graph":
        e = np.ascontiguousarray(np.asarray(edges, dtype=np.uint32).reshape(-1, 2))
        h = C.c_void_p()
        _check(lib().sge_graph_from_edges(n, _ptr(e), e.shape[0], _graph_flags(degree_order, device_build),
                                          C.byref(h)), "sge_graph_from_edges")
        return cls(h)

    @classmethod
    def chung_lu(cls, n, m, gamma, cap=0.0, seed=0, degree_order=True, device_build=False) -> "Graph":
        h = C.c_void_p()
        _check(lib().sge_graph_chung_lu(n, m, gamma, cap, seed, _graph_flags(degree_order, device_build),
                                        C.byref(h)), "sge_graph_chung_lu")
        return cls(h)

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=0, degree_order=True,
             device_build=False) -> "Graph":
        h = C.c_void_p()
        _check(lib().sge_graph_rmat(scale, edge_factor, a, b, c, seed, _graph_flags(degree_order, device_build),
                                    C.byref(h)), "sge_graph_rmat")
        return cls(h)

    @property
    def stats(self) -> GraphStats:
        s = GraphStats()
        _check(lib().sge_graph_stats_get(self._h, C.byref(s)), "sge_graph_stats_get")
        return s

    @property
    def n(self) -> int:
        return self.stats.n

    @property
    def m(self) -> int:
        return self.stats.m

    def csr(self):
        s = self.stats
        off = np.zeros(s.n + 1, dtype=np.uint32)
        nbr = np.zeros(max(2 * s.m, 1), dtype=np.uint32)
        _check(lib().sge_graph_csr(self._h, _ptr(off), _ptr(nbr)), "sge_graph_csr")
        return off, nbr[: 2 * s.m]

    def perm(self):
        p = np.zeros(max(self.n, 1), dtype=np.uint32)
        _check(lib().sge_graph_perm(self._h, _ptr(p)), "sge_graph_perm")
        return p[: self.n]

    def write(self, path: str, binary: bool = False):
        _check(lib().sge_graph_write(self._h, path.encode(), 1 if binary else 0), "sge_graph_write")

    def to_device(self):
        _check(lib().sge_graph_to_device(self._h), "sge_graph_to_device")

    def close(self):
        if self._h and self._h.value:
            lib().sge_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pattern:
    def __init__(self, c: CPattern):
        self.c = c

    @classmethod
    def load(cls, path: str) -> "Pattern":
        c = CPattern()
        _check(lib().sge_pattern_load(path.encode(), C.byref(c)), "sge_pattern_load")
        return cls(c)

    @classmethod
    def from_edges(cls, k: int, edges) -> "Pattern":
        e = np.ascontiguousarray(np.asarray(edges, dtype=np.uint32).reshape(-1, 2))
        c = CPattern()
        _check(lib().sge_pattern_from_edges(k, _ptr(e), e.shape[0], C.byref(c)), "sge_pattern_from_edges")
        return cls(c)

    @property
    def k(self) -> int:
        return self.c.k

    def edges(self) -> List[tuple]:
        return [(i, j) for i in range(self.k) for j in range(i + 1, self.k) if (self.c.adj[i] >> j) & 1]

    def info(self) -> PatternInfo:
        i = PatternInfo()
        _check(lib().sge_pattern_analyze(C.byref(self.c), C.byref(i)), "sge_pattern_analyze")
        return i

    def orders(self) -> List[List[int]]:
        n = C.c_uint32()
        _check(lib().sge_plan_orders(C.byref(self.c), None, 0, C.byref(n)), "sge_plan_orders")
        buf = np.zeros(max(n.value * self.k, 1), dtype=np.uint32)
        _check(lib().sge_plan_orders(C.byref(self.c), _ptr(buf), n.value, C.byref(n)), "sge_plan_orders")
        return buf[: n.value * self.k].reshape(-1, self.k).tolist()

    def __repr__(self):
        return f"Pattern(k={self.k}, edges={self.edges()})"


def rank_orders(pattern: Pattern, graph: Optional[Graph], mode: int):
    """All connected orders ranked like the planner: (orders, costs, tails)."""
    n = C.c_uint32()
    _check(lib().sge_plan_rank_orders(C.byref(pattern.c), graph._h if graph else None, mode, None, None, None, 0,
                                      C.byref(n)), "sge_plan_rank_orders")
    k = pattern.k
    o = np.zeros(max(n.value * k, 1), dtype=np.uint32)
    c = np.zeros(max(n.value, 1), dtype=np.float64)
    t = np.zeros(max(n.value, 1), dtype=np.uint32)
    _check(lib().sge_plan_rank_orders(C.byref(pattern.c), graph._h if graph else None, mode, _ptr(o), _ptr(c), _ptr(t),
                                      n.value, C.byref(n)), "sge_plan_rank_orders")
    return o[: n.value * k].reshape(-1, k).tolist(), c[: n.value].tolist(), t[: n.value].tolist()


def sample_patterns(k, emin, emax, want, seed, exclude: Sequence[int] = ()) -> List[Pattern]:
    out = (CPattern * want)()
    got = C.c_uint32()
    ex = np.ascontiguousarray(np.asarray(list(exclude), dtype=np.uint64))
    _check(lib().sge_pattern_sample(k, emin, emax, want, seed, _ptr(ex) if len(ex) else None, len(ex), out,
                                    C.byref(got)), "sge_pattern_sample")
    return [Pattern(out[i]) for i in range(got.value)]


def build_plan(pattern: Pattern, graph: Optional[Graph] = None, mode: int = MODE_WCOJ,
               order: Optional[Sequence[int]] = None) -> Plan:
    p = Plan()
    o = None
    if order is not None:
        o = np.ascontiguousarray(np.asarray(order, dtype=np.uint32))
    _check(lib().sge_plan_build(C.byref(pattern.c), graph._h if graph else None, mode,
                                _ptr(o) if o is not None else None, C.byref(p)), "sge_plan_build")
    return p


def order_cost(pattern: Pattern, graph: Optional[Graph], order: Sequence[int], mode: int = MODE_WCOJ) -> float:
    o = np.ascontiguousarray(np.asarray(order, dtype=np.uint32))
    c = C.c_double()
    _check(lib().sge_plan_cost(C.byref(pattern.c), graph._h if graph else None, mode, _ptr(o), C.byref(c)),
           "sge_plan_cost")
    return c.value


def order_cost_sampled(pattern: Pattern, graph: Graph, order: Sequence[int], mode: int = MODE_WCOJ,
                       seed: int = 1) -> float:
    """The reference's sampled cost model (computeOrderCost over sampled
    cardinalities) of one order on `graph`, under the order's symmetry bounds."""
    o = np.ascontiguousarray(order, dtype=np.uint32)
    c = C.c_double()
    _check(lib().sge_plan_cost_sampled(C.byref(pattern.c), graph._h, mode, _ptr(o), seed, C.byref(c)),
           "sge_plan_cost_sampled")
    return c.value


def set_bag(plan: Plan, vertices) -> Plan:
    """Copy of `plan` with the pattern vertices `vertices` (iterable or bit
    mask; 0 / empty clears) as its materialised bag.  Raises SgeError when the
    bag is not valid for the plan's order."""
    mask = vertices if isinstance(vertices, int) else sum(1 << int(v) for v in vertices)
    q = Plan.from_buffer_copy(plan)
    _check(lib().sge_plan_set_bag(C.byref(q), mask), "sge_plan_set_bag")
    return q


def add_bag(plan: Plan, vertices) -> Plan:
    """Copy of `plan` (which has exactly one bag) with `vertices` as its second
    materialised bag.  The two bags may share levels; each is joined through
    its own trie.  Raises SgeError when the pair is not valid for the plan."""
    mask = vertices if isinstance(vertices, int) else sum(1 << int(v) for v in vertices)
    q = Plan.from_buffer_copy(plan)
    _check(lib().sge_plan_add_bag(C.byref(q), mask), "sge_plan_add_bag")
    return q


def plan_bags(pattern: Pattern, plan: Plan) -> List[int]:
    """Vertex masks of the decomposition's bags that are valid bags of `plan`."""
    out = np.zeros(16, dtype=np.uint32)
    n = C.c_uint32()
    _check(lib().sge_plan_bags(C.byref(pattern.c), C.byref(plan), _ptr(out), 16, C.byref(n)), "sge_plan_bags")
    return [int(x) for x in out[: min(n.value, 16)]]


def valid_bags(plan: Plan) -> List[int]:
    """Every vertex mask that sge_plan_set_bag accepts for `plan`."""
    out = []
    for m in range(1, 1 << plan.k):
        if bin(m).count("1") < 3:
            continue
        q = Plan.from_buffer_copy(plan)
        if lib().sge_plan_set_bag(C.byref(q), m) == 0:
            out.append(m)
    return out


def valid_bag_pairs(plan: Plan) -> List[tuple]:
    """Every (first, second) pair of vertex masks that sge_plan_set_bag then
    sge_plan_add_bag accept for `plan`, first < second."""
    single = valid_bags(plan)
    out = []
    for i, a in enumerate(single):
        q = Plan.from_buffer_copy(plan)
        if lib().sge_plan_set_bag(C.byref(q), a) != 0:
            continue
        for b in single[i + 1:]:
            r = Plan.from_buffer_copy(q)
            if lib().sge_plan_add_bag(C.byref(r), b) == 0:
                out.append((a, b))
    return out


class Trie:
    """Level-order trie of a plan's bag relation, built on the device.
    values[vb[j]:vb[j+1]] are depth j's node values; the children of the node
    at value index a of depth j are values[child[a + j]:child[a + j + 1]]."""

    def __init__(self, graph: Graph, plan: Plan):
        self._h = C.c_void_p()
        _check(lib().sge_trie_build(graph._h, C.byref(plan), C.byref(self._h)), "sge_trie_build")
        depth, tuples, ms = C.c_uint32(), C.c_uint64(), C.c_double()
        nodes = np.zeros(MAX_K, dtype=np.uint64)
        _check(lib().sge_trie_shape(self._h, C.byref(depth), C.byref(tuples), _ptr(nodes), C.byref(ms)),
               "sge_trie_shape")
        self.depth = depth.value
        self.tuples = tuples.value
        self.nodes = [int(x) for x in nodes[: self.depth]]
        self.build_ms = ms.value
        self.vb = [0]
        for x in self.nodes:
            self.vb.append(self.vb[-1] + x)

    def arrays(self):
        """(values, child) as numpy u32 arrays."""
        p = self.depth
        values = np.zeros(max(self.vb[p], 1), dtype=np.uint32)
        child = np.zeros(max(self.vb[p - 1] + p - 1, 1), dtype=np.uint32)
        _check(lib().sge_trie_copy(self._h, _ptr(values), _ptr(child)), "sge_trie_copy")
        return values[: self.vb[p]], child[: self.vb[p - 1] + p - 1]

    def close(self):
        if self._h:
            lib().sge_trie_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plan_text(plan: Plan) -> str:
    buf = C.create_string_buffer(8192)
    _check(lib().sge_plan_text(C.byref(plan), buf, 8192), "sge_plan_text")
    return buf.value.decode()


def default_tuning() -> Tuning:
    t = Tuning()
    lib().sge_tuning_default(C.byref(t))
    return t


def count(graph: Graph, plans, batched: bool = False, tuning: Optional[Tuning] = None, per_plan: bool = False):
    """Exact counts for each plan (list of int), plus Stats (and per-plan Stats)."""
    if isinstance(plans, Plan):
        plans = [plans]
    n = len(plans)
    arr = (Plan * n)(*plans)
    counts = np.zeros(n, dtype=np.uint64)
    st = Stats()
    pp = (Stats * n)() if per_plan else None
    _check(lib().sge_count(graph._h, arr, n, COUNT_BATCHED if batched else 0,
                           C.byref(tuning) if tuning else None, _ptr(counts), C.byref(st), pp), "sge_count")
    res = [int(x) for x in counts]
    if per_plan:
        return res, st, [pp[i] for i in range(n)]
    return res, st


def count_range(graph: Graph, plan: Plan, begin: int, end: int, tuning: Optional[Tuning] = None):
    c = C.c_uint64()
    st = Stats()
    _check(lib().sge_count_range(graph._h, C.byref(plan), begin, end, C.byref(tuning) if tuning else None,
                                 C.byref(c), C.byref(st)), "sge_count_range")
    return int(c.value), st


def count_ranges(graph: Graph, plans, begins, ends, batched: bool = False, tuning: Optional[Tuning] = None):
    """Each plan over its own root-task range [begins[i], ends[i]) (CSR positions);
    plain (one launch per plan) or batched.  Returns counts, Stats, per-plan Stats."""
    if isinstance(plans, Plan):
        plans = [plans]
    n = len(plans)
    arr = (Plan * n)(*plans)
    b = np.ascontiguousarray(begins, dtype=np.uint64)
    e = np.ascontiguousarray(ends, dtype=np.uint64)
    if len(b) != n or len(e) != n:
        raise ValueError("count_ranges: one range per plan")
    counts = np.zeros(n, dtype=np.uint64)
    st = Stats()
    pp = (Stats * n)()
    _check(lib().sge_count_ranges(graph._h, arr, n, _ptr(b), _ptr(e), COUNT_BATCHED if batched else 0,
                                  C.byref(tuning) if tuning else None, _ptr(counts), C.byref(st), pp),
           "sge_count_ranges")
    return [int(x) for x in counts], st, [pp[i] for i in range(n)]


def count_noalloc(graph: Graph, plans, tuning: Optional[Tuning] = None):
    """All plans in one launch with no device allocation: every buffer comes from
    the library's static per-device array (the graph must already be on the
    device).  Returns counts and Stats."""
    if isinstance(plans, Plan):
        plans = [plans]
    n = len(plans)
    arr = (Plan * n)(*plans)
    counts = np.zeros(n, dtype=np.uint64)
    st = Stats()
    _check(lib().sge_count_noalloc(graph._h, arr, n, C.byref(tuning) if tuning else None, _ptr(counts),
                                   C.byref(st)), "sge_count_noalloc")
    return [int(x) for x in counts], st


def count_sampled(graph: Graph, plans, stride: int, offset: int = 0, tuning: Optional[Tuning] = None):
    """Plain executor over the strided root sample; returns counts and per-plan Stats."""
    if isinstance(plans, Plan):
        plans = [plans]
    n = len(plans)
    arr = (Plan * n)(*plans)
    counts = np.zeros(n, dtype=np.uint64)
    pp = (Stats * n)()
    _check(lib().sge_count_sampled(graph._h, arr, n, stride, offset, C.byref(tuning) if tuning else None,
                                   _ptr(counts), pp), "sge_count_sampled")
    return [int(x) for x in counts], [pp[i] for i in range(n)]


def work_steps_bound(st: "Stats") -> int:
    """An upper bound on the wave steps (the unit of the donation budget and the
    node cap) behind a call's Stats."""
    return int(st.nodes + st.intersections + st.bytes // 256 + 1)


def time_capped(graph: Graph, plans, stride: int, node_cap_k: int, tuning: Optional[Tuning] = None):
    """Plain executor over the strided root sample with a work cap per plan (in
    units of 1024 wave steps): per-plan Stats and whether each plan hit the cap
    (its count is then partial and its time a lower bound)."""
    if isinstance(plans, Plan):
        plans = [plans]
    out = []
    for p in plans:
        t = Tuning()
        if tuning is not None:
            C.memmove(C.byref(t), C.byref(tuning), C.sizeof(Tuning))
        else:
            lib().sge_tuning_default(C.byref(t))
        t.node_cap_k = node_cap_k
        arr = (Plan * 1)(p)
        counts = np.zeros(1, dtype=np.uint64)
        pp = (Stats * 1)()
        rc = lib().sge_count_sampled(graph._h, arr, 1, stride, 0, C.byref(t), _ptr(counts), pp)
        if rc not in (0, ELIMIT):
            _check(rc, "sge_count_sampled")
        out.append((pp[0], rc == ELIMIT))
    return out


def range_capped(graph: Graph, plan: Plan, begin: int, end: int, node_cap_k: int):
    """Plain executor over root edges [begin, end) with a work cap (units of
    1024 wave steps): (Stats, whether the cap stopped it)."""
    t = Tuning()
    lib().sge_tuning_default(C.byref(t))
    t.node_cap_k = node_cap_k
    arr = (Plan * 1)(plan)
    b = np.array([begin], dtype=np.uint64)
    e = np.array([end], dtype=np.uint64)
    counts = np.zeros(1, dtype=np.uint64)
    st = Stats()
    pp = (Stats * 1)()
    rc = lib().sge_count_ranges(graph._h, arr, 1, _ptr(b), _ptr(e), 0, C.byref(t), _ptr(counts), C.byref(st), pp)
    if rc not in (0, ELIMIT):
        _check(rc, "sge_count_ranges")
    return pp[0], rc == ELIMIT


def order_features(pattern: Pattern, graph: Optional[Graph], order, mode: int, nsamples: int = 1024,
                   seed: int = 0, root_range=None) -> np.ndarray:
    """root_range = (begin, end): random walks rooted in those CSR positions only."""
    o = np.ascontiguousarray(np.asarray(order, dtype=np.uint32))
    f = np.zeros(64, dtype=np.float32)
    if root_range is None:
        _check(lib().sge_order_features(C.byref(pattern.c), graph._h if graph else None, _ptr(o), mode, nsamples,
                                        seed, _ptr(f)), "sge_order_features")
    else:
        _check(lib().sge_order_features_range(C.byref(pattern.c), graph._h if graph else None, _ptr(o), mode,
                                              nsamples, seed, int(root_range[0]), int(root_range[1]), _ptr(f)),
               "sge_order_features_range")
    return f


class MLP:
    """The learned cost model's network on the GPU (include/sge.h sge_mlp_*)."""

    def __init__(self, seed: int = 0):
        self._h = C.c_void_p()
        _check(lib().sge_mlp_create(seed, C.byref(self._h)), "sge_mlp_create")
        self.nparam = lib().sge_mlp_num_params()
        self.in_dim = lib().sge_mlp_input_dim()

    def params(self) -> np.ndarray:
        p = np.zeros(self.nparam, dtype=np.float32)
        _check(lib().sge_mlp_get_params(self._h, _ptr(p), self.nparam), "sge_mlp_get_params")
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().sge_mlp_set_params(self._h, _ptr(p), self.nparam), "sge_mlp_set_params")

    def forward(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.zeros(X.shape[0], dtype=np.float32)
        _check(lib().sge_mlp_forward(self._h, _ptr(X), X.shape[0], _ptr(y)), "sge_mlp_forward")
        return y

    def gradients(self, X: np.ndarray, t: np.ndarray):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = C.c_float()
        _check(lib().sge_mlp_gradients(self._h, _ptr(X), _ptr(t), X.shape[0], _ptr(g), C.byref(loss)),
               "sge_mlp_gradients")
        return g, loss.value

    def train_step(self, X: np.ndarray, t: np.ndarray, lr: float) -> float:
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        _check(lib().sge_mlp_train_step(self._h, _ptr(X), _ptr(t), X.shape[0], lr, C.byref(loss)),
               "sge_mlp_train_step")
        return loss.value

    def close(self):
        if self._h and self._h.value:
            lib().sge_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def num_tasks(graph: Graph, plan: Plan) -> int:
    n = C.c_uint64()
    _check(lib().sge_num_tasks(graph._h, C.byref(plan), C.byref(n)), "sge_num_tasks")
    return int(n.value)


def enumerate_embeddings(graph: Graph, plan: Plan, pattern: Pattern, cap: int, canonical: bool = True):
    """One row per counted subgraph (images per pattern vertex, original ids)."""
    out = np.zeros(max(cap, 1) * plan.k, dtype=np.uint32)
    n = C.c_uint64()
    _check(lib().sge_enumerate(graph._h, C.byref(plan), _ptr(out), cap, C.byref(n)), "sge_enumerate")
    rows = out[: n.value * plan.k].reshape(-1, plan.k).copy()
    if canonical:
        canonicalize(pattern, rows)
    return rows


def canonicalize(pattern: Pattern, rows: np.ndarray) -> np.ndarray:
    assert rows.dtype == np.uint32 and rows.flags.c_contiguous
    _check(lib().sge_canonicalize_embeddings(C.byref(pattern.c), _ptr(rows), rows.shape[0]),
           "sge_canonicalize_embeddings")
    return rows


class Batch:
    """Batched executor: all plans in one persistent launch, captured in a hipGraph.

    Root tasks: the strided sample {offset, offset + stride, ...}, or with
    `ranges=(begins, ends)` plan i over its own range [begins[i], ends[i])."""

    def __init__(self, graph: Graph, plans: Sequence[Plan], tuning: Optional[Tuning] = None, stride: int = 1,
                 offset: int = 0, ranges=None):
        self.graph = graph
        self.n = len(plans)
        self._plans = (Plan * self.n)(*plans)
        self._h = C.c_void_p()
        tun = C.byref(tuning) if tuning else None
        if ranges is not None:
            b = np.ascontiguousarray(ranges[0], dtype=np.uint64)
            e = np.ascontiguousarray(ranges[1], dtype=np.uint64)
            if len(b) != self.n or len(e) != self.n:
                raise ValueError("Batch: one range per plan")
            _check(lib().sge_batch_create_ranges(graph._h, self._plans, self.n, _ptr(b), _ptr(e), tun,
                                                 C.byref(self._h)), "sge_batch_create_ranges")
        else:
            _check(lib().sge_batch_create_sampled(graph._h, self._plans, self.n, stride, offset, tun,
                                                  C.byref(self._h)), "sge_batch_create_sampled")

    def run(self):
        counts = np.zeros(self.n, dtype=np.uint64)
        st = Stats()
        _check(lib().sge_batch_run(self._h, _ptr(counts), C.byref(st)), "sge_batch_run")
        return [int(x) for x in counts], st

    def close(self):
        if self._h and self._h.value:
            lib().sge_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def calibration_copy(nbytes: int) -> float:
    ms = C.c_double()
    _check(lib().sge_calibration_copy(nbytes, C.byref(ms)), "sge_calibration_copy")
    return ms.value


def device_info():
    cu, xcd = C.c_int(), C.c_int()
    name = C.create_string_buffer(128)
    _check(lib().sge_device_info(C.byref(cu), C.byref(xcd), name, 128), "sge_device_info")
    return {"num_cu": cu.value, "xcd": xcd.value, "arch": name.value.decode()}


EXPORTED_SYMBOLS = sorted(_SIGS)
