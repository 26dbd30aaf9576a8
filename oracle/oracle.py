"""ctypes wrapper for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: import only from tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The reference binary built by oracle/ref.mk lives
at oracle/_ref/graph_mining_system (present only where it was built).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
REF_BIN = os.path.join(_HERE, "_ref", "graph_mining_system")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/liboracle.so missing: run make -C oracle")
        L = C.CDLL(LIB)
        L.oracle_set_begin_pos.restype = C.c_uint32
        L.oracle_set_begin_pos.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
        L.oracle_set_end_pos.restype = C.c_uint32
        L.oracle_set_end_pos.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
        L.oracle_intersect.restype = C.c_uint32
        L.oracle_intersect.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p]
        L.oracle_choosec.restype = C.c_uint64
        L.oracle_choosec.argtypes = [C.c_uint64, C.c_int]
        L.oracle_count_range.restype = C.c_uint64
        L.oracle_count_range.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint64,
                                         C.c_int, C.POINTER(C.c_uint64)]
        L.oracle_count_strided.restype = C.c_uint64
        L.oracle_count_strided.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint64,
                                           C.c_int]
        L.oracle_count.restype = C.c_uint64
        L.oracle_count.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int]
        L.oracle_enumerate.restype = C.c_uint64
        L.oracle_enumerate.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64]
        L.oracle_bag_trie.restype = C.c_int
        L.oracle_bag_trie.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                      C.c_uint64, C.c_void_p, C.c_uint64]
        L.oracle_count_brute.restype = C.c_uint64
        L.oracle_count_brute.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def set_begin_pos(arr, mx):
    a = _u32(arr)
    return lib().oracle_set_begin_pos(_p(a), len(a), mx)


def set_end_pos(arr, mn):
    a = _u32(arr)
    return lib().oracle_set_end_pos(_p(a), len(a), mn)


def intersect(a, b):
    a, b = _u32(a), _u32(b)
    out = np.zeros(max(1, min(len(a), len(b))), dtype=np.uint32)
    n = lib().oracle_intersect(_p(a), len(a), _p(b), len(b), _p(out))
    return out[:n]


def choosec(n, k):
    return lib().oracle_choosec(n, k)


def count(plan, off, nbr, threads=0, begin=0, end=None):
    off, nbr = _u32(off), _u32(nbr)
    n = len(off) - 1
    if end is None:
        end = int(off[-1])
    inter = C.c_uint64()
    return int(lib().oracle_count_range(C.byref(plan), _p(off), _p(nbr), n, begin, end, threads, C.byref(inter)))


def count_strided(plan, off, nbr, stride, offset=0, threads=0):
    off, nbr = _u32(off), _u32(nbr)
    return int(lib().oracle_count_strided(C.byref(plan), _p(off), _p(nbr), len(off) - 1, stride, offset, threads))


def enumerate_embeddings(plan, off, nbr, cap):
    off, nbr = _u32(off), _u32(nbr)
    out = np.zeros(max(cap, 1) * plan.k, dtype=np.uint32)
    n = lib().oracle_enumerate(C.byref(plan), _p(off), _p(nbr), len(off) - 1, _p(out), cap)
    return out[: min(n, cap) * plan.k].reshape(-1, plan.k).copy(), int(n)


def bag_trie(plan, off, nbr):
    """Trie of plan.bag's relation: (vb, values, child), depth = len(vb) - 1."""
    off, nbr = _u32(off), _u32(nbr)
    vb = np.zeros(9, dtype=np.uint32)
    p = lib().oracle_bag_trie(C.byref(plan), _p(off), _p(nbr), len(off) - 1, _p(vb), None, 0, None, 0)
    if p < 0:
        raise RuntimeError("oracle_bag_trie failed")
    nval, nchild = int(vb[p]), int(vb[p - 1]) + p - 1
    values = np.zeros(max(nval, 1), dtype=np.uint32)
    child = np.zeros(max(nchild, 1), dtype=np.uint32)
    lib().oracle_bag_trie(C.byref(plan), _p(off), _p(nbr), len(off) - 1, _p(vb), _p(values), nval, _p(child), nchild)
    return [int(x) for x in vb[: p + 1]], values[:nval], child[:nchild]


def count_brute(k, adj, off, nbr):
    off, nbr = _u32(off), _u32(nbr)
    a = _u32(list(adj)[:k] + [0] * (8 - k))
    return int(lib().oracle_count_brute(k, _p(a), _p(off), _p(nbr), len(off) - 1))
