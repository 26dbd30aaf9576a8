"""Named patterns, synthetic data graphs and pattern sets of the benchmark configs.

Patterns follow the reference's query files: diamond = dataset/light_query/p2.txt,
house = dataset/light_query/p4.txt; near-5-clique here is K5 minus one edge (the
reference's near_five_clique.out counts a different shape).
"""
from __future__ import annotations

from typing import Dict, List

from . import Graph, Pattern, sample_patterns

DIAMOND = [(0, 1), (0, 2), (0, 3), (1, 2), (2, 3)]                      # K4 minus one edge
HOUSE = [(0, 1), (0, 2), (1, 3), (2, 3), (2, 4), (3, 4)]                 # 4-cycle + triangle roof
NEAR_5_CLIQUE = [(i, j) for i in range(5) for j in range(i + 1, 5) if (i, j) != (3, 4)]  # K5 - e
TRIANGLE = [(0, 1), (0, 2), (1, 2)]


def clique(k: int):
    return [(i, j) for i in range(k) for j in range(i + 1, k)]


NAMED: Dict[str, tuple] = {
    "triangle": (3, TRIANGLE),
    "diamond": (4, DIAMOND),
    "house": (5, HOUSE),
    "near5clique": (5, NEAR_5_CLIQUE),
    "k4": (4, clique(4)),
    "k5": (5, clique(5)),
}


def named(name: str) -> Pattern:
    k, e = NAMED[name]
    return Pattern.from_edges(k, e)


# ---- data graphs (benchmark configurations) -------------------------------------------------
def graph_c1(degree_order=True) -> Graph:
    """Chung-Lu n=379, m=914, gamma=2.5, no cap, seed 0."""
    return Graph.chung_lu(379, 914, 2.5, 0.0, 0, degree_order)


def graph_c2(degree_order=True) -> Graph:
    """Chung-Lu n=7115, m=100762, gamma=2.1, expected degree capped at 1100, seed 1."""
    return Graph.chung_lu(7115, 100762, 2.1, 1100.0, 1, degree_order)


def graph_rmat(scale: int, seed: int, degree_order=True, device_build=False) -> Graph:
    """Undirected deduplicated R-MAT, edge factor 16, (a,b,c,d)=(.57,.19,.19,.05)."""
    return Graph.rmat(scale, 16, 0.57, 0.19, 0.19, seed, degree_order, device_build)


def graph_c4(degree_order=True, device_build=False) -> Graph:
    return graph_rmat(20, 4, degree_order, device_build)


def graph_c5(degree_order=True, device_build=False) -> Graph:
    return graph_rmat(23, 6, degree_order, device_build)


# ---- pattern sets ---------------------------------------------------------------------
def six_vertex_set() -> List[Pattern]:
    """All 16 connected 6-vertex patterns with 11-13 edges (sampler seed 2)."""
    return sample_patterns(6, 11, 13, 16, 2)


def seven_vertex_set() -> List[Pattern]:
    """All 78 connected 7-vertex patterns with 15-19 edges (sampler seed 3).

    The 16-19 edge range named for this set holds only 38 classes; the 78-pattern
    benchmark set of the reference (dataset/7vertex) is the 15-19 edge class."""
    return sample_patterns(7, 15, 19, 78, 3)


def training_set(n: int = 400) -> List[Pattern]:
    """Connected 7-vertex patterns disjoint from the 78-pattern set (seed 5): 11-14 edges."""
    return sample_patterns(7, 11, 14, n, 5)
