"""Time the CSR build on the host and on the GPU for R-MAT graphs.

    python -m sge.graph_build_bench [--scales 18 20 21]

Both builders deduplicate the undirected edge list, drop self-loops, relabel
vertices by ascending degree and build ascending rows; the GPU build
(SGE_GRAPH_BUILD_DEVICE) must give the same arrays.  Edge generation (host) is
not part of either figure.
"""
import argparse

import numpy as np

import sge


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scales", type=int, nargs="+", default=[18, 20, 21])
    ap.add_argument("--seed", type=int, default=4)
    args = ap.parse_args()
    for scale in args.scales:
        h = sge.Graph.rmat(scale, 16, seed=args.seed)
        d = sge.Graph.rmat(scale, 16, seed=args.seed, device_build=True)
        same = all(np.array_equal(a, b) for a, b in zip(h.csr(), d.csr())) and np.array_equal(h.perm(), d.perm())
        print(f"rmat scale {scale}: n {h.n} m {h.m} | host build {h.stats.build_ms:.1f} ms | "
              f"device build {d.stats.build_ms:.1f} ms (incl. upload of the edge list and download of the CSR) | "
              f"identical arrays: {same}", flush=True)
        if not same:
            raise SystemExit("device and host CSR differ")
        del h, d


if __name__ == "__main__":
    main()
