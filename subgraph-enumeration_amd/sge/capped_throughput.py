"""Throughput of the plain executor on a large R-MAT graph, measured under the
work cap (the exact jobs there are far beyond a benchmark window).

    python -m sge.capped_throughput [--scale 23] [--seed 6] [--out results.json]

The graph's CSR is built on the GPU and stays resident in HBM (scale 23:
8.4 M vertices, 129 M edges, about 1 GB).  For K5-e and the first six-vertex
pattern (IEP mode, cost-model plan, whole root set, heavy-first) the executor
runs until its work cap stops it; the cap grows until a run takes a few
seconds.  Reported per run: kernel time, the kernel's algorithmic bytes and
their rate against the 8 TB/s HBM peak.  The counts of capped runs are partial
and are not reported.
"""
import argparse
import json
import time

import sge
import sge.workloads as W

HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=23)
    ap.add_argument("--seed", type=int, default=6)
    ap.add_argument("--target-ms", type=float, default=4000.0)
    ap.add_argument("--patterns", default="near5clique,sixvertex0")
    ap.add_argument("--calibrate", action="store_true",
                    help="run the 256 MiB copy kernel first (scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    if args.calibrate:
        ms = sge.calibration_copy(256 << 20)
        print(f"calibration copy: 256 MiB read + written in {ms:.3f} ms", flush=True)
    print(f"building R-MAT scale {args.scale} (host edges, CSR on the GPU)", flush=True)
    t0 = time.time()
    g = W.graph_rmat(args.scale, args.seed, device_build=True)
    print(f"graph n={g.n} m={g.m} max degree {g.stats.max_degree} in {time.time() - t0:.1f} s", flush=True)
    res = {"graph": f"rmat(scale={args.scale}, ef=16, a=.57, b=.19, c=.19, seed={args.seed})", "n": g.n, "m": g.m,
           "max_degree": g.stats.max_degree, "runs": []}
    pats = {"near5clique": W.named("near5clique"), "sixvertex0": W.six_vertex_set()[0]}
    for name in args.patterns.split(","):
        p = pats[name]
        plan = sge.build_plan(p, g, sge.MODE_IEP)
        cap = 1 << 12
        for _ in range(3):
            st, capped = sge.time_capped(g, [plan], 1, cap)[0]
            gbs = st.bytes / st.kernel_ms / 1e6
            print(f"{name}: cap {cap}k wave steps, capped {capped}, kernel {st.kernel_ms:.1f} ms, "
                  f"{st.bytes / 1e9:.1f} GB algorithmic -> {gbs:.0f} GB/s ({gbs / HBM_PEAK_GBS:.3f} of peak), "
                  f"{st.intersections} intersections", flush=True)
            res["runs"].append({"pattern": name, "cap_k": cap, "capped": capped, "kernel_ms": st.kernel_ms,
                                "algorithmic_bytes": st.bytes, "gb_per_s": gbs, "intersections": st.intersections})
            if not capped or st.kernel_ms > args.target_ms:
                break
            cap = int(cap * min(16.0, max(2.0, args.target_ms / max(st.kernel_ms, 1.0))))
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
