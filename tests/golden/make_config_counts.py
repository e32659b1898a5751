"""Writes tests/golden/config_counts.json: exact counts of the benchmark
configurations' seeded graphs computed by the CPU oracle (oracle/oracle.cpp,
itself pinned by closed forms, networkx and brute-force census in
tests/test_oracle.py).  The GPU tests compare the HIP engine with these at the
configurations' full sizes, where the oracle takes minutes to hours.

    python tests/golden/make_config_counts.py [--threads T] [--only NAME,...]

Each entry records the graph (generator, scale, seed), what was counted and
the oracle's wall time on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "gpm-artifact_amd"))
sys.path.insert(0, ROOT)

import gpm_amd as G  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "config_counts.json")


def load():
    if os.path.exists(OUT):
        with open(OUT) as f:
            return json.load(f)
    return {}


def save(d):
    with open(OUT + ".tmp", "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
    os.replace(OUT + ".tmp", OUT)


def timed(fn, *a):
    t0 = time.perf_counter()
    v = fn(*a)
    return v, round(time.perf_counter() - t0, 1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--only", default="", help="comma list of entry names")
    args = ap.parse_args()
    if args.threads:
        O.set_threads(args.threads)
    d = load()
    want = set(args.only.split(",")) if args.only else None

    def run(name, fn):
        if name in d or (want and name not in want):
            return
        print("computing", name, flush=True)
        d[name] = fn()
        d[name]["oracle_threads"] = O.num_threads()
        save(d)
        print(name, d[name], flush=True)

    graphs = {}

    def dag(scale, seed):
        if (scale, seed) not in graphs:
            graphs.clear()
            g = G.rmat_graph(scale, 16, seed)
            graphs[(scale, seed)] = (g, G.orient(g))
        return graphs[(scale, seed)]

    def clique_entry(scale, seed, k):
        def f():
            _, dd = dag(scale, seed)
            c, t = timed(O.clique, dd.offsets, dd.adj, k)
            return {"graph": f"R-MAT scale {scale}, edge factor 16, seed {seed}, degree DAG",
                    "k": k, "count": int(c), "oracle_seconds": t}
        return f

    def motif_entry(scale, seed):
        def f():
            g, _ = dag(scale, seed)
            c, t = timed(O.motif4, g.offsets, g.adj)
            return {"graph": f"R-MAT scale {scale}, edge factor 16, seed {seed}, symmetric",
                    "motif4_induced": [int(x) for x in c], "oracle_seconds": t}
        return f

    def ceci_each(g, labels, queries):
        # One query at a time, its roots shared over every thread (a batch
        # call runs each query on one thread, and one query dominates).
        out = []
        for i, (qn, qe, ql) in enumerate(queries):
            out.append(int(O.ceci(g.offsets, g.adj, labels, qn, qe, ql)))
            print(f"  query {i}: {out[-1]}", flush=True)
        return out

    def named_entry(scale):
        def f():
            import workloads as W
            g, labels = W.labeled_graph(scale)
            named = dict(W.named_queries())
            names = ["triangle", "4-cycle", "diamond", "three-star", "pyramid-6"]
            c, t = timed(ceci_each, g, labels, [named[q] for q in names])
            return {"graph": f"labeled R-MAT scale {scale} (workloads.labeled_graph, 8 labels)",
                    "queries": names, "counts": c, "oracle_seconds": t}
        return f

    def batch_entry(scale):
        def f():
            import workloads as W
            g, labels = W.labeled_graph(scale)
            c, t = timed(ceci_each, g, labels, W.batch_queries())
            return {"graph": f"labeled R-MAT scale {scale} (workloads.labeled_graph, 8 labels)",
                    "queries": "workloads.batch_queries()", "counts": c, "oracle_seconds": t}
        return f

    run("rmat20_k3", clique_entry(20, 20240520, 3))
    run("rmat20_k4", clique_entry(20, 20240520, 4))
    run("rmat16_s20240522_motif4", motif_entry(16, 20240522))
    run("rmat22_k3", clique_entry(22, 20240522, 3))
    run("rmat22_k4", clique_entry(22, 20240522, 4))
    run("rmat20_k5", clique_entry(20, 20240520, 5))
    run("rmat22_k5", clique_entry(22, 20240522, 5))
    run("labeled18_named", named_entry(18))
    run("labeled11_batch256", batch_entry(11))


if __name__ == "__main__":
    main()
