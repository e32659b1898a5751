"""Trains the matching root-cost model on work logged by the exact miner.

For each training graph (seeded labeled R-MAT, 8 labels, seeds disjoint from
the evaluation seeds) and each training query (the named shapes, seeded
random connected queries and dense queries cut out of the graph), the GPU
logs the weighted DFS steps the exact miner spends below every root
candidate (gpm_match_root_work) under both matching orders, and the
candidates' features (gpm_match_root_features); the MLP is fit to log1p of
the steps on the host and written as an .npz of float32 arrays.  The model
only picks matching orders and sizes tasks (gpm_tuning.root_model); counts
never depend on it.

    python gpm-artifact_amd/train_match_model.py --out gpm-artifact_amd/weights/match_root.npz
"""
import argparse
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import gpm_amd as G  # noqa: E402
import workloads as W  # noqa: E402
from cost_model import default_path, fit  # noqa: E402

TRAIN_SEEDS = (7001, 7002, 7003)
HELD_OUT_SEED = 7100


def queries_for(g, labels, seed):
    rng = np.random.default_rng(seed)
    qs = []
    for _, (n, edges) in W.NAMED_SHAPES.items():
        if n <= 5:
            qs.append((n, np.asarray(edges, dtype=np.uint32), rng.integers(0, 8, n).astype(np.uint32)))
    qs += W.random_queries(12, seed, min_size=3, max_size=5, extra_edge_prob=0.5)
    qs += W.dense_queries(g, labels, 12, seed, min_size=4, max_size=5)
    return qs


def collect(scale, seed, max_rows=8000, log=print):
    g, labels = W.labeled_graph(scale, seed=seed)
    feats, works = [], []
    rng = np.random.default_rng(seed)
    with G.DeviceGraph(g, oriented=False, labels=labels, n_labels=8) as dg:
        for qi, q in enumerate(queries_for(g, labels, seed)):
            for policy in (G.SCHED_DEGREE, G.SCHED_LEARNED):
                t0 = time.time()
                f = dg.match_root_features(q, policy)
                w = dg.match_root_work(q, policy)
                print(f"  s{scale} seed {seed} query {qi} ({q[0]} vertices) policy {policy}: "
                      f"{int(w.sum())} steps, {time.time() - t0:.2f} s", flush=True)
                rows = np.nonzero(f[:, 3] > 0)[0]  # root candidates (query size > 0)
                if len(rows) > max_rows:
                    rows = rng.choice(rows, max_rows, replace=False)
                feats.append(f[rows])
                works.append(w[rows])
    log(f"logged s{scale} seed {seed}: {sum(len(w) for w in works)} roots")
    return np.concatenate(feats), np.concatenate(works)


def spearman(a, b):
    ra = np.argsort(np.argsort(a))
    rb = np.argsort(np.argsort(b))
    return float(np.corrcoef(ra, rb)[0, 1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scales", type=int, nargs="+", default=[11, 12, 13])
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    feats, works = [], []
    for scale in a.scales:
        for seed in TRAIN_SEEDS:
            t = time.time()
            f, w = collect(scale, seed, log=lambda s, t=t: print(s, f"({time.time() - t:.1f}s)",
                                                                flush=True))
            feats.append(f)
            works.append(w)
    feat = np.concatenate(feats)
    work = np.concatenate(works)
    model = fit(feat, work, 0, epochs=a.epochs, log=lambda s: print(s, flush=True))
    out = a.out or default_path(0)
    model.save(out)
    f, w = collect(max(a.scales), HELD_OUT_SEED)
    pred = model.predict_log(f)
    print(f"saved {out}; held-out spearman {spearman(pred, w.astype(np.float64)):.4f} "
          f"over {len(w)} roots", flush=True)


if __name__ == "__main__":
    main()
