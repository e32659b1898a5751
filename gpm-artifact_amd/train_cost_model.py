"""Trains the k-clique root-cost model on work logged by the exact miner.

For each training graph (seeded R-MAT, disjoint from the evaluation seeds) the
GPU computes per-vertex features (gpm_vertex_features) and the per-root plan
entries of the exact k-clique miner (gpm_clique_root_work); the MLP is fit to
log1p(work) on the host and written as an .npz of float32 arrays.

    python gpm-artifact_amd/train_cost_model.py --k 5 --out gpm-artifact_amd/weights/clique_k5.npz
"""
import argparse
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import gpm_amd as G  # noqa: E402
from cost_model import fit, default_path  # noqa: E402

TRAIN_SEEDS = (9001, 9002, 9003)


def collect(scale: int, seed: int, k: int):
    g = G.rmat_graph(scale, 16, seed)
    d = G.orient(g)
    with G.DeviceGraph(g, oriented=False) as sym, G.DeviceGraph(d, oriented=True) as dag:
        feat = G.vertex_features(sym, dag)
        work = dag.clique_root_work(k)
    return feat, work


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--k", type=int, default=5)
    ap.add_argument("--scales", type=int, nargs="+", default=[15, 16, 17])
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    feats, works = [], []
    for scale in a.scales:
        for seed in TRAIN_SEEDS:
            t = time.time()
            f, w = collect(scale, seed, a.k)
            feats.append(f)
            works.append(w)
            print(f"logged s{scale} seed {seed}: {len(w)} roots, "
                  f"work {int(w.sum())} in {time.time() - t:.1f}s", flush=True)
    feat = np.concatenate(feats)
    work = np.concatenate(works)
    model = fit(feat, work, a.k, log=lambda s: print(s, flush=True))
    out = a.out or default_path(a.k)
    model.save(out)
    # Held-out check on a fresh seed: rank correlation of cost vs work.
    f, w = collect(max(a.scales), 9100, a.k)
    pred = model.predict_log(f)
    keep = w > 0
    rp = np.argsort(np.argsort(pred[keep]))
    rw = np.argsort(np.argsort(w[keep]))
    print(f"saved {out}; held-out spearman {np.corrcoef(rp, rw)[0, 1]:.4f}", flush=True)


if __name__ == "__main__":
    main()
