"""Learned work estimator for root scheduling.

A 3-layer MLP (4 -> 64 -> 64 -> 1, ReLU) maps per-vertex features
(degree, oriented out-degree, 2-hop size bound, label) to log1p of the work a
root vertex costs the exact miner (plan entries logged by
``gpm_clique_root_work``).  Inference runs in the library on matrix cores
(``gpm_mlp_predict``: BF16 operands, FP32 accumulation); training is plain
numpy fp32 on the host over logged work, see ``train_cost_model.py``.

The prediction only orders and splits work (``gpm_sched.root_cost`` with the
LEARNED policy); counts never depend on it.  The reference has no such model:
its scheduling is the fixed degree threshold of
src/gpu_kernels/tuning_schedules.cuh:3-30 and src/TCSolver.cuh:83-104, which
remains the DEGREE policy here and the fallback when no model is given.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

import gpm_amd as G

FIELDS = ("mean", "std", "w1", "b1", "w2", "b2", "w3", "b3")
SHAPES = {"mean": (4,), "std": (4,), "w1": (64, 4), "b1": (64,), "w2": (64, 64),
          "b2": (64,), "w3": (1, 64), "b3": (1,)}
HERE = os.path.dirname(os.path.abspath(__file__))


def default_path(k: int) -> str:
    """k-clique root-cost model for k >= 3; k = 0 is the matching root-cost
    model (features: degree, depth-1 label-slice length, degrees summed over
    that slice, query size; target: weighted DFS steps, gpm_match_root_work)."""
    if k == 0:
        return os.path.join(HERE, "weights", "match_root.npz")
    return os.path.join(HERE, "weights", f"clique_k{k}.npz")


@dataclass
class CostModel:
    params: dict
    k: int = 5

    def __post_init__(self):
        for f in FIELDS:
            a = np.ascontiguousarray(self.params[f], dtype=np.float32)
            if a.shape != SHAPES[f]:
                raise ValueError(f"{f}: shape {a.shape}, want {SHAPES[f]}")
            self.params[f] = a

    @classmethod
    def load(cls, path: str) -> "CostModel":
        with np.load(path, allow_pickle=False) as z:
            return cls({f: z[f] for f in FIELDS}, int(z["k"]))

    @classmethod
    def load_default(cls, k: int = 5) -> "CostModel":
        return cls.load(default_path(k))

    def save(self, path: str) -> None:
        os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
        np.savez(path, k=np.int64(self.k), **self.params)

    def packed(self) -> np.ndarray:
        """The parameters as one float32 vector in gpm_mlp_predict's argument
        order (mean, std, W1, b1, W2, b2, W3, b3): gpm_tuning.root_model."""
        return np.concatenate([self.params[f].reshape(-1) for f in FIELDS]).astype(np.float32)

    def predict_log(self, feat: np.ndarray) -> np.ndarray:
        """log1p(work) per row of feat (n x 4), evaluated on the GPU."""
        p = self.params
        return G.mlp_predict(feat, *(p[f] for f in FIELDS))

    def predict(self, sym: G.DeviceGraph, dag: Optional[G.DeviceGraph] = None) -> np.ndarray:
        """Predicted work per vertex (linear scale, >= 0) for the roots of dag."""
        feat = G.vertex_features(sym, dag)
        return np.expm1(np.clip(self.predict_log(feat), 0.0, 60.0)).astype(np.float32)


def fit(feat: np.ndarray, work: np.ndarray, k: int, seed: int = 0, epochs: int = 6,
        batch: int = 4096, lr: float = 2e-3, log=print) -> CostModel:
    """Fits the MLP to log1p(work) with Adam (fp32, mean squared error)."""
    rng = np.random.default_rng(seed)
    x = np.log1p(feat.astype(np.float64))
    mean = x.mean(axis=0)
    std = x.std(axis=0)
    std[std < 1e-3] = 1.0
    x = ((x - mean) / std).astype(np.float32)
    y = np.log1p(work.astype(np.float64)).astype(np.float32)
    he = lambda fan_in, shape: (rng.normal(0, np.sqrt(2.0 / fan_in), shape)).astype(np.float32)
    P = {"w1": he(4, (64, 4)), "b1": np.zeros(64, np.float32),
         "w2": he(64, (64, 64)), "b2": np.zeros(64, np.float32),
         "w3": he(64, (1, 64)), "b3": np.array([y.mean()], np.float32)}
    m = {kk: np.zeros_like(v) for kk, v in P.items()}
    v2 = {kk: np.zeros_like(v) for kk, v in P.items()}
    b1, b2, eps, t = 0.9, 0.999, 1e-8, 0
    n = len(y)
    for ep in range(epochs):
        perm = rng.permutation(n)
        tot = 0.0
        for s in range(0, n, batch):
            idx = perm[s:s + batch]
            xb, yb = x[idx], y[idx]
            z1 = xb @ P["w1"].T + P["b1"]
            h1 = np.maximum(z1, 0)
            z2 = h1 @ P["w2"].T + P["b2"]
            h2 = np.maximum(z2, 0)
            out = (h2 @ P["w3"].T)[:, 0] + P["b3"][0]
            err = out - yb
            tot += float((err * err).sum())
            g_out = (2.0 / len(yb)) * err[:, None]
            grads = {"w3": g_out.T @ h2, "b3": g_out.sum(axis=0)}
            g_h2 = (g_out @ P["w3"]) * (z2 > 0)
            grads["w2"] = g_h2.T @ h1
            grads["b2"] = g_h2.sum(axis=0)
            g_h1 = (g_h2 @ P["w2"]) * (z1 > 0)
            grads["w1"] = g_h1.T @ xb
            grads["b1"] = g_h1.sum(axis=0)
            t += 1
            for kk in P:
                gk = grads[kk].astype(np.float32)
                m[kk] = b1 * m[kk] + (1 - b1) * gk
                v2[kk] = b2 * v2[kk] + (1 - b2) * gk * gk
                mh = m[kk] / (1 - b1 ** t)
                vh = v2[kk] / (1 - b2 ** t)
                P[kk] = (P[kk] - lr * mh / (np.sqrt(vh) + eps)).astype(np.float32)
        log(f"epoch {ep}: mse {tot / n:.4f}")
    P["mean"] = mean.astype(np.float32)
    P["std"] = std.astype(np.float32)
    return CostModel(P, k)


def learned_sched(model: CostModel, sym: G.DeviceGraph, dag: G.DeviceGraph) -> np.ndarray:
    """Root costs for ``DeviceGraph.clique(k, policy=LEARNED, cost=...)``."""
    return model.predict(sym, dag)
