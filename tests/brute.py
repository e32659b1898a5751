"""Brute-force induced-subgraph census for tiny graphs (test reference).

Every k-subset of vertices is classified by the canonical form of its
induced subgraph (minimum adjacency code over all k! relabellings); the
census maps canonical code -> number of induced copies.  Independent of
the set-operation formulas it checks.
"""
import itertools

import numpy as np


def _pairs(k):
    return list(itertools.combinations(range(k), 2))


def _perm_maps(k):
    pairs = _pairs(k)
    index = {p: i for i, p in enumerate(pairs)}
    maps = []
    for perm in itertools.permutations(range(k)):
        m = []
        for (a, b) in pairs:
            x, y = perm[a], perm[b]
            m.append(index[(min(x, y), max(x, y))])
        maps.append(m)
    return np.asarray(maps, dtype=np.int64)


def canonical_codes(bits, k):
    """bits: (S, C(k,2)) 0/1 array -> canonical code per row."""
    maps = _perm_maps(k)
    w = (1 << np.arange(bits.shape[1], dtype=np.int64))
    best = None
    for m in maps:
        code = bits[:, m] @ w
        best = code if best is None else np.minimum(best, code)
    return best


def code_of(k, edges):
    pairs = _pairs(k)
    bits = np.zeros((1, len(pairs)), dtype=np.int64)
    es = {(min(a, b), max(a, b)) for a, b in edges}
    for i, p in enumerate(pairs):
        bits[0, i] = 1 if p in es else 0
    return int(canonical_codes(bits, k)[0])


def census(A, k):
    n = A.shape[0]
    pairs = _pairs(k)
    subs = np.asarray(list(itertools.combinations(range(n), k)), dtype=np.int64)
    if len(subs) == 0:
        return {}
    bits = np.stack([A[subs[:, a], subs[:, b]] for a, b in pairs], axis=1).astype(np.int64)
    codes = canonical_codes(bits, k)
    vals, cnts = np.unique(codes, return_counts=True)
    return {int(v): int(c) for v, c in zip(vals, cnts)}


# Named shapes (vertex labels arbitrary).
SHAPES4 = {
    "claw": [(0, 1), (0, 2), (0, 3)],
    "path4": [(0, 1), (1, 2), (2, 3)],
    "paw": [(0, 1), (1, 2), (0, 2), (2, 3)],          # tailed triangle
    "cycle4": [(0, 1), (1, 2), (2, 3), (3, 0)],
    "diamond": [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3)],
    "k4": list(itertools.combinations(range(4), 2)),
}
SHAPES5 = {
    "chair": [(0, 1), (1, 2), (2, 3), (2, 4)],
    "k14": [(0, 1), (0, 2), (0, 3), (0, 4)],
    "bull": [(0, 1), (1, 2), (0, 2), (0, 3), (1, 4)],
    "diamond_pendant3": [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (0, 4)],
    "k113": [(0, 1), (0, 2), (0, 3), (0, 4), (1, 2), (1, 3), (1, 4)],
    "k4_pendant": list(itertools.combinations(range(4), 2)) + [(3, 4)],
}


def induced_counts(A):
    c4, c5 = census(A, 4), census(A, 5)
    out = {name: c4.get(code_of(4, e), 0) for name, e in SHAPES4.items()}
    out.update({name: c5.get(code_of(5, e), 0) for name, e in SHAPES5.items()})
    return out
