"""Pins the CPU oracle: closed forms, networkx, brute-force census.

These run on the CPU only; the GPU tests then compare the HIP engine with
the pinned oracle.
"""
import itertools
from collections import Counter
from math import comb

import numpy as np
import pytest

import brute
import graphs


def dag(gpm, g):
    return gpm.orient(g)


@pytest.mark.parametrize("n", [3, 5, 9, 17])
def test_tc_clique_complete_graph(gpm, orc, n):
    d = dag(gpm, graphs.complete(n))
    assert orc.tc(d.offsets, d.adj) == comb(n, 3)
    for k in (3, 4, 5):
        assert orc.clique(d.offsets, d.adj, k) == comb(n, k)


def test_tc_zero_cases(gpm, orc):
    for g in (graphs.complete_bipartite(7, 9), graphs.from_edges(6, [(0, 1), (1, 2), (2, 3)])):
        d = dag(gpm, g)
        assert orc.tc(d.offsets, d.adj) == 0
        assert orc.clique(d.offsets, d.adj, 4) == 0


@pytest.mark.parametrize("n", [5, 6, 11, 30])
def test_tc_wheel(gpm, orc, n):
    d = dag(gpm, graphs.wheel(n))
    # hub + C_{n-1}: n-1 triangles (one per rim edge); no K4 once the rim is
    # longer than a triangle.
    assert orc.tc(d.offsets, d.adj) == n - 1
    assert orc.clique(d.offsets, d.adj, 4) == (1 if n == 4 else 0)


@pytest.mark.parametrize("n,r", [(12, 3), (17, 4), (20, 5)])
def test_clique_turan(gpm, orc, n, r):
    d = dag(gpm, graphs.turan(n, r))
    for k in (3, 4, 5):
        assert orc.clique(d.offsets, d.adj, k) == graphs.turan_cliques(n, r, k)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_clique_vs_networkx(gpm, orc, seed):
    nx = pytest.importorskip("networkx")
    g = graphs.er(120, 1400, seed)
    d = dag(gpm, g)
    x = graphs.to_nx(g)
    cl = Counter(len(c) for c in nx.enumerate_all_cliques(x))
    assert orc.tc(d.offsets, d.adj) == cl[3]
    for k in (3, 4, 5):
        assert orc.clique(d.offsets, d.adj, k) == cl[k]


def test_clique_root_ranges_partition(gpm, orc):
    g = gpm.rmat_graph(10, 16, 5)
    d = dag(gpm, g)
    for k in (3, 4, 5):
        total = orc.clique(d.offsets, d.adj, k)
        cuts = [0, 100, 101, 517, 1024]
        parts = [orc.clique(d.offsets, d.adj, k, a, b) for a, b in zip(cuts, cuts[1:])]
        assert sum(parts) == total


def _plan_bytes_python(off, adj, k):
    """Recounts the canonical plan's set sizes with Python sets."""
    n = len(off) - 1
    N = [set(adj[off[v]:off[v + 1]].tolist()) for v in range(n)]
    total = 4 * (n + 1) + 8 * int(off[-1])

    def rec(W, r):
        nonlocal total
        if r == 2:
            for v in W:
                total += 4 * (len(W) + len(N[v]))
            return
        for v in W:
            W2 = W & N[v]
            total += 4 * (len(W) + len(N[v]) + len(W2))
            rec(W2, r - 1)

    for v0 in range(n):
        for v1 in N[v0]:
            if k == 3:
                total += 4 * (len(N[v0]) + len(N[v1]))
                continue
            W = N[v0] & N[v1]
            total += 4 * (len(N[v0]) + len(N[v1]) + len(W))
            rec(W, k - 2)
    return total


@pytest.mark.parametrize("k", [3, 4, 5])
def test_plan_bytes_vs_python(gpm, orc, k):
    g = graphs.er(60, 500, 11)
    d = dag(gpm, g)
    assert orc.plan_bytes(d.offsets, d.adj, k) == _plan_bytes_python(d.offsets, d.adj, k)


@pytest.fixture(scope="module")
def census_graphs(gpm):
    out = []
    for seed, (n, m) in enumerate([(14, 40), (16, 52), (18, 60), (17, 90)]):
        out.append(graphs.er(n, m, 100 + seed))
    out.append(graphs.wheel(9))
    out.append(graphs.complete_bipartite(3, 5))
    return out


def test_motif4_vs_census(gpm, orc, census_graphs):
    for g in census_graphs:
        c = brute.induced_counts(graphs.adjacency_matrix(g))
        got = orc.motif4(g.offsets, g.adj)
        assert got == [c["claw"], c["path4"], c["paw"], c["cycle4"], c["diamond"], c["k4"]]


def test_patterns_vs_census(gpm, orc, census_graphs):
    for g in census_graphs:
        c = brute.induced_counts(graphs.adjacency_matrix(g))
        expect = {
            1: c["claw"],
            2: c["chair"],
            3: 3 * c["claw"] + 4 * c["k14"],
            4: c["path4"],
            5: 2 * c["bull"],
            6: 2 * c["paw"] + 2 * c["diamond_pendant3"],
            7: 2 * c["k113"],
            8: 3 * c["k4_pendant"],
        }
        for pid, want in expect.items():
            assert orc.pattern(g.offsets, g.adj, pid) == want, pid


def test_census_sanity():
    # The census itself: K5 holds C(5,4) induced K4s and one K5-code subset.
    A = np.ones((5, 5), dtype=np.uint8) - np.eye(5, dtype=np.uint8)
    c4 = brute.census(A, 4)
    assert c4 == {brute.code_of(4, brute.SHAPES4["k4"]): 5}
    # A path on 4 vertices is one induced path4, nothing else of size 4.
    P = np.zeros((4, 4), dtype=np.uint8)
    for a, b in [(0, 1), (1, 2), (2, 3)]:
        P[a, b] = P[b, a] = 1
    assert brute.census(P, 4) == {brute.code_of(4, brute.SHAPES4["path4"]): 1}


@pytest.mark.parametrize("n", [6, 11, 20])
def test_clique_list_complete_graph_per_root(gpm, orc, n):
    # K_n oriented by (degree, id): every degree is equal, so v -> w iff w > v
    # and root v roots exactly the cliques whose smallest id is v:
    # C(n-1-v, k-1) of them.
    d = dag(gpm, graphs.complete(n))
    for k in (3, 4, 5):
        for v in range(n):
            assert orc.clique_list(d.offsets, d.adj, k, [v]) == comb(n - 1 - v, k - 1)


def test_clique_list_matches_ranges_and_partitions(gpm, orc):
    g = gpm.rmat_graph(10, 16, 5)
    d = dag(gpm, g)
    rng = np.random.default_rng(3)
    for k in (3, 4, 5):
        total = orc.clique(d.offsets, d.adj, k)
        assert orc.clique_list(d.offsets, d.adj, k, np.arange(d.n)) == total
        assert orc.clique_list(d.offsets, d.adj, k, np.arange(100, 517)) == \
            orc.clique(d.offsets, d.adj, k, 100, 517)
        # A seeded random split of the roots into three sets sums to the total.
        part = rng.integers(0, 3, d.n)
        parts = [orc.clique_list(d.offsets, d.adj, k, np.nonzero(part == i)[0]) for i in range(3)]
        assert sum(parts) == total
        assert orc.clique_list(d.offsets, d.adj, k, []) == 0
