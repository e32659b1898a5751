"""GPU parity: triangle and k-clique counts of the HIP engine vs the pinned
oracle and closed forms (bit-exact uint64)."""
from math import comb

import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu

ER_N, ER_M, ER_SEED = 16384, 131072, 20240601  # triangle-counting workload


def dev_dag(gpm, g):
    return gpm.DeviceGraph(gpm.orient(g), oriented=True)


def test_tc_er_workload(gpm, orc):
    g = gpm.er_graph(ER_N, ER_M, ER_SEED)
    d = gpm.orient(g)
    assert d.m == ER_M
    want = orc.tc(d.offsets, d.adj)
    with gpm.DeviceGraph(d, oriented=True) as dg:
        assert dg.tc(gpm.TC_EDGE) == want
        assert dg.tc(gpm.TC_VERTEX) == want
        assert dg.clique(3) == want
        for reps in (1, 31, 32, 33, 100):
            assert dg.tc_repeat_fused(reps) == want
            assert dg.tc_repeat(gpm.TC_EDGE, reps) == want


@pytest.mark.parametrize("per_vertex", [5, 10, 20, 40, 80, 160])
def test_edge_group_sizes(gpm, orc, per_vertex):
    """The edge kernel's group size G follows m / n (4, 8, 16, 32, 64, 64
    here); lists of up to 2G entries take the lane-exchange path, longer ones
    the binary search, and an Erdos-Renyi graph mixes both in every wave.
    Single counts, hipGraph repeats and fused repeats (4 edges per group)
    against the oracle."""
    n = 3000
    d = gpm.orient(gpm.er_graph(n, n * per_vertex, 500 + per_vertex))
    want = orc.tc(d.offsets, d.adj)
    assert want > 0
    with gpm.DeviceGraph(d, oriented=True) as dg:
        assert dg.tc(gpm.TC_EDGE) == want
        assert dg.tc_repeat(gpm.TC_EDGE, 33) == want
        assert dg.tc_repeat_fused(33) == want
        assert dg.tc_repeat_fused(7) == want


@pytest.mark.parametrize("n", [3, 4, 40, 65, 300, 700])
def test_complete_graphs(gpm, n):
    # K_n exercises every tier: out-degrees reach n-1 (wave <= 64, LDS block
    # <= 512, HBM block above).
    with dev_dag(gpm, graphs.complete(n)) as dg:
        assert dg.tc(gpm.TC_EDGE) == comb(n, 3)
        assert dg.tc(gpm.TC_VERTEX) == comb(n, 3)
        for k in (3, 4):
            assert dg.clique(k) == comb(n, k)
        if n <= 300:
            assert dg.clique(5) == comb(n, 5)


def test_zero_and_wheels(gpm):
    for g, tri in ((graphs.complete_bipartite(30, 40), 0), (graphs.wheel(100), 99)):
        with dev_dag(gpm, g) as dg:
            assert dg.tc(gpm.TC_EDGE) == tri
            assert dg.tc(gpm.TC_VERTEX) == tri
            assert dg.clique(4) == 0


@pytest.mark.parametrize("n,r", [(90, 3), (130, 6)])
def test_turan(gpm, n, r):
    with dev_dag(gpm, graphs.turan(n, r)) as dg:
        for k in (3, 4, 5):
            assert dg.clique(k) == graphs.turan_cliques(n, r, k)


@pytest.mark.parametrize("scale", [12, 15])
def test_rmat_vs_oracle(gpm, orc, scale):
    g = gpm.rmat_graph(scale, 16, 1000 + scale)
    d = gpm.orient(g)
    with gpm.DeviceGraph(d, oriented=True) as dg:
        tri = orc.tc(d.offsets, d.adj)
        assert dg.tc(gpm.TC_EDGE) == tri
        assert dg.tc(gpm.TC_VERTEX) == tri
        for k in (3, 4, 5):
            assert dg.clique(k) == orc.clique(d.offsets, d.adj, k), k


def test_union_with_big_clique(gpm, orc):
    # Random part checked by the oracle, dense part by the closed form.
    r = gpm.rmat_graph(11, 16, 77)
    u = graphs.disjoint_union(r, graphs.complete(600))
    d = gpm.orient(u)
    rd = gpm.orient(r)
    with gpm.DeviceGraph(d, oriented=True) as dg:
        for k in (3, 4):
            assert dg.clique(k) == orc.clique(rd.offsets, rd.adj, k) + comb(600, k)


def test_relabel_invariance(gpm, orc):
    g = gpm.rmat_graph(13, 16, 5)
    d = gpm.orient(g)
    s, _ = gpm.sort_by_degree(d)
    want = orc.tc(d.offsets, d.adj)
    with gpm.DeviceGraph(s, oriented=True) as dg:
        assert dg.tc(gpm.TC_VERTEX) == want
        assert dg.tc(gpm.TC_EDGE) == want
        assert dg.clique(4) == orc.clique(d.offsets, d.adj, 4)


def test_root_ranges_partition(gpm, orc):
    g = gpm.rmat_graph(13, 16, 6)
    d = gpm.orient(g)
    cuts = [0, 1, 37, 1000, 5000, d.n]
    with gpm.DeviceGraph(d, oriented=True) as dg:
        for k in (4, 5):
            total = dg.clique(k)
            parts = [dg.clique(k, root_begin=a, root_end=b) for a, b in zip(cuts, cuts[1:])]
            assert sum(parts) == total
            assert parts[2] == orc.clique(d.offsets, d.adj, k, cuts[2], cuts[3])


def test_learned_policy_same_counts(gpm):
    g = gpm.rmat_graph(13, 16, 8)
    d = gpm.orient(g)
    rng = np.random.default_rng(0)
    cost = rng.random(d.n).astype(np.float32) * d.degrees().astype(np.float32) ** 2
    with gpm.DeviceGraph(d, oriented=True) as dg:
        for k in (3, 4, 5):
            plain = dg.clique(k)
            assert dg.clique(k, policy=gpm.SCHED_LEARNED, cost=cost) == plain
            assert dg.clique(k, policy=gpm.SCHED_LEARNED, cost=cost, chunk=1) == plain


@pytest.mark.parametrize("k", [3, 4, 5])
def test_plan_bytes_and_root_work(gpm, orc, k):
    g = gpm.rmat_graph(11, 16, 9)
    d = gpm.orient(g)
    with gpm.DeviceGraph(d, oriented=True) as dg:
        b = dg.clique_plan_bytes(k)
        assert b == orc.plan_bytes(d.offsets, d.adj, k)
        w = dg.clique_root_work(k)
        assert int(w.sum()) * 4 + 4 * (d.n + 1) + 8 * d.m == b


def test_device_orientation_matches_host(gpm):
    g = gpm.rmat_graph(13, 16, 10)
    d = gpm.orient(g)
    with gpm.DeviceGraph(g, oriented=False) as sg:
        dd = sg.orient()
        try:
            assert dd.info()[:2] == (d.n, d.m)
            assert dd.tc(gpm.TC_VERTEX) == gpm.DeviceGraph(d, oriented=True).tc(gpm.TC_VERTEX)
        finally:
            dd.close()


def test_error_paths(gpm):
    g = graphs.complete(10)
    with gpm.DeviceGraph(g, oriented=False) as sg:
        with pytest.raises(gpm.GpmError):
            sg.tc()  # needs an oriented graph
    with dev_dag(gpm, g) as dg:
        with pytest.raises(gpm.GpmError):
            dg.clique(6)
        with pytest.raises(gpm.GpmError):
            dg.clique(4, policy=gpm.SCHED_LEARNED)  # no costs given
        with pytest.raises(gpm.GpmError):
            dg.tc(7)
    bad = gpm.HostCSR(np.array([0, 1], dtype=np.uint32), np.array([5], dtype=np.uint32))
    with pytest.raises(gpm.GpmError):
        gpm.DeviceGraph(bad, oriented=True)


def test_tc_repeat_graph_replays(gpm, orc):
    """Repeated counts (a hipGraph of independent counts for the edge variant)
    equal the eager count on every replay, with eager work in between."""
    g = gpm.er_graph(4096, 40000, 11)
    d = gpm.orient(g)
    want = orc.tc(d.offsets, d.adj)
    with gpm.DeviceGraph(d, oriented=True) as dg:
        for _ in range(3):
            assert dg.tc_repeat(gpm.TC_EDGE, 64) == want
            assert dg.tc(gpm.TC_EDGE) == want
            assert dg.tc_repeat(gpm.TC_EDGE, 5) == want
            assert dg.tc_repeat(gpm.TC_VERTEX, 3) == want
            assert dg.clique(3) == want
        with pytest.raises(gpm.GpmError):
            dg.tc_repeat(gpm.TC_EDGE, 0)


def test_probe_paths_closed_form(gpm):
    """Roots whose out-neighbours have long lists (> max(512, 8 d)) build
    their rows by binary search instead of streaming: small vertices (wave
    tier) and medium vertices (LDS tier) hanging off a 700-clique.  With no
    edges between non-clique vertices the count is closed-form:
    C(700, k) + sum over non-clique v of C(t_v, k - 1), t_v = hubs of v."""
    rng = np.random.default_rng(2024)
    h, n_med, n_small = 700, 100, 2000
    src, dst = [], []
    for a in range(h):
        for b in range(a + 1, h):
            src.append(a)
            dst.append(b)
    ts = []
    for v in range(h, h + n_med + n_small):
        t = int(rng.integers(70, 90)) if v < h + n_med else int(rng.integers(4, 13))
        for hub in rng.choice(h, size=t, replace=False):
            src.append(v)
            dst.append(int(hub))
        ts.append(t)
    g = gpm.csr_from_edges(h + n_med + n_small, src, dst, gpm.CSR_CLEAN)
    with gpm.DeviceGraph(gpm.orient(g), oriented=True) as dg:
        for k in (3, 4, 5):
            want = comb(h, k) + sum(comb(t, k - 1) for t in ts)
            assert dg.clique(k) == want
            deg = np.diff(gpm.orient(g).offsets.astype(np.int64)).astype(np.float32)
            assert dg.clique(k, policy=gpm.SCHED_LEARNED, cost=deg ** 2) == want
        assert dg.tc(gpm.TC_EDGE) == comb(h, 3) + sum(comb(t, 2) for t in ts)


def test_empty_and_tiny_graphs(gpm):
    """Edge cases: no edges, a single vertex, isolated vertices, one edge."""
    cases = [
        (gpm.HostCSR(np.array([0], dtype=np.uint32), np.zeros(0, dtype=np.uint32)), 0),
        (gpm.HostCSR(np.array([0, 0], dtype=np.uint32), np.zeros(0, dtype=np.uint32)), 0),
        (gpm.HostCSR(np.zeros(101, dtype=np.uint32), np.zeros(0, dtype=np.uint32)), 0),
        (gpm.csr_from_edges(5, [0], [1], gpm.CSR_CLEAN), 0),
        (graphs.complete(3), 1),
    ]
    for g, tri in cases:
        with gpm.DeviceGraph(gpm.orient(g), oriented=True) as dg:
            assert dg.tc(gpm.TC_EDGE) == tri
            assert dg.tc(gpm.TC_VERTEX) == tri
            for k in (3, 4, 5):
                assert dg.clique(k) == (tri if k == 3 else 0)
            assert dg.tc_repeat(gpm.TC_EDGE, 40) == tri
        with gpm.DeviceGraph(g, oriented=False) as sg:
            counts, _ = sg.pattern("Motif4", gpm.FORM_PATTERN)
            assert counts == [0, 0, 0, 0, 0, 0]


def test_clique_all_matches_single_k(gpm, orc):
    """The fused k = 3, 4, 5 pass equals the per-k counts (every tier,
    both policies, probe paths, K_n closed forms)."""
    graphs_ = [gpm.orient(gpm.rmat_graph(14, 16, 31)), gpm.orient(graphs.complete(300)),
               gpm.orient(graphs.complete(1100)), gpm.orient(graphs.turan(130, 6))]
    for d in graphs_:
        with gpm.DeviceGraph(d, oriented=True) as dg:
            want = tuple(dg.clique(k) for k in (3, 4, 5)) if d.n != 1100 else \
                (comb(1100, 3), comb(1100, 4), comb(1100, 5))
            got = dg.clique_all()
            assert got == want
            deg = d.degrees().astype(np.float32)
            assert dg.clique_all(gpm.SCHED_LEARNED, deg ** 3) == got
    g = gpm.rmat_graph(12, 16, 1012)
    d = gpm.orient(g)
    with gpm.DeviceGraph(d, oriented=True) as dg:
        assert dg.clique_all() == tuple(orc.clique(d.offsets, d.adj, k) for k in (3, 4, 5))
