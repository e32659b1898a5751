"""GPU parity at the benchmark configurations' full sizes.

Config 1 (triangle counting on Erdos-Renyi n=16384, M=131072) is
tests/test_gpu_tc_clique.py::test_tc_er_workload.  Here:

  config 2  k-clique k=3,4,5 on R-MAT scale 20 (seed 20240520): the fused
            pass equals the single-k passes; k=3 and k=4 equal the oracle's
            full counts (tests/golden/config_counts.json, written by
            tests/golden/make_config_counts.py), k=5 too when that count is
            present; k=4 and k=5 equal the oracle on a seeded uniform random
            sample of roots; the learned schedule gives the same counts.
  config 5  R-MAT scale 22 (seed 20240522): 5-clique learned == plain ==
            degree-split == the oracle's full count (15 548 s on 6 threads);
            triangles and 4-cliques equal the oracle's full counts; the induced Motif4 counts satisfy the non-induced
            identities checked against independently computed totals, and at
            scale 16 (the largest the oracle finishes in minutes) they equal
            the oracle's counts.
  config 3  the five named queries (triangle, 4-cycle, diamond, three-star,
            pyramid-6) on labeled R-MAT: oracle parity at scale 12; plain,
            learned and fixed-chunk graph replay agree at scale 18 for all
            five and equal the oracle's full counts there (pyramid-6 there: 6.0e10 embeddings, about 70 s per run) and
            at scale 16.
  config 4  256 seeded connected queries of 4-10 vertices (8 labels) as one
            batch on labeled R-MAT: oracle parity per query at scale 10
            under the plain and learned orders, both probe variants and
            fixed-chunk graph replay; at the bench's scale (11) the orders
            and pass budgets agree.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD_PATH = os.path.join(ROOT, "tests", "golden", "config_counts.json")


def gold():
    with open(GOLD_PATH) as f:
        return json.load(f)


def test_config2_clique_rmat20(gpm, orc, workloads):
    from cost_model import CostModel
    G = gold()
    g = workloads.clique_graph(20)
    d = gpm.orient(g)
    print(f"config 2: n={d.n} |E+|={d.m}", flush=True)
    with gpm.DeviceGraph(d, oriented=True) as dag:
        k3, k4, k5 = dag.clique_all()
        assert k3 == G["rmat20_k3"]["count"]
        assert k4 == G["rmat20_k4"]["count"]
        if "rmat20_k5" in G:
            assert k5 == G["rmat20_k5"]["count"]
        assert (dag.clique(3), dag.clique(4), dag.clique(5)) == (k3, k4, k5)
        assert dag.tc(gpm.TC_EDGE) == k3 and dag.tc(gpm.TC_VERTEX) == k3
        with gpm.DeviceGraph(g, oriented=False) as sym:
            cost = CostModel.load_default(5).predict(sym, dag)
        assert dag.clique_all(policy=gpm.SCHED_LEARNED, cost=cost) == (k3, k4, k5)
        assert dag.clique(5, policy=gpm.SCHED_LEARNED, cost=cost) == k5
        # Seeded uniform random roots: exact against the oracle on the same roots.
        rng = np.random.default_rng(20240520)
        roots = np.sort(rng.choice(d.n, 4000, replace=False)).astype(np.uint32)
        for k in (4, 5):
            assert dag.clique(k, roots=roots) == orc.clique_list(d.offsets, d.adj, k, roots), k


@pytest.fixture(scope="module")
def rmat22(gpm):
    g = gpm.rmat_graph(22, 16, 20240522)
    return g, gpm.orient(g)


def test_config5_clique_rmat22(gpm, rmat22):
    from cost_model import CostModel
    G = gold()
    g, d = rmat22
    print(f"config 5: n={d.n} |E+|={d.m}", flush=True)
    with gpm.DeviceGraph(g, oriented=False) as sym, gpm.DeviceGraph(d, oriented=True) as dag:
        cost = CostModel.load_default(5).predict(sym, dag)
        plain = dag.clique(5)
        assert dag.clique(5, policy=gpm.SCHED_LEARNED, cost=cost) == plain
        assert dag.clique_all(policy=gpm.SCHED_LEARNED, cost=cost)[2] == plain
        if "rmat22_k5" in G:
            assert plain == G["rmat22_k5"]["count"]
        assert dag.clique(3) == G["rmat22_k3"]["count"]
        if "rmat22_k4" in G:
            assert dag.clique(4) == G["rmat22_k4"]["count"]


def test_config5_motif4_rmat22_identities(gpm, rmat22):
    """Induced Motif4 counts at scale 22 against non-induced totals computed
    independently of the Motif4 kernels:
      sum_v C(d_v, 3)                       = star + tailed + 2 diamond + 4 K4
      sum_{uv} (d_u - 1)(d_v - 1) - 3 T     = path + 2 tailed + 4 cycle
                                              + 6 diamond + 12 K4
      4-cliques of the degree DAG (k-clique kernels) = K4
    with T the triangle count of the DAG."""
    g, d = rmat22
    with gpm.DeviceGraph(g, oriented=False) as sym, gpm.DeviceGraph(d, oriented=True) as dag:
        star, path, tailed, cycle, diamond, k4 = sym.pattern("Motif4", gpm.FORM_PATTERN)[0]
        tri = dag.clique(3)
        k4_dag = dag.clique(4)
    dd = g.degrees().astype(np.int64)  # products and sums below fit in int64
    stars = int(np.sum(dd * (dd - 1) * (dd - 2) // 6))
    src = np.repeat(np.arange(g.n, dtype=np.int64), dd)
    dst = g.adj.astype(np.int64)
    keep = src < dst  # each undirected edge once
    paths = int(np.sum((dd[src[keep]] - 1) * (dd[dst[keep]] - 1))) - 3 * tri
    assert stars == star + tailed + 2 * diamond + 4 * k4
    assert paths == path + 2 * tailed + 4 * cycle + 6 * diamond + 12 * k4
    assert k4 == k4_dag


def test_config5_motif4_rmat16_oracle(gpm):
    want = gold()["rmat16_s20240522_motif4"]["motif4_induced"]
    g = gpm.rmat_graph(16, 16, 20240522)
    with gpm.DeviceGraph(g, oriented=False) as sym:
        assert sym.pattern("Motif4", gpm.FORM_PATTERN)[0] == want
        assert sym.pattern("Motif4", gpm.FORM_G2MINER)[0] == want


def test_config3_named_queries_oracle_s12(gpm, orc, workloads):
    g, labels = workloads.labeled_graph(12)
    named = workloads.named_queries()
    qs = [q for _, q in named]
    want = orc.ceci_batch(g.offsets, g.adj, labels, qs)
    with gpm.DeviceGraph(g, oriented=False, labels=labels, n_labels=8) as dg:
        for kw in (dict(policy=0), dict(policy=1), dict(autotune=0, chunk=8, use_graph=1)):
            assert np.array_equal(dg.match_batch(qs, **kw), want), kw
            for i, q in enumerate(qs):  # one query per batch too
                assert int(dg.match_batch([q], **kw)[0]) == int(want[i]), (kw, named[i][0])


def test_config3_named_queries_s18(gpm, workloads):
    g, labels = workloads.labeled_graph(18)
    named = dict(workloads.named_queries())
    four = [named[q] for q in ("triangle", "4-cycle", "diamond", "three-star")]
    with gpm.DeviceGraph(g, oriented=False, labels=labels, n_labels=8) as dg:
        plain = dg.match_batch(four, policy=0)
        print("config 3 s18 counts", plain.tolist(), flush=True)
        G = gold()
        if "labeled18_named" in G:  # the oracle's full counts, when computed
            assert plain.tolist() == G["labeled18_named"]["counts"][:4]
        assert np.array_equal(dg.match_batch(four, policy=1), plain)
        assert np.array_equal(dg.match_batch(four, autotune=0, chunk=8, use_graph=1), plain)
    # pyramid-6 grows ~10x per two scales (3.1e9 embeddings at scale 16).
    g, labels = workloads.labeled_graph(16)
    with gpm.DeviceGraph(g, oriented=False, labels=labels, n_labels=8) as dg:
        pyr = [named["pyramid-6"]]
        p = dg.match_batch(pyr, policy=0)
        print("config 3 pyramid-6 s16", int(p[0]), flush=True)
        assert np.array_equal(dg.match_batch(pyr, policy=1), p)


@pytest.fixture(scope="module")
def pyramid_s18(gpm, workloads):
    """pyramid-6 on labeled R-MAT scale 18 (config 3's nominal scale): the
    graph handle and the plain count (6.0e10 embeddings, ~70 s)."""
    g, labels = workloads.labeled_graph(18)
    pyr = [dict(workloads.named_queries())["pyramid-6"]]
    dg = gpm.DeviceGraph(g, oriented=False, labels=labels, n_labels=8)
    plain = dg.match_batch(pyr, policy=0)
    print("config 3 pyramid-6 s18", int(plain[0]), dg.last_kernel_ms(), "ms", flush=True)
    G = gold()
    if "labeled18_named" in G:  # the oracle's full count, when computed
        assert int(plain[0]) == G["labeled18_named"]["counts"][4]
    yield dg, pyr, plain
    dg.close()


def test_config3_pyramid_s18_learned(pyramid_s18):
    dg, pyr, plain = pyramid_s18
    assert np.array_equal(dg.match_batch(pyr, policy=1), plain)


def test_config3_pyramid_s18_graph_replay(pyramid_s18):
    dg, pyr, plain = pyramid_s18
    assert np.array_equal(dg.match_batch(pyr, autotune=0, chunk=8, use_graph=1), plain)


def test_config4_batch_bench_scale(gpm, workloads):
    """Config 4 at the bench's scale (11: 4.5e10 embeddings, one query
    3.5e10 of them; the oracle would need minutes): the plain and learned
    orders and a small explicit pass budget (tasks cut down several levels)
    give identical per-query counts."""
    g, labels = workloads.labeled_graph(workloads.BATCH_SCALE)
    qs = workloads.batch_queries()
    with gpm.DeviceGraph(g, oriented=False, labels=labels, n_labels=8) as dg:
        ref = dg.match_batch(qs, policy=1)
        print("config 4 at scale", workloads.BATCH_SCALE, int(ref.sum()), flush=True)
        G = gold()
        key = f"labeled{workloads.BATCH_SCALE}_batch256"
        if key in G:  # the oracle's per-query counts, when computed
            assert ref.tolist() == G[key]["counts"]
        assert int(ref.sum()) > 4 * 10 ** 10 and np.count_nonzero(ref) > 200
        for kw in (dict(policy=0), dict(policy=1, pass_budget=64)):
            got = dg.match_batch(qs, **kw)
            print("config 4", kw, dg.last_kernel_ms(), "ms", flush=True)
            assert np.array_equal(got, ref), kw


def test_config4_batch_oracle(gpm, orc, workloads):
    """Config 4's workload (256 seeded connected 4-10-vertex queries, 8
    labels, one batch) at the bench's scale, which the oracle enumerates in
    seconds: every per-query count equals the oracle's under the plain and
    learned orders, both membership-probe variants and fixed-chunk graph
    replay."""
    g, labels = workloads.labeled_graph(workloads.BATCH_ORACLE_SCALE)
    qs = workloads.batch_queries()
    want = orc.ceci_batch(g.offsets, g.adj, labels, qs)
    print("config 4 oracle total", int(want.sum()), flush=True)
    assert np.count_nonzero(want) > 200
    with gpm.DeviceGraph(g, oriented=False, labels=labels, n_labels=8) as dg:
        for kw in (dict(policy=0), dict(policy=1), dict(policy=1, set_variant=1),
                   dict(policy=1, set_variant=2), dict(autotune=0, chunk=8, use_graph=1)):
            got = dg.match_batch(qs, **kw)
            print("config 4", kw, dg.last_kernel_ms(), "ms", flush=True)
            assert np.array_equal(got, want), kw

