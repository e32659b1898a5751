"""Pins the CECI restatement: the reference's 200 known answers, networkx
subgraph monomorphisms, and the reference matcher built from source."""
import os

import numpy as np
import pytest

import graphs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ceci_hprd.npz")


def load_golden(gpm):
    z = np.load(GOLDEN)
    n = int(z["data_n"])
    e = z["data_edges"]
    g = gpm.csr_from_edges(n, e[:, 0], e[:, 1], gpm.CSR_CLEAN)
    qs, eo, lo = [], 0, 0
    for i in range(len(z["q_n"])):
        qn, qm = int(z["q_n"][i]), int(z["q_m"][i])
        qs.append((qn, z["q_edges"][eo:eo + 2 * qm].reshape(-1, 2), z["q_labels"][lo:lo + qn]))
        eo += 2 * qm
        lo += qn
    return g, z["data_labels"], qs, z["expected"].astype(np.uint64), list(z["names"])


def test_ceci_reference_known_answers(gpm, orc):
    g, labels, qs, expected, _ = load_golden(gpm)
    got = orc.ceci_batch(g.offsets, g.adj, labels, qs)
    assert len(got) == 200
    assert np.array_equal(got, expected)
    # Single-query entry point agrees with the batch one.
    for i in (0, 1, 57, 199):
        assert orc.ceci(g.offsets, g.adj, labels, *qs[i]) == expected[i]


def _nx_count(g, labels, q):
    nx = pytest.importorskip("networkx")
    from networkx.algorithms import isomorphism as iso
    G = graphs.to_nx(g)
    for v in range(g.n):
        G.nodes[v]["l"] = int(labels[v])
    qn, qe, ql = q
    Q = nx.Graph()
    Q.add_nodes_from(range(qn))
    Q.add_edges_from([tuple(map(int, e)) for e in qe])
    for v in range(qn):
        Q.nodes[v]["l"] = int(ql[v])
    m = iso.GraphMatcher(G, Q, node_match=lambda a, b: a["l"] == b["l"])
    return sum(1 for _ in m.subgraph_monomorphisms_iter())


@pytest.mark.parametrize("seed", [3, 4])
def test_ceci_vs_networkx(gpm, orc, seed):
    rng = np.random.default_rng(seed)
    g = graphs.er(60, 240, seed)
    labels = rng.integers(0, 3, g.n).astype(np.uint32)
    shapes = [
        (3, [(0, 1), (1, 2), (0, 2)]),
        (4, [(0, 1), (1, 2), (2, 3), (3, 0)]),
        (4, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3)]),
        (4, [(0, 1), (0, 2), (0, 3)]),
        (5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 2)]),
    ]
    for qn, qe in shapes:
        ql = rng.integers(0, 3, qn).astype(np.uint32)
        q = (qn, np.asarray(qe, dtype=np.uint32), ql)
        assert orc.ceci(g.offsets, g.adj, labels, *q) == _nx_count(g, labels, q)


def test_ceci_vs_reference_binary(gpm, orc, tmp_path):
    if not orc.ref_ceci_available():
        pytest.skip("reference CECI matcher not built here (needs the reference checkout)")
    rng = np.random.default_rng(9)
    g = graphs.er(300, 1500, 9)
    labels = rng.integers(0, 4, g.n).astype(np.uint32)

    def write(path, n, edges, lab):
        deg = np.zeros(n, dtype=np.int64)
        for a, b in edges:
            deg[a] += 1
            deg[b] += 1
        with open(path, "w") as f:
            f.write(f"t {n} {len(edges)}\n")
            for v in range(n):
                f.write(f"v {v} {lab[v]} {deg[v]}\n")
            for a, b in edges:
                f.write(f"e {a} {b}\n")

    dedges = [(u, int(v)) for u in range(g.n) for v in g.adj[g.offsets[u]:g.offsets[u + 1]] if u < v]
    dpath = tmp_path / "data.graph"
    write(dpath, g.n, dedges, labels)
    for qi, (qn, qe) in enumerate([(3, [(0, 1), (1, 2), (0, 2)]),
                                   (4, [(0, 1), (1, 2), (2, 3)]),
                                   (5, [(0, 1), (1, 2), (2, 3), (3, 4), (1, 3)])]):
        ql = rng.integers(0, 4, qn).astype(np.uint32)
        qpath = tmp_path / f"q{qi}.graph"
        write(qpath, qn, qe, ql)
        want = orc.ref_ceci(str(dpath), str(qpath))
        assert orc.ceci(g.offsets, g.adj, labels, qn, np.asarray(qe, dtype=np.uint32), ql) == want
