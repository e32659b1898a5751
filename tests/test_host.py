"""Host-side graph preparation of libgpm (no GPU needed): CSR build,
orientation rule, degree relabel, mtx reader, generators."""
import os

import numpy as np
import pytest

import graphs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_csr_from_edges_clean(gpm):
    src = [0, 1, 1, 2, 2, 3, 3]
    dst = [1, 0, 2, 2, 0, 1, 1]
    g = gpm.csr_from_edges(4, src, dst, gpm.CSR_CLEAN)
    lists = [g.adj[g.offsets[v]:g.offsets[v + 1]].tolist() for v in range(4)]
    assert lists == [[1, 2], [0, 2, 3], [0, 1], [1]]


def test_csr_from_edges_input_order(gpm):
    # flags == 0: Graph::Init's CSR, lists in input order, nothing dropped.
    g = gpm.csr_from_edges(3, [2, 0, 2, 0], [1, 2, 0, 2], 0)
    assert g.offsets.tolist() == [0, 2, 2, 4]
    assert g.adj.tolist() == [2, 2, 1, 0]


def test_csr_rejects_bad_ids(gpm):
    with pytest.raises(gpm.GpmError):
        gpm.csr_from_edges(3, [0, 5], [1, 1], gpm.CSR_CLEAN)


def test_orient_rule(gpm):
    g = graphs.er(300, 2000, 4)
    d = gpm.orient(g)
    deg = g.degrees()
    want = []
    for u in range(g.n):
        for v in g.adj[g.offsets[u]:g.offsets[u + 1]]:
            if deg[v] > deg[u] or (deg[v] == deg[u] and v > u):
                want.append((u, int(v)))
    got = [(u, int(v)) for u in range(d.n) for v in d.adj[d.offsets[u]:d.offsets[u + 1]]]
    assert got == want
    assert d.m * 2 == g.m


def test_sort_by_degree(gpm):
    g = graphs.er(200, 900, 8)
    d = gpm.orient(g)
    s, perm = gpm.sort_by_degree(d)
    deg = d.degrees()
    order = sorted(range(d.n), key=lambda v: -deg[v])  # stable, descending
    assert [int(perm[v]) for v in order] == list(range(d.n))
    for i, v in enumerate(order):
        lst = sorted(int(perm[x]) for x in d.adj[d.offsets[v]:d.offsets[v + 1]])
        assert s.adj[s.offsets[i]:s.offsets[i + 1]].tolist() == lst


def test_load_mtx_reference_fixture(gpm):
    # datasets/simple.mtx of the reference: header "4 4 6", six pairs, the
    # last line unterminated; ids rebased by the minimum (1).
    n, s, d = gpm.load_mtx(os.path.join(GOLDEN, "simple.mtx"))
    assert n == 4
    pairs = list(zip(s.tolist(), d.tolist()))
    with open(os.path.join(GOLDEN, "simple.mtx"), "rb") as f:
        ends_with_newline = f.read().endswith(b"\n")
    full = [(0, 1), (1, 2), (2, 3), (1, 0), (2, 1), (3, 2)]
    # io.h:113-116 drops a final pair whose line hits EOF before a newline.
    assert pairs == (full if ends_with_newline else full[:-1])


def test_load_mtx_comments_self_loops(gpm, tmp_path):
    p = tmp_path / "g.mtx"
    p.write_text("%%MatrixMarket matrix coordinate\n% comment\n9 9 4\n5 6\n6 6\n7 5\n8 9\n")
    n, s, d = gpm.load_mtx(str(p))
    assert n == 5  # ids 5..9 rebased to 0..4
    assert list(zip(s.tolist(), d.tolist())) == [(0, 1), (2, 0), (3, 4)]


def test_load_mtx_missing_file(gpm):
    with pytest.raises(gpm.GpmError):
        gpm.load_mtx("/nonexistent/file.mtx")


def test_generators_deterministic(gpm):
    a = gpm.gen_rmat(12, 8, 42)
    b = gpm.gen_rmat(12, 8, 42)
    c = gpm.gen_rmat(12, 8, 43)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])
    assert len(a[0]) == 8 << 12
    s, d = gpm.gen_er(1000, 4000, 3)
    assert len(s) == 4000 and np.all(s < d)
    assert len(set(zip(s.tolist(), d.tolist()))) == 4000
    l1, l2 = gpm.gen_labels(5000, 8, 1), gpm.gen_labels(5000, 8, 1)
    assert np.array_equal(l1, l2) and l1.max() == 7 and l1.min() == 0


def test_rmat_skew(gpm):
    # a = 0.57 puts the heaviest vertex at id 0.
    g = gpm.rmat_graph(12, 16, 7)
    deg = g.degrees()
    assert deg[0] == deg.max()
    assert np.all(np.diff(g.offsets.astype(np.int64)) >= 0)


# ---------------------------------------------------------------- sub_ceci text format

def test_load_labeled_fixture(gpm):
    g, labels, L = gpm.load_labeled(os.path.join(GOLDEN, "tiny_labeled.graph"))
    assert g.offsets.tolist() == [0, 2, 5, 8, 11, 12]
    assert g.adj.tolist() == [1, 2, 0, 2, 3, 0, 1, 3, 1, 2, 4, 3]
    assert labels.tolist() == [1, 0, 2, 1, 0] and L == 3


def test_load_labeled_roundtrip_hprd(gpm, tmp_path):
    from test_ceci import load_golden
    g, labels, _, _, _ = load_golden(gpm)
    p = str(tmp_path / "hprd.graph")
    graphs.write_labeled(p, g, labels)
    h, lab, L = gpm.load_labeled(p)
    assert np.array_equal(h.offsets, g.offsets) and np.array_equal(h.adj, g.adj)
    assert np.array_equal(lab, labels) and L == int(labels.max()) + 1


@pytest.mark.parametrize("text", [
    "x 2 1\nv 0 0 1\nv 1 0 1\ne 0 1\n",          # bad header
    "t 2 1\nv 1 0 1\nv 0 0 1\ne 0 1\n",          # vertices out of order
    "t 2 1\nv 0 0 1\nv 1 0 2\ne 0 1\n",          # declared degree not met
    "t 2 1\nv 0 0 1\nv 1 0 1\ne 0 1\ne 0 1\n",   # more edges than declared
    "t 2 1\nv 0 0 1\nv 1 0 1\ne 0 0\n",          # self loop
    "t 3 1\nv 0 0 1\nv 1 0 1\ne 0 1\n",          # fewer vertices than header
])
def test_load_labeled_rejects(gpm, tmp_path, text):
    p = tmp_path / "bad.graph"
    p.write_text(text)
    with pytest.raises(gpm.GpmError):
        gpm.load_labeled(str(p))
    with pytest.raises(gpm.GpmError):
        gpm.load_labeled(str(tmp_path / "missing.graph"))
