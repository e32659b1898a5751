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


def test_ceci_bin_byte_layout(gpm, tmp_path):
    """The compressed format as sub_ceci/graph/graph.cpp:184-304 reads it:
    int32 4, |V|, |E|, degrees; lists back to back (sorted on load); labels."""
    import struct
    # path 0-1-2 plus triangle 2-3-4: lists written unsorted on purpose
    lists = [[1], [2, 0], [4, 3, 1], [2, 4], [3, 2]]
    deg = str(tmp_path / "g_deg.bin")
    adj = str(tmp_path / "g_adj.bin")
    lab = str(tmp_path / "g_label.bin")
    with open(deg, "wb") as f:
        f.write(struct.pack("<iII", 4, 5, 5) + struct.pack("<5I", *[len(l) for l in lists]))
    with open(adj, "wb") as f:
        for l in lists:
            f.write(struct.pack(f"<{len(l)}I", *l))
    with open(lab, "wb") as f:
        f.write(struct.pack("<5I", 3, 0, 7, 0, 3))
    g, labels, nl = gpm.load_ceci_bin(deg, adj, lab)
    assert g.offsets.tolist() == [0, 1, 3, 6, 8, 10]
    assert g.adj.tolist() == [1, 0, 2, 1, 3, 4, 2, 4, 2, 3]
    assert labels.tolist() == [3, 0, 7, 0, 3] and nl == 8
    # Writer round trip gives the same bytes for sorted input.
    gpm.store_ceci_bin(str(tmp_path / "r_deg.bin"), str(tmp_path / "r_adj.bin"), g, labels,
                       str(tmp_path / "r_label.bin"))
    assert open(str(tmp_path / "r_deg.bin"), "rb").read() == open(deg, "rb").read()
    g2, l2, _ = gpm.load_ceci_bin(str(tmp_path / "r_deg.bin"), str(tmp_path / "r_adj.bin"),
                                  str(tmp_path / "r_label.bin"))
    assert np.array_equal(g2.adj, g.adj) and np.array_equal(l2, labels)


def test_ceci_bin_roundtrip_hprd_and_rejects(gpm, tmp_path):
    from test_ceci import load_golden
    g, labels, _, _, _ = load_golden(gpm)
    d, a, l = (str(tmp_path / x) for x in ("d.bin", "a.bin", "l.bin"))
    gpm.store_ceci_bin(d, a, g, labels, l)
    g2, l2, nl = gpm.load_ceci_bin(d, a, l)
    assert np.array_equal(g2.offsets, g.offsets) and np.array_equal(g2.adj, g.adj)
    assert np.array_equal(l2, labels) and nl == int(labels.max()) + 1
    g3, l3, _ = gpm.load_ceci_bin(d, a)  # no label file
    assert l3 is None and np.array_equal(g3.adj, g.adj)
    # A truncated adjacency file and a wrong int size are rejected.
    raw = open(a, "rb").read()
    open(a, "wb").write(raw[:-4])
    with pytest.raises(gpm.GpmError):
        gpm.load_ceci_bin(d, a, l)
    raw = bytearray(open(d, "rb").read())
    raw[0] = 8
    open(d, "wb").write(bytes(raw))
    with pytest.raises(gpm.GpmError):
        gpm.load_ceci_bin(d, a, l)


def test_gbin_byte_layout_and_roundtrip(gpm, tmp_path):
    """GraphMiner binary (scripts/gbin_reader/greader.h:129-272): int64 row
    pointers, int32 ids in input order, meta |V| nnz 4 max_degree."""
    import struct
    d = tmp_path / "gm"
    d.mkdir()
    rows = [[2, 1], [], [0], [2, 0, 1]]
    rp = [0]
    for r in rows:
        rp.append(rp[-1] + len(r))
    (d / "graph.vertex.bin").write_bytes(struct.pack(f"<{len(rp)}q", *rp))
    flat = [x for r in rows for x in r]
    (d / "graph.edge.bin").write_bytes(struct.pack(f"<{len(flat)}i", *flat))
    (d / "graph.meta.txt").write_text("4\n6\n4\n3\n")
    g, maxd = gpm.load_gbin(str(d))
    assert g.offsets.tolist() == rp and g.adj.tolist() == flat and maxd == 3
    e = tmp_path / "out"
    e.mkdir()
    gpm.store_gbin(str(e), g)
    for name in ("graph.vertex.bin", "graph.edge.bin", "graph.meta.txt"):
        assert (e / name).read_bytes() == (d / name).read_bytes(), name
    big = gpm.rmat_graph(10, 8, 5)
    gpm.store_gbin(str(e), big)
    g2, maxd2 = gpm.load_gbin(str(e))
    assert np.array_equal(g2.offsets, big.offsets) and np.array_equal(g2.adj, big.adj)
    assert maxd2 == int(big.degrees().max())
    (e / "graph.meta.txt").write_text(f"{big.n}\n{big.m}\n4\n{maxd2 + 1}\n")
    with pytest.raises(gpm.GpmError):  # max degree disagrees with the data
        gpm.load_gbin(str(e))
    with pytest.raises(gpm.GpmError):
        gpm.load_gbin(str(tmp_path / "missing"))
