"""GPU graph preparation (gpm_graph_build, gpm_graph_relabel_by_degree)
equals the host passes it replaces (gpm_csr_from_edges, gpm_csr_sort_by_degree,
both pinned in tests/test_host.py) bit for bit, and graphs built on the device
count the same as graphs uploaded from the host."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _edges(seed, n, e, self_frac=0.05):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, n, e).astype(np.uint32)
    d = rng.integers(0, n, e).astype(np.uint32)
    k = rng.random(e) < self_frac
    d[k] = s[k]
    dup = rng.integers(0, e, e // 10)
    return np.concatenate([s, s[dup]]), np.concatenate([d, d[dup]])


FLAGS = [0, 0x1, 0x2, 0x4, 0x8, 0x5, 0x9, 0xF, 0x3, 0xC, 0xD]


@pytest.mark.parametrize("flags", FLAGS)
def test_build_matches_host(gpm, flags):
    for seed, (n, e) in enumerate([(1, 0), (5, 3), (300, 2000), (5000, 60000)]):
        s, d = _edges(seed, n, e)
        want = gpm.csr_from_edges(n, s, d, flags)
        with gpm.DeviceGraph.build(n, s, d, flags) as dg:
            got = dg.to_host()
            assert dg.info()[:2] == (n, want.m)
        assert np.array_equal(got.offsets, want.offsets), (flags, n)
        assert np.array_equal(got.adj, want.adj), (flags, n)


def test_build_rmat_counts(gpm, orc):
    s, d = gpm.gen_rmat(14, 16, 77)
    host = gpm.csr_from_edges(1 << 14, s, d)
    with gpm.DeviceGraph.build(1 << 14, s, d) as sym:
        got = sym.to_host()
        assert np.array_equal(got.offsets, host.offsets) and np.array_equal(got.adj, host.adj)
        hd = gpm.orient(host)
        with sym.orient() as dag:
            assert dag.tc(gpm.TC_EDGE) == orc.tc(hd.offsets, hd.adj)
        assert sym.pattern("Motif4", gpm.FORM_PATTERN)[0] == \
            gpm.DeviceGraph(host, oriented=False).pattern("Motif4", gpm.FORM_PATTERN)[0]


def test_relabel_matches_host(gpm):
    for g in (gpm.rmat_graph(12, 8, 3), gpm.orient(gpm.rmat_graph(11, 16, 4)),
              gpm.csr_from_edges(6, [], [], gpm.CSR_CLEAN)):
        want, want_perm = gpm.sort_by_degree(g)
        with gpm.DeviceGraph(g, oriented=False) as dg:
            r, perm = dg.relabel_by_degree()
            with r:
                got = r.to_host()
        assert np.array_equal(perm, want_perm)
        assert np.array_equal(got.offsets, want.offsets) and np.array_equal(got.adj, want.adj)


def test_build_labeled_matches(gpm):
    s, d = gpm.gen_rmat(11, 8, 5)
    n = 1 << 11
    labels = gpm.gen_labels(n, 4, 9)
    qs = [(3, np.array([[0, 1], [1, 2], [0, 2]], dtype=np.uint32), np.array([0, 1, 2], np.uint32)),
          (4, np.array([[0, 1], [1, 2], [2, 3], [3, 0]], dtype=np.uint32), np.zeros(4, np.uint32))]
    host = gpm.csr_from_edges(n, s, d)
    with gpm.DeviceGraph(host, oriented=False, labels=labels, n_labels=4) as a, \
            gpm.DeviceGraph.build(n, s, d, labels=labels, n_labels=4) as b:
        assert np.array_equal(a.match_batch(qs), b.match_batch(qs))
        r, perm = b.relabel_by_degree()
        with r:  # matching counts do not depend on vertex ids
            assert np.array_equal(r.match_batch(qs), a.match_batch(qs))
