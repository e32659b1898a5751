"""GPU checks of the learned scheduler's pieces: per-vertex features (exact
integer sums held in fp32) and the BF16-MFMA MLP against two references — a
numpy model that rounds the same operands to bf16 (tight: pins the matrix-core
operand layouts, where a transposed or misplaced operand shifts every output)
and a plain fp32 model (loose: bf16 quantisation error only)."""
import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu


def to_bf16(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def mlp_ref(feat, p, bf16, mass=False):
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float32))
    x = r((np.log1p(feat.astype(np.float64)) - p["mean"]) / p["std"]).astype(np.float64)
    h1 = np.maximum(x @ r(p["w1"]).astype(np.float64).T + p["b1"], 0.0)
    h1 = r(h1).astype(np.float64)
    h2 = np.maximum(h1 @ r(p["w2"]).astype(np.float64).T + p["b2"], 0.0)
    w3 = p["w3"].astype(np.float64).reshape(64)
    out = h2 @ w3 + p["b3"][0]
    # mass = sum of |terms| of the last dot product (what rounding errors scale with)
    return (out, np.abs(h2) @ np.abs(w3)) if mass else out


def random_params(seed):
    rng = np.random.default_rng(seed)
    return dict(
        mean=rng.normal(0, 1, 4).astype(np.float32),
        std=rng.uniform(0.5, 2.0, 4).astype(np.float32),
        w1=rng.normal(0, 0.5, (64, 4)).astype(np.float32),
        b1=rng.normal(0, 0.1, 64).astype(np.float32),
        w2=rng.normal(0, 0.125, (64, 64)).astype(np.float32),
        b2=rng.normal(0, 0.1, 64).astype(np.float32),
        w3=rng.normal(0, 0.125, (1, 64)).astype(np.float32),
        b3=rng.normal(0, 0.1, 1).astype(np.float32),
    )


def predict(gpm, feat, p):
    return gpm.mlp_predict(feat, p["mean"], p["std"], p["w1"], p["b1"], p["w2"], p["b2"],
                           p["w3"], p["b3"])


def host_features(sym, dag, labels):
    deg = np.diff(sym.offsets.astype(np.int64))
    dout = np.diff(dag.offsets.astype(np.int64))
    src = np.repeat(np.arange(dag.n), dout)
    hop = np.bincount(src, weights=dout[dag.adj.astype(np.int64)], minlength=dag.n)
    return np.stack([deg, dout, hop, labels], axis=1).astype(np.float32)


def test_vertex_features(gpm):
    g = gpm.rmat_graph(13, 16, 777)
    d = gpm.orient(g)
    labels = gpm.gen_labels(g.n, 8, 5)
    with gpm.DeviceGraph(g, oriented=False, labels=labels, n_labels=8) as sym, \
            gpm.DeviceGraph(d, oriented=True) as dag:
        got = gpm.vertex_features(sym, dag)
        want = host_features(g, d, labels)
        assert np.array_equal(got, want)
        # Without a DAG the 2-hop column sums symmetric degrees.
        got = gpm.vertex_features(sym, None)
        deg = np.diff(g.offsets.astype(np.int64))
        src = np.repeat(np.arange(g.n), deg)
        hop = np.bincount(src, weights=deg[g.adj.astype(np.int64)], minlength=g.n)
        assert np.array_equal(got[:, 1], np.zeros(g.n, np.float32))
        assert np.array_equal(got[:, 2], hop.astype(np.float32))


@pytest.mark.parametrize("n", [1, 15, 16, 37, 4096 + 5])
def test_mlp_matches_references(gpm, n):
    p = random_params(100 + n)
    rng = np.random.default_rng(n)
    feat = np.stack([rng.integers(0, 3000, n), rng.integers(0, 700, n),
                     rng.integers(0, 10 ** 6, n), rng.integers(0, 8, n)], axis=1).astype(np.float32)
    got = predict(gpm, feat, p).astype(np.float64)
    ref16 = mlp_ref(feat, p, bf16=True)
    ref32, mass = mlp_ref(feat, p, bf16=False, mass=True)
    # Same bf16 operands, fp32 accumulation in another order: 1e-4 of the
    # mean output magnitude.
    scale = np.abs(ref32).mean() + 1e-3
    assert np.max(np.abs(got - ref16)) <= 1e-4 * scale + 1e-5
    # vs fp32: x, W1, h1 and W2 are each rounded to bf16 (unit roundoff 2^-8),
    # so every term of the last dot product carries a relative error of a few
    # 2^-8; bound each row by 2^-4 of its term mass.
    assert np.all(np.abs(got - ref32) <= 2.0 ** -4 * mass + 1e-4)


def test_mlp_errors(gpm):
    p = random_params(1)
    bad = dict(p, std=np.array([1, 0, 1, 1], np.float32))
    with pytest.raises(gpm.GpmError):
        predict(gpm, np.zeros((4, 4), np.float32), bad)
    with pytest.raises(gpm.GpmError):
        predict(gpm, np.full((4, 4), np.nan, np.float32), p)
    assert predict(gpm, np.zeros((0, 4), np.float32), p).shape == (0,)


def test_trained_model_orders_roots(gpm):
    """The shipped cost model ranks the roots of an unseen graph by logged
    work: Spearman correlation of predicted cost vs plan entries."""
    from cost_model import CostModel
    cm = CostModel.load_default(5)
    g = gpm.rmat_graph(16, 16, 4242)
    d = gpm.orient(g)
    with gpm.DeviceGraph(g, oriented=False) as sym, gpm.DeviceGraph(d, oriented=True) as dag:
        cost = cm.predict(sym, dag)
        work = dag.clique_root_work(5).astype(np.float64)
    keep = work > 0
    rc = np.argsort(np.argsort(cost[keep]))
    rw = np.argsort(np.argsort(work[keep]))
    rho = np.corrcoef(rc, rw)[0, 1]
    assert rho > 0.8, rho


@pytest.mark.parametrize("k", [3, 4, 5])
def test_learned_split_counts_identical(gpm, k):
    """LEARNED orders roots by cost and deals heavy ones out in parts (every
    n-th row of the root's local DAG per part); counts must not change."""
    g = gpm.rmat_graph(15, 16, 99)
    d = gpm.orient(g)
    deg = np.diff(d.offsets.astype(np.int64)).astype(np.float32)
    rng = np.random.default_rng(k)
    with gpm.DeviceGraph(d, oriented=True) as dag:
        plain = dag.clique(k)
        for cost in (deg ** 3, deg, rng.random(d.n).astype(np.float32) * 1e6,
                     np.zeros(d.n, np.float32)):
            assert dag.clique(k, policy=gpm.SCHED_LEARNED, cost=cost) == plain
        # the degree-estimate split (same splitting, cost d+^2, no model)
        assert dag.clique(k, policy=gpm.SCHED_DEGREE_SPLIT) == plain
        assert dag.clique_all(policy=gpm.SCHED_DEGREE_SPLIT)[k - 3] == plain


@pytest.mark.parametrize("n", [40, 300, 700, 1100])
def test_learned_split_complete(gpm, n):
    # One root of K_n carries the whole count: it is split into the most
    # parts, in every block tier (LDS, wide LDS, HBM rows).
    from math import comb
    with gpm.DeviceGraph(gpm.orient(graphs.complete(n)), oriented=True) as dag:
        cost = np.zeros(n, np.float32)
        cost[np.argmax(np.diff(dag_offsets(gpm, n)))] = 1e9
        for k in (3, 4) + ((5,) if n <= 300 else ()):
            assert dag.clique(k, policy=gpm.SCHED_LEARNED, cost=cost) == comb(n, k)


def dag_offsets(gpm, n):
    return gpm.orient(graphs.complete(n)).offsets.astype(np.int64)
