"""Host side of the learned work estimator: the numpy trainer fits a known
monotone cost law, and weights survive a save/load round trip."""
import numpy as np

from cost_model import CostModel, FIELDS, fit


def synthetic(n, seed):
    rng = np.random.default_rng(seed)
    f = np.stack([rng.integers(1, 3000, n), rng.integers(0, 700, n),
                  rng.integers(0, 10 ** 6, n), np.zeros(n)], axis=1).astype(np.float32)
    w = (1.5 * f[:, 1] ** 2 + 0.1 * f[:, 2]).astype(np.uint64)
    return f, w


def forward(p, feat):
    x = (np.log1p(feat.astype(np.float64)) - p["mean"]) / p["std"]
    h1 = np.maximum(x @ p["w1"].T + p["b1"], 0)
    h2 = np.maximum(h1 @ p["w2"].T + p["b2"], 0)
    return (h2 @ p["w3"].T)[:, 0] + p["b3"][0]


def test_fit_learns_cost_law(tmp_path):
    f, w = synthetic(100000, 1)
    m = fit(f, w, 5, epochs=3, log=lambda s: None)
    ft, wt = synthetic(5000, 2)
    pred = forward(m.params, ft)
    y = np.log1p(wt.astype(np.float64))
    assert np.mean((pred - y) ** 2) < 0.05
    path = tmp_path / "m.npz"
    m.save(str(path))
    m2 = CostModel.load(str(path))
    assert m2.k == 5
    for k in FIELDS:
        assert np.array_equal(m.params[k], m2.params[k])


def test_shape_check():
    f, w = synthetic(2000, 3)
    m = fit(f, w, 4, epochs=1, log=lambda s: None)
    bad = dict(m.params, w2=np.zeros((64, 63), np.float32))
    try:
        CostModel(bad, 4)
    except ValueError:
        return
    raise AssertionError("bad shape accepted")
