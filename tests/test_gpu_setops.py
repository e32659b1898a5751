"""GPU parity of the set primitives (gpm_setops) in every variant against
numpy's set routines (np.intersect1d / np.setdiff1d on strictly increasing
lists), including the edge cases of the reference primitives: empty A or B,
upper bounds at or below min(A) and above max(A), list lengths around the
64-lane wave width, lists past the LDS hash and bitmap capacities (fallback),
and fuse_set's three outputs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VARIANTS = [0, 1, 2, 3, 4]  # binary, two-phase, merge, hash, bitmap


def _pairs(seed):
    rng = np.random.default_rng(seed)
    out = []
    for na, nb, span in [(0, 0, 10), (0, 5, 10), (5, 0, 10), (1, 1, 2), (63, 64, 200),
                         (64, 65, 150), (65, 63, 130), (64, 64, 64), (200, 3, 1000),
                         (3, 2000, 5000), (700, 900, 2000), (1000, 1000, 100000),
                         (129, 4000, 40000), (50, 600, 700)]:
        a = np.sort(rng.choice(span, min(na, span), replace=False)).astype(np.uint32)
        b = np.sort(rng.choice(span, min(nb, span), replace=False)).astype(np.uint32)
        out.append((a, b))
    # identical lists, disjoint interleaved lists, B a prefix of A
    x = np.arange(0, 300, 3, dtype=np.uint32)
    out += [(x, x.copy()), (x, x + 1), (x, x[:17]), (x[:1], x), (x + 4000000000, x)]
    return out


def _ubs(pairs, rng):
    ub = []
    for a, _ in pairs:
        if len(a) == 0:
            ub.append(5)
        else:
            ub.append(int(rng.choice([0, int(a[0]), int(a[0]) + 1, int(a[len(a) // 2]),
                                      int(a[-1]) + 1, 0xFFFFFFFF])))
    return np.asarray(ub, dtype=np.uint32)


@pytest.mark.parametrize("variant", VARIANTS)
def test_intersection_and_difference(gpm, variant):
    pairs = _pairs(variant)
    rng = np.random.default_rng(100 + variant)
    for use_ub in (False, True):
        ub = _ubs(pairs, rng) if use_ub else None
        want_i, want_d = [], []
        for k, (a, b) in enumerate(pairs):
            aa = a[a < ub[k]] if use_ub else a
            want_i.append(np.intersect1d(aa, b))
            want_d.append(np.setdiff1d(aa, b))
        c, lists = gpm.setops(gpm.SETOP_ISECT, variant, pairs, ub)
        assert [int(x) for x in c] == [len(w) for w in want_i]
        assert all(np.array_equal(g, w) for g, w in zip(lists, want_i))
        c, _ = gpm.setops(gpm.SETOP_ISECT_COUNT, variant, pairs, ub)
        assert [int(x) for x in c] == [len(w) for w in want_i]
        c, lists = gpm.setops(gpm.SETOP_DIFF, variant, pairs, ub)
        assert [int(x) for x in c] == [len(w) for w in want_d]
        assert all(np.array_equal(g, w) for g, w in zip(lists, want_d))
        c, _ = gpm.setops(gpm.SETOP_DIFF_COUNT, variant, pairs, ub)
        assert [int(x) for x in c] == [len(w) for w in want_d]


@pytest.mark.parametrize("variant", VARIANTS)
def test_fuse(gpm, variant):
    pairs = _pairs(10 + variant)
    c, lists = gpm.setops(gpm.SETOP_FUSE, variant, pairs)
    for k, (a, b) in enumerate(pairs):
        d_ab, i_ab, d_ba = lists[k]
        assert np.array_equal(d_ab, np.setdiff1d(a, b)), k
        assert np.array_equal(i_ab, np.intersect1d(a, b)), k
        assert np.array_equal(d_ba, np.setdiff1d(b, a)), k
        assert c[k].tolist() == [len(d_ab), len(i_ab), len(d_ba)]


def test_many_pairs_all_variants_agree(gpm):
    rng = np.random.default_rng(7)
    pairs = []
    for _ in range(3000):
        na, nb = rng.integers(0, 300, 2)
        span = int(rng.integers(1, 3000))
        a = np.sort(rng.choice(span, min(int(na), span), replace=False))
        b = np.sort(rng.choice(span, min(int(nb), span), replace=False))
        pairs.append((a, b))
    want = np.array([len(np.intersect1d(a, b)) for a, b in pairs])
    for v in VARIANTS:
        c, _ = gpm.setops(gpm.SETOP_ISECT_COUNT, v, pairs)
        assert np.array_equal(c.astype(np.int64), want), v


def test_setops_rejects_unsorted(gpm):
    with pytest.raises(gpm.GpmError):
        gpm.setops(gpm.SETOP_ISECT, 0, [(np.array([3, 1]), np.array([1]))])
    with pytest.raises(gpm.GpmError):
        gpm.setops(9, 0, [(np.array([1]), np.array([1]))])
