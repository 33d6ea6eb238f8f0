"""CPU checks of the restated selection rule and audit figures (select_ref.py)
against independent definitions: the predecessor-count rank, scipy's
Spearman, and closed cases."""
import numpy as np
import scipy.stats

import select_ref as sr


def _scores(rng, P, ties=True, nans=3):
    x = rng.normal(0, 1, P).astype(np.float32)
    if ties:
        x[rng.integers(0, P, P // 8)] = x[0]  # a block of equal scores
        x[rng.integers(0, P, 4)] = -0.0
        x[rng.integers(0, P, 4)] = 0.0
    x[rng.integers(0, P, nans)] = np.nan
    return x


def test_greedy_order_is_the_count_rank():
    rng = np.random.default_rng(1)
    for P in (1, 7, 64, 300):
        x = _scores(rng, P, nans=min(3, P))
        r = sr.ranks_by_count(x)
        assert sorted(r.tolist()) == list(range(P))  # a permutation: no two equal ranks
        o = sr.greedy_order(x)
        assert np.array_equal(r[o], np.arange(P))


def test_nan_ranks_last_and_ties_by_index():
    x = np.array([np.nan, 1.0, 0.5, 1.0, np.nan, -2.0], dtype=np.float32)
    assert sr.greedy_order(x).tolist() == [5, 2, 1, 3, 0, 4]
    assert sr.ranks_by_count(x).tolist() == [4, 2, 1, 3, 5, 0]


def test_select_warm_and_greedy_and_explore():
    rng = np.random.default_rng(2)
    P, K = 256, 16
    x = rng.permutation(P).astype(np.float32)
    assert sr.select(x, K, 9, 40, warm=1) == list(range(K))
    sel = sr.select(x, K, 9, 40, warm=0)
    G = K - K // 4
    assert [int(x[p]) for p in sel[:G]] == list(range(G))  # the G lowest scores in order
    assert len(set(sel)) == K and all(0 <= p < P for p in sel)
    assert all(x[p] >= G for p in sel[G:])
    # the exploration draws depend on (seed, step) only
    assert sr.select(x, K, 9, 40, 0)[G:] != sr.select(x, K, 9, 41, 0)[G:]


def test_spearman_matches_scipy_with_ties():
    rng = np.random.default_rng(3)
    for _ in range(5):
        a = np.round(rng.normal(0, 1, 500), 1)
        b = np.round(a + rng.normal(0, 1, 500), 0)
        assert abs(sr.spearman(a, b) - scipy.stats.spearmanr(a, b).correlation) < 1e-12
    a = np.arange(10.0)
    assert sr.spearman(a, a) == 1.0
    assert abs(sr.spearman(a, -a) + 1.0) < 1e-15
    assert sr.spearman(np.ones(10), a) == 0.0


def test_miss_rate_closed_cases():
    a = np.arange(100.0)
    assert sr.miss_rate(a, a, 10) == 0.0
    assert sr.miss_rate(a, -a, 10) == 1.0
    b = a.copy()
    b[:5] += 1000  # five of the exact top ten pushed out of the score's top ten
    assert sr.miss_rate(b, a, 10) == 0.5
