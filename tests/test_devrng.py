"""The restatement of the GPU proposal streams (tests/devrng_ref.py), pinned
by splitmix64's published outputs and by properties of Lemire's bounded
mapping that hold exactly."""
import numpy as np

import devrng_ref as R


def test_splitmix64_known_answer():
    # splitmix64 from state 0 (Vigna's reference splitmix64.c)
    s = R.Stream(raw_state=0)
    assert [s.next() for _ in range(4)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                            0x06C45D188009454F, 0xF88BB8A8724C81EC]


def test_below_powers_of_two_are_top_bits():
    # 2^64 mod 2^k = 0: no rejection, the result is the top k bits
    for k in (1, 5, 17, 40, 63):
        a, b = R.Stream(3, 7, 11), R.Stream(3, 7, 11)
        for _ in range(200):
            assert a.below(1 << k) == b.next() >> (64 - k)


def test_below_range_one_and_bounds():
    s = R.Stream(1, 2, 3)
    assert all(s.below(1) == 0 for _ in range(50))
    r = (1 << 63) + 12345  # rejects about half of the draws
    vals = [s.below(r) for _ in range(400)]
    assert all(0 <= v < r for v in vals)
    assert 0.35 < np.mean([v < r // 2 for v in vals]) < 0.65


def test_below_uniform_small_range():
    s = R.Stream(9, 0, 0)
    counts = np.bincount([s.below(3) for _ in range(30000)], minlength=3)
    chi2 = ((counts - 10000) ** 2 / 10000).sum()
    assert chi2 < 20, counts


def test_streams_differ_by_step_and_lane():
    firsts = {R.Stream(5, st, ln).next() for st in range(20) for ln in range(20)}
    assert len(firsts) == 400


def test_proposals_are_simple_swaps():
    import graphs
    adj = graphs.circulant(64, [1, 3, 7])
    for p in range(50):
        a, b, c, d, kind = R.propose(adj, 42, 3, p)
        assert len({a, b, c, d}) == 4 and b in adj[a] and d in adj[c]
        xa, xb = (d, c) if kind else (c, d)
        assert xa not in adj[a] and xb not in adj[b]
