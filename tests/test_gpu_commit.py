"""GPU parity of the in-place commit: committing a candidate of the last
incremental launch rewrites the stored base levels from what the launch kept
(per-level Delta logs, the endpoint columns' words) instead of running a new
BFS.  The stored state after any sequence of such commits must equal the state
a fresh evaluator builds from the same graph (an order-independent checksum of
every stored word, column part and level count), and every later evaluation
must agree with the fresh evaluator and the CPU oracle."""
import numpy as np
import pytest

import oracle_lib as ol
import feature_ref
from test_gpu_batch import _random_swaps
from test_gpu_learn import _moves

pytestmark = pytest.mark.gpu


def _step(gg, be, adj, k, rng, pick="best"):
    cands = _random_swaps(gg, adj, k, rng)
    be.set_option(gg.GG_BATCH_OPT_PATH, gg.PATH_SPARSE)
    rs = be.eval([c[0] for c in cands])
    assert be.get_option(gg.GG_BATCH_INFO_LAST_SPARSE) == 1
    ok = [i for i, r in enumerate(rs) if r.status == gg.GG_OK]
    assert ok
    key = (lambda i: rs[i].total_reps)
    i = min(ok, key=key) if pick == "best" else max(ok, key=key)
    base = be.commit(cands[i][0])
    assert (base.total_reps, base.diameter, base.status) == \
        (rs[i].total_reps, rs[i].diameter, gg.GG_OK)
    return cands[i][1], i


def _same_as_fresh(gg, be, adj, k, rng, n_oracle=2):
    n, d = adj.shape
    with gg.BatchEvaluator(n, d, max_k=k) as fresh:
        fb = fresh.set_base(adj)
        assert be.levels_digest() == fresh.levels_digest()
        cands = _random_swaps(gg, adj, k, rng)
        ra = be.eval([c[0] for c in cands])
        rb = fresh.eval([c[0] for c in cands])
        for a, b in zip(ra, rb):
            assert (a.total_reps, a.diameter, a.status) == (b.total_reps, b.diameter, b.status)
        for (sw, new), a in list(zip(cands, ra))[:n_oracle]:
            o = ol.adj_eval(new)
            assert a.total_reps == o["total"]
        assert np.array_equal(be.base_ecc(), fresh.base_ecc())
        return fb


@pytest.mark.parametrize("n,d,k,steps", [(512, 5, 16, 12), (1024, 6, 16, 10), (4096, 12, 32, 6),
                                         (16384, 24, 32, 4)])
def test_inplace_commit_equals_fresh_base(gpu, n, d, k, steps):
    gg = gpu
    print("commit case", n, d, k, flush=True)
    rng = np.random.default_rng(n + 3)
    adj = gg.random_regular(n, d, 7)
    with gg.BatchEvaluator(n, d, max_k=k) as be:
        be.set_base(adj)
        for _ in range(steps):
            adj, _ = _step(gg, be, adj, k, rng)
            assert be.get_option(gg.GG_BATCH_INFO_LAST_COMMIT_INPLACE) == 1
        base = _same_as_fresh(gg, be, adj, k, rng)
        assert base.total_reps == ol.adj_eval(adj)["total"]


def test_inplace_commit_of_worse_candidates(gpu):
    """Committing the worst connected candidate raises column eccentricities:
    levels that were full in the base must be written whole."""
    gg = gpu
    n, d, k = 600, 4, 24
    rng = np.random.default_rng(17)
    adj = gg.random_regular(n, d, 3)
    with gg.BatchEvaluator(n, d, max_k=k) as be:
        base = be.set_base(adj)
        if be.get_option(gg.GG_BATCH_INFO_SPARSE_OK) != 1:
            pytest.skip(f"base diameter {base.diameter} beyond the stored levels")
        for step in range(8):
            adj, _ = _step(gg, be, adj, k, rng, pick="worst")
            if be.get_option(gg.GG_BATCH_INFO_SPARSE_OK) != 1:
                break  # the diameter outgrew the stored levels: dense from here
            assert be.get_option(gg.GG_BATCH_INFO_LAST_COMMIT_INPLACE) == 1, step
            with gg.BatchEvaluator(n, d, max_k=k) as fresh:
                fresh.set_base(adj)
                assert be.levels_digest() == fresh.levels_digest(), step
        if be.get_option(gg.GG_BATCH_INFO_SPARSE_OK) == 1:
            _same_as_fresh(gg, be, adj, k, rng)


def test_commit_mode_new_bfs_and_chunks(gpu):
    """GG_BATCH_OPT_COMMIT = 1 always runs a new BFS; a candidate outside the
    last launch of a batch split into launches of 64 is committed with a new
    BFS too, one of a single launch in place.  Every path ends in the same
    stored state."""
    gg = gpu
    n, d = 1024, 6
    rng = np.random.default_rng(23)
    adj = gg.random_regular(n, d, 5)
    with gg.BatchEvaluator(n, d, max_k=100) as be:
        be.set_base(adj)
        be.set_option(gg.GG_BATCH_OPT_COMMIT, 1)
        adj, _ = _step(gg, be, adj, 16, rng)
        assert be.get_option(gg.GG_BATCH_INFO_LAST_COMMIT_INPLACE) == 0
        be.set_option(gg.GG_BATCH_OPT_COMMIT, 0)
        _same_as_fresh(gg, be, adj, 16, rng)
        # 100 candidates in launches of at most 64: the logs cover the last one
        be.set_option(gg.GG_BATCH_OPT_LAUNCH_K, 64)
        cands = _random_swaps(gg, adj, 100, rng)
        rs = be.eval([c[0] for c in cands])
        ok = [i for i, r in enumerate(rs) if r.status == gg.GG_OK]
        early = [i for i in ok if i < 36]
        late = [i for i in ok if i >= 64]
        i = early[0]
        be.commit(cands[i][0])
        assert be.get_option(gg.GG_BATCH_INFO_LAST_COMMIT_INPLACE) == 0
        adj = cands[i][1]
        _same_as_fresh(gg, be, adj, 100, rng)
        cands = _random_swaps(gg, adj, 100, rng)
        rs = be.eval([c[0] for c in cands])
        late = [i for i, r in enumerate(rs) if r.status == gg.GG_OK and i >= 64]
        i = late[0]
        base = be.commit(cands[i][0])
        assert be.get_option(gg.GG_BATCH_INFO_LAST_COMMIT_INPLACE) == 1
        assert base.total_reps == rs[i].total_reps
        adj = cands[i][1]
        _same_as_fresh(gg, be, adj, 100, rng)
        # by default all 100 run in one launch and every candidate commits in place
        be.set_option(gg.GG_BATCH_OPT_LAUNCH_K, 0)
        cands = _random_swaps(gg, adj, 100, rng)
        rs = be.eval([c[0] for c in cands])
        i = [i for i, r in enumerate(rs) if r.status == gg.GG_OK][0]
        base = be.commit(cands[i][0])
        assert be.get_option(gg.GG_BATCH_INFO_LAST_COMMIT_INPLACE) == 1
        assert base.total_reps == rs[i].total_reps
        _same_as_fresh(gg, be, cands[i][1], 100, rng)


def test_features_after_inplace_commits(gpu):
    """Features (eccentricities derived from the stored ball rows) follow the
    committed graph."""
    gg = gpu
    n, d, k = 1024, 6, 16
    rng = np.random.default_rng(31)
    adj = gg.random_regular(n, d, 9)
    with gg.BatchEvaluator(n, d, max_k=k) as be:
        be.set_base(adj)
        for _ in range(5):
            adj, _ = _step(gg, be, adj, k, rng)
        mv = _moves(gg, adj, 40, rng)
        X = be.features(mv)
        diam = ol.adj_eval(adj)["diameter"]
    R = feature_ref.features(adj, mv, diam)
    assert np.allclose(X, R, rtol=1e-6, atol=1e-6), np.abs(X - R).max()
