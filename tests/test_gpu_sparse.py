"""GPU parity of the incremental (difference-propagation) batched path: every
candidate's exact total and diameter equal the dense path's and the CPU
oracle's, across shapes, repeated steps (graph replays) and commits."""
import numpy as np
import pytest

import oracle_lib as ol
from test_gpu_batch import _random_swaps

pytestmark = pytest.mark.gpu


def _eval_both(gg, be, swaps):
    be.set_option(gg.GG_BATCH_OPT_PATH, gg.PATH_SPARSE)
    rs = be.eval(swaps)
    assert be.get_option(gg.GG_BATCH_INFO_LAST_SPARSE) == 1
    be.set_option(gg.GG_BATCH_OPT_PATH, gg.PATH_DENSE)
    rd = be.eval(swaps)
    assert be.get_option(gg.GG_BATCH_INFO_LAST_SPARSE) == 0
    return rs, rd


@pytest.mark.parametrize("n,d,k", [(256, 4, 32), (1024, 6, 32), (4096, 12, 32), (8192, 16, 64),
                                   (8192, 16, 256), (16384, 24, 32)])
def test_sparse_matches_dense_and_oracle(gpu, n, d, k):
    gg = gpu
    print("sparse case", n, d, k, flush=True)
    rng = np.random.default_rng(n + k)
    adj = gg.random_regular(n, d, 11)
    cands = _random_swaps(gg, adj, k, rng)
    with gg.BatchEvaluator(n, d, max_k=k) as be:
        be.set_base(adj)
        assert be.get_option(gg.GG_BATCH_INFO_SPARSE_OK) == 1
        rs, rd = _eval_both(gg, be, [c[0] for c in cands])
        for (sw, new), a, b in zip(cands, rs, rd):
            assert (a.total_reps, a.diameter, a.status) == (b.total_reps, b.diameter, b.status)
        # oracle on a few candidates (all for small graphs)
        check = cands if n <= 4096 else cands[:3]
        for (sw, new), a in zip(check, rs):
            o = ol.adj_eval(new)
            assert a.total_reps == o["total"]
            if not o["unconnected"]:
                assert a.diameter == o["diameter"] and a.status == gg.GG_OK


def test_sparse_replays_and_commits(gpu):
    """Several steps through the captured graph with commits in between."""
    gg = gpu
    n, d, k = 2048, 8, 16
    rng = np.random.default_rng(5)
    adj = gg.random_regular(n, d, 2)
    with gg.BatchEvaluator(n, d, max_k=k) as be:
        be.set_base(adj)
        for step in range(5):
            cands = _random_swaps(gg, adj, k, rng)
            be.set_option(gg.GG_BATCH_OPT_PATH, gg.PATH_SPARSE)
            rs = be.eval([c[0] for c in cands])
            for (sw, new), a in zip(cands, rs):
                assert a.total_reps == ol.adj_eval(new)["total"], step
            best = int(np.argmin([r.total_reps for r in rs]))
            base = be.commit(cands[best][0])
            adj = cands[best][1]
            assert base.total_reps == ol.adj_eval(adj)["total"]


def test_sparse_disconnecting_candidates(gpu):
    """Two random halves joined by two bridge edges: the swap that rewires
    both bridges back inside the halves disconnects the graph; random swaps
    around it stay connected or stretch distances."""
    gg = gpu
    import graphs
    h, d = 128, 10
    adj, ((x1, y1), (x2, y2)) = graphs.bridged(gg.random_regular(h, d, 21),
                                               gg.random_regular(h, d, 22))
    n = 2 * h
    ia = int(np.nonzero(adj[x1] == y1)[0][0])
    ic = int(np.nonzero(adj[x2] == y2)[0][0])
    cut, cut_adj = gg.make_swap(adj, x1, ia, x2, ic, 0)  # (x1,x2), (y1,y2)
    assert cut is not None
    rng = np.random.default_rng(9)
    cands = [(cut, cut_adj)] + _random_swaps(gg, adj, 39, rng)
    with gg.BatchEvaluator(n, d, max_k=40) as be:
        base = be.set_base(adj)
        assert be.get_option(gg.GG_BATCH_INFO_SPARSE_OK) == 1, base.diameter
        rs, rd = _eval_both(gg, be, [c[0] for c in cands])
        assert rs[0].status == gg.GG_UNCONNECTED
        for (sw, new), a, b in zip(cands, rs, rd):
            o = ol.adj_eval(new)
            assert a.total_reps == b.total_reps == o["total"]
            assert a.status == b.status


@pytest.mark.parametrize("limit", [1, 8])
def test_sparse_overflow_goes_through_the_dense_queue(gpu, limit):
    """With the Delta limit lowered, most (candidate, column) tasks overflow
    and are evaluated by the dense fallback queue kernel (a bounded grid that
    loops over the queue); every candidate still matches the dense path and
    the oracle, replays included, and committing an overflowed candidate
    re-evaluates the base."""
    gg = gpu
    n, d, k = 4096, 12, 32
    print("overflow case", limit, flush=True)
    rng = np.random.default_rng(77 + limit)
    adj = gg.random_regular(n, d, 13)
    cands = _random_swaps(gg, adj, k, rng)
    with gg.BatchEvaluator(n, d, max_k=k) as be:
        be.set_base(adj)
        be.set_option(gg.GG_BATCH_OPT_DELTA_LIMIT, limit)
        assert be.get_option(gg.GG_BATCH_OPT_DELTA_LIMIT) == limit
        for rep in range(2):
            rs, rd = _eval_both(gg, be, [c[0] for c in cands])
            for (sw, new), a, b in zip(cands, rs, rd):
                assert (a.total_reps, a.diameter, a.status) == (b.total_reps, b.diameter,
                                                                  b.status), rep
        for (sw, new), a in zip(cands[:6], rs):
            o = ol.adj_eval(new)
            assert a.total_reps == o["total"]
        # commit the best sparse result, then the next batch on the new base
        be.set_option(gg.GG_BATCH_OPT_PATH, gg.PATH_SPARSE)
        rs = be.eval([c[0] for c in cands])
        i = min((i for i, r in enumerate(rs) if r.status == gg.GG_OK),
                key=lambda i: rs[i].total_reps)
        be.commit(cands[i][0])
        adj2 = cands[i][1]
        cands2 = _random_swaps(gg, adj2, k, rng)
        rs2, rd2 = _eval_both(gg, be, [c[0] for c in cands2])
        for (sw, new), a, b in zip(cands2, rs2, rd2):
            assert (a.total_reps, a.diameter, a.status) == (b.total_reps, b.diameter, b.status)
        for (sw, new), a in zip(cands2[:4], rs2):
            assert a.total_reps == ol.adj_eval(new)["total"]
        # back to the full capacity
        be.set_option(gg.GG_BATCH_OPT_DELTA_LIMIT, 0)
        rs3, rd3 = _eval_both(gg, be, [c[0] for c in cands2])
        assert [r.total_reps for r in rs3] == [r.total_reps for r in rs2]


@pytest.mark.parametrize("n,d", [(4096, 12), (2048, 8)])
def test_sparse_directed_candidates_match_dense(gpu, n, d):
    """Swaps that rewrite single slots (directed candidate graphs): the
    incremental path follows changed words through the changed vertices' own
    rows, which is exact only on undirected graphs, so such a batch runs on
    the dense path even where the incremental one is asked for."""
    gg = gpu
    rng = np.random.default_rng(n)
    adj = gg.random_regular(n, d, 3)
    swaps = []
    for _ in range(24):
        sw = gg.Swap()
        for j in range(4):
            sw.v[j], sw.slot[j], sw.nbr[j] = -1, 0, 0
        for j in range(int(rng.integers(1, 3))):
            v = int(rng.integers(0, n))
            sw.v[j], sw.slot[j], sw.nbr[j] = v, int(rng.integers(0, d)), int(rng.integers(0, n))
        swaps.append(sw)
    with gg.BatchEvaluator(n, d, max_k=24) as be:
        be.set_base(adj)
        be.set_option(gg.GG_BATCH_OPT_PATH, gg.PATH_SPARSE)
        rs = be.eval(swaps)
        assert be.get_option(gg.GG_BATCH_INFO_LAST_SPARSE) == 0
        be.set_option(gg.GG_BATCH_OPT_PATH, gg.PATH_DENSE)
        rd = be.eval(swaps)
        for i, (a, b) in enumerate(zip(rs, rd)):
            assert (a.total_reps, a.diameter, a.status) == (b.total_reps, b.diameter, b.status), i
        # one undirected swap among them: the batch still runs densely; a
        # batch of undirected swaps alone runs incrementally again
        cands = _random_swaps(gg, adj, 8, rng)
        be.set_option(gg.GG_BATCH_OPT_PATH, gg.PATH_SPARSE)
        be.eval([c[0] for c in cands])
        assert be.get_option(gg.GG_BATCH_INFO_LAST_SPARSE) == 1
