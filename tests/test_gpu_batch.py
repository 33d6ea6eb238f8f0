"""GPU parity of the plain single-move path (resident base + one swap) and of
the batched evaluator (K swaps per launch) against the CPU oracle on the
materialised candidate graphs."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _random_swaps(gg, adj, k, rng):
    out = []
    n, d = adj.shape
    while len(out) < k:
        a, c = (int(x) for x in rng.integers(0, n, 2))
        ia, ic = (int(x) for x in rng.integers(0, d, 2))
        sw, new = gg.make_swap(adj, a, ia, c, ic, int(rng.integers(0, 2)))
        if sw is not None:
            out.append((sw, new))
    return out


@pytest.mark.parametrize("n,d", [(256, 4), (2000, 6), (4096, 12)])
def test_swap_path_matches_oracle(gpu, n, d):
    gg = gpu
    rng = np.random.default_rng(n)
    adj = gg.random_regular(n, d, 7)
    with gg.SwapEvaluator(n, d) as ev:
        base = ev.set_base(adj)
        assert base.total_reps == ol.adj_eval(adj)["total"]
        for sw, new in _random_swaps(gg, adj, 6, rng):
            r = ev.eval_swap(sw)
            o = ol.adj_eval(new)
            assert r.total_reps == o["total"]
            if not o["unconnected"]:
                assert r.diameter == o["diameter"]
        # commit the last one, then the base evaluates as the new graph
        ev.commit(sw)
        r = ev.eval_swap(gg.Swap(v=(-1, -1, -1, -1)))
        assert r.total_reps == ol.adj_eval(new)["total"]


@pytest.mark.parametrize("n,d,k", [(256, 4, 16), (1024, 8, 37), (8192, 16, 8)])
def test_batch_matches_oracle(gpu, n, d, k):
    gg = gpu
    rng = np.random.default_rng(k)
    adj = gg.random_regular(n, d, 3)
    cands = _random_swaps(gg, adj, k, rng)
    with gg.BatchEvaluator(n, d, max_k=64) as be:
        be.set_option(gg.GG_BATCH_OPT_PATH, gg.PATH_DENSE)
        base = be.set_base(adj)
        assert base.total_reps == ol.adj_eval(adj)["total"]
        res = be.eval([c[0] for c in cands])
        for (sw, new), r in zip(cands, res):
            o = ol.adj_eval(new)
            assert r.total_reps == o["total"]
            if not o["unconnected"]:
                assert r.diameter == o["diameter"] and r.status == gg.GG_OK
        # commit candidate 0; base re-evaluates to the new graph
        b2 = be.commit(cands[0][0])
        assert b2.total_reps == ol.adj_eval(cands[0][1])["total"]


def test_batch_disconnecting_swap(gpu):
    """A swap that splits a graph of two joined cycles is reported unconnected."""
    gg = gpu
    import graphs
    # swapping two edges of a cycle can split it into two cycles
    adj = graphs.cycle(10)
    sw, new = gg.make_swap(adj, 0, 1, 5, 1, 1)  # edges (0,1),(5,6) -> (0,6),(1,5)
    assert sw is not None
    o = ol.adj_eval(new)
    with gg.BatchEvaluator(10, 2, 4) as be:
        be.set_option(gg.GG_BATCH_OPT_PATH, gg.PATH_DENSE)
        be.set_base(adj)
        r = be.eval([sw])[0]
    assert o["unconnected"] and r.status == gg.GG_UNCONNECTED and r.total_reps == o["total"]


def test_swap_graph_matches_launched_path(gpu):
    """GG_OPT_SWAP_GRAPH: the captured single-swap evaluation (swap read from
    and result written to pinned memory) gives the results of the launched
    path on swaps that keep the graph undirected and on swaps that do not
    (the two push settings are two captured graphs), before and after
    commits; with GG_OPT_LEVELS 0 gg_eval_levels reports GG_UNSUPPORTED, and
    kernel-time events exist only for launched evaluations."""
    gg = gpu
    n, d = 2048, 8
    adj = gg.random_regular(n, d, 12)
    rng = np.random.default_rng(3)
    swaps = [c[0] for c in _random_swaps(gg, adj, 24, rng)]
    # a directed swap: one slot rewritten without its partner
    bad = gg.Swap()
    for j in range(4):
        bad.v[j], bad.slot[j], bad.nbr[j] = -1, 0, 0
    bad.v[0], bad.slot[0], bad.nbr[0] = 5, 0, int(adj[5][0]) ^ 1
    swaps.insert(7, bad)
    with gg.SwapEvaluator(n, d) as a, gg.SwapEvaluator(n, d) as b, \
            gg.SwapEvaluator(n, d) as c:
        for ev in (a, b, c):
            ev.set_base(adj)
        b.set_option(gg.GG_OPT_SWAP_GRAPH, 1)
        b.set_option(gg.GG_OPT_LEVELS, 0)
        c.set_option(gg.GG_OPT_SWAP_GRAPH, 2)
        for i, sw in enumerate(swaps):
            ra, rb, rc = a.eval_swap(sw), b.eval_swap(sw), c.eval_swap(sw)
            assert ra.as_dict() == rb.as_dict() == rc.as_dict(), i
            if i in (3, 11, 12):  # commits (two in a row) of undirected swaps
                for ev in (a, b, c):
                    ev.commit(sw)
        assert np.array_equal(a.levels()[0], c.levels()[0])
        with pytest.raises(gg.GGError) as e:
            b.levels()
        assert e.value.code == gg.GG_UNSUPPORTED
        with pytest.raises(gg.GGError):
            b.last_kernel_ms()
        assert a.last_kernel_ms() > 0
