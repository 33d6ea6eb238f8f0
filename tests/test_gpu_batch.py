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
