"""Annealing trajectories on the GPU evaluator match, move for move, the same
driver run on the CPU oracle (the reference's compile-time evaluator swap,
cppcode/annealing.cpp:296 vs annealing_cpu.cpp:292): config C2 (random
12-regular n=4096, 2-opt, geometric T from 1e-3 to 1e-6 over 10^4 moves),
a 500-move prefix."""
import numpy as np
import pytest

import oracle_eval
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _cfg(gg, mode, K=1, total=10000):
    return gg.search_cfg(mode=mode, rule=gg.RULE_ANNEAL, K=K, total_steps=total, T0=1e-3,
                         T1=1e-6, seed=12345)


def test_c2_plain_trajectory_prefix_matches_cpu(gpu):
    gg = gpu
    adj = gg.random_regular(4096, 12, 2)
    steps = 500
    g = gg.Search(adj, _cfg(gg, gg.MODE_PLAIN), log=steps)
    g.run(steps)
    cb = gg.PyEvaluatorCallbacks(oracle_eval.OracleSwapEvaluator(), *adj.shape)
    c = gg.Search(adj, _cfg(gg, gg.MODE_PLAIN), callbacks=cb, log=steps)
    c.run(steps)
    lg, lc = g.log(), c.log()
    assert lg.shape == (steps, 10)
    assert np.array_equal(lg, lc)
    assert np.array_equal(g.graph(), c.graph())
    print("C2 prefix: accepted", int(lg[:, 9].sum()), "of", steps)


def test_c2_batched_trajectory_prefix_matches_cpu(gpu):
    gg = gpu
    adj = gg.random_regular(4096, 12, 2)
    steps = 40
    g = gg.Search(adj, _cfg(gg, gg.MODE_BATCHED, K=8, total=1000), log=steps)
    g.run(steps)
    cb = gg.PyEvaluatorCallbacks(oracle_eval.OracleSwapEvaluator(), *adj.shape)
    c = gg.Search(adj, _cfg(gg, gg.MODE_BATCHED, K=8, total=1000), callbacks=cb, log=steps)
    c.run(steps)
    assert np.array_equal(g.log(), c.log())
    assert np.array_equal(g.graph(), c.graph())


def test_c2_full_run_consistent(gpu):
    """All 10^4 moves on the GPU; the final and best graphs re-evaluated by the
    oracle give the driver's totals."""
    gg = gpu
    adj = gg.random_regular(4096, 12, 2)
    g = gg.Search(adj, _cfg(gg, gg.MODE_PLAIN))
    st = g.run(10000)
    assert st.steps == 10000 and st.exact_evals == 10000
    cur, best = g.graph(), g.graph(best=True)
    oc, ob = ol.adj_eval(cur), ol.adj_eval(best)
    assert oc["total"] == st.total_cur and oc["diameter"] == st.diam_cur
    assert ob["total"] == st.total_best and ob["diameter"] == st.diam_best
    o0 = ol.adj_eval(adj)
    assert (st.diam_best, st.total_best) <= (o0["diameter"], o0["total"])
    print("C2 10^4 moves: ASPL", o0["total"] / (4096 * 4095), "->", st.total_best / (4096 * 4095),
          "accepted", st.accepted)
