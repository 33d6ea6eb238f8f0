"""The WVC acceptance rules on the GPU (cppcode/annealingWVC.cpp:412-417,
HillClimbWVC.cpp:240): the plain search over an explicit graph and the orbit
search over a part table match, move for move, the same drivers run on the
CPU oracle with WVC from per-source eccentricities; the adjacency
evaluator's WVC equals the oracle's."""
import numpy as np
import pytest

import oracle_eval
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def test_adjacency_evaluator_wvc(gpu):
    gg = gpu
    for n, d, seed in ((1024, 8, 3), (300, 5, 9), (40, 3, 2)):
        adj = gg.random_regular(n, d, seed)
        o = ol.adj_eval(adj)
        with gg.AdjEvaluator(n, d) as ev:
            ev.set_option(gg.GG_OPT_WVC, 1)
            r = ev.eval(adj)
        assert r.wvc == ol.wvc(n, o["ecc"], o["diameter"])
        assert r.total_reps == o["total"] and r.diameter == o["diameter"]


@pytest.mark.parametrize("rule", ["anneal_wvc", "hill_wvc"])
def test_plain_wvc_search_matches_cpu_twin(gpu, rule):
    gg = gpu
    adj = gg.random_regular(1024, 8, 5)
    steps = 150
    r = gg.RULE_ANNEAL_WVC if rule == "anneal_wvc" else gg.RULE_HILL_WVC
    cfg = gg.search_cfg(mode=gg.MODE_PLAIN, rule=r, total_steps=steps, T0=2e-3, T1=1e-6,
                        seed=17)
    g = gg.Search(adj, cfg, log=steps)
    sg = g.run(steps)
    cb = gg.PyEvaluatorCallbacks(oracle_eval.OracleSwapEvaluator(wvc=True), *adj.shape)
    c = gg.Search(adj, cfg, callbacks=cb, log=steps)
    sc = c.run(steps)
    assert np.array_equal(g.log(), c.log())
    assert np.array_equal(g.graph(), c.graph())
    assert np.array_equal(g.graph(best=True), c.graph(best=True))
    assert (sg.total_best, sg.diam_best) == (sc.total_best, sc.diam_best)
    print(f"{rule}: accepted {sg.accepted} of {steps}")
    g.close()
    c.close()


def test_orbit_wvc_search_matches_cpu_twin(gpu):
    gg = gpu
    g_, m, d = 32, 128, 16
    N = g_ * m
    tab = gg.orbit_part(g_, m, d, 3)
    steps = 120
    cfg = gg.search_cfg(mode=gg.MODE_PLAIN, rule=gg.RULE_ANNEAL_WVC, total_steps=steps,
                        T0=1e-3, T1=1e-6, seed=5)

    def ev(t):
        o = ol.part_eval(N, m, t)
        ok = not o["unconnected"]
        return o["total_scaled"], o["diameter"] if ok else 0, 0 if ok else 1, o["wvc"]
    with gg.OrbitSearch(N, tab, cfg, log=steps) as s:
        s.run(steps)
        lg, tg = s.log(), s.table()
    with gg.OrbitSearch(N, tab, cfg, table_eval=ev, log=steps) as s:
        s.run(steps)
        lc, tc = s.log(), s.table()
    assert np.array_equal(lg, lc)
    assert np.array_equal(tg, tc)
