"""Orbit-graph search on the GPU evaluator, move for move against the same
driver on the CPU oracle: the cyclically symmetric configuration n = g*m with
g = 32 copies of m = 128 base vertices, degree 16 (BFS from the 128
representatives, total scaled by 32), hill climbing over 10^3 orbit moves
(edge-length changes and orbit 2-opt, cppcode/annealing.cpp:112-197), and a
shorter annealing run."""
import numpy as np
import pytest

import oracle_eval

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rule,steps", [("hill", 1000), ("anneal", 300)])
def test_orbit_search_gpu_matches_cpu_twin(gpu, rule, steps):
    gg = gpu
    g, m, d = 32, 128, 16
    N = g * m
    tab = gg.orbit_part(g, m, d, 3)
    cfg = gg.search_cfg(mode=gg.MODE_PLAIN,
                        rule=gg.RULE_HILL if rule == "hill" else gg.RULE_ANNEAL,
                        total_steps=steps, T0=1e-3, T1=1e-6, seed=3)
    print("orbit search", rule, steps, flush=True)
    with gg.OrbitSearch(N, tab, cfg, log=steps) as s:
        st = s.run(steps)
        log, final = s.log(), s.table()
    with gg.OrbitSearch(N, tab, cfg, log=steps,
                        table_eval=lambda t: oracle_eval.evaluate_part(N, t)) as c:
        sc = c.run(steps)
        clog, cfinal = c.log(), c.table()
    assert np.array_equal(log, clog)
    assert np.array_equal(final, cfinal)
    assert (st.total_cur, st.diam_cur, st.accepted) == (sc.total_cur, sc.diam_cur, sc.accepted)
    assert st.exact_evals == steps
    tot, dia, status = oracle_eval.evaluate_part(N, final)
    assert status == 0 and tot == st.total_cur and dia == st.diam_cur
    assert log[:, 9].sum() > 0  # the run moved
