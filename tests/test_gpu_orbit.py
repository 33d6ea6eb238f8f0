"""Orbit-graph searches on the GPU evaluators, move for move against the same
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


def test_orbit_search_rows_of_different_lengths_gpu_matches_cpu_twin(gpu):
    """annealingNR's rows of different lengths (cppcode/annealingNR.cpp:117,
    146-147) on the g = 32, m = 128 orbit graph with 300 edge orbits removed:
    the GPU evaluator sees short rows padded with offset 0 (self-loops, as
    cudaASPLconv.cu:199-201 pads them) and agrees move for move with the CPU
    oracle."""
    import driver_ref
    gg = gpu
    g, m, d = 32, 128, 16
    N = g * m
    tab, lens = driver_ref.irregular_part(gg.orbit_part(g, m, d, 3), N, 300, 9)
    steps = 300
    cfg = gg.search_cfg(mode=gg.MODE_PLAIN, rule=gg.RULE_ANNEAL, total_steps=steps,
                        T0=1e-3, T1=1e-6, seed=4)
    with gg.OrbitSearch(N, tab, cfg, log=steps, row_lengths=lens) as s:
        st = s.run(steps)
        log, final = s.log(), s.table()
    with gg.OrbitSearch(N, tab, cfg, log=steps, row_lengths=lens,
                        table_eval=lambda t: oracle_eval.evaluate_part(N, t)) as c:
        c.run(steps)
        clog, cfinal = c.log(), c.table()
    assert np.array_equal(log, clog) and np.array_equal(final, cfinal)
    tot, dia, status = oracle_eval.evaluate_part(N, final)
    assert status == 0 and tot == st.total_cur and dia == st.diam_cur
    assert log[:, 9].sum() > 0


def _lattice_eval(Ns, Ms):
    import oracle_lib as ol

    def f(t):
        if len(Ns) == 2:
            r = ol.piece_eval(Ns[0], Ns[1], Ms[0], Ms[1], t)
        else:
            r = ol.box_eval(Ns[0], Ns[1], Ns[2], Ms[0], Ms[1], Ms[2], t)
        ok = not r["unconnected"]
        return (r["total_scaled"] if ok else 0), (r["diameter"] if ok else 0), (0 if ok else 1)
    return f


@pytest.mark.parametrize("Ns,Ms,d,rule,steps", [
    ((16, 64), (4, 8), 8, "hill", 300),        # piece (pieceHillClimb.cpp)
    ((16, 64), (4, 8), 8, "anneal", 200),      # piece (pieceannealing.cpp)
    ((32, 128), (1, 128), 16, "hill", 200),    # the g = 32, m = 128 orbit graph as a piece
    ((8, 8, 8), (2, 2, 2), 6, "hill", 200),    # box (boxHillClimb.cpp)
])
def test_lattice_search_gpu_matches_cpu_twin(gpu, Ns, Ms, d, rule, steps):
    """Piece / box search (gg_piece_search_create / gg_box_search_create) on
    the GPU evaluators, move for move against the same driver on the oracle."""
    import driver_ref
    gg = gpu
    tab = driver_ref.random_lattice_table(Ns, Ms, d, 7)
    cfg = gg.search_cfg(mode=gg.MODE_PLAIN,
                        rule=gg.RULE_HILL if rule == "hill" else gg.RULE_ANNEAL,
                        total_steps=steps, T0=1e-3, T1=1e-6, seed=11)
    print("lattice search", Ns, Ms, d, rule, steps, flush=True)
    with gg.LatticeSearch(Ns, Ms, tab, cfg, log=steps) as s:
        st = s.run(steps)
        log, final = s.log(), s.table()
    ev = _lattice_eval(Ns, Ms)
    with gg.LatticeSearch(Ns, Ms, tab, cfg, log=steps, table_eval=ev) as c:
        sc = c.run(steps)
        clog, cfinal = c.log(), c.table()
    assert np.array_equal(log, clog)
    assert np.array_equal(final, cfinal)
    assert (st.total_cur, st.diam_cur, st.accepted) == (sc.total_cur, sc.diam_cur, sc.accepted)
    tot, dia, status = ev(final)
    assert status == 0 and tot == st.total_cur and dia == st.diam_cur
    assert log[:, 9].sum() > 0
