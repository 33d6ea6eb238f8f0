"""The C++ orbit-graph search driver (symmetric `part` tables) against a
pure-Python restatement of its semantics -- createNeighbour's edge-length
change and orbit 2-opt (cppcode/annealing.cpp:112-197), RNG draws, acceptance
-- both on the CPU oracle: identical per-step logs and final tables.  Every
accepted table stays a symmetric simple regular graph."""
import numpy as np
import pytest

import driver_ref
import oracle_eval
import oracle_lib as ol


def _tab_eval(N):
    def f(t):
        tot, dia, st = oracle_eval.evaluate_part(N, t)
        return tot, dia, st
    return f


def _check_symmetric_simple(N, t):
    M, d = t.shape
    for r in range(M):
        xs = [int(x) % N for x in t[r]]
        assert 0 not in xs and len(set(xs)) == d, (r, xs)
        for x in xs:
            partner = (r + x) % N % M
            assert (-x) % N in [int(y) % N for y in t[partner]], (r, x)
    adj = ol.expand_part(N, M, t)
    # undirected: the expanded edge multiset equals its reverse
    e = {(v, int(u)) for v in range(N) for u in adj[v]}
    assert all((u, v) in e for v, u in e)


@pytest.mark.parametrize("g,m,d,rule", [(8, 32, 6, "hill"), (8, 32, 6, "anneal"),
                                        (1, 64, 4, "hill"), (5, 20, 4, "anneal")])
def test_orbit_driver_matches_python_restatement(gg, g, m, d, rule):
    N = g * m
    tab = gg.orbit_part(g, m, d, 21) if g > 1 else ol_table_plain(gg, m, d)
    steps = 80
    cfg = gg.search_cfg(mode=gg.MODE_PLAIN,
                        rule=gg.RULE_HILL if rule == "hill" else gg.RULE_ANNEAL,
                        total_steps=steps, T0=2e-2, T1=1e-4, seed=99)
    ev = _tab_eval(N)
    with gg.OrbitSearch(N, tab, cfg, table_eval=ev, log=steps) as s:
        st = s.run(steps)
        log = s.log()
        final = s.table()
    ref_log, ref_tab = driver_ref.run_orbit(
        N, tab, lambda t: (lambda r: (r[0], r[1], r[2] == 0))(ev(t)), steps, rule=rule,
        T0=2e-2, T1=1e-4, seed=99)
    assert np.array_equal(log, ref_log)
    assert np.array_equal(final, ref_tab)
    assert st.steps == steps and st.exact_evals == steps
    assert st.accepted == int(ref_log[:, 9].sum())
    if g > 1:
        assert set(ref_log[:, 5].tolist()) == {0, 1}  # both move kinds drawn
    _check_symmetric_simple(N, final)
    tot, dia, status = ev(final)
    assert status == 0 and st.total_cur == tot and st.diam_cur == dia
    if rule == "hill":
        t0, d0, _ = ev(tab)
        assert (st.diam_cur, st.total_cur) <= (d0, t0)


@pytest.mark.parametrize("rule", ["hill", "anneal"])
def test_orbit_driver_rows_of_different_lengths(gg, rule):
    """annealingNR's input: rows with different numbers of entries
    (cppcode/annealingNR.cpp:117,146-147).  Slots are drawn among a row's
    entries, the padding stays offset 0 and the per-step logs and final table
    equal the restatement's."""
    g, m, d = 8, 32, 6
    N = g * m
    tab, lens = driver_ref.irregular_part(gg.orbit_part(g, m, d, 21), N, 20, 5)
    assert len(set(lens.tolist())) > 1
    steps = 120
    cfg = gg.search_cfg(mode=gg.MODE_PLAIN,
                        rule=gg.RULE_HILL if rule == "hill" else gg.RULE_ANNEAL,
                        total_steps=steps, T0=2e-2, T1=1e-4, seed=7)
    ev = _tab_eval(N)
    with gg.OrbitSearch(N, tab, cfg, table_eval=ev, log=steps, row_lengths=lens) as s:
        st = s.run(steps)
        log = s.log()
        final = s.table()
    ref_log, ref_tab = driver_ref.run_orbit(
        N, tab, lambda t: (lambda r: (r[0], r[1], r[2] == 0))(ev(t)), steps, rule=rule,
        T0=2e-2, T1=1e-4, seed=7, lens=lens)
    assert np.array_equal(log, ref_log)
    assert np.array_equal(final, ref_tab)
    assert st.accepted == int(ref_log[:, 9].sum()) and st.accepted > 0
    # every drawn slot lies inside its row; padding untouched; still symmetric
    for a_slot in np.concatenate([log[:, 1], log[:, 2]]):
        r, i = divmod(int(a_slot), d)
        assert i < lens[r]
    for r in range(m):
        assert not final[r, lens[r]:].any()
        xs = [int(x) % N for x in final[r, :lens[r]]]
        assert 0 not in xs and len(set(xs)) == len(xs)
        for x in xs:
            p = (r + x) % N % m
            assert (-x) % N in [int(y) % N for y in final[p, :lens[p]]]


def test_row_lengths_are_validated(gg):
    g, m, d = 8, 32, 6
    N = g * m
    tab = gg.orbit_part(g, m, d, 21)
    cfg = gg.search_cfg(mode=gg.MODE_PLAIN, rule=gg.RULE_HILL, total_steps=10, seed=1)
    ev = _tab_eval(N)
    for bad in (np.full(m, 0), np.full(m, d + 1), np.full(m, d - 1)):  # last: padding not 0
        with pytest.raises(gg.GGError):
            gg.OrbitSearch(N, tab, cfg, table_eval=ev, row_lengths=bad.astype(np.int32))


def ol_table_plain(gg, m, d):
    """A plain graph (g = 1) as a part table: offsets v -> u as (u - v)."""
    adj = gg.random_regular(m, d, 4)
    return (adj - np.arange(m)[:, None]).astype(np.int32)


def test_orbit_search_rejects_other_modes(gg):
    tab = gg.orbit_part(4, 16, 4, 2)
    cfg = gg.search_cfg(mode=gg.MODE_BATCHED, K=4, total_steps=10)
    with pytest.raises(gg.GGError):
        gg.OrbitSearch(64, tab, cfg, table_eval=_tab_eval(64))


def _lattice_eval(Ns, Ms):
    def f(t):
        if len(Ns) == 2:
            r = ol.piece_eval(Ns[0], Ns[1], Ms[0], Ms[1], t)
        else:
            r = ol.box_eval(Ns[0], Ns[1], Ns[2], Ms[0], Ms[1], Ms[2], t)
        ok = not r["unconnected"]
        return (r["total_scaled"] if ok else 0), (r["diameter"] if ok else 0), (0 if ok else 1)
    return f


def _check_lattice_symmetric_simple(L, t):
    for r in range(L.Mtot):
        assert L.row_simple(t, r) and L.row_symmetric(t, r), r


@pytest.mark.parametrize("Ns,Ms,d,rule", [
    ((8, 12), (2, 3), 4, "hill"),      # piece, both axes lifted: length changes too
    ((8, 12), (2, 3), 4, "anneal"),
    ((6, 9), (1, 9), 4, "hill"),       # piece with one axis unlifted: 2-opt only
    ((4, 4, 6), (2, 1, 2), 4, "hill"),  # box
    ((4, 6, 4), (2, 2, 2), 3, "anneal"),
])
def test_lattice_driver_matches_python_restatement(gg, Ns, Ms, d, rule):
    """pieceannealing.cpp:113-260 / boxHillClimb.cpp:113-273 moves and
    acceptance: the C++ driver and the restatement log the same moves and
    results step for step on the CPU oracle."""
    L = driver_ref.Lattice(Ns, Ms, d)
    tab = driver_ref.random_lattice_table(Ns, Ms, d, 5)
    steps = 60
    cfg = gg.search_cfg(mode=gg.MODE_PLAIN,
                        rule=gg.RULE_HILL if rule == "hill" else gg.RULE_ANNEAL,
                        total_steps=steps, T0=2e-2, T1=1e-4, seed=17)
    ev = _lattice_eval(Ns, Ms)
    N = int(np.prod(Ns))
    with gg.LatticeSearch(Ns, Ms, tab, cfg, table_eval=ev, log=steps) as s:
        st = s.run(steps)
        log = s.log()
        final = s.table()
    ref_log, ref_tab = driver_ref.run_orbit(
        N, tab, lambda t: (lambda r: (r[0], r[1], r[2] == 0))(ev(t)), steps, rule=rule,
        T0=2e-2, T1=1e-4, seed=17, propose=lambda t, rng: driver_ref.propose_lattice(t, L, rng))
    assert np.array_equal(log, ref_log)
    assert np.array_equal(final, ref_tab)
    assert st.steps == steps and st.exact_evals == steps
    kinds = set(ref_log[:, 5].tolist())
    assert kinds == ({0, 1} if all(g > 1 for g in L.g) else {1})
    _check_lattice_symmetric_simple(L, final)
    tot, dia, status = ev(final)
    assert status == 0 and st.total_cur == tot and st.diam_cur == dia
    if rule == "hill":
        t0, d0, _ = ev(tab)
        assert (st.diam_cur, st.total_cur) <= (d0, t0)


def test_lattice_generator_tables_are_symmetric_simple():
    for Ns, Ms, d in (((8, 12), (2, 3), 4), ((4, 4, 6), (2, 1, 2), 4)):
        L = driver_ref.Lattice(Ns, Ms, d)
        t = driver_ref.random_lattice_table(Ns, Ms, d, 1)
        _check_lattice_symmetric_simple(L, t)
        # expanded graph: every vertex has d distinct neighbours, undirected
        if len(Ns) == 2:
            adj = ol.expand_piece(Ns[0], Ns[1], Ms[0], Ms[1], t)
        else:
            adj = ol.expand_box(*Ns, *Ms, t)
        e = {(v, int(u)) for v in range(adj.shape[0]) for u in adj[v]}
        assert all((u, v) in e for v, u in e)
        assert all(len(set(adj[v].tolist())) == d and v not in adj[v] for v in range(adj.shape[0]))


@pytest.mark.parametrize("rule", ["anneal_wvc", "hill_wvc"])
def test_orbit_wvc_rules_match_python_restatement(gg, rule):
    g, m, d = 8, 32, 6
    N = g * m
    tab = gg.orbit_part(g, m, d, 21)
    steps = 80
    r = gg.RULE_ANNEAL_WVC if rule == "anneal_wvc" else gg.RULE_HILL_WVC
    cfg = gg.search_cfg(mode=gg.MODE_PLAIN, rule=r, total_steps=steps, T0=2e-2, T1=1e-4,
                        seed=99)

    def ev(t):
        o = ol.part_eval(N, m, t)
        ok = not o["unconnected"]
        return o["total_scaled"], o["diameter"] if ok else 0, 0 if ok else 1, o["wvc"]
    with gg.OrbitSearch(N, tab, cfg, table_eval=ev, log=steps) as s:
        st = s.run(steps)
        log = s.log()
        final = s.table()
    ref_log, ref_tab = driver_ref.run_orbit(
        N, tab, lambda t: (lambda x: (x[0], x[1], x[2] == 0, x[3]))(ev(t)), steps, rule=rule,
        T0=2e-2, T1=1e-4, seed=99)
    assert np.array_equal(log, ref_log)
    assert np.array_equal(final, ref_tab)
    assert st.accepted == int(ref_log[:, 9].sum())
