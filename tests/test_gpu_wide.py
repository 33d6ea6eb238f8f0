"""GPU parity of the level-synchronous path: a single evaluation whose column
bitset does not fit in LDS runs one launch per level over the whole grid
(GG_INFO_VARIANT 4), in passes of 256 columns when it has more.  Checked bit for bit
against the CPU oracle and against the persistent one-workgroup-per-column
global kernel (GG_OPT_WIDE 0, GG_INFO_VARIANT 3): distance totals, diameters,
per-source eccentricities, per-level counts, WVC, status codes."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _both(gg, make, evaluate, lds_limit=1024, pushes=True):
    """(result dict, ecc, reach) of the level-synchronous path with push levels
    allowed, with pull only, and of the persistent global kernel; all three
    must agree.  lds_limit forces the global kernels; `pushes`: whether the
    graph is one the host lets push (verified undirected)."""
    out = []
    for wide, push in ((1, 1), (1, 0), (0, 1)):
        with make() as ev:
            ev.set_option(gg.GG_OPT_LDS_LIMIT, lds_limit)
            ev.set_option(gg.GG_OPT_WIDE, wide)
            ev.set_option(gg.GG_OPT_PUSH, push)
            r = evaluate(ev)
            ecc, reach = ev.levels()
            out.append((r.as_dict(), ecc, reach, ev.get_option(gg.GG_INFO_VARIANT),
                        ev.get_option(gg.GG_INFO_PUSH_CAP)))
    (rw, ew, cw, vw, pw), (rq, eq, cq, vq, pq), (rp, ep, cp, vp, pp) = out
    assert (vw, vq, vp) == (4, 4, 3)
    assert (pw > 0) == pushes and pq == 0 and pp == 0
    assert rw == rq == rp
    assert np.array_equal(ew, eq) and np.array_equal(ew, ep)
    assert np.array_equal(cw, cq) and np.array_equal(cw, cp)
    return rw, ew, cw


@pytest.mark.parametrize("n,d,seed", [(2048, 8, 17), (3000, 5, 2), (1000, 6, 3)])
def test_adjacency_matches_oracle(gpu, n, d, seed):
    """Full and ragged last columns (3000 and 1000 sources), W = 47 pads to 64."""
    gg = gpu
    adj = gg.random_regular(n, d, seed)
    o = ol.adj_eval(adj)
    r, ecc, reach = _both(gg, lambda: gg.AdjEvaluator(n, d), lambda ev: ev.eval(adj))
    assert r["status"] == gg.GG_OK
    assert r["total_reps"] == o["total"] and r["diameter"] == o["diameter"]
    assert r["levels"] == o["diameter"] + 1
    assert np.array_equal(ecc, o["ecc"])
    assert np.array_equal(reach[:len(o["reach"])], o["reach"][:len(reach)])


def test_adjacency_in_several_passes_matches_oracle(gpu):
    """17000 sources = 266 columns: a pass of 256 columns, then one of 10
    with a ragged last column."""
    gg = gpu
    n, d = 17000, 6
    adj = gg.random_regular(n, d, 21)
    o = ol.adj_eval(adj)
    r, ecc, reach = _both(gg, lambda: gg.AdjEvaluator(n, d), lambda ev: ev.eval(adj))
    assert r["status"] == gg.GG_OK
    assert r["total_reps"] == o["total"] and r["diameter"] == o["diameter"]
    assert np.array_equal(ecc, o["ecc"])
    assert np.array_equal(reach[:len(o["reach"])], o["reach"][:len(reach)])


def test_plain_large_n_every_vertex_a_source(gpu):
    """The plain evaluator on a graph of N = 131840 > 2^17 vertices, every
    vertex a source (2060 columns: eight passes of 256 and one of 12), given
    as an explicit adjacency: an orbit graph of g = 1030 copies of 128
    representatives, so that translation by 128 is an automorphism and the
    oracle's BFS from the 128 representatives fixes everything: the total
    over all N sources is g times theirs, the diameter is theirs, vertex v's
    eccentricity is that of representative v mod 128, and the per-level
    counts are g times theirs.  The one-workgroup-per-column kernel agrees."""
    gg = gpu
    g, m, d = 1030, 128, 16
    N = g * m
    tab = gg.orbit_part(g, m, d, 3)
    o = ol.part_eval(N, m, tab)
    adj = ol.expand_part(N, m, tab)
    got = []
    for wide in (1, 0):
        with gg.AdjEvaluator(N, d) as ev:
            ev.set_option(gg.GG_OPT_WIDE, wide)
            print(f"large plain N={N} wide={wide}", flush=True)
            r = ev.eval(adj)
            assert ev.get_option(gg.GG_INFO_VARIANT) == (4 if wide else 3)
            ecc, reach = ev.levels()
            got.append((r.as_dict(), ecc, reach))
    (r, ecc, reach), (rg, eccg, reachg) = got
    assert r == rg and np.array_equal(ecc, eccg) and np.array_equal(reach, reachg)
    assert r["status"] == gg.GG_OK
    assert r["total_reps"] == g * o["total"] and r["diameter"] == o["diameter"]
    assert np.array_equal(ecc, np.tile(o["ecc"], g))
    L = len(o["reach"])
    assert np.array_equal(reach[:L], g * o["reach"][:len(reach)])


def test_part_piece_box_match_oracle(gpu):
    gg = gpu
    g, m, d = 64, 96, 6
    N = g * m
    tab = gg.orbit_part(g, m, d, 5)
    o = ol.part_eval(N, m, tab)
    r, ecc, _ = _both(gg, lambda: gg.PartEvaluator(N, m, d), lambda ev: ev.eval(tab))
    assert r["total_scaled"] == o["total_scaled"] and r["diameter"] == o["diameter"]
    assert r["wvc"] == o["wvc"]
    assert np.array_equal(ecc, o["ecc"])

    Nx, Ny, Mx, My, d = 32, 64, 2, 64, 6
    tab = gg.gen_lattice((Nx, Ny), (Mx, My), d, 7)
    o = ol.piece_eval(Nx, Ny, Mx, My, tab)
    r, ecc, _ = _both(gg, lambda: gg.PieceEvaluator(Nx, Ny, Mx, My, d), lambda ev: ev.eval(tab),
                      pushes=False)
    assert r["total_scaled"] == o["total_scaled"] and r["diameter"] == o["diameter"]
    assert np.array_equal(ecc, o["ecc"])

    Ns, Ms, d = (8, 8, 16), (2, 2, 16), 6
    tab = gg.gen_lattice(Ns, Ms, d, 9)
    o = ol.box_eval(*Ns, *Ms, tab)
    r, ecc, _ = _both(gg, lambda: gg.BoxEvaluator(*Ns, *Ms, d), lambda ev: ev.eval(tab),
                      pushes=False)
    assert r["total_scaled"] == o["total_scaled"] and r["diameter"] == o["diameter"]
    assert np.array_equal(ecc, o["ecc"])


def test_swap_unconnected_and_level_limit(gpu):
    gg = gpu
    n, d = 2048, 8
    adj = gg.random_regular(n, d, 6)
    c = 700
    sw, new = gg.make_swap(adj, 5, 1, c, 2, 0)
    while sw is None:
        c += 1
        sw, new = gg.make_swap(adj, 5, 1, c, 2, 0)
    o = ol.adj_eval(new)

    def swap_eval(ev):
        ev.set_base(adj)
        return ev.eval_swap(sw)
    r, ecc, _ = _both(gg, lambda: gg.SwapEvaluator(n, d), swap_eval)
    assert r["total_reps"] == o["total"] and r["diameter"] == o["diameter"]
    assert np.array_equal(ecc, o["ecc"])

    # two disjoint random regular graphs of 600 vertices
    h = gg.random_regular(600, 4, 3)
    two = np.concatenate([h, h + 600]).astype(np.int32)
    o = ol.adj_eval(two)
    r, _, _ = _both(gg, lambda: gg.AdjEvaluator(1200, 4), lambda ev: ev.eval(two))
    assert o["unconnected"] and r["status"] == gg.GG_UNCONNECTED
    assert r["total_reps"] == o["total"]

    def limited(ev):
        ev.set_option(gg.GG_OPT_LEVEL_LIMIT, 2)
        return ev.eval(adj)
    r, _, _ = _both(gg, lambda: gg.AdjEvaluator(n, d), limited)
    assert r["status"] == gg.GG_DIAMETER_LIMIT


def test_directed_part_table_pulls(gpu):
    """Random offsets (a directed table, which no reference driver makes): the
    level-synchronous path must not push there, and agrees with the
    persistent kernel (both pull from listed neighbours)."""
    gg = gpu
    g, m, d = 32, 96, 5
    N = g * m
    rng = np.random.default_rng(11)
    tab = rng.integers(1, N, size=(m, d)).astype(np.int32)
    _both(gg, lambda: gg.PartEvaluator(N, m, d), lambda ev: ev.eval(tab), pushes=False)


@pytest.mark.parametrize("g,m,d,directed", [(64, 64, 6, False), (48, 128, 5, True),
                                             (2, 2048, 4, False)])
def test_copy_major_word_order(gpu, g, m, d, directed):
    """Tables of 2^k rows repeated over the vertices keep their words copy by
    copy (row r's copies side by side); push, pull and the persistent kernel
    agree with each other and the oracle, for orbit graphs and for a
    directed table (pull only)."""
    gg = gpu
    N = g * m
    if directed:
        rng = np.random.default_rng(g)
        tab = rng.integers(1, N, size=(m, d)).astype(np.int32)
    else:
        tab = gg.orbit_part(g, m, d, 7)
    r, ecc, reach = _both(gg, lambda: gg.PartEvaluator(N, m, d), lambda ev: ev.eval(tab),
                          pushes=not directed)
    if not directed:
        o = ol.part_eval(N, m, tab)
        assert r["total_scaled"] == o["total_scaled"] and r["diameter"] == o["diameter"]
        assert np.array_equal(ecc, o["ecc"])
        assert np.array_equal(reach[:len(o["reach"])], o["reach"][:len(reach)])


def test_large_orbit_graph_matches_oracle(gpu):
    """N = 2^18 with 128 representatives (two columns): the level-synchronous
    path is the automatic choice and matches the oracle."""
    gg = gpu
    g, m, d = 2048, 128, 16
    N = g * m
    tab = gg.orbit_part(g, m, d, 3)
    o = ol.part_eval(N, m, tab)
    with gg.PartEvaluator(N, m, d) as ev:
        r = ev.eval(tab)
        assert ev.get_option(gg.GG_INFO_VARIANT) == 4
        assert ev.get_option(gg.GG_INFO_PUSH_CAP) > 0
        ecc, _ = ev.levels()
    assert r.status == gg.GG_OK
    assert r.total_scaled == o["total_scaled"] and r.diameter == o["diameter"]
    assert r.wvc == o["wvc"]
    assert np.array_equal(ecc, o["ecc"])


def test_push_continues_while_lists_fit(gpu):
    """A ring of 256 vertices with the diameter chord, as a part table of 16
    rows (g = 16): from 16 consecutive sources every level adds about four
    frontier words, so no list of the push levels overflows and every level
    up to the diameter pushes (level 3 onwards reads the list level 2 wrote).
    Push, pull and the persistent kernel agree with the oracle."""
    gg = gpu
    N, M = 256, 16
    tab = np.tile(np.array([1, N - 1, N // 2], dtype=np.int32), (M, 1))
    o = ol.part_eval(N, M, tab)
    r, ecc, reach = _both(gg, lambda: gg.PartEvaluator(N, M, 3), lambda ev: ev.eval(tab))
    assert o["diameter"] > 8
    assert r["total_scaled"] == o["total_scaled"] and r["diameter"] == o["diameter"]
    assert np.array_equal(ecc, o["ecc"])
    assert np.array_equal(reach[:len(o["reach"])], o["reach"][:len(reach)])


def test_part_table_with_partial_copy_pulls(gpu):
    """M does not divide N: the rows of the vertices past the last full copy
    wrap differently, so a table whose rows 0..M-1 pass the undirected test
    is still treated as directed (push capacity 0); push-allowed, pull-only
    and persistent runs agree."""
    gg = gpu
    g, m, d = 32, 96, 6
    N = g * m + 24
    tab = gg.orbit_part(g, m, d, 5)
    _both(gg, lambda: gg.PartEvaluator(N, m, d), lambda ev: ev.eval(tab), pushes=False)
