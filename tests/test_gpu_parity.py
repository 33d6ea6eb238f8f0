"""GPU parity: the HIP evaluators against the CPU oracle, bit-exact on the
integer distance total, the diameter, per-source eccentricities and the
reference WVC, on the seeded config shapes and on edge cases."""
import json
import os

import numpy as np
import pytest

import graphs
import oracle_lib as ol

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "golden.json")


def _check_adj(gg, adj, r, o):
    assert r.total_reps == o["total"], (r.as_dict(), o["total"])
    if o["unconnected"]:
        assert r.status == gg.GG_UNCONNECTED
    else:
        assert r.status == gg.GG_OK, r.as_dict()
        assert r.diameter == o["diameter"]
        assert r.levels == o["diameter"] + 1


@pytest.mark.parametrize("name,adj", [
    ("C5cycle", graphs.cycle(5)), ("C64", graphs.cycle(64)), ("C65", graphs.cycle(65)),
    ("K7", graphs.complete(7)), ("Q6", graphs.hypercube(6)), ("Q10", graphs.hypercube(10)),
    ("petersen", graphs.petersen()), ("K33", graphs.kdd(3)),
    ("circ", graphs.circulant(300, [1, 7, 31])), ("two_cycles", graphs.disjoint_cycles(30, 40)),
])
def test_closed_form_graphs(gpu, name, adj):
    gg = gpu
    n, d = adj.shape
    o = ol.adj_eval(adj)
    with gg.AdjEvaluator(n, d) as ev:
        r = ev.eval(adj)
    _check_adj(gg, adj, r, o)
    # the same graph as a 'part' with M = N
    with gg.PartEvaluator(n, n, d) as ev:
        r2 = ev.eval(graphs.part_of_adj(adj))
    assert r2.total_reps == o["total"] and r2.status == r.status


def test_golden_configs(gpu):
    """Every config graph against the values of the reference's own CPU oracle."""
    gg = gpu
    gold = json.load(open(GOLDEN))
    for case in gold["cases"]:
        N = case["N"]
        print("golden case", case["name"], N, flush=True)
        if case["kind"] == "regular":
            adj = gg.random_regular(N, case["d"], case["seed"])
            with gg.AdjEvaluator(N, case["d"]) as ev:
                r = ev.eval(adj)
            assert r.total_scaled == case["total_scaled"]
        else:
            tab = gg.orbit_part(case["g"], case["m"], case["d"], case["seed"])
            with gg.PartEvaluator(N, case["m"], case["d"]) as ev:
                r = ev.eval(tab)
            assert r.total_scaled == case["total_scaled"]
        assert r.diameter == case["diameter"] and r.status == gg.GG_OK
        assert gg.aspl(r, N) == case["aspl"]


@pytest.mark.parametrize("n,d,seed", [(256, 4, 1), (256, 4, 2), (1000, 6, 3), (4096, 12, 2),
                                      (8192, 16, 4)])
def test_random_regular_parity_with_levels(gpu, n, d, seed):
    gg = gpu
    adj = gg.random_regular(n, d, seed)
    o = ol.adj_eval(adj)
    with gg.AdjEvaluator(n, d) as ev:
        r = ev.eval(adj)
        _check_adj(gg, adj, r, o)
        ecc, reach = ev.levels()
    assert np.array_equal(ecc, o["ecc"])
    assert np.array_equal(reach[:len(o["reach"])], o["reach"][:len(reach)])


@pytest.mark.parametrize("block", [256, 1024])
@pytest.mark.parametrize("lds_limit", [0, 1024])
@pytest.mark.parametrize("generic", [0, 1])
@pytest.mark.parametrize("push", [1, 0])
def test_kernel_variants_agree(gpu, block, lds_limit, generic, push):
    """256/1024-thread LDS kernels, the global-memory kernel, fixed-degree
    (prefetching) vs runtime-degree gathers, push vs pull first levels."""
    gg = gpu
    adj = gg.random_regular(2048, 8, 17)
    o = ol.adj_eval(adj)
    with gg.AdjEvaluator(2048, 8) as ev:
        ev.set_option(gg.GG_OPT_BLOCK, block)
        ev.set_option(gg.GG_OPT_LDS_LIMIT, lds_limit)
        ev.set_option(gg.GG_OPT_GENERIC, generic)
        ev.set_option(gg.GG_OPT_PUSH, push)
        r = ev.eval(adj)
        variant = ev.get_option(gg.GG_INFO_VARIANT)
        fixed = ev.get_option(gg.GG_INFO_FIXED)
        push_cap = ev.get_option(gg.GG_INFO_PUSH_CAP)
        ecc, _ = ev.levels()
    # the global kernels: a single evaluation of 32 columns runs level-synchronously
    assert variant == (4 if lds_limit else (1 if block == 256 else 2))
    assert fixed == (8 if (variant == 2 and not generic) else 0)
    # both the LDS kernels and the level-synchronous path push on this undirected graph
    assert (push_cap > 0) == bool(push)
    _check_adj(gg, adj, r, o)
    assert np.array_equal(ecc, o["ecc"])


def test_push_only_on_undirected(gpu):
    """The push levels run only where they equal the pull BFS: on an
    undirected adjacency, for swaps that keep it undirected; directed
    adjacencies and reference offset tables stay pull-only."""
    gg = gpu
    n, d = 1024, 6
    adj = gg.random_regular(n, d, 3)
    with gg.SwapEvaluator(n, d) as ev:
        ev.set_base(adj)
        assert ev.get_option(gg.GG_INFO_PUSH_CAP) > 0
        c = 700
        sw, new = gg.make_swap(adj, 5, 1, c, 2, 0)
        while sw is None:
            c += 1
            sw, new = gg.make_swap(adj, 5, 1, c, 2, 0)
        r = ev.eval_swap(sw)
        assert ev.get_option(gg.GG_INFO_PUSH_CAP) > 0
        o = ol.adj_eval(new)
        assert r.total_reps == o["total"] and r.diameter == o["diameter"]
    directed = adj.copy()
    directed[0, 0] = next(v for v in range(1, n) if v not in adj[0])
    o = ol.adj_eval(directed)
    with gg.AdjEvaluator(n, d) as ev:
        r = ev.eval(directed)
        assert ev.get_option(gg.GG_INFO_PUSH_CAP) == 0
    assert r.total_reps == o["total"]
    tab = gg.orbit_part(8, 128, 6, 4)
    with gg.PartEvaluator(8 * 128, 128, 6) as ev:
        ev.eval(tab)
        assert ev.get_option(gg.GG_INFO_PUSH_CAP) == 0


@pytest.mark.parametrize("g,m,d,seed", [(32, 128, 16, 3), (6, 50, 4, 9), (4, 20, 3, 1),
                                        (16, 70, 6, 2), (2, 300, 5, 8)])
def test_symmetric_part_parity(gpu, g, m, d, seed):
    gg = gpu
    if (m * d) % 2:
        d += 1
    tab = gg.orbit_part(g, m, d, seed)
    N = g * m
    o = ol.part_eval(N, m, tab)
    with gg.PartEvaluator(N, m, d) as ev:
        r = ev.eval(tab)
        ecc, _ = ev.levels()
    assert r.total_reps == o["total"] and r.total_scaled == o["total_scaled"]
    if not o["unconnected"]:
        assert r.diameter == o["diameter"] and r.status == gg.GG_OK
        assert r.wvc == o["wvc"]
        assert np.array_equal(ecc, o["ecc"])
    else:
        assert r.status == gg.GG_UNCONNECTED


def test_part_with_irregular_offsets(gpu):
    """Random offset tables, offsets beyond +-N included.  Such tables describe
    directed graphs: the reference GPU kernel pulls from out-neighbours
    (cudacode/cudaASPLconv.cu:35-48), so its per-source eccentricities are
    in-eccentricities, while the CPU queue BFS (cpuASPLqueue.cpp:80-84) yields
    out-eccentricities.  The totals and the diameter agree by the cyclic
    symmetry of 'part' graphs; WVC (a per-source quantity) is only compared on
    undirected graphs, the only kind the reference's drivers produce."""
    gg = gpu
    rng = np.random.default_rng(3)
    for trial in range(6):
        M = int(rng.integers(1, 200))
        g = int(rng.integers(1, 6))
        N = M * g
        D = int(rng.integers(1, 7))
        tab = rng.integers(-N, 3 * N, size=(M, D)).astype(np.int32)
        o = ol.part_eval(N, M, tab)
        with gg.PartEvaluator(N, M, D) as ev:
            r = ev.eval(tab)
        assert r.total_reps == o["total"], trial
        if not o["unconnected"]:
            assert r.diameter == o["diameter"]


def _sym_piece(rng, Nx, Ny, Mx, My, D):
    """Undirected piece table by stub pairing over base vertices with random
    lifts (the construction of cppcode/RNDpiecegen.cpp:89-127)."""
    stubs = [(x, y) for x in range(Mx) for y in range(My) for _ in range(D)]
    rng.shuffle(stubs)
    rows = [[] for _ in range(Mx * My)]
    for i in range(0, len(stubs) - 1, 2):
        (ax, ay), (bx, by) = stubs[i], stubs[i + 1]
        dx = (bx - ax) + Mx * int(rng.integers(0, Nx // Mx))
        dy = (by - ay) + My * int(rng.integers(0, Ny // My))
        rows[ax * My + ay].append((dx, dy))
        rows[bx * My + by].append((-dx, -dy))
    return np.array(rows, dtype=np.int32)


def _sym_box(rng, N3, M3, D):
    Nx, Ny, Nz = N3
    Mx, My, Mz = M3
    stubs = [(x, y, z) for x in range(Mx) for y in range(My) for z in range(Mz) for _ in range(D)]
    rng.shuffle(stubs)
    rows = [[] for _ in range(Mx * My * Mz)]
    for i in range(0, len(stubs) - 1, 2):
        a, b = stubs[i], stubs[i + 1]
        dv = [(b[t] - a[t]) + M3[t] * int(rng.integers(0, N3[t] // M3[t])) for t in range(3)]
        rows[(a[0] * My + a[1]) * Mz + a[2]].append(tuple(dv))
        rows[(b[0] * My + b[1]) * Mz + b[2]].append(tuple(-x for x in dv))
    return np.array(rows, dtype=np.int32)


def _reverse(adj):
    """Reversed explicit adjacency (-1 padded)."""
    n = adj.shape[0]
    rows = [[] for _ in range(n)]
    for v in range(n):
        for u in adj[v]:
            if u >= 0:
                rows[u].append(v)
    dm = max(1, max(len(r) for r in rows))
    out = -np.ones((n, dm), dtype=np.int32)
    for v, r in enumerate(rows):
        out[v, :len(r)] = r
    return out


def test_piece_and_box_parity(gpu):
    gg = gpu
    rng = np.random.default_rng(7)
    # C3 as a piece: Nx = 32, Mx = 1, Ny = My = 128, offset (s, j - i)
    tab = gg.orbit_part(32, 128, 16, 3)
    s = np.floor_divide(tab + np.arange(128)[:, None], 128)
    piece = np.stack([s, tab - 128 * s], axis=2)
    o = ol.piece_eval(32, 128, 1, 128, piece)
    with gg.PieceEvaluator(32, 128, 1, 128, 16) as ev:
        r = ev.eval(piece)
    assert r.total_scaled == o["total_scaled"] == ol.part_eval(4096, 128, tab)["total_scaled"]
    assert r.diameter == o["diameter"]
    # undirected random pieces / boxes (periods dividing the torus)
    for trial in range(6):
        Mx, My = int(rng.integers(1, 5)), int(rng.integers(1, 5))
        Nx, Ny = Mx * int(rng.integers(1, 5)), My * int(rng.integers(1, 5))
        D = 2 * int(rng.integers(1, 4))
        t = _sym_piece(rng, Nx, Ny, Mx, My, D)
        o = ol.piece_eval(Nx, Ny, Mx, My, t)
        with gg.PieceEvaluator(Nx, Ny, Mx, My, D) as ev:
            r = ev.eval(t)
        assert r.total_scaled == o["total_scaled"], trial
        if not o["unconnected"]:
            assert r.diameter == o["diameter"]
        M3 = [int(x) for x in rng.integers(1, 4, size=3)]
        N3 = [m * int(rng.integers(1, 4)) for m in M3]
        D = 2 * int(rng.integers(1, 3))
        t = _sym_box(rng, N3, M3, D)
        o = ol.box_eval(*N3, *M3, t)
        with gg.BoxEvaluator(*N3, *M3, D) as ev:
            r = ev.eval(t)
        assert r.total_scaled == o["total_scaled"], trial
        if not o["unconnected"]:
            assert r.diameter == o["diameter"]


def test_piece_directed_pull_semantics(gpu):
    """A directed piece table with Mx not dividing Nx: the kernel pulls from
    listed neighbours (cudacode/cudaASPLpiece.cu:35-54), i.e. it runs the BFS
    of the reversed graph from the representatives."""
    gg = gpu
    rng = np.random.default_rng(11)
    for trial in range(4):
        Nx, Ny = int(rng.integers(3, 12)), int(rng.integers(3, 12))
        Mx, My = int(rng.integers(1, Nx + 1)), int(rng.integers(1, Ny + 1))
        D = int(rng.integers(2, 6))
        t = rng.integers(-2 * Nx, 2 * Ny, size=(Mx * My, D, 2))
        adj = ol.expand_piece(Nx, Ny, Mx, My, t)
        srcs = np.array([x * Ny + y for x in range(Mx) for y in range(My)], dtype=np.int32)
        o = ol.bfs(_reverse(adj), srcs)
        with gg.PieceEvaluator(Nx, Ny, Mx, My, D) as ev:
            r = ev.eval(t)
        assert r.total_reps == o["total"], trial
        if not o["unconnected"]:
            assert r.diameter == o["diameter"]


def test_diameter_limit_and_unconnected(gpu):
    gg = gpu
    # diameter 98 is evaluated; 99 and 100 hit the reference's level limit
    for n, want in [(196, gg.GG_OK), (198, gg.GG_DIAMETER_LIMIT), (200, gg.GG_DIAMETER_LIMIT)]:
        with gg.AdjEvaluator(n, 2) as ev:
            r = ev.eval(graphs.cycle(n))
        assert r.status == want, (n, r.as_dict())
        if want == gg.GG_OK:
            assert r.diameter == 98 and r.total_reps == graphs.cycle_total(n)
    adj = graphs.disjoint_cycles(100, 101)
    o = ol.adj_eval(adj)
    with gg.AdjEvaluator(201, 2) as ev:
        r = ev.eval(adj)
    assert r.status == gg.GG_UNCONNECTED and r.total_reps == o["total"]


@pytest.mark.parametrize("seed", range(8))
def test_fuzz_small_irregular(gpu, seed):
    gg = gpu
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(2, 300))
    p = float(rng.uniform(0.005, 0.1))
    m = np.triu(rng.random((n, n)) < p, 1)
    m = m | m.T
    deg = max(1, int(m.sum(1).max()))
    adj = -np.ones((n, deg), dtype=np.int32)
    for v in range(n):
        nb = np.nonzero(m[v])[0]
        adj[v, :len(nb)] = nb
    o = ol.adj_eval(adj)
    with gg.AdjEvaluator(n, deg) as ev:
        r = ev.eval(adj)
    _check_adj(gg, adj, r, o)


def test_repeated_calls_reset_state(gpu):
    gg = gpu
    adjs = [gg.random_regular(512, 6, s) for s in range(4)]
    with gg.AdjEvaluator(512, 6) as ev:
        for _ in range(3):
            for a in adjs:
                r = ev.eval(a)
                assert r.total_reps == ol.adj_eval(a)["total"]


def test_autotuner_picks_and_caches(gpu, tmp_path, monkeypatch):
    gg = gpu
    monkeypatch.setenv("GG_TUNE_CACHE", str(tmp_path / "tune.txt"))
    adj = gg.random_regular(4096, 12, 2)
    o = ol.adj_eval(adj)
    with gg.AdjEvaluator(4096, 12) as ev:
        rep = ev.autotune(adj, reps=3)
        assert rep.count == 4 and rep.best_ms == min(rep.ms[:4])
        assert (ev.get_option(gg.GG_OPT_BLOCK), ev.get_option(gg.GG_OPT_GENERIC)) == \
            (rep.best_block, rep.best_generic)
        r = ev.eval(adj)
        assert r.total_reps == o["total"]
        print("autotune n=4096 d=12:", [(rep.block[i], rep.generic[i], round(rep.ms[i], 4))
                                        for i in range(4)])
    text = (tmp_path / "tune.txt").read_text()
    assert " 4096 4096 12 " in text
    with gg.AdjEvaluator(4096, 12) as ev2:  # starts from the cached choice
        assert ev2.get_option(gg.GG_OPT_BLOCK) == rep.best_block
        assert ev2.get_option(gg.GG_OPT_GENERIC) == rep.best_generic


def test_time_bfs_relaunches_leave_results_intact(gpu):
    """gg_eval_time_bfs re-runs the last evaluation's BFS kernel back to back;
    the kernel is idempotent, so later evaluations are unchanged."""
    gg = gpu
    n, d = 2048, 8
    adj = gg.random_regular(n, d, 3)
    rng = np.random.default_rng(5)
    ev = gg.SwapEvaluator(n, d)
    ev.set_base(adj)
    sw = None
    while sw is None:
        a, c = (int(x) for x in rng.integers(0, n, 2))
        ia, ic = (int(x) for x in rng.integers(0, d, 2))
        sw, new = gg.make_swap(adj, a, ia, c, ic, 1)
    r1 = ev.eval_swap(sw).as_dict()
    t = ev.time_bfs(7)
    assert t > 0
    r2 = ev.eval_swap(sw).as_dict()
    assert r1 == r2
    o = ol.adj_eval(new)
    assert r1["total_reps"] == o["total"] and r1["diameter"] == o["diameter"]
    ev.close()
