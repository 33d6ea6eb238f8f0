"""Pins the CPU oracle (test infrastructure) before it is trusted as the checker.

Each check is independent of the oracle's own BFS: closed forms, Floyd-Warshall
brute force, invariants between encodings, the reference's own cpuASPL*
templates compiled from the reference tree (oracle/_ref), and committed golden
vectors written by tests/golden/make_golden.py from that reference build.
"""
import json
import os

import numpy as np
import pytest

import graphs
import oracle_lib as ol

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "golden.json")


@pytest.mark.parametrize("n", [3, 4, 5, 8, 11, 16, 31, 40])
def test_cycle_closed_form(n):
    r = ol.adj_eval(graphs.cycle(n))
    assert r["total"] == graphs.cycle_total(n)
    assert r["diameter"] == n // 2


@pytest.mark.parametrize("n", [2, 3, 7, 20])
def test_complete_closed_form(n):
    r = ol.adj_eval(graphs.complete(n))
    assert r["total"] == n * (n - 1) and r["diameter"] == 1


@pytest.mark.parametrize("k", [1, 2, 3, 5, 7])
def test_hypercube_closed_form(k):
    r = ol.adj_eval(graphs.hypercube(k))
    assert r["total"] == graphs.hypercube_total(k) and r["diameter"] == k


def test_petersen_and_kdd():
    r = ol.adj_eval(graphs.petersen())
    assert r["total"] == 150 and r["diameter"] == 2
    for d in (2, 3, 6):
        r = ol.adj_eval(graphs.kdd(d))
        assert r["total"] == 2 * d * (3 * d - 2) and r["diameter"] == 2


def _random_graph(n, p, rng):
    m = rng.random((n, n)) < p
    m = np.triu(m, 1)
    m = m | m.T
    deg = m.sum(1).max() if n else 0
    adj = -np.ones((n, max(1, deg)), dtype=np.int32)
    for v in range(n):
        nb = np.nonzero(m[v])[0]
        adj[v, :len(nb)] = nb
    return adj


@pytest.mark.parametrize("seed", range(12))
def test_bfs_matches_floyd_warshall(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 60))
    adj = _random_graph(n, float(rng.uniform(0.03, 0.4)), rng)
    r = ol.adj_eval(adj)
    fw = ol.floyd_warshall(adj)
    assert r["unconnected"] == fw["unconnected"]
    assert r["total"] == fw["total"]
    if not fw["unconnected"]:
        assert r["diameter"] == fw["diameter"]
        # ecc(s) = max_v d(s,v); reach[l] = #pairs at distance l
        assert np.array_equal(r["ecc"], fw["dist"].max(1))
        for l in range(fw["diameter"] + 1):
            assert r["reach"][l] == (fw["dist"] == l).sum()
    else:
        assert r["diameter"] == ol.INT_MAX


def test_symmetric_part_scales_to_expanded():
    # orbit graph: BFS from M reps, times g == BFS from all N vertices
    import graphgolf_mi355x as gg
    g, m, d = 8, 24, 6
    tab = gg.orbit_part(g, m, d, 11)
    N = g * m
    sym = ol.part_eval(N, m, tab)
    full = ol.adj_eval(ol.expand_part(N, m, tab))
    assert sym["total"] * g == full["total"]
    assert sym["diameter"] == full["diameter"]


def test_part_and_piece_encodings_agree():
    import graphgolf_mi355x as gg
    g, m, d = 32, 128, 16
    tab = gg.orbit_part(g, m, d, 3)
    # orbit offset (j - i) + m*s  <->  piece offset (dx, dy) = (s, j - i), Nx = g, Mx = 1
    s = np.floor_divide(tab + np.arange(m)[:, None], m)  # = s for j - i + m*s with 0 <= j < m
    dy = tab - m * s
    piece = np.stack([s, dy], axis=2)
    p = ol.piece_eval(g, m, 1, m, piece)
    q = ol.part_eval(g * m, m, tab)
    assert p["total"] == q["total"] and p["diameter"] == q["diameter"]


@pytest.mark.skipif(not ol.ref_available(), reason="reference tree not present")
@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5])
def test_oracle_vs_reference_build_c1(seed):
    import graphgolf_mi355x as gg
    adj = gg.random_regular(256, 4, seed)
    r = ol.adj_eval(adj)
    d, a = ol.ref_part(256, 256, graphs.part_of_adj(adj))
    assert d == r["diameter"]
    assert a == ol.aspl(r["total"], 256)


@pytest.mark.skipif(not ol.ref_available(), reason="reference tree not present")
def test_oracle_vs_reference_build_symmetric():
    import graphgolf_mi355x as gg
    rng = np.random.default_rng(5)
    # part with M < N, M not a multiple of 64
    tab = gg.orbit_part(6, 50, 4, 9)
    r = ol.part_eval(300, 50, tab)
    assert ol.ref_part(300, 50, tab) == (r["diameter"], ol.aspl(r["total_scaled"], 300))
    # piece / box with random offsets; per-axis offsets stay >= -N (the
    # reference indexes out of bounds below that, cpuASPLpiece.cpp:20-21)
    for _ in range(4):
        Nx, Ny, Mx, My, D = 6, 10, 2, 5, 3
        t = np.stack([rng.integers(-Nx, 2 * Nx, size=(Mx * My, D)),
                      rng.integers(-Ny, 2 * Ny, size=(Mx * My, D))], axis=2)
        r = ol.piece_eval(Nx, Ny, Mx, My, t)
        dref, aref = ol.ref_piece(Nx, Ny, Mx, My, t)
        if not r["unconnected"]:
            assert (dref, aref) == (r["diameter"], ol.aspl(r["total_scaled"], Nx * Ny))
        Nx, Ny, Nz, Mx, My, Mz, D = 4, 6, 3, 2, 3, 1, 3
        t = np.stack([rng.integers(-n, 2 * n, size=(Mx * My * Mz, D)) for n in (Nx, Ny, Nz)],
                     axis=2)
        r = ol.box_eval(Nx, Ny, Nz, Mx, My, Mz, t)
        dref, aref = ol.ref_box(Nx, Ny, Nz, Mx, My, Mz, t)
        if not r["unconnected"]:
            assert (dref, aref) == (r["diameter"], ol.aspl(r["total_scaled"], Nx * Ny * Nz))


@pytest.mark.skipif(not ol.ref_available(), reason="reference tree not present")
def test_oracle_vs_reference_build_unconnected():
    adj = graphs.disjoint_cycles(7, 9)
    r = ol.adj_eval(adj)
    d, a = ol.ref_part(16, 16, graphs.part_of_adj(adj))
    assert d == ol.INT_MAX and r["diameter"] == ol.INT_MAX
    assert a == ol.aspl(r["total"], 16)  # partial total over reached pairs


def _wvc_by_reference_process(adj, M):
    """WVC by replaying the reference GPU process with 32-bit words
    (cudacode/cudaASPLconv.cu:244-308): AND over all words of all vertices each
    level, WVC = popcount(mask ^ complete) at each level with updates."""
    N = adj.shape[0]
    width = 1
    while 32 * width < M:
        width *= 2
    bits = np.zeros((N, width), dtype=np.uint64)
    for b in range(width):
        lo = 32 * b
        if M >= lo + 32:
            pad = 0
        elif lo >= M:
            pad = 0xFFFFFFFF
        else:
            pad = (~((1 << (M - lo)) - 1)) & 0xFFFFFFFF
        bits[:, b] = pad
    for v in range(min(N, 32 * width)):
        bits[v, v // 32] |= np.uint64(1 << (v % 32))
    mask = int(np.bitwise_and.reduce(bits.ravel()))
    wvc = None
    for step in range(1, 101):
        new = np.zeros_like(bits)
        for i in range(adj.shape[1]):
            new |= bits[adj[:, i]]
        new &= ~bits & np.uint64(0xFFFFFFFF)
        bits |= new
        upd = int(np.bitwise_or.reduce(new.ravel()))
        comp = int(np.bitwise_and.reduce(bits.ravel()))
        if upd:
            wvc = bin(mask ^ comp).count("1")
            mask = comp
        else:
            return wvc, step - 1
    return None, None


@pytest.mark.parametrize("N,M", [(40, 40), (72, 72), (96, 24), (130, 65)])
def test_wvc_definition(N, M):
    import graphgolf_mi355x as gg
    if N == M:
        adj = gg.random_regular(N, 3 if N % 2 == 0 else 4, N)
        tab = graphs.part_of_adj(adj)
    else:
        tab = gg.orbit_part(N // M, M, 4, N)
        adj = ol.expand_part(N, M, tab)
    r = ol.part_eval(N, M, tab)
    w, diam = _wvc_by_reference_process(adj, M)
    assert diam == r["diameter"]
    assert w == r["wvc"]


def test_golden_vectors():
    """Values written by the reference's own cpuASPLqueue build (make_golden.py)."""
    import hashlib
    import graphgolf_mi355x as gg
    with open(GOLDEN) as f:
        gold = json.load(f)
    for case in gold["cases"]:
        if case["N"] > 8192:
            continue  # large cases: hash only here; totals checked on the GPU box
        if case["kind"] == "regular":
            adj = gg.random_regular(case["N"], case["d"], case["seed"])
            assert hashlib.sha256(adj.tobytes()).hexdigest() == case["sha256"]
            r = ol.adj_eval(adj)
            assert r["total"] == case["total_scaled"] and r["diameter"] == case["diameter"]
        else:
            tab = gg.orbit_part(case["g"], case["m"], case["d"], case["seed"])
            assert hashlib.sha256(tab.tobytes()).hexdigest() == case["sha256"]
            r = ol.part_eval(case["N"], case["m"], tab)
            assert r["total_scaled"] == case["total_scaled"]
            assert r["diameter"] == case["diameter"]
