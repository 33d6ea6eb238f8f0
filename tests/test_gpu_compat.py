"""The drop-in evaluator classes (include/graphgolf_compat.hpp) on the GPU,
called from C++ the way the reference's drivers call cudaASPLconv,
cudaASPLpiece, cudaASPLbox and cudaASPLgraph, against the oracle: values,
row padding and truncation, and every sentinel the reference returns
(cudacode/cudaASPLconv.cu:182,229,293; cudaASPLgraph.cu:182,190), with the
level limits at their boundaries (diameter 98/99/100 for the part
evaluator, 198/199 for the graph evaluator).  Also the one-shot evaluator
CLI's piece and box output (cppcode/piece2aspl.cpp:102-109,
box2aspl.cpp:102-109)."""
import os
import re
import subprocess

import numpy as np
import pytest

import graphs
import oracle_lib as ol

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(REPO, "tests", "compat", "compat_check")
BIN = os.path.join(REPO, "graphgolf-cuda_amd", "bin")
BAD = 100000000


def _run(args):
    print("run", " ".join(str(a) for a in args), flush=True)
    p = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


def _call(mode, path):
    return _run([CHECK, mode, path]).split()


def _write_part(path, N, rows):
    with open(path, "w") as f:
        f.write(f"{N} {len(rows)}\n")
        for r in rows:
            f.write(" ".join(str(int(x)) for x in [len(r)] + list(r)) + "\n")


def _write_edges(path, adj):
    with open(path, "w") as f:
        for v, row in enumerate(adj):
            for u in row:
                if v < int(u):
                    f.write(f"{v} {int(u)}\n")


def _ring_rows(n):
    return [[1, -1]] * n


@pytest.fixture(scope="module")
def built(gpu):
    assert os.path.exists(CHECK), "tests/compat/compat_check not built: run __graft_entry__.build()"
    return gpu


def test_conv_methods_match_oracle_with_short_rows(built, tmp_path):
    gg = built
    n, d = 96, 4
    adj = gg.random_regular(n, d, 3)
    # drop the edge {0, u}: two rows of length 3, which the shim pads with
    # offset 0 (a self-loop) as cudaASPLconv.cu:199-201 does
    u = int(adj[0][0])
    nbrs = [[int(x) for x in adj[v] if {v, int(x)} != {0, u}] for v in range(n)]
    rows = [[x - v for x in nb] for v, nb in enumerate(nbrs)]
    assert len(rows[0]) == len(rows[u]) == d - 1
    pad = np.array([nb + [v] * (d - len(nb)) for v, nb in enumerate(nbrs)], dtype=np.int32)
    o = ol.adj_eval(pad)
    assert not o["unconnected"]
    f = tmp_path / "g.part"
    _write_part(f, n, rows)
    assert float(_call("conv-calc", f)[0]) == o["total"] / (n * (n - 1))
    dia, aspl = _call("conv-diam", f)
    assert int(dia) == o["diameter"] and float(aspl) == o["total"] / (n * (n - 1))
    wvc, dia, aspl = _call("conv-wvc", f)
    assert int(dia) == o["diameter"] and float(aspl) == o["total"] / (n * (n - 1))
    assert int(wvc) == ol.wvc(n, o["ecc"], o["diameter"])


def test_conv_symmetric_part(built, tmp_path):
    gg = built
    g, m, d = 32, 128, 16
    N = g * m
    tab = gg.orbit_part(g, m, d, 3)
    o = ol.part_eval(N, m, tab)
    f = tmp_path / "c3.part"
    _write_part(f, N, tab.tolist())
    dia, aspl = _call("conv-diam", f)
    assert int(dia) == o["diameter"] and float(aspl) == o["total_scaled"] / (N * (N - 1))
    if ol.ref_available():
        assert (int(dia), float(aspl)) == ol.ref_part(N, m, tab)
    wvc, dia, _ = _call("conv-wvc", f)
    assert int(wvc) == o["wvc"]


@pytest.mark.parametrize("n", [196, 198, 200])
def test_conv_level_limit_boundaries(built, tmp_path, n):
    """Rings of n vertices have diameter n/2.  diameterASPL and calc loop to
    level 99 (cudaASPLconv.cu:161,208): diameter 98 is evaluated, 99 and 100
    are the sentinel / 1e9.  WVCdiameterASPL loops to 100 (:264): 98 and 99
    are evaluated, 100 is the sentinel (:296-298)."""
    f = tmp_path / "ring.part"
    _write_part(f, n, _ring_rows(n))
    dia = n // 2
    aspl = graphs.cycle_total(n) / (n * (n - 1))
    calc = float(_call("conv-calc", f)[0])
    d1, a1 = _call("conv-diam", f)
    w, d2, a2 = _call("conv-wvc", f)
    if dia <= 98:
        assert calc == aspl and (int(d1), float(a1)) == (dia, aspl)
    else:
        assert calc == 1e9 and (int(d1), float(a1)) == (BAD, 1e8)
    if dia <= 99:
        # every source of a ring has eccentricity n/2: all 32 classes count
        assert (int(w), int(d2), float(a2)) == (32, dia, aspl)
    else:
        assert (int(w), int(d2), float(a2)) == (BAD, BAD, 1e8)


def test_conv_unconnected_sentinels(built, tmp_path):
    adj = graphs.disjoint_cycles(10, 14)
    n = adj.shape[0]
    f = tmp_path / "two.part"
    _write_part(f, n, (adj - np.arange(n)[:, None]).tolist())
    assert float(_call("conv-calc", f)[0]) == 1e9  # cudaASPLconv.cu:182
    assert [int(x) if i == 0 else float(x) for i, x in enumerate(_call("conv-diam", f))] == \
        [BAD, 1e8]  # :229
    w, d, a = _call("conv-wvc", f)
    assert (int(w), int(d), float(a)) == (BAD, BAD, 1e8)  # :293
    e = tmp_path / "two.edges"
    _write_edges(e, adj)
    assert _call("graph", e) == ["-1", "-1"]  # cudaASPLgraph.cu:182


def test_graph_calc_values_and_limit(built, tmp_path):
    gg = built
    n, d = 300, 5
    adj = gg.random_regular(n, d, 9)
    o = ol.adj_eval(adj)
    e = tmp_path / "g.edges"
    _write_edges(e, adj)
    assert _call("graph", e) == [str(o["diameter"]), str(o["total"])]
    # cudaASPLgraph::calc loops to level 199 (cudaASPLgraph.cu:163,187-190):
    # a ring of 396 (diameter 198) is evaluated, one of 398 (199) is (-2, -2)
    for n, want in ((396, ["198", str(graphs.cycle_total(396))]), (398, ["-2", "-2"])):
        _write_edges(e, graphs.cycle(n))
        assert _call("graph", e) == want


def _write_lattice(path, head, table, dims):
    rows = np.asarray(table).reshape(len(table), -1, dims)
    with open(path, "w") as f:
        f.write(" ".join(str(x) for x in head) + "\n")
        for r in rows:
            f.write(" ".join(str(int(x)) for x in [len(r)] + r.ravel().tolist()) + "\n")


def test_piece_and_box_match_oracle_and_reference(built, tmp_path):
    gg = built
    Nx, Ny, Mx, My, d = 32, 128, 1, 128, 16
    tab = gg.gen_lattice((Nx, Ny), (Mx, My), d, 3)
    o = ol.piece_eval(Nx, Ny, Mx, My, tab)
    N = Nx * Ny
    f = tmp_path / "c3.piece"
    _write_lattice(f, (Nx, Ny, Mx, My), tab, 2)
    dia, aspl = _call("piece", f)
    assert (int(dia), float(aspl)) == (o["diameter"], o["total_scaled"] / (N * (N - 1)))
    if ol.ref_available():
        assert (int(dia), float(aspl)) == ol.ref_piece(Nx, Ny, Mx, My, tab)
    Ns, Ms, d = (8, 8, 16), (2, 2, 16), 6
    tab = gg.gen_lattice(Ns, Ms, d, 9)
    o = ol.box_eval(*Ns, *Ms, tab)
    N = int(np.prod(Ns))
    f = tmp_path / "b.box"
    _write_lattice(f, Ns + Ms, tab, 3)
    dia, aspl = _call("box", f)
    assert (int(dia), float(aspl)) == (o["diameter"], o["total_scaled"] / (N * (N - 1)))
    if ol.ref_available():
        assert (int(dia), float(aspl)) == ol.ref_box(*Ns, *Ms, tab)


def test_piece_rows_truncated_to_first_row_degree(built, tmp_path):
    """degree is the first row's length (piece.hpp:28-31) and longer rows are
    cut to it when packed (cudaASPLpiece.cu:163-174): an extra offset on a
    later row changes nothing."""
    gg = built
    Nx, Ny, Mx, My, d = 8, 16, 2, 16, 4
    tab = gg.gen_lattice((Nx, Ny), (Mx, My), d, 5).reshape(Mx * My, d, 2)
    o = ol.piece_eval(Nx, Ny, Mx, My, tab)
    N = Nx * Ny
    with open(tmp_path / "t.piece", "w") as f:
        f.write(f"{Nx} {Ny} {Mx} {My}\n")
        for i, r in enumerate(tab):
            vals = r.ravel().tolist() + ([3, 5] if i == 3 else [])
            f.write(" ".join(str(int(x)) for x in [len(vals) // 2] + vals) + "\n")
    dia, aspl = _call("piece", tmp_path / "t.piece")
    assert (int(dia), float(aspl)) == (o["diameter"], o["total_scaled"] / (N * (N - 1)))


@pytest.mark.parametrize("fmt", ["piece", "box"])
def test_gg_eval_piece_and_box_output(built, tmp_path, fmt):
    """gg-eval -f piece|box prints piece2aspl / box2aspl's three lines: ASPL
    with 11 significant digits, the diameter, the time in whole
    milliseconds; -o writes the same lines to the file."""
    gg = built
    if fmt == "piece":
        Ns, Ms, d = (16, 64), (2, 64), 6
    else:
        Ns, Ms, d = (8, 8, 16), (2, 2, 16), 6
    tab = gg.gen_lattice(Ns, Ms, d, 7)
    o = (ol.piece_eval if fmt == "piece" else ol.box_eval)(*Ns, *Ms, tab)
    N = int(np.prod(Ns))
    f, out = tmp_path / f"g.{fmt}", tmp_path / "out.txt"
    _write_lattice(f, Ns + Ms, tab, len(Ns))
    text = _run([os.path.join(BIN, "gg-eval"), "-f", fmt, "-i", f, "-o", out])
    lines = text.strip().splitlines()
    assert len(lines) == 3
    assert lines[0] == "ASPL: " + f"{o['total_scaled'] / (N * (N - 1)):.11g}"
    assert lines[1] == f"Diameter: {o['diameter']}"
    assert re.fullmatch(r"TIME: \d+ msec\.", lines[2])
    assert open(out).read().strip().splitlines() == lines
