"""The command-line drivers on the GPU: gg-eval prints the reference's
evaluator lines (cppcode/main.cpp:99-116, aspl_cuda.cpp:82-105) with the
oracle's values, and gg-anneal reads a `part` table, writes the log lines of
cppcode/annealing.cpp:321-324 and :415-427, and prints a best solution that
the oracle finds connected and no worse than the input, with an ASPL no
larger than any logged ASPL(x_best) when the diameter did not change, both for
an explicit graph (M = N, 2-opt swaps) and a cyclically symmetric one (M < N,
orbit moves); with -f piece / -f box it reads and writes the 2-D / 3-D forms."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_eval

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "graphgolf-cuda_amd", "bin")


def _write_part(path, N, table):
    M, d = table.shape
    with open(path, "w") as f:
        f.write(f"{N} {M}\n")
        for r in table:
            f.write(f"{d} " + " ".join(str(int(x)) for x in r) + "\n")


def _read_part(path):
    toks = open(path).read().split()
    N, M = int(toks[0]), int(toks[1])
    rows, k = [], 2
    for _ in range(M):
        E = int(toks[k])
        rows.append([int(x) for x in toks[k + 1:k + 1 + E]])
        k += 1 + E
    return N, np.array(rows, dtype=np.int32)


def _run(args, timeout=120):
    print("run", " ".join(args), flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


def _part_of_adj(adj):
    n = adj.shape[0]
    return (adj - np.arange(n, dtype=np.int32)[:, None]).astype(np.int32)


def test_gg_eval_part_and_graph(gpu, tmp_path):
    gg = gpu
    n, d = 512, 6
    adj = gg.random_regular(n, d, 11)
    tot, dia, st = oracle_eval.evaluate(adj)
    assert st == 0
    f = tmp_path / "g.part"
    _write_part(f, n, _part_of_adj(adj))
    out = _run([os.path.join(BIN, "gg-eval"), "-f", "part", "-i", str(f)])
    aspl = float(re.search(r"ASPL: (\S+)", out).group(1))
    assert abs(aspl - tot / (n * (n - 1))) < 1e-9
    assert int(re.search(r"Diameter: (\d+)", out).group(1)) == dia
    assert re.search(r"WVC: \d+", out) and re.search(r"TIME: \S+ msec\.", out)
    e = tmp_path / "g.edges"
    with open(e, "w") as fh:
        for v in range(n):
            for u in adj[v]:
                if v < u:
                    fh.write(f"{v} {u}\n")
    out = _run([os.path.join(BIN, "gg-eval"), "-f", "graph", "-i", str(e)])
    assert f"Nodes: {n} Degrees_max: {d}" in out
    assert int(re.search(r"Diameter: (\d+)", out).group(1)) == dia
    m = re.search(r"ASPL: (\S+) \((\d+)/(\d+)\)", out)
    assert int(m.group(2)) == tot // 2 and int(m.group(3)) == n * (n - 1) // 2


def _check_anneal(tmp_path, N, table, extra):
    f, o, lg = tmp_path / "in.part", tmp_path / "out.part", tmp_path / "run.log"
    _write_part(f, N, table)
    tot0, dia0, st0 = oracle_eval.evaluate_part(N, table)
    assert st0 == 0
    out = _run([os.path.join(BIN, "gg-anneal"), "-c", "600", "-s", "7", "-i", str(f),
                "-o", str(o), "-l", str(lg), "-t", "40"] + extra)
    init = float(re.search(r"ASPL\(init_x\): (\S+)", out).group(1))
    assert abs(init - tot0 / (N * (N - 1))) < 1e-9
    assert "total elapsed time:" in out
    lines = open(lg).read().splitlines()
    assert lines[0] == f"#input file: {f}"
    assert lines[1] == f"#output file: {o}"
    assert lines[2] == "#iteration ASPL(x_best) ASPL(x) ASPL(y) temp"
    z = lines[3].split()
    assert z[0] == "0" and len(z) == 5 and float(z[4]) == 40.0
    assert abs(float(z[1]) - init) < 1e-9 and z[1] == z[2] == z[3]
    recs = [l.split() for l in lines[4:]]
    its = [int(r[0]) for r in recs]
    assert its == sorted(its) and all(1 <= i <= 600 for i in its)
    assert all(len(r) == 4 for r in recs)
    # the step schedule holds the initial temperature for 10^4 iterations
    assert all(abs(float(r[3]) - 40.0) < 1e-6 for r in recs)
    assert len(recs) > 0, "the run accepted no change"
    N2, best = _read_part(o)
    assert N2 == N and best.shape == table.shape
    tot, dia, st = oracle_eval.evaluate_part(N, best)
    assert st == 0
    best_aspl = tot / (N * (N - 1))
    # the printed solution is the best by (diameter, ASPL); with the diameter
    # unchanged it is at least as good as every logged ASPL(x_best)
    assert (dia, tot) <= (dia0, tot0)
    if dia == dia0:
        assert all(float(r[1]) >= best_aspl - 1e-9 for r in recs)
    return best_aspl, init


def test_gg_anneal_explicit_graph(gpu, tmp_path):
    gg = gpu
    n, d = 256, 6
    adj = gg.random_regular(n, d, 5)
    best, init = _check_anneal(tmp_path, n, _part_of_adj(adj), [])
    assert best < init


def test_gg_anneal_orbit_part(gpu, tmp_path):
    gg = gpu
    g, m, d = 16, 64, 8
    tab = gg.orbit_part(g, m, d, 5)
    best, init = _check_anneal(tmp_path, g * m, tab, [])
    assert best < init


def test_gg_anneal_learned_mode(gpu, tmp_path):
    gg = gpu
    n, d = 1024, 8
    adj = gg.random_regular(n, d, 9)
    _check_anneal(tmp_path, n, _part_of_adj(adj), ["--mode", "learned", "-K", "8", "-P", "64"])


def _write_lattice(path, Ns, Ms, table):
    M, d, D = table.shape
    with open(path, "w") as f:
        f.write(" ".join(str(x) for x in (*Ns, *Ms)) + "\n")
        for r in range(M):
            f.write(f"{d} " + " ".join(str(int(x)) for x in table[r].ravel()) + "\n")


def _read_lattice(path, D):
    with open(path) as f:
        head = [int(x) for x in f.readline().split()]
        rows = [[int(x) for x in line.split()] for line in f if line.strip()]
    d = rows[0][0]
    t = np.array([r[1:] for r in rows], dtype=np.int32).reshape(len(rows), d, D)
    return head[:D], head[D:], t


@pytest.mark.parametrize("fmt,Ns,Ms,d", [("piece", (16, 64), (4, 8), 8),
                                         ("box", (8, 8, 8), (2, 2, 2), 6)])
def test_gg_anneal_piece_and_box(gpu, tmp_path, fmt, Ns, Ms, d):
    """-f piece / -f box: the 2-D / 3-D text formats in and out
    (piece.hpp:18-52, box.hpp:37-80) and the pieceHillClimb / boxHillClimb
    moves; the printed best is connected and no worse than the input."""
    import driver_ref
    import oracle_lib as ol
    D = len(Ns)
    tab = driver_ref.random_lattice_table(Ns, Ms, d, 3)
    f, o = tmp_path / f"in.{fmt}", tmp_path / f"out.{fmt}"
    _write_lattice(f, Ns, Ms, tab)

    def ev(t):
        r = (ol.piece_eval(*Ns, *Ms, t) if D == 2 else ol.box_eval(*Ns, *Ms, t))
        assert not r["unconnected"]
        return r["total_scaled"], r["diameter"]
    N = int(np.prod(Ns))
    tot0, dia0 = ev(tab)
    out = _run([os.path.join(BIN, "gg-anneal"), "-f", fmt, "--hill", "-c", "400", "-s", "5",
                "-i", str(f), "-o", str(o)])
    init = float(re.search(r"ASPL\(init_x\): (\S+)", out).group(1))
    assert abs(init - tot0 / (N * (N - 1))) < 1e-9
    Ns2, Ms2, best = _read_lattice(o, D)
    assert tuple(Ns2) == tuple(Ns) and tuple(Ms2) == tuple(Ms) and best.shape == tab.shape
    L = driver_ref.Lattice(Ns, Ms, d)
    assert all(L.row_simple(best, r) and L.row_symmetric(best, r) for r in range(L.Mtot))
    tot, dia = ev(best)
    assert (dia, tot) < (dia0, tot0)


@pytest.mark.parametrize("N,M,d", [(512, 64, 6), (256, 256, 6)])
def test_gg_inittemp_prints_the_library_estimate(gpu, N, M, d):
    """gg-inittemp: initTemp.cpp's output lines (cnt, avg, inittemp =
    -avg / ln 0.4) with the library's values for the same seed, for orbit
    graphs (M < N) and random regular graphs (M = N)."""
    gg = gpu
    out = _run([os.path.join(BIN, "gg-inittemp"), "-n", str(N), "-m", str(M), "-d", str(d),
                "-s", "4", "--graphs", "6", "--neighbours", "20"])
    vals = dict(re.findall(r"^(cnt|avg|inittemp): (\S+)$", out, flags=re.M))
    r = gg.init_temp(N, d, graphs=6, neighbours=20, seed=4, M=M)
    assert int(vals["cnt"]) == r.count > 0
    assert abs(float(vals["avg"]) - r.mean_increase_total) <= 1e-8 * abs(r.mean_increase_total)
    assert abs(float(vals["inittemp"]) - r.temp_total) <= 1e-8 * abs(r.temp_total)
    assert abs(r.temp_total + r.mean_increase_total / np.log(0.4)) <= 1e-9 * abs(r.temp_total)


@pytest.mark.parametrize("extra", [["--wvc", "--hill"], ["--wvc"]])
def test_gg_anneal_wvc(gpu, tmp_path, extra):
    """--wvc: annealingWVC / HillClimbWVC acceptance on (diameter, WVC, ASPL)
    for an explicit graph and for an orbit graph; the printed best is no
    worse than the input in that order (WVC from the oracle)."""
    import oracle_lib as ol
    gg = gpu
    for N, tab in ((256, _part_of_adj(gg.random_regular(256, 6, 5))),
                   (16 * 64, gg.orbit_part(16, 64, 8, 5))):
        M = tab.shape[0]
        f, o = tmp_path / "in.part", tmp_path / "out.part"
        _write_part(f, N, tab)
        r0 = ol.part_eval(N, M, tab)
        _run([os.path.join(BIN, "gg-anneal"), "-c", "300", "-s", "3", "-i", str(f), "-o", str(o),
              "-t", "20"] + extra)
        N2, best = _read_part(o)
        r1 = ol.part_eval(N, M, best)
        assert not r1["unconnected"]
        assert (r1["diameter"], r1["wvc"], r1["total_scaled"]) <= \
            (r0["diameter"], r0["wvc"], r0["total_scaled"])


def test_gg_anneal_nr_rows_of_different_lengths(gpu, tmp_path):
    """--nr (annealingNR): a symmetric part table whose rows have different
    numbers of entries is searched with the rows' lengths kept; the output
    table has the input's row lengths and the oracle's ASPL and diameter
    match the printed best; without --nr the input is refused as
    non-regular."""
    import driver_ref
    import oracle_lib as ol
    gg = gpu
    g, m, d = 16, 64, 8
    N = g * m
    tab, lens = driver_ref.irregular_part(gg.orbit_part(g, m, d, 5), N, 40, 3)
    f, o = tmp_path / "nr.part", tmp_path / "out.part"
    with open(f, "w") as fh:
        fh.write(f"{N} {m}\n")
        for r in range(m):
            fh.write(" ".join(str(int(x)) for x in [lens[r]] + tab[r, :lens[r]].tolist()) + "\n")
    p = subprocess.run([os.path.join(BIN, "gg-anneal"), "-c", "50", "-i", str(f)],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode != 0 and "non-regular" in p.stdout
    out = _run([os.path.join(BIN, "gg-anneal"), "--nr", "--hill", "-c", "400", "-s", "5",
                "-i", str(f), "-o", str(o)])
    toks = open(o).read().split()
    assert (int(toks[0]), int(toks[1])) == (N, m)
    k, rows = 2, []
    for r in range(m):
        E = int(toks[k])
        assert E == lens[r]
        rows.append([int(x) for x in toks[k + 1:k + 1 + E]] + [0] * (d - E))
        k += 1 + E
    best = np.array(rows, dtype=np.int32)
    r1 = ol.part_eval(N, m, best)
    r0 = ol.part_eval(N, m, tab)
    assert not r1["unconnected"]
    assert (r1["diameter"], r1["total_scaled"]) <= (r0["diameter"], r0["total_scaled"])
    assert "total elapsed time" in out
