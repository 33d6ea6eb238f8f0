"""Seeded symmetric instance generator (gg_gen_lattice, the gg-gen CLI) and the
edge-list converter (gg-convert): generated tables are simple, symmetric and
d-regular for 1, 2 and 3 axes; the CLI prints the library's table in the
reference's text formats; the converter's edge list is the expanded graph's,
each undirected edge once (cppcode/piece2graph.cpp:84-103).  CPU only."""
import os
import subprocess

import numpy as np
import pytest

import driver_ref
import oracle_lib as ol

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "graphgolf-cuda_amd", "bin")

CASES = [((64,), (1,), 6), ((256,), (32,), 6), ((8, 12), (2, 3), 4), ((32, 128), (1, 128), 16),
         ((4, 4, 6), (2, 1, 2), 4), ((8, 8, 8), (2, 2, 2), 6), ((6, 6, 6), (1, 1, 1), 8)]


def _expand(Ns, Ms, t):
    if len(Ns) == 1:
        return ol.expand_part(Ns[0], Ms[0], t.reshape(Ms[0], -1))
    if len(Ns) == 2:
        return ol.expand_piece(*Ns, *Ms, t)
    return ol.expand_box(*Ns, *Ms, t)


@pytest.mark.parametrize("Ns,Ms,d", CASES)
def test_generated_tables_are_simple_symmetric_regular(gg, Ns, Ms, d):
    t = gg.gen_lattice(Ns, Ms, d, 7)
    assert np.array_equal(t, gg.gen_lattice(Ns, Ms, d, 7))  # seeded
    assert not np.array_equal(t, gg.gen_lattice(Ns, Ms, d, 8))
    L = driver_ref.Lattice(Ns, Ms, d)
    for r in range(L.Mtot):
        assert L.row_simple(t, r) and L.row_symmetric(t, r), r
    adj = _expand(Ns, Ms, t)
    n = adj.shape[0]
    e = {(v, int(u)) for v in range(n) for u in adj[v]}
    assert all((u, v) in e for v, u in e)
    assert all(len(set(adj[v].tolist())) == d and v not in adj[v] for v in range(n))


def test_generator_rejects_inconsistent_sizes(gg):
    with pytest.raises(gg.GGError):
        gg.gen_lattice((10, 12), (3, 3), 4, 1)  # 3 does not divide 10
    with pytest.raises(gg.GGError):
        gg.gen_lattice((9,), (3,), 3, 1)  # M*d odd


def _parse(text, dims):
    lines = [l.split() for l in text.strip().splitlines()]
    head = [int(x) for x in lines[0]]
    rows = [[int(x) for x in l] for l in lines[1:]]
    d = rows[0][0]
    t = np.array([r[1:] for r in rows], dtype=np.int32).reshape(len(rows), d, dims)
    return head, t


@pytest.mark.parametrize("Ns,Ms,d", [((256,), (32,), 6), ((8, 12), (2, 3), 4),
                                     ((8, 8, 8), (2, 2, 2), 6)])
def test_cli_generate_and_convert(gg, Ns, Ms, d, tmp_path):
    dims = len(Ns)
    fmt = {1: "part", 2: "piece", 3: "box"}[dims]
    if dims == 1:
        args = ["-n", str(Ns[0]), "-m", str(Ms[0])]
    else:
        args = []
        for k, ax in enumerate("xyz"[:dims]):
            args += [f"--n{ax}", str(Ns[k]), f"--m{ax}", str(Ms[k])]
    f = tmp_path / f"g.{fmt}"
    subprocess.run([os.path.join(BIN, "gg-gen"), "-f", fmt, *args, "-d", str(d), "-s", "9",
                    "-o", str(f)], check=True)
    head, t = _parse(f.read_text(), dims)
    assert head == list(Ns) + list(Ms) if dims > 1 else head == [Ns[0], Ms[0]]
    assert np.array_equal(t, gg.gen_lattice(Ns, Ms, d, 9))
    out = subprocess.run([os.path.join(BIN, "gg-convert"), "-f", fmt, "-i", str(f)],
                         check=True, capture_output=True, text=True).stdout
    edges = [tuple(int(x) for x in l.split()) for l in out.strip().splitlines()]
    adj = _expand(Ns, Ms, t)
    want = sorted({(min(v, int(u)), max(v, int(u))) for v in range(adj.shape[0]) for u in adj[v]})
    assert edges == want
    assert len(edges) == adj.shape[0] * d // 2
