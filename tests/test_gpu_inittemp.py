"""Initial-temperature estimate (cppcode/initTemp.cpp:399-430) on the GPU
evaluators against the same computation on the CPU oracle: identical graphs
(seeded generators), identical moves (the restated mt19937_64 stream), exact
counts and increases of the distance total."""
import math

import numpy as np
import pytest

import driver_ref
import oracle_eval

pytestmark = pytest.mark.gpu


def _expected(total, cnt, N):
    mean = total / cnt
    return cnt, mean, -mean / math.log(0.4), -mean / math.log(0.4) / (N * (N - 1))


def test_init_temp_explicit_graphs(gpu):
    gg = gpu
    N, d, graphs, nb, seed = 512, 6, 4, 50, 77
    rng = driver_ref.MT64(seed)
    total, cnt = 0, 0
    for g in range(graphs):
        adj = gg.random_regular(N, d, seed + 1 + g)
        moves = []
        for _ in range(nb):
            while True:
                p = driver_ref.propose(adj, rng)
                if p is not None:
                    break
            moves.append(p[1])
        t0, _, st0 = oracle_eval.evaluate(adj)
        if st0 != 0:
            continue
        for ents in moves:
            new = adj.copy()
            for v, s, x in ents:
                new[v, s] = x
            t, _, st = oracle_eval.evaluate(new)
            if st == 0 and t > t0:
                total += t - t0
                cnt += 1
    r = gg.init_temp(N, d, graphs=graphs, neighbours=nb, seed=seed)
    c, mean, tt, ta = _expected(total, cnt, N)
    assert r.count == c and cnt > 0
    assert r.mean_increase_total == pytest.approx(mean, rel=1e-12)
    assert r.temp_total == pytest.approx(tt, rel=1e-12)
    assert r.temp_aspl == pytest.approx(ta, rel=1e-12)


def test_init_temp_orbit_graphs(gpu):
    gg = gpu
    g, m, d, graphs, nb, seed = 8, 64, 6, 4, 30, 5
    N = g * m
    rng = driver_ref.MT64(seed)
    total, cnt = 0, 0
    for i in range(graphs):
        tab = gg.orbit_part(g, m, d, seed + 1 + i)
        t0, _, st0 = oracle_eval.evaluate_part(N, tab)
        for _ in range(nb):
            while True:
                p = driver_ref.propose_orbit(np.asarray(tab, dtype=np.int32), N, rng)
                if p is not None:
                    break
            t, _, st = oracle_eval.evaluate_part(N, p[1])
            if st0 == 0 and st == 0 and t > t0:
                total += t - t0
                cnt += 1
    r = gg.init_temp(N, d, graphs=graphs, neighbours=nb, seed=seed, M=m)
    c, mean, tt, ta = _expected(total, cnt, N)
    assert r.count == c and cnt > 0
    assert r.mean_increase_total == pytest.approx(mean, rel=1e-12)
    assert r.temp_total == pytest.approx(tt, rel=1e-12)
