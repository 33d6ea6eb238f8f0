"""gg_eval_part_noalloc, the part evaluator that allocates nothing, called
through the C ABI: every field of its result equals gg_create_part +
gg_eval_part's on the same table, on calls in a row whose second input is
smaller than the first (the fixed scratch keeps the larger call's contents),
and the small cases equal the oracle.  Inputs past the fixed scratch are
refused with GG_SCRATCH_LIMIT before any device work."""
import ctypes as C

import numpy as np
import pytest

import graphs
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _fields(r):
    return (r.total_reps, r.total_scaled, r.diameter, r.wvc, r.levels, r.status)


def _noalloc(gg, N, M, d, tab):
    t = np.ascontiguousarray(np.asarray(tab, dtype=np.int32).reshape(M, d))
    r = gg.Result()
    st = gg.lib().gg_eval_part_noalloc(N, M, d, 0, t.ctypes.data_as(C.POINTER(C.c_int32)),
                                       C.byref(r))
    return st, r


def _handle(gg, N, M, d, tab):
    ev = gg.PartEvaluator(N, M, d)
    try:
        return ev.eval(tab)
    finally:
        ev.close()


def _check(gg, N, M, d, tab, oracle=False):
    print(f"noalloc N={N} M={M} d={d}", flush=True)
    st, r = _noalloc(gg, N, M, d, tab)
    assert st == gg.GG_OK, gg.strerror(st)
    want = _handle(gg, N, M, d, tab)
    assert _fields(r) == _fields(want), (_fields(r), _fields(want))
    if oracle:
        o = ol.part_eval(N, M, np.asarray(tab, dtype=np.int32).reshape(M, d))
        assert (r.total_reps, r.diameter, r.wvc) == (o["total"], o["diameter"], o["wvc"])
        assert r.total_scaled == o["total_scaled"]
    return r


def test_noalloc_matches_part_evaluator_on_shrinking_calls(gpu):
    gg = gpu
    # the largest accepted shape first: 16384 vertices, 1024 rows of degree 24
    _check(gg, 16384, 1024, 24, gg.orbit_part(16, 1024, 24, 3))
    # then smaller ones on the same scratch: fewer rows, lower degree, a row
    # count that is no multiple of 64 and does not divide N, every vertex its
    # own row, an unconnected table and a ring past the level limit
    _check(gg, 4096, 256, 8, gg.orbit_part(16, 256, 8, 5), oracle=True)
    _check(gg, 1000, 100, 4, gg.orbit_part(10, 100, 4, 7), oracle=True)
    # 96 rows on 390 vertices: the rows wrap unevenly, the graph is directed,
    # and both evaluators pull along listed neighbours as the reference's
    # kernel does (cudaASPLconv.cu:35-48) where the oracle follows them
    # outward, so only the two entry points are compared
    _check(gg, 390, 96, 6, gg.orbit_part(4, 96, 6, 9))
    adj = gg.random_regular(640, 5, 11)
    rows = (adj - np.arange(640)[:, None]).astype(np.int32)
    _check(gg, 640, 640, 5, rows, oracle=True)
    two = graphs.disjoint_cycles(10, 14)
    n = two.shape[0]
    r = _check(gg, n, n, two.shape[1], (two - np.arange(n)[:, None]).astype(np.int32))
    assert r.status == gg.GG_UNCONNECTED
    r = _check(gg, 200, 1, 2, np.array([[1, -1]], dtype=np.int32))
    assert r.status == gg.GG_DIAMETER_LIMIT


def test_noalloc_refuses_inputs_past_its_scratch(gpu):
    gg = gpu
    tab = np.zeros((1, 2), dtype=np.int32)
    assert _noalloc(gg, 16385, 1, 2, tab)[0] == gg.GG_SCRATCH_LIMIT
    big = np.zeros((16384, 33), dtype=np.int32)
    assert _noalloc(gg, 16384, 16384, 33, big)[0] == gg.GG_SCRATCH_LIMIT
    assert _noalloc(gg, 16, 32, 2, np.zeros((32, 2), dtype=np.int32))[0] == gg.GG_BAD_ARG  # M > N
    assert "scratch" in gg.strerror(gg.GG_SCRATCH_LIMIT)
    # a refused call leaves the next one correct
    _check(gg, 2048, 128, 6, gg.orbit_part(16, 128, 6, 13), oracle=True)
