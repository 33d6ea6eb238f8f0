"""Exact evaluator on the CPU oracle with the search-callback interface
(set_base / eval / commit).  Test infrastructure only."""
import ctypes as C
import os

import numpy as np

import oracle_lib as ol


def evaluate(adj, threads=None):
    """(total, diameter, status) of an explicit adjacency via the oracle BFS."""
    a = np.ascontiguousarray(adj, dtype=np.int32)
    n, d = a.shape
    srcs = np.arange(n, dtype=np.int32)
    tot = C.c_int64()
    dia = C.c_int32()
    th = threads or min(16, os.cpu_count() or 1)
    rc = ol.oracle().oracle_bfs_omp(n, d, ol._p(a), n, ol._p(srcs), th, C.byref(tot),
                                    C.byref(dia))
    assert rc >= 0
    return tot.value, (dia.value if rc == 0 else 0), (0 if rc == 0 else 1)


def evaluate_wvc(adj):
    """(total, diameter, status, WVC) of an explicit adjacency: queue BFS from
    every vertex with per-source eccentricities, WVC by its definition over
    the 32 source classes (cudacode/cudaASPLconv.cu:256-285)."""
    r = ol.adj_eval(adj)
    ok = not r["unconnected"]
    w = ol.wvc(adj.shape[0], r["ecc"], r["diameter"]) if ok else 0
    return r["total"], (r["diameter"] if ok else 0), (0 if ok else 1), w


class OracleSwapEvaluator:
    """The search-callback evaluator on the oracle; with wvc=True the results
    carry WVC (single-threaded BFS with eccentricities)."""

    def __init__(self, wvc=False):
        self.adj = None
        self.fn = evaluate_wvc if wvc else evaluate

    def set_base(self, adj):
        self.adj = np.array(adj, dtype=np.int32)
        return self.fn(self.adj)

    def _apply(self, sw, adj):
        new = adj.copy()
        for j in range(4):
            if sw.v[j] >= 0:
                new[sw.v[j], sw.slot[j]] = sw.nbr[j]
        return new

    def eval(self, swaps):
        return [self.fn(self._apply(sw, self.adj)) for sw in swaps]

    def commit(self, sw):
        self.adj = self._apply(sw, self.adj)
        return self.fn(self.adj)


def evaluate_part(N, table, threads=None):
    """(total scaled by N/M, diameter, status) of a symmetric part table via
    the oracle BFS from the M representatives (cpuASPLqueue.cpp:54-96)."""
    t = np.ascontiguousarray(table, dtype=np.int32)
    M = t.shape[0]
    adj = ol.expand_part(N, M, t)
    d = adj.shape[1]
    srcs = np.arange(M, dtype=np.int32)
    tot = C.c_int64()
    dia = C.c_int32()
    th = threads or min(16, os.cpu_count() or 1)
    rc = ol.oracle().oracle_bfs_omp(N, d, ol._p(adj), M, ol._p(srcs), th, C.byref(tot),
                                    C.byref(dia))
    assert rc >= 0
    return tot.value * (N // M), (dia.value if rc == 0 else 0), (0 if rc == 0 else 1)
