"""ctypes loader for the CPU oracle (oracle/liboracle.so) and the reference's
own CPU evaluators compiled from the reference tree (oracle/_ref/libref.so).

Test infrastructure only: the product never loads these.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libref.so")
INT_MAX = 2147483647

_o = None
_r = None


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a, ct=C.c_int32):
    return a.ctypes.data_as(C.POINTER(ct))


def oracle() -> C.CDLL:
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run __graft_entry__.build()")
        _o = C.CDLL(ORACLE_SO)
    return _o


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref() -> C.CDLL:
    global _r
    if _r is None:
        _r = C.CDLL(REF_SO)
        for n in ("ref_part_diameter_aspl",):
            getattr(_r, n).restype = C.c_int
    return _r


# ---- expansion to explicit adjacency ------------------------------------------------

def expand_part(N, M, table):
    t = _i32(table).reshape(M, -1)
    D = t.shape[1]
    lens = np.full(M, D, dtype=np.int32)
    adj = np.empty((N, D), dtype=np.int32)
    rc = oracle().oracle_expand_part(N, M, D, _p(lens), _p(t), D, _p(adj))
    assert rc == 0
    return adj


def expand_piece(Nx, Ny, Mx, My, table):
    t = _i32(table).reshape(Mx * My, -1, 2)
    D = t.shape[1]
    lens = np.full(Mx * My, D, dtype=np.int32)
    adj = np.empty((Nx * Ny, D), dtype=np.int32)
    rc = oracle().oracle_expand_piece(Nx, Ny, Mx, My, D, _p(lens), _p(t), D, _p(adj))
    assert rc == 0
    return adj


def expand_box(Nx, Ny, Nz, Mx, My, Mz, table):
    t = _i32(table).reshape(Mx * My * Mz, -1, 3)
    D = t.shape[1]
    lens = np.full(Mx * My * Mz, D, dtype=np.int32)
    adj = np.empty((Nx * Ny * Nz, D), dtype=np.int32)
    rc = oracle().oracle_expand_box(Nx, Ny, Nz, Mx, My, Mz, D, _p(lens), _p(t), D, _p(adj))
    assert rc == 0
    return adj


# ---- BFS ------------------------------------------------------------------------------

def bfs(adj, srcs, reach_cap=100):
    """Queue BFS from each source (cppcode/cpuASPLqueue.cpp:65-91 semantics).
    Returns dict(total, diameter, ecc, reach, unconnected)."""
    a = _i32(adj)
    N, D = a.shape
    s = _i32(srcs)
    tot = C.c_int64(0)
    dia = C.c_int32(0)
    ecc = np.zeros(len(s), dtype=np.int32)
    reach = np.zeros(reach_cap, dtype=np.int64)
    rc = oracle().oracle_bfs(N, D, _p(a), len(s), _p(s), C.byref(tot), C.byref(dia), _p(ecc),
                             _p(reach, C.c_int64), reach_cap)
    assert rc >= 0
    return dict(total=tot.value, diameter=dia.value, ecc=ecc, reach=reach, unconnected=bool(rc))


def part_eval(N, M, table):
    """cpuASPLqueue::diameterASPL restated: BFS from vertices 0..M-1."""
    adj = expand_part(N, M, table)
    r = bfs(adj, np.arange(M))
    r["total_scaled"] = r["total"] * (N // M)
    r["wvc"] = wvc(M, r["ecc"], r["diameter"])
    return r


def piece_eval(Nx, Ny, Mx, My, table):
    """cpuASPLpiece::diameterASPL restated: sources x < Mx, y < My (vertex x*Ny+y)."""
    adj = expand_piece(Nx, Ny, Mx, My, table)
    srcs = np.array([x * Ny + y for x in range(Mx) for y in range(My)], dtype=np.int32)
    r = bfs(adj, srcs)
    r["total_scaled"] = r["total"] * ((Nx * Ny) // (Mx * My))
    return r


def box_eval(Nx, Ny, Nz, Mx, My, Mz, table):
    """cpuASPLbox::diameterASPL restated."""
    adj = expand_box(Nx, Ny, Nz, Mx, My, Mz, table)
    srcs = np.array([(x * Ny + y) * Nz + z for x in range(Mx) for y in range(My)
                     for z in range(Mz)], dtype=np.int32)
    r = bfs(adj, srcs)
    r["total_scaled"] = r["total"] * ((Nx * Ny * Nz) // (Mx * My * Mz))
    return r


def adj_eval(adj):
    a = _i32(adj)
    return bfs(a, np.arange(a.shape[0]))


def wvc(M, ecc, diameter):
    e = _i32(ecc)
    return int(oracle().oracle_wvc(M, _p(e), diameter))


def floyd_warshall(adj, nsrc=None):
    a = _i32(adj)
    N, D = a.shape
    nsrc = N if nsrc is None else nsrc
    tot = C.c_int64(0)
    dia = C.c_int32(0)
    unc = C.c_int32(0)
    dist = np.zeros((N, N), dtype=np.int32)
    rc = oracle().oracle_apsp_fw(N, D, _p(a), nsrc, C.byref(tot), C.byref(dia), C.byref(unc),
                                 _p(dist))
    assert rc == 0
    return dict(total=tot.value, diameter=dia.value, unconnected=bool(unc.value), dist=dist)


# ---- the reference's own evaluators (compiled from the reference tree) ----------------

def ref_part(N, M, table):
    t = _i32(table).reshape(M, -1)
    D = t.shape[1]
    lens = np.full(M, D, dtype=np.int32)
    dia = C.c_int(0)
    asp = C.c_double(0)
    rc = ref().ref_part_diameter_aspl(N, M, D, _p(lens), _p(t), C.byref(dia), C.byref(asp))
    assert rc == 0
    return dia.value, asp.value


def ref_piece(Nx, Ny, Mx, My, table):
    t = _i32(table).reshape(Mx * My, -1, 2)
    D = t.shape[1]
    lens = np.full(Mx * My, D, dtype=np.int32)
    dia = C.c_int(0)
    asp = C.c_double(0)
    rc = ref().ref_piece_diameter_aspl(Nx, Ny, Mx, My, D, _p(lens), _p(t), C.byref(dia),
                                       C.byref(asp))
    assert rc == 0
    return dia.value, asp.value


def ref_box(Nx, Ny, Nz, Mx, My, Mz, table):
    t = _i32(table).reshape(Mx * My * Mz, -1, 3)
    D = t.shape[1]
    lens = np.full(Mx * My * Mz, D, dtype=np.int32)
    dia = C.c_int(0)
    asp = C.c_double(0)
    rc = ref().ref_box_diameter_aspl(Nx, Ny, Nz, Mx, My, Mz, D, _p(lens), _p(t), C.byref(dia),
                                     C.byref(asp))
    assert rc == 0
    return dia.value, asp.value


def aspl(total_scaled, N):
    return float(total_scaled) / (N * (N - 1))
