"""CPU checks of the C-ABI library: it loads, exports every symbol the public
headers declare, validates arguments before touching a device, and its seeded
generators are deterministic simple regular graphs."""
import glob
import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")) + \
            glob.glob(os.path.join(REPO, "graphgolf-cuda_amd", "csrc", "gg_host.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"//[^\n]*", "", txt)
        for m in re.finditer(r"\b(gg_\w+)\s*\(", txt):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol(gg):
    L = gg.lib()
    names = _declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_strerror(gg):
    assert gg.strerror(gg.GG_OK) == "ok"
    assert "unconnected" in gg.strerror(gg.GG_UNCONNECTED)
    assert gg.strerror(gg.GG_BAD_ARG) == "bad argument"


def test_bad_args_rejected_before_device(gg):
    import ctypes as C
    L = gg.lib()
    h = C.c_void_p()
    assert L.gg_create_part(10, 20, 3, 0, C.byref(h)) == gg.GG_BAD_ARG  # M > N
    assert L.gg_create_part(0, 1, 3, 0, C.byref(h)) == gg.GG_BAD_ARG
    assert L.gg_create_piece(4, 4, 8, 1, 2, 0, C.byref(h)) == gg.GG_BAD_ARG
    assert L.gg_create_box(2, 2, 2, 1, 1, 3, 2, 0, C.byref(h)) == gg.GG_BAD_ARG
    assert L.gg_create_adj(-5, 3, 0, C.byref(h)) == gg.GG_BAD_ARG
    assert L.gg_create_piece(70000, 1, 1, 1, 2, 0, C.byref(h)) == gg.GG_UNSUPPORTED
    assert L.gg_create_part(10, 5, 3, 0, None) == gg.GG_BAD_ARG
    r = gg.Result()
    t = np.zeros((1, 2), dtype=np.int32)
    tp = t.ctypes.data_as(C.POINTER(C.c_int32))
    assert L.gg_eval_part_noalloc(16385, 1, 2, 0, tp, C.byref(r)) == gg.GG_SCRATCH_LIMIT
    assert L.gg_eval_part_noalloc(16384, 16384, 33, 0, tp, C.byref(r)) == gg.GG_SCRATCH_LIMIT
    assert L.gg_eval_part_noalloc(16, 32, 2, 0, tp, C.byref(r)) == gg.GG_BAD_ARG
    assert L.gg_eval_part_noalloc(16, 1, 2, 0, tp, None) == gg.GG_BAD_ARG


def test_no_device_is_an_error_not_a_fallback(gg):
    import ctypes as C
    if gg.device_count() > 0:
        pytest.skip("device present")
    h = C.c_void_p()
    assert gg.lib().gg_create_part(16, 16, 2, 0, C.byref(h)) == gg.GG_NO_DEVICE
    with pytest.raises(gg.GGError):
        gg.PartEvaluator(16, 16, 2)


@pytest.mark.parametrize("n,d,seed", [(256, 4, 1), (100, 3, 7), (4096, 12, 2), (8192, 16, 4),
                                      (16384, 24, 5)])
def test_random_regular_is_simple_and_regular(gg, n, d, seed):
    a = gg.random_regular(n, d, seed)
    assert a.shape == (n, d)
    assert (a >= 0).all() and (a < n).all()
    assert (np.diff(a, axis=1) > 0).all()  # sorted rows, no multi-edges
    assert not (a == np.arange(n)[:, None]).any()  # no loops
    # symmetric: count of (v,u) equals count of (u,v)
    e1 = np.sort((a.astype(np.int64) * n + np.arange(n)[:, None]).ravel())
    e2 = np.sort((np.arange(n)[:, None].astype(np.int64) * n + a).ravel())
    assert np.array_equal(e1, e2)
    b = gg.random_regular(n, d, seed)
    assert np.array_equal(a, b)


def test_orbit_part_is_regular_simple(gg):
    import oracle_lib as ol
    g, m, d = 32, 128, 16
    t = gg.orbit_part(g, m, d, 3)
    adj = ol.expand_part(g * m, m, t)
    n = g * m
    assert not (adj == np.arange(n)[:, None]).any()
    s = np.sort(adj, axis=1)
    assert (np.diff(s, axis=1) > 0).all()
    assert np.array_equal(t, gg.orbit_part(g, m, d, 3))
