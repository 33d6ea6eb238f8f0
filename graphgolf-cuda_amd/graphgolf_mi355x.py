"""ctypes binding of libgg_mi355x.so (the C ABI in include/graphgolf_mi355x.h).

The evaluator classes mirror the reference's evaluator objects
(cudacode/cudaASPLconv.hpp:6-23, cudaASPLpiece.hpp, cudaASPLbox.hpp,
cudaASPLgraph.hpp): a constructor with the graph dimensions and a device id,
and one call that returns the exact integer distance total and the diameter.
Every call runs on the GPU; there is no CPU fallback, and a missing library or
device raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgg_mi355x.so")

GG_OK = 0
GG_UNCONNECTED = 1
GG_DIAMETER_LIMIT = 2
GG_BAD_ARG = -1
GG_HIP_ERROR = -2
GG_OOM = -3
GG_NO_DEVICE = -4
GG_UNSUPPORTED = -5
GG_SCRATCH_LIMIT = -6
GG_LEVEL_LIMIT = 99

GG_OPT_BLOCK = 1
GG_OPT_LDS_LIMIT = 2
GG_OPT_GENERIC = 3
GG_OPT_LEVEL_LIMIT = 4
GG_OPT_PUSH = 5
GG_OPT_WIDE = 6
GG_OPT_SWAP_GRAPH = 7
GG_OPT_LEVELS = 8
GG_OPT_WVC = 9
GG_INFO_VARIANT = 100
GG_INFO_VPT = 101
GG_INFO_FIXED = 102
GG_INFO_PUSH_CAP = 103
GG_LEVEL_LIMIT_MAX = 199

GG_BATCH_OPT_PATH = 1
GG_BATCH_OPT_GRAPH = 2
GG_BATCH_OPT_COMMIT = 3
GG_BATCH_OPT_DELTA_LIMIT = 4
GG_BATCH_INFO_LAST_SPARSE = 100
GG_BATCH_INFO_SPARSE_OK = 101
GG_BATCH_INFO_LAST_COMMIT_INPLACE = 102
GG_BATCH_OPT_LAUNCH_K = 5
GG_BATCH_OPT_GROUP_K = 6
GG_BATCH_OPT_SPARSE_BLOCK = 7
GG_BATCH_OPT_STAGED_KEYS = 8
GG_BATCH_OPT_BASE_BLOCK = 9
GG_BATCH_OPT_BASE_PUSH = 10
GG_BATCH_INFO_LAUNCH_MAX = 103
GG_BATCH_INFO_STAGED_MAX = 104
GG_BATCH_TUNE_MAX = 64
PATH_AUTO, PATH_DENSE, PATH_SPARSE = 0, 1, 2


class GGError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{msg} (status {code})")
        self.code = code


class Result(C.Structure):
    _fields_ = [
        ("total_reps", C.c_int64),
        ("total_scaled", C.c_int64),
        ("diameter", C.c_int32),
        ("wvc", C.c_int32),
        ("levels", C.c_int32),
        ("status", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneReport(C.Structure):
    _fields_ = [("count", C.c_int32), ("block", C.c_int32 * 4), ("generic", C.c_int32 * 4),
                ("variant", C.c_int32 * 4), ("ms", C.c_float * 4), ("best_block", C.c_int32),
                ("best_generic", C.c_int32), ("best_ms", C.c_float)]


class BatchTuneReport(C.Structure):
    _T = GG_BATCH_TUNE_MAX
    _fields_ = [("count", C.c_int32), ("launch_k", C.c_int32 * _T), ("group_k", C.c_int32 * _T),
                ("block", C.c_int32 * _T), ("staged_keys", C.c_int32 * _T),
                ("ms", C.c_float * _T), ("mismatches", C.c_int32),
                ("best_launch_k", C.c_int32), ("best_group_k", C.c_int32),
                ("best_block", C.c_int32), ("best_staged_keys", C.c_int32),
                ("best_ms", C.c_float), ("default_ms", C.c_float), ("base_count", C.c_int32),
                ("base_block", C.c_int32 * 4), ("base_push", C.c_int32 * 4),
                ("base_ms", C.c_float * 4), ("best_base_block", C.c_int32),
                ("best_base_push", C.c_int32), ("best_base_ms", C.c_float)]

    def as_dict(self):
        n, nb = self.count, self.base_count
        return {
            "configs": [{"launch_k": self.launch_k[i], "group_k": self.group_k[i],
                         "block": self.block[i], "staged_keys": self.staged_keys[i],
                         "ms": round(self.ms[i], 4)} for i in range(n)],
            "mismatches": self.mismatches,
            "best": {"launch_k": self.best_launch_k, "group_k": self.best_group_k,
                     "block": self.best_block, "staged_keys": self.best_staged_keys,
                     "ms": round(self.best_ms, 4)},
            "default_ms": round(self.default_ms, 4),
            "base": [{"block": self.base_block[i], "push": self.base_push[i],
                      "ms": round(self.base_ms[i], 4)} for i in range(nb)],
            "best_base": {"block": self.best_base_block, "push": self.best_base_push,
                          "ms": round(self.best_base_ms, 4)},
        }


class LdsProbe(C.Structure):
    _fields_ = [("gbs", C.c_double), ("kernel_ms", C.c_double), ("clock_mhz", C.c_double),
                ("bytes_per_clk_cu", C.c_double), ("cus", C.c_int32)]


class Swap(C.Structure):
    _fields_ = [("v", C.c_int32 * 4), ("slot", C.c_int32 * 4), ("nbr", C.c_int32 * 4)]


class SearchCfg(C.Structure):
    _fields_ = [
        ("mode", C.c_int32), ("rule", C.c_int32), ("K", C.c_int32), ("P", C.c_int32),
        ("total_steps", C.c_int32), ("warmup_pairs", C.c_int32), ("train_every", C.c_int32),
        ("train_batch", C.c_int32), ("T0", C.c_double), ("T1", C.c_double), ("lr", C.c_double),
        ("seed", C.c_uint64), ("schedule", C.c_int32),
    ]


class SearchStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64), ("exact_evals", C.c_int64), ("proposals", C.c_int64),
        ("accepted", C.c_int64), ("total_cur", C.c_int64), ("total_best", C.c_int64),
        ("diam_cur", C.c_int32), ("diam_best", C.c_int32), ("seconds", C.c_double),
        ("train_pairs", C.c_int64), ("train_steps", C.c_int64), ("train_loss", C.c_double),
        ("spearman", C.c_double), ("miss_rate", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SET_BASE_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_int32), C.POINTER(Result))
EVAL_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(Swap), C.c_int, C.POINTER(Result))
COMMIT_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(Swap), C.POINTER(Result))


class EvalCallbacks(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("set_base", SET_BASE_CB), ("eval", EVAL_CB),
                ("commit", COMMIT_CB)]


TABLE_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_int32), C.POINTER(Result))


class InitTempResult(C.Structure):
    _fields_ = [("count", C.c_int64), ("mean_increase_total", C.c_double),
                ("temp_total", C.c_double), ("temp_aspl", C.c_double)]


class TableCallbacks(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("eval", TABLE_CB)]


MODE_PLAIN, MODE_BATCHED, MODE_LEARNED = 0, 1, 2
MLP_FEATURES, MLP_HIDDEN = 32, 128
MLP_PARAMS = MLP_HIDDEN * MLP_FEATURES + MLP_HIDDEN + MLP_HIDDEN + 1
MLP_MAX_BATCH = 512
RULE_ANNEAL, RULE_HILL, RULE_ANNEAL_WVC, RULE_HILL_WVC = 0, 1, 2, 3
SCHED_GEOMETRIC, SCHED_STEP = 0, 1

_lib = None


def lib() -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GGError(GG_UNSUPPORTED, f"{LIB_PATH} not built: run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    i32p, i64p = P(C.c_int32), P(C.c_int64)
    vp = C.c_void_p
    sig = {
        "gg_strerror": (C.c_char_p, [C.c_int]),
        "gg_device_count": (C.c_int, [P(C.c_int)]),
        "gg_aspl": (C.c_double, [P(Result), C.c_int]),
        "gg_create_part": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, P(vp)]),
        "gg_create_piece": (C.c_int, [C.c_int] * 6 + [P(vp)]),
        "gg_create_box": (C.c_int, [C.c_int] * 8 + [P(vp)]),
        "gg_create_adj": (C.c_int, [C.c_int, C.c_int, C.c_int, P(vp)]),
        "gg_destroy": (None, [vp]),
        "gg_eval_part": (C.c_int, [vp, i32p, P(Result)]),
        "gg_eval_part_noalloc": (C.c_int, [C.c_int] * 4 + [i32p, P(Result)]),
        "gg_eval_piece": (C.c_int, [vp, i32p, P(Result)]),
        "gg_eval_box": (C.c_int, [vp, i32p, P(Result)]),
        "gg_eval_adj": (C.c_int, [vp, i32p, P(Result)]),
        "gg_eval_levels": (C.c_int, [vp, i32p, i64p, C.c_int]),
        "gg_eval_last_kernel_ms": (C.c_int, [vp, P(C.c_float)]),
        "gg_eval_time_bfs": (C.c_int, [vp, C.c_int, P(C.c_float)]),
        "gg_eval_set_option": (C.c_int, [vp, C.c_int, C.c_int]),
        "gg_eval_get_option": (C.c_int, [vp, C.c_int, P(C.c_int)]),
        "gg_eval_adj_swap": (C.c_int, [vp, P(Swap), P(Result)]),
        "gg_eval_autotune": (C.c_int, [vp, i32p, C.c_int, P(TuneReport)]),
        "gg_eval_adj_commit": (C.c_int, [vp, P(Swap)]),
        "gg_probe_lds_read": (C.c_int, [C.c_int, C.c_int, P(LdsProbe)]),
        "gg_batch_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, P(vp)]),
        "gg_batch_destroy": (None, [vp]),
        "gg_batch_set_base": (C.c_int, [vp, i32p, P(Result)]),
        "gg_batch_eval": (C.c_int, [vp, P(Swap), C.c_int, P(Result)]),
        "gg_batch_commit": (C.c_int, [vp, P(Swap), P(Result)]),
        "gg_batch_last_kernel_ms": (C.c_int, [vp, P(C.c_float)]),
        "gg_batch_base_ecc": (C.c_int, [vp, i32p]),
        "gg_batch_levels_digest": (C.c_int, [vp, P(C.c_uint64)]),
        "gg_batch_set_option": (C.c_int, [vp, C.c_int, C.c_int]),
        "gg_batch_get_option": (C.c_int, [vp, C.c_int, P(C.c_int)]),
        "gg_batch_autotune": (C.c_int, [vp, P(Swap), C.c_int, C.c_int, P(BatchTuneReport)]),
        "gg_search_create": (C.c_int, [C.c_int, C.c_int, i32p, P(SearchCfg), C.c_int, P(vp)]),
        "gg_search_create_cb": (C.c_int, [C.c_int, C.c_int, i32p, P(SearchCfg),
                                          P(EvalCallbacks), P(vp)]),
        "gg_search_destroy": (None, [vp]),
        "gg_search_run": (C.c_int, [vp, C.c_int, P(SearchStats)]),
        "gg_search_graph": (C.c_int, [vp, i32p, C.c_int]),
        "gg_search_set_log": (C.c_int, [vp, C.c_int]),
        "gg_search_get_log": (C.c_int, [vp, i64p, C.c_int, P(C.c_int)]),
        "gg_search_audit": (C.c_int, [vp, C.c_uint64, P(C.c_double), P(C.c_double),
                                      P(C.c_float), P(C.c_double)]),
        "gg_search_last_selection": (C.c_int, [vp, i64p, i32p, P(C.c_float), i32p, i32p]),
        "gg_orbit_search_create": (C.c_int, [C.c_int, C.c_int, C.c_int, i32p, P(SearchCfg),
                                             C.c_int, P(vp)]),
        "gg_orbit_search_create_cb": (C.c_int, [C.c_int, C.c_int, C.c_int, i32p, P(SearchCfg),
                                                P(TableCallbacks), P(vp)]),
        "gg_piece_search_create": (C.c_int, [C.c_int] * 5 + [i32p, P(SearchCfg), C.c_int, P(vp)]),
        "gg_box_search_create": (C.c_int, [C.c_int] * 7 + [i32p, P(SearchCfg), C.c_int, P(vp)]),
        "gg_lattice_search_create_cb": (C.c_int, [C.c_int, i32p, i32p, C.c_int, i32p,
                                                  P(SearchCfg), P(TableCallbacks), P(vp)]),
        "gg_orbit_search_set_row_lengths": (C.c_int, [vp, i32p]),
        "gg_orbit_search_destroy": (None, [vp]),
        "gg_orbit_search_run": (C.c_int, [vp, C.c_int, P(SearchStats)]),
        "gg_orbit_search_table": (C.c_int, [vp, i32p, C.c_int]),
        "gg_orbit_search_set_log": (C.c_int, [vp, C.c_int]),
        "gg_orbit_search_get_log": (C.c_int, [vp, i64p, C.c_int, P(C.c_int)]),
        "gg_init_temp": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int,
                                   P(InitTempResult)]),
        "gg_init_temp_part": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64,
                                        C.c_int, P(InitTempResult)]),
        "gg_mlp_create": (C.c_int, [C.c_int, C.c_uint64, P(vp)]),
        "gg_mlp_destroy": (None, [vp]),
        "gg_mlp_set_params": (C.c_int, [vp, P(C.c_float)]),
        "gg_mlp_get_params": (C.c_int, [vp, P(C.c_float)]),
        "gg_mlp_forward": (C.c_int, [vp, P(C.c_float), C.c_int, P(C.c_float)]),
        "gg_mlp_grad": (C.c_int, [vp, P(C.c_float), P(C.c_float), C.c_int, P(C.c_float),
                                  P(C.c_float)]),
        "gg_mlp_train": (C.c_int, [vp, P(C.c_float), P(C.c_float), C.c_int, C.c_float,
                                   P(C.c_float)]),
        "gg_batch_features": (C.c_int, [vp, i32p, C.c_int, P(C.c_float)]),
        "gg_batch_score": (C.c_int, [vp, vp, i32p, C.c_int, P(C.c_float), P(C.c_float)]),
        "gg_batch_propose": (C.c_int, [vp, C.c_uint64, C.c_int64, C.c_int, i32p]),
        "gg_gen_random_regular": (C.c_int, [C.c_int, C.c_int, C.c_uint64, i32p]),
        "gg_gen_orbit_part": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint64, i32p]),
        "gg_gen_lattice": (C.c_int, [C.c_int, i32p, i32p, C.c_int, C.c_uint64, i32p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def strerror(code: int) -> str:
    return lib().gg_strerror(code).decode()


def check(code: int, what: str = "call"):
    if code < 0:
        raise GGError(code, f"{what}: {strerror(code)}")
    return code


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a: np.ndarray, ct=C.c_int32):
    return a.ctypes.data_as(C.POINTER(ct))


def device_count() -> int:
    n = C.c_int(0)
    code = lib().gg_device_count(C.byref(n))
    return n.value if code == GG_OK else 0


def probe_lds_read(bytes_per_read: int = 16, device: int = 0) -> dict:
    """Measured LDS read bandwidth of the device (wave-consecutive reads)."""
    r = LdsProbe()
    check(lib().gg_probe_lds_read(device, bytes_per_read, C.byref(r)), "gg_probe_lds_read")
    return {k: getattr(r, k) for k, _ in r._fields_}


# ---- generators (host code inside the library) ----------------------------------

def random_regular(n: int, d: int, seed: int) -> np.ndarray:
    """Seeded simple d-regular graph as an int32 [n, d] adjacency with sorted rows."""
    out = np.empty((n, d), dtype=np.int32)
    check(lib().gg_gen_random_regular(n, d, seed, _ptr(out)), "gg_gen_random_regular")
    return out


def orbit_part(g: int, m: int, d: int, seed: int) -> np.ndarray:
    """Seeded orbit graph (N = g*m, M = m) as an int32 [m, d] 'part' offset table."""
    out = np.empty((m, d), dtype=np.int32)
    check(lib().gg_gen_orbit_part(g, m, d, seed, _ptr(out)), "gg_gen_orbit_part")
    return out


def gen_lattice(Ns, Ms, d: int, seed: int) -> np.ndarray:
    """Seeded symmetric d-regular table [prod(Ms), d, len(Ns)] (part, piece or
    box form) by stub pairing with a lift per axis, self-loops and repeated
    offsets repaired."""
    dims = len(Ns)
    ns, ms = _i32(list(Ns)), _i32(list(Ms))
    out = np.empty((int(np.prod(Ms)), d, dims), dtype=np.int32)
    check(lib().gg_gen_lattice(dims, _ptr(ns), _ptr(ms), d, seed, _ptr(out)), "gg_gen_lattice")
    return out


# ---- evaluators ----------------------------------------------------------------------

class _Evaluator:
    kind = ""

    def __init__(self):
        self._h = C.c_void_p()

    def close(self):
        if self._h:
            lib().gg_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _run(self, fn, table: np.ndarray) -> Result:
        r = Result()
        check(fn(self._h, _ptr(table), C.byref(r)), fn.__name__)
        return r

    def levels(self, reach_cap: int = GG_LEVEL_LIMIT + 1):
        ecc = np.empty(self.M, dtype=np.int32)
        reach = np.zeros(reach_cap, dtype=np.int64)
        check(lib().gg_eval_levels(self._h, _ptr(ecc), _ptr(reach, C.c_int64), reach_cap),
              "gg_eval_levels")
        return ecc, reach

    def last_kernel_ms(self) -> float:
        ms = C.c_float(0)
        check(lib().gg_eval_last_kernel_ms(self._h, C.byref(ms)), "gg_eval_last_kernel_ms")
        return ms.value

    def time_bfs(self, reps: int = 10) -> float:
        """Mean ms per launch of the last evaluation's BFS kernel, `reps`
        launches back to back between one event pair."""
        ms = C.c_float(0)
        check(lib().gg_eval_time_bfs(self._h, reps, C.byref(ms)), "gg_eval_time_bfs")
        return ms.value

    def set_option(self, opt: int, value: int):
        check(lib().gg_eval_set_option(self._h, opt, value), "gg_eval_set_option")

    def autotune(self, table, reps: int = 5) -> TuneReport:
        t = _i32(table)
        rep = TuneReport()
        check(lib().gg_eval_autotune(self._h, _ptr(t), reps, C.byref(rep)), "gg_eval_autotune")
        return rep

    def get_option(self, opt: int) -> int:
        v = C.c_int(0)
        check(lib().gg_eval_get_option(self._h, opt, C.byref(v)), "gg_eval_get_option")
        return v.value


class PartEvaluator(_Evaluator):
    """1-D symmetric graph (cudaASPLconv): table = [M, degree] offsets."""

    def __init__(self, N: int, M: int, degree: int, device: int = 0):
        super().__init__()
        self.N, self.M, self.degree = N, M, degree
        check(lib().gg_create_part(N, M, degree, device, C.byref(self._h)), "gg_create_part")

    def eval(self, table) -> Result:
        t = _i32(table).reshape(self.M, self.degree)
        return self._run(lib().gg_eval_part, t)


class AdjEvaluator(_Evaluator):
    """Explicit adjacency (cudaASPLgraph): adj = [N, degree_max], -1 = padding."""

    def __init__(self, N: int, degree_max: int, device: int = 0):
        super().__init__()
        self.N, self.M, self.degree = N, N, degree_max
        check(lib().gg_create_adj(N, degree_max, device, C.byref(self._h)), "gg_create_adj")

    def eval(self, adj) -> Result:
        a = _i32(adj).reshape(self.N, self.degree)
        return self._run(lib().gg_eval_adj, a)


class SwapEvaluator(AdjEvaluator):
    """Plain single-move path: resident base adjacency, one swap per launch."""

    def set_base(self, adj) -> Result:
        return self.eval(adj)

    def eval_swap(self, sw: "Swap") -> Result:
        r = Result()
        check(lib().gg_eval_adj_swap(self._h, C.byref(sw), C.byref(r)), "gg_eval_adj_swap")
        return r

    def commit(self, sw: "Swap"):
        check(lib().gg_eval_adj_commit(self._h, C.byref(sw)), "gg_eval_adj_commit")


class BatchEvaluator:
    """K swap candidates against a resident base in one launch."""

    def __init__(self, N: int, degree: int, max_k: int, device: int = 0):
        self.N, self.degree, self.max_k = N, degree, max_k
        self._h = C.c_void_p()
        check(lib().gg_batch_create(N, degree, max_k, device, C.byref(self._h)),
              "gg_batch_create")

    def close(self):
        if self._h:
            lib().gg_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_base(self, adj) -> Result:
        a = _i32(adj).reshape(self.N, self.degree)
        r = Result()
        check(lib().gg_batch_set_base(self._h, _ptr(a), C.byref(r)), "gg_batch_set_base")
        return r

    def eval(self, swaps) -> list:
        K = len(swaps)
        arr = (Swap * K)(*swaps)
        out = (Result * K)()
        check(lib().gg_batch_eval(self._h, arr, K, out), "gg_batch_eval")
        return list(out)

    def commit(self, sw: "Swap") -> Result:
        r = Result()
        check(lib().gg_batch_commit(self._h, C.byref(sw), C.byref(r)), "gg_batch_commit")
        return r

    def last_kernel_ms(self) -> float:
        ms = C.c_float(0)
        check(lib().gg_batch_last_kernel_ms(self._h, C.byref(ms)), "gg_batch_last_kernel_ms")
        return ms.value

    def autotune(self, swaps, reps: int = 3) -> BatchTuneReport:
        """Time the incremental step's launch shapes and the base BFS variants
        on these candidates, apply and cache the fastest."""
        K = len(swaps)
        arr = (Swap * K)(*swaps)
        rep = BatchTuneReport()
        check(lib().gg_batch_autotune(self._h, arr, K, reps, C.byref(rep)), "gg_batch_autotune")
        return rep

    def shape(self) -> dict:
        """The incremental step's current launch shape."""
        names = {"launch_k": GG_BATCH_OPT_LAUNCH_K, "group_k": GG_BATCH_OPT_GROUP_K,
                 "block": GG_BATCH_OPT_SPARSE_BLOCK, "staged_keys": GG_BATCH_OPT_STAGED_KEYS,
                 "base_block": GG_BATCH_OPT_BASE_BLOCK, "base_push": GG_BATCH_OPT_BASE_PUSH}
        return {k: self.get_option(v) for k, v in names.items()}

    def base_ecc(self) -> np.ndarray:
        e = np.empty(self.N, dtype=np.int32)
        check(lib().gg_batch_base_ecc(self._h, _ptr(e)), "gg_batch_base_ecc")
        return e

    def levels_digest(self) -> int:
        h = C.c_uint64(0)
        check(lib().gg_batch_levels_digest(self._h, C.byref(h)), "gg_batch_levels_digest")
        return h.value

    def features(self, moves) -> np.ndarray:
        """32 features per move row (a, b, c, d, kind) on the current base."""
        m = _i32(moves).reshape(-1, 5)
        X = np.empty((m.shape[0], MLP_FEATURES), dtype=np.float32)
        check(lib().gg_batch_features(self._h, _ptr(m), m.shape[0], _ptr(X, C.c_float)),
              "gg_batch_features")
        return X

    def score(self, mlp, moves):
        """(features [P, 32], MLP scores [P]) of P moves in the learned
        search's fused launch."""
        m = _i32(moves).reshape(-1, 5)
        X = np.empty((m.shape[0], MLP_FEATURES), dtype=np.float32)
        sc = np.empty(m.shape[0], dtype=np.float32)
        check(lib().gg_batch_score(self._h, mlp._h, _ptr(m), m.shape[0], _ptr(X, C.c_float),
                                   _ptr(sc, C.c_float)), "gg_batch_score")
        return X, sc

    def propose(self, seed: int, step: int, P: int) -> np.ndarray:
        """P device-drawn 2-opt proposals {a, b, c, d, kind} on the current base."""
        m = np.empty((P, 5), dtype=np.int32)
        check(lib().gg_batch_propose(self._h, seed, step, P, _ptr(m)), "gg_batch_propose")
        return m

    def set_option(self, opt: int, value: int):
        check(lib().gg_batch_set_option(self._h, opt, value), "gg_batch_set_option")

    def get_option(self, opt: int) -> int:
        v = C.c_int(0)
        check(lib().gg_batch_get_option(self._h, opt, C.byref(v)), "gg_batch_get_option")
        return v.value


def make_swap(adj, a, ia, c, ic, kind=1):
    """2-opt swap of edges (a, adj[a][ia]) and (c, adj[c][ic]) as four slot
    replacements: kind 1 -> (a,d),(b,c); kind 0 -> (a,c),(b,d).  Returns
    (Swap, new_adjacency) or (None, None) when the result is not simple."""
    adj = np.asarray(adj)
    b = int(adj[a, ia])
    d = int(adj[c, ic])
    if len({int(a), b, int(c), d}) < 4:
        return None, None
    x, y = (d, c) if kind == 1 else (c, d)  # a gets x, b gets y
    if x in adj[a] or y in adj[b]:
        return None, None
    ib = int(np.nonzero(adj[b] == a)[0][0])
    idd = int(np.nonzero(adj[d] == c)[0][0])
    # a: b->x ; b: a->y ; c: d->(b if kind==1 else a) ; d: c->(a if kind==1 else b)
    cx = b if kind == 1 else a
    dx = a if kind == 1 else b
    sw = Swap()
    for j, (v, s, n) in enumerate([(a, ia, x), (b, ib, y), (c, ic, cx), (d, idd, dx)]):
        sw.v[j], sw.slot[j], sw.nbr[j] = int(v), int(s), int(n)
    new = adj.copy()
    for j in range(4):
        new[sw.v[j], sw.slot[j]] = sw.nbr[j]
    return sw, new


class PieceEvaluator(_Evaluator):
    """2-D symmetric graph (cudaASPLpiece): table = [Mx*My, degree, 2]."""

    def __init__(self, Nx, Ny, Mx, My, degree, device=0):
        super().__init__()
        self.N, self.M, self.degree = Nx * Ny, Mx * My, degree
        check(lib().gg_create_piece(Nx, Ny, Mx, My, degree, device, C.byref(self._h)),
              "gg_create_piece")

    def eval(self, table) -> Result:
        t = _i32(table).reshape(self.M, self.degree, 2)
        return self._run(lib().gg_eval_piece, t)


class BoxEvaluator(_Evaluator):
    """3-D symmetric graph (cudaASPLbox): table = [Mx*My*Mz, degree, 3]."""

    def __init__(self, Nx, Ny, Nz, Mx, My, Mz, degree, device=0):
        super().__init__()
        self.N, self.M, self.degree = Nx * Ny * Nz, Mx * My * Mz, degree
        check(lib().gg_create_box(Nx, Ny, Nz, Mx, My, Mz, degree, device, C.byref(self._h)),
              "gg_create_box")

    def eval(self, table) -> Result:
        t = _i32(table).reshape(self.M, self.degree, 3)
        return self._run(lib().gg_eval_box, t)


def aspl(result: Result, N: int) -> float:
    """ASPL as the reference derives it (cudacode/cudaASPLconv.cu:242)."""
    return float(result.total_scaled) / (N * (N - 1))


# ---- search drivers ---------------------------------------------------------------------

def search_cfg(mode=MODE_PLAIN, rule=RULE_ANNEAL, K=1, P=1, total_steps=1000, T0=1e-3,
               T1=1e-6, seed=12345, warmup_pairs=2048, train_every=1, train_batch=256,
               lr=1e-3, schedule=0) -> SearchCfg:
    return SearchCfg(mode=mode, rule=rule, K=K, P=P, total_steps=total_steps,
                     warmup_pairs=warmup_pairs, train_every=train_every, train_batch=train_batch,
                     T0=T0, T1=T1, lr=lr, seed=seed, schedule=schedule)


class Search:
    """Annealing / hill-climbing driver over 2-opt swaps on the exact evaluator.
    With `callbacks`, the exact evaluator is the caller's (as the reference
    swaps cudaASPLconv for cpuASPLqueue, cppcode/annealing_cpu.cpp:292)."""

    def __init__(self, adj, cfg: SearchCfg, device: int = 0, callbacks=None, log: int = 0):
        a = _i32(adj)
        self.N, self.d = a.shape
        self.P, self.K = cfg.P, cfg.K
        self._h = C.c_void_p()
        self._cb = callbacks  # keep the ctypes callbacks alive
        if callbacks is None:
            check(lib().gg_search_create(self.N, self.d, _ptr(a), C.byref(cfg), device,
                                         C.byref(self._h)), "gg_search_create")
        else:
            check(lib().gg_search_create_cb(self.N, self.d, _ptr(a), C.byref(cfg),
                                            C.byref(callbacks.struct), C.byref(self._h)),
                  "gg_search_create_cb")
        if log:
            check(lib().gg_search_set_log(self._h, log), "gg_search_set_log")

    def run(self, steps: int) -> SearchStats:
        st = SearchStats()
        check(lib().gg_search_run(self._h, steps, C.byref(st)), "gg_search_run")
        return st

    def set_log(self, capacity: int):
        """Log the next `capacity` steps (clears what was logged)."""
        check(lib().gg_search_set_log(self._h, capacity), "gg_search_set_log")

    def graph(self, best=False) -> np.ndarray:
        out = np.empty((self.N, self.d), dtype=np.int32)
        check(lib().gg_search_graph(self._h, _ptr(out), 1 if best else 0), "gg_search_graph")
        return out

    def log(self) -> np.ndarray:
        n = C.c_int(0)
        check(lib().gg_search_get_log(self._h, None, 0, C.byref(n)), "gg_search_get_log")
        buf = np.zeros((n.value, 10), dtype=np.int64)
        check(lib().gg_search_get_log(self._h, _ptr(buf, C.c_int64), n.value, C.byref(n)),
              "gg_search_get_log")
        return buf

    def audit(self, seed: int):
        """Learned mode: (spearman, miss_rate, scores[P], exact_deltas[P])."""
        P = self.P
        sc = np.empty(P, dtype=np.float32)
        dl = np.empty(P, dtype=np.float64)
        rho, miss = C.c_double(), C.c_double()
        check(lib().gg_search_audit(self._h, seed, C.byref(rho), C.byref(miss),
                                    _ptr(sc, C.c_float), _ptr(dl, C.c_double)), "gg_search_audit")
        return rho.value, miss.value, sc, dl

    def last_selection(self):
        """Learned mode (proposals drawn on the GPU): the last step's
        (step index, warm flag, scores[P], move rows[P][5], chosen indices[K])."""
        P, K = self.P, self.K
        step, warm = C.c_int64(), C.c_int32()
        sc = np.empty(P, dtype=np.float32)
        mv = np.empty((P, 5), dtype=np.int32)
        sel = np.empty(K, dtype=np.int32)
        check(lib().gg_search_last_selection(self._h, C.byref(step), C.byref(warm),
                                             _ptr(sc, C.c_float), _ptr(mv), _ptr(sel)),
              "gg_search_last_selection")
        return step.value, warm.value, sc, mv, sel

    def close(self):
        if self._h:
            lib().gg_search_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def init_temp(N, degree, graphs=100, neighbours=100, seed=1, device=0, M=None) -> InitTempResult:
    """Initial temperature as cppcode/initTemp.cpp:399-430 computes it: explicit
    random regular graphs (M None or N), or symmetric part tables with M rows."""
    r = InitTempResult()
    if M is None or M == N:
        check(lib().gg_init_temp(N, degree, graphs, neighbours, seed, device, C.byref(r)),
              "gg_init_temp")
    else:
        check(lib().gg_init_temp_part(N, M, degree, graphs, neighbours, seed, device, C.byref(r)),
              "gg_init_temp_part")
    return r


class OrbitSearch:
    """Annealing / hill climbing over cyclically symmetric graphs in the
    reference's `part` form (M rows of offsets, N = g*M): createNeighbour's
    edge-length change and orbit 2-opt (cppcode/annealing.cpp:112-197), each
    candidate table evaluated exactly from the M representatives.  With
    `table_eval` (a Python function table[M][d] -> (total_scaled, diameter,
    status[, wvc])), the exact evaluator is the caller's."""

    def __init__(self, N, table, cfg: SearchCfg, device: int = 0, table_eval=None, log: int = 0,
                 row_lengths=None):
        t = _i32(table)
        self.N = N
        self.M, self.d = t.shape
        self._h = C.c_void_p()
        if table_eval is None:
            self._cb = None
            check(lib().gg_orbit_search_create(N, self.M, self.d, _ptr(t), C.byref(cfg), device,
                                               C.byref(self._h)), "gg_orbit_search_create")
        else:
            M, d = self.M, self.d

            def ev(ctx, tab, out):
                a = np.ctypeslib.as_array(tab, shape=(M * d,)).reshape(M, d).copy()
                res = tuple(table_eval(a))
                total, diam, status = res[:3]
                r = out.contents
                r.total_reps, r.total_scaled, r.diameter, r.wvc, r.levels, r.status = (
                    total // (N // M), total, diam, res[3] if len(res) > 3 else 0, diam + 1,
                    status)
                return 0

            self._fn = TABLE_CB(ev)
            self._cb = TableCallbacks(None, self._fn)
            check(lib().gg_orbit_search_create_cb(N, self.M, self.d, _ptr(t), C.byref(cfg),
                                                  C.byref(self._cb), C.byref(self._h)),
                  "gg_orbit_search_create_cb")
        if row_lengths is not None:
            lens = _i32(row_lengths)
            check(lib().gg_orbit_search_set_row_lengths(self._h, _ptr(lens)),
                  "gg_orbit_search_set_row_lengths")
        if log:
            check(lib().gg_orbit_search_set_log(self._h, log), "gg_orbit_search_set_log")

    def run(self, steps: int) -> SearchStats:
        st = SearchStats()
        check(lib().gg_orbit_search_run(self._h, steps, C.byref(st)), "gg_orbit_search_run")
        return st

    def table(self, best=False) -> np.ndarray:
        out = np.empty((self.M, self.d), dtype=np.int32)
        check(lib().gg_orbit_search_table(self._h, _ptr(out), 1 if best else 0),
              "gg_orbit_search_table")
        return out

    def log(self) -> np.ndarray:
        n = C.c_int(0)
        check(lib().gg_orbit_search_get_log(self._h, None, 0, C.byref(n)), "get_log")
        buf = np.zeros((n.value, 10), dtype=np.int64)
        check(lib().gg_orbit_search_get_log(self._h, _ptr(buf, C.c_int64), n.value, C.byref(n)),
              "get_log")
        return buf

    def close(self):
        if self._h:
            lib().gg_orbit_search_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LatticeSearch(OrbitSearch):
    """The orbit search over 2-D (`piece`, Ns = (Nx, Ny)) or 3-D (`box`) symmetric
    graphs: table [Mx*My(*Mz), degree, dims] of offset vectors, moves of
    cppcode/pieceannealing.cpp:113-204 / boxHillClimb.cpp:113-220, candidates
    evaluated by gg_eval_piece / gg_eval_box (or `table_eval`, a Python
    function table[M, d, dims] -> (total_scaled, diameter, status))."""

    def __init__(self, Ns, Ms, table, cfg: SearchCfg, device: int = 0, table_eval=None,
                 log: int = 0):
        dims = len(Ns)
        assert dims in (2, 3) and len(Ms) == dims
        M = int(np.prod(Ms))
        t = _i32(table)
        d = t.size // (M * dims)
        t = t.reshape(M, d, dims)
        self.N, self.M, self.d, self.dims = int(np.prod(Ns)), M, d, dims
        self._h = C.c_void_p()
        ns, ms = _i32(list(Ns)), _i32(list(Ms))
        if table_eval is None:
            self._cb = None
            if dims == 2:
                check(lib().gg_piece_search_create(Ns[0], Ns[1], Ms[0], Ms[1], d, _ptr(t),
                                                   C.byref(cfg), device, C.byref(self._h)),
                      "gg_piece_search_create")
            else:
                check(lib().gg_box_search_create(Ns[0], Ns[1], Ns[2], Ms[0], Ms[1], Ms[2], d,
                                                 _ptr(t), C.byref(cfg), device,
                                                 C.byref(self._h)), "gg_box_search_create")
        else:
            N = self.N

            def ev(ctx, tab, out):
                a = np.ctypeslib.as_array(tab, shape=(M * d * dims,)).reshape(M, d, dims).copy()
                total, diam, status = table_eval(a)
                r = out.contents
                r.total_reps, r.total_scaled, r.diameter, r.wvc, r.levels, r.status = (
                    total // (N // M), total, diam, 0, diam + 1, status)
                return 0

            self._fn = TABLE_CB(ev)
            self._cb = TableCallbacks(None, self._fn)
            check(lib().gg_lattice_search_create_cb(dims, _ptr(ns), _ptr(ms), d, _ptr(t),
                                                    C.byref(cfg), C.byref(self._cb),
                                                    C.byref(self._h)),
                  "gg_lattice_search_create_cb")
        if log:
            check(lib().gg_orbit_search_set_log(self._h, log), "gg_orbit_search_set_log")

    def table(self, best=False) -> np.ndarray:
        out = np.empty((self.M, self.d, self.dims), dtype=np.int32)
        check(lib().gg_orbit_search_table(self._h, _ptr(out), 1 if best else 0),
              "gg_orbit_search_table")
        return out


class PyEvaluatorCallbacks:
    """Wraps a Python evaluator object exposing set_base(adj) -> (total, diam,
    status[, wvc]), eval(list of Swap) -> list of such tuples and
    commit(Swap) -> one as gg_eval_callbacks."""

    def __init__(self, impl, N, d):
        self.impl, self.N, self.d = impl, N, d

        def _fill(r, t):
            r.total_reps, r.total_scaled, r.diameter, r.wvc, r.levels, r.status = (
                t[0], t[0], t[1], t[3] if len(t) > 3 else 0, t[1] + 1, t[2])

        def set_base(ctx, adj, out):
            a = np.ctypeslib.as_array(adj, shape=(N * d,)).reshape(N, d).copy()
            _fill(out.contents, impl.set_base(a))
            return 0

        def ev(ctx, swaps, K, out):
            for k, t in enumerate(impl.eval([swaps[k] for k in range(K)])):
                _fill(out[k], t)
            return 0

        def commit(ctx, sw, out):
            _fill(out.contents, impl.commit(sw.contents))
            return 0

        self._fns = (SET_BASE_CB(set_base), EVAL_CB(ev), COMMIT_CB(commit))
        self.struct = EvalCallbacks(None, *self._fns)


class MLP:
    """32-128-1 proposal scorer (bf16 weights, fp32 accumulation) on the GPU."""

    def __init__(self, seed: int = 1, device: int = 0):
        self._h = C.c_void_p()
        check(lib().gg_mlp_create(device, seed, C.byref(self._h)), "gg_mlp_create")

    def close(self):
        if self._h:
            lib().gg_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        w = np.empty(MLP_PARAMS, dtype=np.float32)
        check(lib().gg_mlp_get_params(self._h, _ptr(w, C.c_float)), "gg_mlp_get_params")
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == MLP_PARAMS
        check(lib().gg_mlp_set_params(self._h, _ptr(w, C.c_float)), "gg_mlp_set_params")

    def forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, MLP_FEATURES)
        out = np.empty(X.shape[0], dtype=np.float32)
        check(lib().gg_mlp_forward(self._h, _ptr(X, C.c_float), X.shape[0], _ptr(out, C.c_float)),
              "gg_mlp_forward")
        return out

    def grad(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, MLP_FEATURES)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        g = np.empty(MLP_PARAMS, dtype=np.float32)
        loss = C.c_float()
        check(lib().gg_mlp_grad(self._h, _ptr(X, C.c_float), _ptr(Y, C.c_float), X.shape[0],
                                C.byref(loss), _ptr(g, C.c_float)), "gg_mlp_grad")
        return loss.value, g

    def train(self, X, Y, lr=1e-3) -> float:
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, MLP_FEATURES)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        loss = C.c_float()
        check(lib().gg_mlp_train(self._h, _ptr(X, C.c_float), _ptr(Y, C.c_float), X.shape[0], lr,
                                 C.byref(loss)), "gg_mlp_train")
        return loss.value
