#!/usr/bin/env python3
"""Benchmark: exact moves evaluated per second on one MI355X.

Workload: random 24-regular graph, n = 16384 (seeded), simulated annealing over
2-opt swaps with the exact ASPL/diameter evaluator.  A step is one search step:
  plain    one proposal, one exact evaluation (one BFS launch);
  learned  1024 proposals scored by the MLP, the top 32 evaluated exactly in
           one incremental batched launch (HIP graph), MLP trained online.
`value` is exact evaluations/s of the batched+learned search (proposals that
are only scored are not counted); the plain single-move GPU search is measured
beside it.  Inputs (graph, tables, base BFS levels) are resident in HBM.

Prints one JSON line.  Only the cpu_baseline leg loads the CPU oracle and the
reference's CPU evaluator built from its sources (oracle/_ref), to time them.
"""
import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(REPO, "graphgolf-cuda_amd")]

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# LDS: 256 CUs x 256 B/clk x 2.4 GHz peak engine clock.  256 B/clk/CU is the
# smallest power of two above the 16-byte read rate the LDS probe measures on
# gfx950 (about 230 B/clk/CU); the probe's figure is reported beside it as
# what is achievable.
LDS_PEAK_GBS = 256 * 256 * 2.4
# exact (features, delta) pairs the scorer trains on before the timed window
TRAIN_PAIRS = 100_000


def host_cpu():
    """Model name and logical CPU count of the host the CPU legs run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"cpu_model": model, "nproc": avail, "logical_cpus": os.cpu_count()}


def cpu_baseline(adj, seconds_target=12.0):
    """Oracle queue BFS (OpenMP over sources) on a bounded sample of sources,
    scaled to exact evaluations/s of the whole graph."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as ol
    n, d = adj.shape
    threads = min(os.cpu_count() or 1, 16)
    a = np.ascontiguousarray(adj, dtype=np.int32)
    tot, dia = C.c_int64(), C.c_int32()
    # grow the source sample until one pass takes a measurable time, then
    # repeat passes until about seconds_target of CPU work has been timed
    nsrc = 256
    while True:
        srcs = np.arange(nsrc, dtype=np.int32)
        t0 = time.perf_counter()
        ol.oracle().oracle_bfs_omp(n, d, ol._p(a), nsrc, ol._p(srcs), threads, C.byref(tot),
                                   C.byref(dia))
        dt = time.perf_counter() - t0
        if dt > seconds_target / 4 or nsrc >= n:
            break
        nsrc = min(n, nsrc * 4)
    passes, t_all = 1, dt
    while t_all < seconds_target:
        t0 = time.perf_counter()
        ol.oracle().oracle_bfs_omp(n, d, ol._p(a), nsrc, ol._p(srcs), threads, C.byref(tot),
                                   C.byref(dia))
        t_all += time.perf_counter() - t0
        passes += 1
    return {"value": passes * (nsrc / n) / t_all, "unit": "exact evaluations/s",
            "cores": threads, "kind": "port",
            "sample": f"oracle queue BFS, OpenMP {threads} threads, {passes} passes over {nsrc} "
                      f"of {n} sources of the n={n} d={d} graph in {t_all:.2f} s, "
                      f"scaled by {n}/{nsrc}"}


def cpu_baseline_reference(gg, n, d, seconds_target=10.0, g=16):
    """The reference's own CPU evaluator (cppcode/cpuASPLqueue.cpp
    diameterASPL, compiled from the reference sources into oracle/_ref;
    single-threaded as the reference runs it).  A whole evaluation at
    n = 16384, d = 24 takes about a minute on one thread, so the sample is the
    same evaluator on an n-vertex d-regular orbit graph with g copies: a queue
    BFS from n/g representatives over all n vertices (the per-source work of
    the explicit graph), scaled by n/g / n.  None when oracle/_ref is not
    built."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as ol
    if not ol.ref_available() or n % g:
        return None
    m = n // g
    # a simple d-regular graph on the m representatives, every edge {i, j}
    # lifted by a random shift s: offsets (j - i) + m*s and its reverse
    base = gg.random_regular(m, d, 11)
    rng = np.random.default_rng(11)
    table = np.zeros((m, d), dtype=np.int32)
    for i in range(m):
        for si, j in enumerate(base[i]):
            j = int(j)
            if j < i:
                continue
            off = (j - i) + m * int(rng.integers(0, g))
            table[i, si] = off
            table[j, int(np.nonzero(base[j] == i)[0][0])] = -off
    passes, t_all = 0, 0.0
    while t_all < seconds_target or passes == 0:
        t0 = time.perf_counter()
        ol.ref_part(n, m, table)
        t_all += time.perf_counter() - t0
        passes += 1
    return {"value": passes * (m / n) / t_all, "unit": "exact evaluations/s", "cores": 1,
            "kind": "reference",
            "sample": f"cpuASPLqueue<16384>::diameterASPL of the reference, built from its "
                      f"sources (oracle/_ref), on one host thread: {passes} evaluation(s) of an "
                      f"n={n} d={d} orbit graph (BFS from {m} representatives over all {n} "
                      f"vertices) in {t_all:.2f} s, scaled by {m}/{n}"}


def plain_kernel_roofline(gg, adj, reps=30):
    """Dense column-BFS kernel (the plain evaluator) timed with HIP events on
    its own stream: per swap, ten back-to-back launches of that swap's BFS
    between one event pair; algorithmic bytes 8*(d+2) per word-level."""
    n, d = adj.shape
    rng = np.random.default_rng(7)
    ev = gg.SwapEvaluator(n, d)
    ev.set_base(adj)
    info = {k: ev.get_option(v) for k, v in (("variant", gg.GG_INFO_VARIANT),
                                               ("vpt", gg.GG_INFO_VPT),
                                               ("fixed", gg.GG_INFO_FIXED),
                                               ("push_cap", gg.GG_INFO_PUSH_CAP))}
    kms, lev, single = [], [], []
    while len(kms) < reps:
        a, c = (int(x) for x in rng.integers(0, n, 2))
        ia, ic = (int(x) for x in rng.integers(0, d, 2))
        sw, _ = gg.make_swap(adj, a, ia, c, ic, int(rng.integers(0, 2)))
        if sw is None:
            continue
        r = ev.eval_swap(sw)
        single.append(ev.last_kernel_ms())
        # the same launch repeated back to back between one event pair: the
        # per-launch time without the event overhead of a single bracket
        kms.append(ev.time_bfs(10))
        lev.append(r.levels)
    ev.close()
    W = (n + 63) // 64
    L = float(np.mean(lev))
    alg = 8.0 * (d + 2) * n * W * L
    t = float(np.mean(kms[3:])) * 1e-3
    ach = alg / t / 1e9
    bs = 256 if info["variant"] == 1 else 1024
    name = (f"k_dense_lds<TopoRingQ, {bs}, {info['vpt']}, {info['fixed']}>"
            if info["variant"] in (1, 2) else "k_dense_global<TopoRingQ>")
    # LDS ceiling: the fixed width-times-clock figure; the measured
    # conflict-free read rates of the device are what is achievable
    p16 = gg.probe_lds_read(16)
    p8 = gg.probe_lds_read(8)
    peak = LDS_PEAK_GBS
    return {"bound": "lds", "achieved": ach, "peak": peak, "unit": "GB/s",
            "frac": ach / peak, "traffic": hbm_traffic(n, d), "lds_counters": lds_counters(n, d),
            "kernel": name,
            "kernel_ms": t * 1e3, "kernel_ms_single_bracket": float(np.mean(single[3:])),
            "levels": L, "alg_bytes_per_launch": alg,
            "push_cap": info["push_cap"],
            "peak_basis": "256 CU x 256 B/clk x 2.4 GHz",
            "achievable": {"gbs": p16["gbs"], "frac": ach / p16["gbs"],
                           "basis": "measured conflict-free 16-byte LDS read probe"},
            "lds_probe": {"b128_gbs": p16["gbs"], "b64_gbs": p8["gbs"],
                          "clock_mhz": p16["clock_mhz"],
                          "b128_bytes_per_clk_cu": p16["bytes_per_clk_cu"],
                          "b64_bytes_per_clk_cu": p8["bytes_per_clk_cu"]},
            "alg_rate_over_hbm_peak": ach / HBM_PEAK_GBS}


def pmc_summary(name):
    """The `key: value` figures of a committed rocprofv3 --pmc summary under
    profiles/ (counters scaled by the copy-kernel calibration recorded in
    it), or {} when there is none."""
    path = os.path.join(REPO, "profiles", name)
    vals = {}
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                if ":" in line and not line.startswith("#"):
                    k, v = line.split(":", 1)
                    try:
                        vals[k.strip()] = float(v)
                    except ValueError:
                        pass
    return vals


def hbm_traffic(n, d):
    """HBM bytes per launch of the dense kernel (FETCH_SIZE + WRITE_SIZE) from
    the committed summary for this shape, or None."""
    return pmc_summary(f"pmc_k_dense_lds_n{n}_d{d}.txt").get("hbm_bytes_per_launch")


def lds_counters(n, d):
    """LDS instructions and cycles per launch of the dense kernel from the
    committed summary, or None."""
    v = pmc_summary(f"pmc_k_dense_lds_n{n}_d{d}.txt")
    if "lds_insts_per_launch" not in v:
        return None
    return {"lds_insts_per_launch": v["lds_insts_per_launch"],
            "bank_conflict_cycles": v["lds_bank_conflict_cycles_per_launch"],
            "lds_active_cycles": v["lds_active_cycles_per_launch"],
            "bank_conflict_frac_of_active": v["lds_bank_conflict_cycles_per_launch"]
            / v["lds_active_cycles_per_launch"]}


def sparse_roofline(n, d, K):
    """HBM roofline of the incremental kernel at the learned search's shape
    from its committed counter summary: bytes fetched and written per launch
    over the kernel time of the same profiled command."""
    v = pmc_summary(f"pmc_k_sparse_n{n}_d{d}_k{K}.txt")
    if "hbm_bytes_per_launch" not in v:
        return None
    t = v["kernel_us"] * 1e-6
    ach = v["hbm_bytes_per_launch"] / t / 1e9
    return {"bound": "hbm", "kernel": "k_sparse<1024, 24, 1>", "achieved": ach,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": v["hbm_bytes_per_launch"], "kernel_ms": v["kernel_us"] * 1e-3,
            "lds_insts_per_launch": v.get("lds_insts_per_launch"),
            "source": f"profiles/pmc_k_sparse_n{n}_d{d}_k{K}.txt (FETCH_SIZE + WRITE_SIZE and "
                      "kernel time of bench.py --only sparse under rocprofv3)"}


def copy_calibration(nbytes=1 << 30, reps=5):
    """Device-to-device hipMemcpy of a known size (the runtime's blit copy
    kernel), the reference point that scales the FETCH_SIZE / WRITE_SIZE
    counters to bytes."""
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        assert hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0, "hipMalloc"
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0, "hipMemset"
    for _ in range(reps):
        assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0, "hipMemcpy"  # device to device
    assert hip.hipDeviceSynchronize() == 0
    for p in (src, dst):
        hip.hipFree(p)
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps}),
          flush=True)


def batched_vs_sequential(gg, n=8192, d=16, K=256, steps=10, seed=4):
    """K candidate 2-opt swaps on a random d-regular graph, one batched launch
    per step (incremental path, HIP graph) against K sequential single-move
    launches of the plain evaluator on the same candidates."""
    rng = np.random.default_rng(seed)
    adj = gg.random_regular(n, d, seed)
    cands = []
    while len(cands) < K:
        a, c = (int(x) for x in rng.integers(0, n, 2))
        ia, ic = (int(x) for x in rng.integers(0, d, 2))
        sw, _ = gg.make_swap(adj, a, ia, c, ic, int(rng.integers(0, 2)))
        if sw is not None:
            cands.append(sw)
    def timed(be):
        r = be.eval(cands)  # warm-up (captures the graph)
        t0 = time.perf_counter()
        for _ in range(steps):
            r = be.eval(cands)
        return r, (time.perf_counter() - t0) / steps

    with gg.BatchEvaluator(n, d, max_k=K) as be:
        be.set_base(adj)
        # the automatic launch shape (all K in one launch), then the tuned one
        for opt in (gg.GG_BATCH_OPT_LAUNCH_K, gg.GG_BATCH_OPT_GROUP_K,
                    gg.GG_BATCH_OPT_SPARSE_BLOCK, gg.GG_BATCH_OPT_STAGED_KEYS):
            be.set_option(opt, 0)
        rd, td = timed(be)
        default_shape = be.shape()
        rep = be.autotune(cands, reps=3)
        rb, tb = timed(be)
        tuned_shape = be.shape()
    assert [(a.total_reps, a.diameter) for a in rd] == [(a.total_reps, a.diameter) for a in rb]
    tune = rep.as_dict()
    ev = gg.SwapEvaluator(n, d)
    ev.set_base(adj)
    ev.eval_swap(cands[0])
    t0 = time.perf_counter()
    rs = [ev.eval_swap(sw) for sw in cands]
    ts = time.perf_counter() - t0
    ev.close()
    assert all(a.total_reps == b.total_reps and a.diameter == b.diameter for a, b in zip(rb, rs))
    return {"workload": f"K={K} 2-opt candidates, random {d}-regular graph n={n}",
            "batched_ms_per_step": tb * 1e3, "sequential_ms": ts * 1e3,
            "speedup": ts / tb, "batched_evals_per_s": K / tb, "sequential_evals_per_s": K / ts,
            "launches_per_step": -(-K // tuned_shape["launch_k"]),
            "default_shape": default_shape, "default_ms_per_step": td * 1e3,
            "tuned_shape": tuned_shape,
            "autotune": {"configs": len(tune["configs"]), "mismatches": tune["mismatches"],
                         "best": tune["best"], "default_ms": tune["default_ms"],
                         "base": tune["base"], "best_base": tune["best_base"]}}


def sparse_steps(gg, adj, K, steps=20, seed=13):
    """Incremental steps of K random 2-opt candidates on the start graph, each
    launched without a HIP graph (so a counter run can attribute every
    kernel), the results checked against the dense path once."""
    n, d = adj.shape
    rng = np.random.default_rng(seed)
    cands = []
    while len(cands) < K:
        a, c = (int(x) for x in rng.integers(0, n, 2))
        ia, ic = (int(x) for x in rng.integers(0, d, 2))
        sw, _ = gg.make_swap(adj, a, ia, c, ic, int(rng.integers(0, 2)))
        if sw is not None:
            cands.append(sw)
    with gg.BatchEvaluator(n, d, max_k=K) as be:
        be.set_base(adj)
        be.set_option(gg.GG_BATCH_OPT_GRAPH, 0)
        r = be.eval(cands)
        t0 = time.perf_counter()
        for _ in range(steps):
            r = be.eval(cands)
        ms = (time.perf_counter() - t0) / steps * 1e3
        be.set_option(gg.GG_BATCH_OPT_PATH, gg.PATH_DENSE)
        rd = be.eval(cands)
    assert [(x.total_reps, x.diameter) for x in r] == [(x.total_reps, x.diameter) for x in rd]
    return {"workload": f"K={K} random 2-opt candidates, random {d}-regular graph n={n}",
            "ms_per_step": ms, "steps": steps}


def tune_step_shape(gg, adj, K, seed=11):
    """Autotunes the incremental step's launch shape for (n, d, K) on K random
    2-opt candidates of the search's start graph; the cached choice applies
    to every batch evaluator of that shape created afterwards (the learned
    search's included)."""
    n, d = adj.shape
    rng = np.random.default_rng(seed)
    cands = []
    while len(cands) < K:
        a, c = (int(x) for x in rng.integers(0, n, 2))
        ia, ic = (int(x) for x in rng.integers(0, d, 2))
        sw, _ = gg.make_swap(adj, a, ia, c, ic, int(rng.integers(0, 2)))
        if sw is not None:
            cands.append(sw)
    with gg.BatchEvaluator(n, d, max_k=K) as be:
        be.set_base(adj)
        rep = be.autotune(cands, reps=3).as_dict()
        shape = be.shape()
    return {"shape": shape, "configs": len(rep["configs"]), "mismatches": rep["mismatches"],
            "best_ms": rep["best"]["ms"], "default_ms": rep["default_ms"],
            "best_base": rep["best_base"]}


def run_search(gg, adj, cfg, warmup, steps):
    """`warmup` untimed steps, then `steps` timed ones (wall time measured by
    the driver around its own loop).  Returns the search, the stats before
    and after the timed window, the exact evaluations and the seconds."""
    s = gg.Search(adj, cfg)
    if warmup:
        s.run(warmup)
    st0 = s.run(0)
    st = s.run(steps)
    return s, st0, st, st.exact_evals - st0.exact_evals, st.seconds - st0.seconds


def large_n_orbit(gg, g=8192, m=128, d=16, seed=3, reps=5):
    """Large-N evaluation: an orbit graph of N = g*m = 2^20 vertices with m = 128
    representatives (two 64-source columns; the reference's own large-N design
    point, cudaASPLconv.cu:157).  The level-synchronous path (one launch per
    level over the whole grid) against the persistent one-workgroup-per-column
    global kernel; algorithmic bytes 8*(d+2)*N*W*L as for the dense kernel."""
    N = g * m
    tab = gg.orbit_part(g, m, d, seed)
    out = {"workload": f"orbit graph N={N}, {m} representatives, d={d}"}
    with gg.PartEvaluator(N, m, d) as ev:
        r = ev.eval(tab)
        t0 = time.perf_counter()
        for _ in range(reps):
            r2 = ev.eval(tab)
        wide_ms = (time.perf_counter() - t0) / reps * 1e3
        assert r2.total_scaled == r.total_scaled
        bfs_ms = ev.time_bfs(reps)
        W = (m + 63) // 64
        alg = 8 * (d + 2) * N * W * r.levels
        out.update({"diameter": r.diameter, "levels": r.levels, "total_scaled": r.total_scaled,
                    "level_sync_ms_per_eval": wide_ms, "level_sync_bfs_ms": bfs_ms,
                    "level_sync_alg_GB_per_s": alg / (bfs_ms * 1e-3) / 1e9,
                    "level_sync_alg_frac_of_hbm_peak": alg / (bfs_ms * 1e-3) / 1e9 / HBM_PEAK_GBS})
        v = pmc_summary("pmc_k_wide_level_n1048576_orbit_n131072_adjacency_d16.txt")
        if N == 1 << 20 and "orbit_fetch_bytes_per_eval" in v:
            miss = v["orbit_fetch_bytes_per_eval"] + v["orbit_write_bytes_per_eval"]
            out.update({"l2_miss_bytes_per_eval": miss, "l2_hit_rate": v["orbit_tcc_hit_rate"],
                        "l2_miss_GB_per_s": miss / (bfs_ms * 1e-3) / 1e9})
        ev.set_option(gg.GG_OPT_WIDE, 0)
        t0 = time.perf_counter()
        r3 = ev.eval(tab)
        pers_ms = (time.perf_counter() - t0) * 1e3
        assert ev.get_option(gg.GG_INFO_VARIANT) == 3 and r3.total_scaled == r.total_scaled
        out.update({"persistent_ms_per_eval": pers_ms, "speedup": pers_ms / wide_ms})
    return out


def large_n_plain(gg, g=1024, m=128, d=16, seed=3, reps=3):
    """Plain large-N evaluation with every vertex a source: the orbit graph of
    N = g*m = 2^17 vertices as an explicit adjacency (the reference's
    cudaASPLgraph, 1024-source passes, cudaASPLgraph.cu:155-160); here passes
    of 256 columns, each level one launch over the whole grid, against the
    one-workgroup-per-column global kernel.  Algorithmic bytes 8*(d+2)*N*W*L."""
    N = g * m
    tab = gg.orbit_part(g, m, d, seed)
    # the orbit graph's explicit adjacency: vertex v's neighbours are
    # (v + o) mod N for the offsets o of row v mod m
    v = np.arange(N, dtype=np.int64)[:, None]
    adj = ((v + tab[v[:, 0] % m]) % N).astype(np.int32)
    out = {"workload": f"explicit adjacency N={N} (all sources, {N // 64} columns), d={d}"}
    with gg.AdjEvaluator(N, d) as ev:
        r = ev.eval(adj)
        bfs_ms = ev.time_bfs(reps)
        W = N // 64
        alg = 8 * (d + 2) * N * W * r.levels
        out.update({"diameter": r.diameter, "levels": r.levels, "total": r.total_reps,
                    "passes": -(-W // 256), "level_sync_bfs_ms": bfs_ms,
                    "level_sync_alg_GB_per_s": alg / (bfs_ms * 1e-3) / 1e9})
        ev.set_option(gg.GG_OPT_WIDE, 0)
        r2 = ev.eval(adj)
        assert ev.get_option(gg.GG_INFO_VARIANT) == 3 and r2.total_reps == r.total_reps
        pers_ms = ev.time_bfs(1)
        out.update({"persistent_bfs_ms": pers_ms, "speedup": pers_ms / bfs_ms})
    return out


def equal_time(gg, adj, cfgs, seconds):
    """Each search for `seconds` of wall time from the same graph (its
    temperature schedule sized to the steps it makes in that time), best
    ASPL and diameter reached."""
    n = adj.shape[0]
    res = {}
    for name, cfg, chunk in cfgs:
        s = gg.Search(adj, cfg)
        t_end = time.perf_counter() + seconds
        st = s.run(0)
        while time.perf_counter() < t_end:
            st = s.run(chunk)
        res[name] = {"aspl_best": st.total_best / (n * (n - 1)), "diam_best": st.diam_best,
                     "exact_evals": st.exact_evals, "steps": st.steps, "seconds": st.seconds,
                     "schedule_steps": cfg.total_steps}
        s.close()
    res["learned_aspl_le_plain"] = ((res["learned"]["diam_best"], res["learned"]["aspl_best"])
                                    <= (res["plain"]["diam_best"], res["plain"]["aspl_best"]))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed steps of each search")
    ap.add_argument("--warmup", type=int, default=100,
                    help="untimed steps of each search (the learned search also runs until "
                         "its scorer has trained on %d exact pairs)" % TRAIN_PAIRS)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--d", type=int, default=24)
    ap.add_argument("--K", type=int, default=32)
    ap.add_argument("--P", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the automatic launch shape of the batched step")
    ap.add_argument("--only", choices=["roofline", "copy", "learned", "plain", "large",
                                       "sparse", "batched"],
                    default=None,
                    help="run only the dense-kernel roofline launches, only a device copy of "
                         "known size (the two legs of the rocprofv3 --pmc traffic runs), "
                         "only the learned or the plain search (for kernel profiles of their "
                         "steps), only the large-N evaluations, only incremental steps of K "
                         "random candidates at the learned search's shape launched without "
                         "a graph (the incremental kernel's counter runs), or only the "
                         "K = 256 batched step")
    ap.add_argument("--equal-time", type=float, default=10.0,
                    help="seconds each of plain and learned search from the same graph for "
                         "the equal-wall-time ASPL comparison (0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: run with --gpus 1"}))
        sys.exit(2)
    # launch shapes the autotuner picks are cached per (device, n, d, K)
    os.environ.setdefault("GG_TUNE_CACHE",
                          os.path.join(tempfile.gettempdir(), "gg_mi355x_tune_bench.txt"))
    import graphgolf_mi355x as gg
    assert gg.device_count() >= 1, "no HIP device"
    n, d = args.n, args.d
    if args.only == "copy":
        copy_calibration()
        return
    if args.only == "large":
        print(json.dumps({"large_n": large_n_orbit(gg), "large_n_plain": large_n_plain(gg)}),
              flush=True)
        return
    if args.only == "batched":
        print(json.dumps({"batched_vs_sequential": batched_vs_sequential(gg)}), flush=True)
        return
    adj = gg.random_regular(n, d, args.seed)
    if args.only == "sparse":
        print(json.dumps({"sparse_steps": sparse_steps(gg, adj, args.K)}), flush=True)
        return
    if args.only == "roofline":
        print(json.dumps({"roofline": plain_kernel_roofline(gg, adj)}), flush=True)
        return
    T0, T1 = 1e-3, 1e-6
    K = args.K
    # the learned search trains its scorer on TRAIN_PAIRS exact pairs of its own
    # run before the timed window (the warm-up pairs are 2048 of them), and
    # keeps training (one optimiser step per search step) inside it
    learn_warm = max(args.warmup, -(-TRAIN_PAIRS // K))
    learn_cfg = gg.search_cfg(mode=gg.MODE_LEARNED, K=K, P=args.P,
                              total_steps=learn_warm + args.steps, T0=T0, T1=T1, seed=12345,
                              warmup_pairs=2048, train_every=1, train_batch=256, lr=3e-3)
    tune = None if args.no_tune else tune_step_shape(gg, adj, K)
    if args.only == "learned":
        sl, _, _, el, tl = run_search(gg, adj, learn_cfg, learn_warm, args.steps)
        print(json.dumps({"learned_evals_per_s": el / tl, "ms_per_step": tl / args.steps * 1e3,
                          "step_shape": tune}), flush=True)
        return

    plain_cfg = gg.search_cfg(mode=gg.MODE_PLAIN, total_steps=(args.warmup + args.steps) * K,
                              T0=T0, T1=T1, seed=12345)
    # the plain search times as many exact evaluations as the learned one
    # (steps * K moves, after warmup * K untimed ones)
    if args.only == "plain":
        sp, _, _, ep, tp = run_search(gg, adj, plain_cfg, args.warmup * K, args.steps * K)
        print(json.dumps({"plain_evals_per_s": ep / tp, "ms_per_move": tp / ep * 1e3,
                          "moves_timed": ep}), flush=True)
        return
    sp, _, stp, ep, tp = run_search(gg, adj, plain_cfg, args.warmup * K, args.steps * K)
    plain_rate = ep / tp
    sp.close()

    sl, stl0, stl, el, tl = run_search(gg, adj, learn_cfg, learn_warm, args.steps)
    learn_rate = el / tl
    trained_in_window = stl.train_steps - stl0.train_steps
    # the scorer of the timed search, audited after the window: all P of a
    # fresh proposal draw evaluated exactly (mean over three draws)
    aud = [sl.audit(777 + i)[:2] for i in range(3)]
    sl.close()
    rho = float(np.mean([a[0] for a in aud]))
    miss = float(np.mean([a[1] for a in aud]))
    uniform_miss = 1.0 - K / args.P
    valid = trained_in_window > 0 and stl0.train_steps > 0 and miss < uniform_miss

    out = {
        "metric": "exact moves evaluated/sec (batched+learned search)",
        "value": learn_rate if valid else None, "unit": "exact evaluations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": tl / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic: seeded random regular graph, seeded 2-opt proposals",
        "config": {"workload": f"annealing over 2-opt swaps, random {d}-regular graph n={n}",
                   "n": n, "d": d, "K_exact_per_step": K, "P_proposals_per_step": args.P},
        "learned_warmup_steps": learn_warm,
        "step_shape_autotune": tune,
        "plain": {"value": plain_rate, "unit": "exact evaluations/s",
                  "moves_timed": ep, "ms_per_step": tp / ep * 1e3},
        "speedup_vs_plain": learn_rate / plain_rate,
        "scorer": {"spearman": rho, "top_k_miss_rate": miss, "audits": len(aud),
                   "uniform_top_k_miss_rate": uniform_miss,
                   "train_pairs": stl.train_pairs, "train_steps": stl.train_steps,
                   "train_steps_before_window": stl0.train_steps,
                   "train_steps_in_window": trained_in_window, "train_loss": stl.train_loss},
    }
    if not valid:
        out["invalid_reason"] = ("the scorer did not train inside the timed window or does not "
                                 "beat uniform selection")
    print(json.dumps({"progress": "searches timed", "value": out["value"]}), file=sys.stderr,
          flush=True)
    if not args.no_roofline:
        out["roofline"] = plain_kernel_roofline(gg, adj)
        # the plain path against its own BFS kernel: a move costs at least one
        # launch of it, so this is the plain rate no host overhead could beat
        kms = out["roofline"]["kernel_ms"]
        out["plain"]["kernel_ms"] = kms
        out["plain"]["kernel_bound_value"] = 1e3 / kms
        out["plain"]["ms_per_step_over_kernel_ms"] = out["plain"]["ms_per_step"] / kms
        out["speedup_vs_plain_kernel"] = learn_rate / (1e3 / kms)
    out["roofline_sparse"] = sparse_roofline(n, d, K)
    out["batched_vs_sequential"] = batched_vs_sequential(gg)
    out["large_n"] = large_n_orbit(gg)
    out["large_n_plain"] = large_n_plain(gg)
    print(json.dumps({"progress": "kernels timed"}), file=sys.stderr, flush=True)
    if args.equal_time > 0:
        sec = args.equal_time
        cfgs = (("plain", gg.search_cfg(mode=gg.MODE_PLAIN,
                                        total_steps=max(1, int(sec * plain_rate)),
                                        T0=T0, T1=T1, seed=12345), 200),
                ("learned", gg.search_cfg(mode=gg.MODE_LEARNED, K=K, P=args.P,
                                          total_steps=max(1, int(sec * learn_rate / K)),
                                          T0=T0, T1=T1, seed=12345, warmup_pairs=2048,
                                          train_every=1, train_batch=256, lr=3e-3), 10))
        out["equal_time"] = equal_time(gg, adj, cfgs, sec)
    if not args.no_cpu:
        host = host_cpu()
        port = cpu_baseline(adj)
        ref = cpu_baseline_reference(gg, n, d)
        out["cpu_baseline"] = dict(ref if ref is not None else port, **host)
        if ref is not None:
            out["cpu_baseline_port"] = dict(port, **host)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
