"""The incremental step's launch shapes and their autotuner (gg_batch_autotune):
every shape gives the dense path's results bit for bit, K = 256 candidates run
as one launch by default, and the tuned shape is applied, cached per (device,
N, degree, max_k) and picked up by the next evaluator of that key.  Random
subsets of the candidates are checked on the oracle."""
import os

import numpy as np
import pytest

import oracle_lib as ol
from test_gpu_batch import _random_swaps

pytestmark = pytest.mark.gpu


def _key(r):
    return (r.total_reps, r.diameter, r.levels, r.status)


def _dense(gg, n, d, adj, swaps):
    with gg.BatchEvaluator(n, d, max_k=len(swaps)) as be:
        be.set_option(gg.GG_BATCH_OPT_PATH, gg.PATH_DENSE)
        be.set_base(adj)
        return [_key(r) for r in be.eval(swaps)]


@pytest.mark.parametrize("shape", [
    dict(launch_k=7, group_k=3, block=512, staged_keys=16),
    dict(launch_k=40, group_k=5, block=1024, staged_keys=0),
    dict(launch_k=20, group_k=64, block=512, staged_keys=0),
    dict(launch_k=1, group_k=1, block=1024, staged_keys=1),
])
def test_launch_shapes_match_dense(gpu, tmp_path, monkeypatch, shape):
    """Groups with a remainder, several launches, the 512-thread instances
    and Delta keys read from memory instead of LDS all give the dense
    results; the last launch's candidates still commit in place."""
    gg = gpu
    monkeypatch.setenv("GG_TUNE_CACHE", str(tmp_path / "tune.txt"))
    n, d, K = 2048, 8, 40
    rng = np.random.default_rng(5)
    adj = gg.random_regular(n, d, 17)
    cands = _random_swaps(gg, adj, K, rng)
    swaps = [c[0] for c in cands]
    want = _dense(gg, n, d, adj, swaps)
    print("shape", shape, flush=True)
    with gg.BatchEvaluator(n, d, max_k=K) as be:
        be.set_base(adj)
        assert be.get_option(gg.GG_BATCH_INFO_LAUNCH_MAX) == K
        be.set_option(gg.GG_BATCH_OPT_LAUNCH_K, shape["launch_k"])
        be.set_option(gg.GG_BATCH_OPT_GROUP_K, shape["group_k"])
        be.set_option(gg.GG_BATCH_OPT_SPARSE_BLOCK, shape["block"])
        be.set_option(gg.GG_BATCH_OPT_STAGED_KEYS, shape["staged_keys"])
        for _ in range(2):  # capture, then replay
            got = [_key(r) for r in be.eval(swaps)]
            assert be.get_option(gg.GG_BATCH_INFO_LAST_SPARSE) == 1
            assert got == want
        # commit a candidate of the last launch: in place, equal to the oracle
        k = K - 1
        r = be.commit(swaps[k])
        assert be.get_option(gg.GG_BATCH_INFO_LAST_COMMIT_INPLACE) == 1
        o = ol.adj_eval(cands[k][1])
        assert (r.total_reps, r.diameter) == (o["total"], o["diameter"])
    for k in rng.choice(K, 3, replace=False):
        o = ol.adj_eval(cands[k][1])
        assert want[k][0] == o["total"]


def test_out_of_range_shapes_are_refused(gpu):
    gg = gpu
    with gg.BatchEvaluator(512, 4, max_k=8) as be:
        for opt, bad in ((gg.GG_BATCH_OPT_LAUNCH_K, 9), (gg.GG_BATCH_OPT_GROUP_K, 65),
                         (gg.GG_BATCH_OPT_SPARSE_BLOCK, 256), (gg.GG_BATCH_OPT_BASE_BLOCK, 512),
                         (gg.GG_BATCH_OPT_BASE_PUSH, 2),
                         (gg.GG_BATCH_OPT_STAGED_KEYS,
                          be.get_option(gg.GG_BATCH_INFO_STAGED_MAX) + 1)):
            with pytest.raises(gg.GGError):
                be.set_option(opt, bad)


def test_autotune_at_n8192_d16_k256(gpu, tmp_path, monkeypatch):
    """The batched step of a random 16-regular graph on 8192 vertices with
    K = 256 candidates: one launch by default, every tuned configuration
    equal to the dense results, the winner applied and cached."""
    gg = gpu
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("GG_TUNE_CACHE", str(cache))
    n, d, K = 8192, 16, 256
    rng = np.random.default_rng(256)
    adj = gg.random_regular(n, d, 4)
    cands = _random_swaps(gg, adj, K, rng)
    swaps = [c[0] for c in cands]
    want = _dense(gg, n, d, adj, swaps)
    with gg.BatchEvaluator(n, d, max_k=K) as be:
        base = be.set_base(adj)
        assert be.get_option(gg.GG_BATCH_INFO_LAUNCH_MAX) == K  # a single launch
        assert [_key(r) for r in be.eval(swaps)] == want
        digest = be.levels_digest()
        rep = be.autotune(swaps, reps=2)
        info = rep.as_dict()
        print("tune n=8192 d=16 K=256:", info["best"], "default", info["default_ms"],
              "base", info["best_base"], flush=True)
        assert rep.mismatches == 0
        assert rep.count >= 6 and rep.base_count == 4
        assert rep.best_ms <= rep.default_ms
        shape = be.shape()
        assert (shape["launch_k"], shape["group_k"], shape["block"], shape["staged_keys"]) == \
            (rep.best_launch_k, rep.best_group_k, rep.best_block, rep.best_staged_keys)
        # the base is unchanged and the tuned shape gives the same results
        assert be.levels_digest() == digest
        assert [_key(r) for r in be.eval(swaps)] == want
    assert cache.exists() and f"batch " in cache.read_text()
    with gg.BatchEvaluator(n, d, max_k=K) as be2:
        assert be2.shape() == shape
        assert _key(be2.set_base(adj)) == _key(base)
        assert [_key(r) for r in be2.eval(swaps)] == want
    for k in rng.choice(K, 4, replace=False):
        o = ol.adj_eval(cands[k][1])
        assert want[k][:2] == (o["total"], o["diameter"])
