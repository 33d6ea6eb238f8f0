"""GPU checks of the learned proposal scorer: MLP forward/gradient against the
FP64 reference (bf16-rounded weights, tolerance for fp32 accumulation), Adam
training, swap features against their set-based definition, and the learned
search (exact acceptance, audit, gain over uniform selection)."""
import os
import sys

import numpy as np
import pytest

import feature_ref
import oracle_lib as ol

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu


def _random_params(rng):
    return (rng.normal(0, 0.25, mo.NP)).astype(np.float32)


def test_mlp_forward_and_grad_match_fp64(gpu):
    gg = gpu
    rng = np.random.default_rng(1)
    m = gg.MLP(seed=3)
    w = _random_params(rng)
    m.set_params(w)
    X = rng.normal(0, 1, (256, 32)).astype(np.float32)
    Y = rng.normal(0, 1, 256).astype(np.float32)
    wq = mo.bf16_round(w)
    out_ref, _ = mo.forward(wq, X.astype(np.float64))
    out = m.forward(X)
    # fp32 accumulation of 32 + 128 terms against fp64
    assert np.allclose(out, out_ref, rtol=1e-4, atol=1e-4 * np.abs(out_ref).max())
    loss, g = m.grad(X, Y)
    loss_ref, g_ref = mo.loss_and_grad(wq, X.astype(np.float64), Y.astype(np.float64))
    assert abs(loss - loss_ref) <= 1e-4 * abs(loss_ref)
    scale = np.abs(g_ref).max()
    assert np.allclose(g, g_ref, rtol=1e-3, atol=1e-4 * scale), np.abs(g - g_ref).max() / scale
    m.close()


def test_mlp_adam_training_reduces_loss(gpu):
    gg = gpu
    rng = np.random.default_rng(2)
    m = gg.MLP(seed=5)
    A = rng.normal(0, 1, 32)
    X = rng.normal(0, 1, (4096, 32)).astype(np.float32)
    Y = (np.tanh(X @ A / 4) + 0.3 * np.maximum(X[:, 0], 0)).astype(np.float32)
    first = None
    for it in range(300):
        idx = rng.integers(0, 4096, 256)
        loss = m.train(X[idx], Y[idx], lr=3e-3)
        first = loss if first is None else first
    final = float(np.mean((m.forward(X) - Y) ** 2))
    assert final < 0.2 * first, (first, final)
    m.close()


def _moves(gg, adj, k, rng):
    out = []
    n, d = adj.shape
    while len(out) < k:
        a, c = (int(x) for x in rng.integers(0, n, 2))
        ia, ic = (int(x) for x in rng.integers(0, d, 2))
        kind = int(rng.integers(0, 2))
        sw, new = gg.make_swap(adj, a, ia, c, ic, kind)
        if sw is not None:
            out.append((a, int(adj[a, ia]), c, int(adj[c, ic]), kind))
    return np.array(out, dtype=np.int32)


# n = 1000: the last column holds padding sources; n = 17000: rows wider than
# the LDS-staged radius-1 rows (the kernel reads the stored levels instead);
# d = 6: degrees off the 16-byte row loads of the all-loads-first pass; d = 32:
# its largest degree (4d = 128 second-layer entries, two per lane)
@pytest.mark.parametrize("n,d", [(1000, 6), (1024, 6), (4096, 12), (17000, 8), (1000, 8),
                                 (2048, 24), (1024, 32)])
def test_features_match_set_definition(gpu, n, d):
    gg = gpu
    rng = np.random.default_rng(n)
    adj = gg.random_regular(n, d, 4)
    mv = _moves(gg, adj, 40, rng)
    with gg.BatchEvaluator(n, d, 8) as be:
        base = be.set_base(adj)
        X = be.features(mv)
    R = feature_ref.features(adj, mv, base.diameter)
    assert np.allclose(X, R, rtol=1e-6, atol=1e-6), np.abs(X - R).max()


def test_learned_search_exact_acceptance_and_gain(gpu):
    gg = gpu
    n, d = 2048, 8
    adj = gg.random_regular(n, d, 6)
    cfg = gg.search_cfg(mode=gg.MODE_LEARNED, K=16, P=256, total_steps=400, T0=1e-4, T1=1e-6,
                        seed=9, warmup_pairs=1024, train_batch=256, lr=2e-3)
    s = gg.Search(adj, cfg, log=400)
    st = s.run(300)
    assert st.exact_evals == 300 * 16 and st.proposals == 300 * 256
    assert st.train_steps > 100
    # accepted moves were decided on exact values: the current graph's oracle
    # total equals the driver's
    cur = s.graph()
    o = ol.adj_eval(cur)
    assert o["total"] == st.total_cur
    rho, miss, sc, dl = s.audit(12345)
    K = 16
    top = np.argsort(sc)[:K]
    # the scorer's top-K is better on the exact objective than a uniform K
    assert dl[top].mean() < dl.mean()
    assert rho > 0.05, rho
    print(f"learned audit: spearman {rho:.3f} miss {miss:.3f} "
          f"top-K mean delta {dl[top].mean():.1f} vs uniform {dl.mean():.1f}")
    s.close()


def test_learned_search_at_n16384_d24_k32(gpu):
    """The learned configuration at full size (random 24-regular graph,
    n = 16384, 1024 proposals scored per step, the top 32 evaluated exactly):
    past the 2048-pair warm-up and 300 steps of online training, the driver's
    current total and diameter equal the oracle's on the final graph, so every
    acceptance was decided on exact values (annealing.cpp:399-412)."""
    import ctypes as C
    gg = gpu
    n, d, K, P = 16384, 24, 32, 1024
    adj = gg.random_regular(n, d, 5)
    cfg = gg.search_cfg(mode=gg.MODE_LEARNED, K=K, P=P, total_steps=400, T0=1e-3, T1=1e-6,
                        seed=12345, warmup_pairs=2048, train_every=1, train_batch=256, lr=3e-3)
    s = gg.Search(adj, cfg)
    warm = -(-2048 // K)
    st0 = s.run(warm)
    st = s.run(300)
    cur = s.graph()
    rho, miss, sc, dl = s.audit(4242)
    s.close()
    assert st.exact_evals == (warm + 300) * K and st.proposals == (warm + 300) * P
    assert st.train_steps - st0.train_steps >= 290, (st0.train_steps, st.train_steps)
    assert st.accepted > 0
    a = np.ascontiguousarray(cur, dtype=np.int32)
    assert a.shape == (n, d)
    srcs = np.arange(n, dtype=np.int32)
    tot, dia = C.c_int64(), C.c_int32()
    rc = ol.oracle().oracle_bfs_omp(n, d, ol._p(a), n, ol._p(srcs), min(16, os.cpu_count() or 1),
                                    C.byref(tot), C.byref(dia))
    assert rc == 0
    assert (tot.value, dia.value) == (st.total_cur, st.diam_cur)
    # the trained scorer ranks proposals better than a uniform choice of K
    assert miss < 1.0 - K / P, miss
    print(f"n=16384 d=24 K=32: {st.accepted} accepted, total {st.total_cur} diam {st.diam_cur}, "
          f"spearman {rho:.3f}, top-32 miss {miss:.3f}")


def test_device_proposals_match_restatement(gpu):
    """The learned search's GPU draws equal the restated streams, before and
    after commits change the base."""
    gg = gpu
    import devrng_ref
    from test_gpu_batch import _random_swaps
    n, d = 512, 6
    adj = gg.random_regular(n, d, 8)
    rng = np.random.default_rng(4)
    with gg.BatchEvaluator(n, d, 16) as be:
        be.set_base(adj)
        for step, seed in ((0, 77), (5, 77), (123456, 2 ** 63 + 9)):
            mv = be.propose(seed, step, 300)
            ref = [devrng_ref.propose(adj, seed, step, p) for p in range(300)]
            assert [tuple(int(x) for x in r) for r in mv] == ref, step
        cands = _random_swaps(gg, adj, 16, rng)
        rs = be.eval([c[0] for c in cands])
        i = min((i for i, r in enumerate(rs) if r.status == gg.GG_OK),
                key=lambda i: rs[i].total_reps)
        be.commit(cands[i][0])
        adj = cands[i][1]
        mv = be.propose(77, 6, 300)
        ref = [devrng_ref.propose(adj, 77, 6, p) for p in range(300)]
        assert [tuple(int(x) for x in r) for r in mv] == ref


def test_learned_searches_back_to_back_are_independent(gpu):
    """Learned searches created one after another in one process (their device
    buffers reuse freed memory) run without stale control words, and a
    repeated configuration repeats its trajectory exactly."""
    gg = gpu
    n, d = 2048, 8
    adj = gg.random_regular(n, d, 6)
    runs = []
    for tb in (256, 512, 256):
        cfg = gg.search_cfg(mode=gg.MODE_LEARNED, K=16, P=256, total_steps=200, T0=1e-4,
                            T1=1e-6, seed=9, warmup_pairs=512, train_batch=tb, lr=2e-3)
        s = gg.Search(adj, cfg)
        st = s.run(80)
        assert st.exact_evals == 80 * 16
        runs.append((st.total_cur, st.total_best, st.accepted, s.graph().tobytes()))
        s.close()
    assert runs[0] == runs[2]


@pytest.mark.parametrize("n,d", [(1000, 6), (4096, 12), (2048, 24)])
def test_fused_features_and_scores(gpu, n, d):
    """The learned search's fused launch (features, then the MLP score in the
    same wave) gives the set-defined features and exactly the scores of the
    standalone forward kernel on them."""
    gg = gpu
    rng = np.random.default_rng(n + 1)
    adj = gg.random_regular(n, d, 4)
    mv = _moves(gg, adj, 70, rng)  # not a multiple of four: a partial workgroup
    m = gg.MLP(seed=5)
    m.set_params(_random_params(rng))
    with gg.BatchEvaluator(n, d, 8) as be:
        base = be.set_base(adj)
        X, sc = be.score(m, mv)
        X2 = be.features(mv)
    assert np.array_equal(X, X2)
    R = feature_ref.features(adj, mv, base.diameter)
    assert np.allclose(X, R, rtol=1e-6, atol=1e-6)
    ref = m.forward(X)
    assert np.array_equal(sc, ref), np.abs(sc - ref).max()
    m.close()


@pytest.mark.parametrize("graph_seed,search_seed", [(6, 21), (7, 22), (8, 23)])
def test_learned_search_beats_uniform_batches_at_equal_evaluations(gpu, graph_seed, search_seed):
    """What the trained scorer gains: hill climbing from the same graph with
    the same number of exact evaluations per step (K = 16), the learned mode
    (top 16 of 256 scored proposals) reaches a lower distance total than the
    batched mode without the scorer (16 uniform proposals), on three graphs
    and search seeds."""
    gg = gpu
    n, d, steps = 2048, 8, 600
    adj = gg.random_regular(n, d, graph_seed)
    common = dict(rule=gg.RULE_HILL, K=16, total_steps=steps, T0=1e-4, T1=1e-6,
                  seed=search_seed)
    s = gg.Search(adj, gg.search_cfg(mode=gg.MODE_BATCHED, P=16, **common))
    su = s.run(steps)
    s.close()
    s = gg.Search(adj, gg.search_cfg(mode=gg.MODE_LEARNED, P=256, warmup_pairs=1024,
                                     train_batch=256, lr=3e-3, **common))
    sl = s.run(steps)
    s.close()
    assert su.exact_evals == sl.exact_evals == steps * 16
    print(f"hill climbing, {steps} steps x 16 exact: uniform total {su.total_best} "
          f"diam {su.diam_best}, learned total {sl.total_best} diam {sl.diam_best}")
    assert (sl.diam_best, sl.total_best) < (su.diam_best, su.total_best)


def test_scorer_learns_when_warmup_sample_holds_a_diameter_change(gpu):
    """The targets are normalised by the median and MAD of the warm-up deltas.
    A delta that changes the diameter carries a (+-1) x N^2 term; on this seeded
    run (n = 16384, d = 24, K = 16) the warm-up sample holds one, and a mean /
    std scale would flatten every other target to ~0 (the scorer then stays
    constant: Spearman 0). With the robust scale the scorer ranks proposals."""
    gg = gpu
    n, d = 16384, 24
    adj = gg.random_regular(n, d, 5)
    cfg = gg.search_cfg(mode=gg.MODE_LEARNED, K=16, P=1024, total_steps=300, T0=1e-3, T1=1e-6,
                        seed=12345, warmup_pairs=2048, train_every=1, train_batch=256, lr=3e-3)
    s = gg.Search(adj, cfg)
    s.run(1000)
    rho, miss, sc, dl = s.audit(777)
    s.close()
    print(f"K=16 after 1000 steps: spearman {rho:.3f}, top-16 miss {miss:.3f}, "
          f"score std {sc.std():.3f}")
    assert sc.std() > 0.1
    assert rho > 0.5


def test_mlp_adam_step_matches_fp64(gpu):
    """Three Adam steps of the GPU optimiser, element by element, against the
    fp64 restatement driven by the GPU's own gradients (the gradient kernel
    is checked against fp64 above; its reduction order is fixed, so the
    gradient the training step uses equals the one read back)."""
    gg = gpu
    rng = np.random.default_rng(6)
    m = gg.MLP(seed=7)
    w = _random_params(rng)
    m.set_params(w)
    ref_w = w.astype(np.float64)
    ref_m = np.zeros(mo.NP)
    ref_v = np.zeros(mo.NP)
    lr = 2e-3
    for t in range(1, 4):
        X = rng.normal(0, 1, (256, 32)).astype(np.float32)
        Y = rng.normal(0, 1, 256).astype(np.float32)
        _, g = m.grad(X, Y)
        m.train(X, Y, lr=lr)
        ref_w, ref_m, ref_v = mo.adam_step(ref_w, ref_m, ref_v, g, t, lr)
        got = m.params()
        # fp32 arithmetic of an update of size <= lr on weights of size ~1
        err = np.abs(got - ref_w)
        assert err.max() < 1e-6, (t, err.max(), int(err.argmax()))
        ref_w = got.astype(np.float64)  # continue from the GPU's fp32 masters
    m.close()
