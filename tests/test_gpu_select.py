"""GPU checks of the learned search's candidate selection and audit against
their restatements (select_ref.py, devrng_ref.py), and of the learned search
at the n = 16384, d = 24 shape (K = 32 of P = 1024) past its warm-up, its
final graph checked on the oracle."""
import numpy as np
import pytest

import devrng_ref
import oracle_eval
import select_ref as sr

pytestmark = pytest.mark.gpu


def _check_step(s, seed, adj_before, K, P, check_moves=True):
    step, warm, sc, mv, sel = s.last_selection()
    if check_moves:
        ref = [devrng_ref.propose(adj_before, seed, step, p) for p in range(P)]
        assert [tuple(int(x) for x in r) for r in mv] == ref, step
    want = sr.select(sc, K, seed, step, warm)
    assert sel.tolist() == want, (step, warm)
    return step, warm, sc


# K = 32 of P = 40: the 8 exploration picks come from the 16 proposals ranked
# 24 and below, so a step often needs more than one batch of 32 draws
@pytest.mark.parametrize("K,P,warm_steps,greedy", [(16, 256, 32, 16), (32, 40, 8, 32)])
def test_selection_matches_restated_rule(gpu, K, P, warm_steps, greedy):
    gg = gpu
    n, d, seed = 2048, 8, 9
    adj = gg.random_regular(n, d, 6)
    cfg = gg.search_cfg(mode=gg.MODE_LEARNED, K=K, P=P, total_steps=200, T0=1e-4, T1=1e-6,
                        seed=seed, warmup_pairs=warm_steps * K, train_batch=256, lr=2e-3)
    s = gg.Search(adj, cfg)
    warm_seen, greedy_seen = 0, 0
    for t in range(warm_steps + greedy):  # the warm-up steps, then selection by score
        before = s.graph()
        s.run(1)
        step, warm, sc = _check_step(s, seed, before, K, P, check_moves=(t % 4 == 0))
        assert step == t
        warm_seen += warm
        greedy_seen += 1 - warm
        assert np.isfinite(sc).all()
    assert warm_seen == warm_steps and greedy_seen == greedy
    s.close()


def test_audit_figures_recomputed(gpu):
    gg = gpu
    n, d, K, P = 2048, 8, 16, 256
    adj = gg.random_regular(n, d, 6)
    cfg = gg.search_cfg(mode=gg.MODE_LEARNED, K=K, P=P, total_steps=400, T0=1e-4, T1=1e-6,
                        seed=9, warmup_pairs=1024, train_batch=256, lr=2e-3)
    s = gg.Search(adj, cfg)
    s.run(150)
    for seed in (1, 2):
        rho, miss, sc, dl = s.audit(seed)
        assert abs(rho - sr.spearman(sc, dl)) < 1e-9, (rho, sr.spearman(sc, dl))
        assert miss == sr.miss_rate(sc, dl, K)
    s.close()


def test_learned_search_at_n16384_d24(gpu):
    """Learned search at its benchmark shape (K = 32 of P = 1024, 2048 warm-up
    pairs = 64 steps) for 300 steps: training runs in every step after the
    warm-up, one step's selection follows the restated rule, and the driver's
    current and best totals and diameters are the oracle's on its graphs."""
    gg = gpu
    n, d, K, P, seed = 16384, 24, 32, 1024, 12345
    adj = gg.random_regular(n, d, 5)
    cfg = gg.search_cfg(mode=gg.MODE_LEARNED, K=K, P=P, total_steps=300, T0=1e-3, T1=1e-6,
                        seed=seed, warmup_pairs=2048, train_every=1, train_batch=256, lr=3e-3)
    s = gg.Search(adj, cfg)
    st1 = s.run(99)
    before = s.graph()
    s.run(1)
    step, warm, _ = _check_step(s, seed, before, K, P)
    assert step == 99 and warm == 0
    st = s.run(200)
    assert st.exact_evals == 300 * K and st.proposals == 300 * P
    assert st1.train_steps > 0 and st.train_steps - st1.train_steps == 201
    cur = oracle_eval.evaluate(s.graph())
    assert (cur[0], cur[1], cur[2]) == (st.total_cur, st.diam_cur, 0)
    best = oracle_eval.evaluate(s.graph(best=True))
    assert (best[0], best[1]) == (st.total_best, st.diam_best)
    assert st.accepted > 0
    rho, miss, sc, dl = s.audit(777)
    assert abs(rho - sr.spearman(sc, dl)) < 1e-9 and miss == sr.miss_rate(sc, dl, K)
    print(f"n=16384 d=24 learned, 300 steps: accepted {st.accepted}, total {st.total_cur}, "
          f"train steps {st.train_steps}, audit spearman {rho:.3f} miss {miss:.3f}")
    s.close()


@pytest.mark.parametrize("hill", [False, True])
def test_resident_loop_acceptance_matches_restated_rule(gpu, hill):
    """After the warm-up the learned search decides acceptance on the GPU: every
    logged step's decision equals the restated rule applied to the logged best
    candidate, the current graph following the accepted ones, and the final
    current graph's oracle total equals the driver's."""
    gg = gpu
    # few candidates per step, so some steps hold no improving one and the
    # annealing rule draws its uniform
    n, d, K, P, seed = 2048, 8, 4, 16, 31
    T0, T1, steps = 5e-3, 1e-5, 400
    adj = gg.random_regular(n, d, 6)
    cfg = gg.search_cfg(mode=gg.MODE_LEARNED, rule=gg.RULE_HILL if hill else gg.RULE_ANNEAL,
                        K=K, P=P, total_steps=steps, T0=T0, T1=T1, seed=seed,
                        warmup_pairs=64, train_batch=64, lr=2e-3)
    s = gg.Search(adj, cfg)
    s.run(40)  # past the warm-up (64 pairs = 16 steps): the loop runs from here on
    st0 = s.run(0)
    s.set_log(1000)
    st = s.run(300)
    log = s.log()
    assert len(log) == 300 and int(log[0, 0]) == st0.steps + 1
    cur_total, cur_diam = st0.total_cur, st0.diam_cur
    uphill = drawn = 0
    for row in log:
        step, total, diam, status, acc = (int(row[0]), int(row[6]), int(row[7]), int(row[8]),
                                          int(row[9]))
        ok = status == 0
        want = sr.accept(hill, n, cur_total, cur_diam, ok, total, diam, step, seed, T0, T1,
                         steps)
        assert bool(acc) == want, (step, total, diam, cur_total, cur_diam)
        drawn += ok and (diam, total) >= (cur_diam, cur_total)
        if acc:
            uphill += (diam, total) > (cur_diam, cur_total)
            cur_total, cur_diam = total, diam
    assert (cur_total, cur_diam) == (st.total_cur, st.diam_cur)
    assert drawn > 0
    if not hill:
        assert uphill > 0  # the uniform branch accepted
    o = oracle_eval.evaluate(s.graph())
    assert (o[0], o[1]) == (st.total_cur, st.diam_cur)
    s.close()


def _apply_move(G, m):
    """The 2-opt rewiring of move (a, b, c, d, kind) applied in place, each
    removed neighbour's slot taking the new one (the rows as the swap
    rewrites them)."""
    a, b, c, d, kind = (int(x) for x in m)
    xa, xb, xc, xd = (d, c, b, a) if kind else (c, d, a, b)
    for v, old, new in ((a, b, xa), (b, a, xb), (c, d, xc), (d, c, xd)):
        G[v][list(G[v]).index(old)] = new


# d = 8 and 12: the four-lane draw beside the commit (two and three row
# quads); d = 6: degrees it does not take, drawn after the commit instead
@pytest.mark.parametrize("n,d", [(2048, 8), (1536, 12), (1024, 6)])
def test_resident_loop_draws_on_the_committed_graph(gpu, n, d):
    """In the resident loop every step but a chunk's first draws its
    proposals in the commit launch of the step before, on the adjacency with
    that step's swap applied while the commit writes it.  Each chunk's last
    step's draw equals the restated draw on the graph after the step before
    it, and the driver's graph after 100 such steps is the initial graph with
    the logged accepted moves applied: symmetric, every swap written at the
    right slots; its oracle total is the driver's."""
    gg = gpu
    K, P, seed = 16, 256, 22
    adj0 = np.array(gg.random_regular(n, d, 7))
    cfg = gg.search_cfg(mode=gg.MODE_LEARNED, rule=gg.RULE_HILL, K=K, P=P, total_steps=600,
                        T0=1e-4, T1=1e-6, seed=seed, warmup_pairs=1024, train_batch=256,
                        lr=3e-3)
    s = gg.Search(adj0, cfg)
    s.set_log(1000)
    s.run(70)  # 64 warm-up steps, then the loop
    for _ in range(20):
        G = np.array(s.graph())
        s.run(5)
        for row in s.log()[-5:-1]:
            if int(row[9]):
                _apply_move(G, row[1:6])
        step, warm, sc, mv, sel = s.last_selection()
        ref = [devrng_ref.propose(G, seed, step, p) for p in range(P)]
        assert [tuple(int(x) for x in r) for r in mv] == ref, step
    st = s.run(0)
    G = adj0.copy()
    for row in s.log():
        if int(row[9]):
            _apply_move(G, row[1:6])
    H = np.array(s.graph())
    assert np.array_equal(H, G)
    assert all(v in H[u] for v in range(n) for u in H[v])
    o = oracle_eval.evaluate(H)
    assert (o[0], o[1]) == (st.total_cur, st.diam_cur)
    s.close()
