"""Restatement of the learned search's candidate selection and of its audit
figures.  Test infrastructure.

Selection (k_rank / k_select in graphgolf-cuda_amd/csrc/gg_learn_search.hip):
during the warm-up the K candidates are proposals 0..K-1 (the proposals are
uniform draws already); afterwards the G = K - K/4 proposals with the lowest
predicted change of the objective are taken in score order (NaN ranks as
+inf, ties go to the lower proposal index), then K/4 more are drawn uniformly
among the rest from the exploration stream (seed, step, 2^32 - 1) of the
counter-based generator, skipping proposals already chosen.

Audit (gg_search_audit): Spearman's rho = Pearson correlation of the average
ranks of scores and exact deltas; miss rate = 1 - |top-K by score intersect
top-K by exact delta| / K, ties in either top-K broken by index.
"""
import numpy as np

import devrng_ref

EXPLORE_LANE = 0xFFFFFFFF


def ranks_by_count(scores):
    """rank[p] = #{q : s_q < s_p or (s_q == s_p and q < p)}, NaN as +inf: the
    definition the rank kernel evaluates, one predecessor count per proposal."""
    x = np.where(np.isnan(scores), np.inf, np.asarray(scores, dtype=np.float64))
    P = len(x)
    idx = np.arange(P)
    less = x[None, :] < x[:, None]
    tie = (x[None, :] == x[:, None]) & (idx[None, :] < idx[:, None])
    return (less | tie).sum(axis=1)


def greedy_order(scores):
    """Proposal indices in selection order (stable sort, NaN as +inf)."""
    x = np.where(np.isnan(scores), np.inf, np.asarray(scores, dtype=np.float64))
    return np.argsort(x, kind="stable")


def select(scores, K, seed, step, warm):
    """The K chosen proposal indices in evaluation order."""
    P = len(scores)
    if warm:
        return list(range(K))
    G = K - K // 4
    sel = [int(p) for p in greedy_order(scores)[:G]]
    used = set(sel)
    st = devrng_ref.Stream(seed, step, EXPLORE_LANE)
    while len(sel) < K:
        x = int(st.below(P))
        if x not in used:
            used.add(x)
            sel.append(x)
    return sel


def average_ranks(x):
    """Ranks 0..n-1 with ties sharing the mean of their positions."""
    x = np.asarray(x, dtype=np.float64)
    o = np.argsort(x, kind="stable")
    r = np.empty(len(x))
    xs = x[o]
    i = 0
    while i < len(x):
        j = i
        while j + 1 < len(x) and xs[j + 1] == xs[i]:
            j += 1
        r[o[i:j + 1]] = 0.5 * (i + j)
        i = j + 1
    return r


def spearman(sc, dl):
    a, b = average_ranks(sc), average_ranks(dl)
    a, b = a - a.mean(), b - b.mean()
    den = np.sqrt((a * a).sum() * (b * b).sum())
    return float((a * b).sum() / den) if den > 0 else 0.0


def miss_rate(sc, dl, K):
    A = set(np.argsort(np.asarray(sc, dtype=np.float64), kind="stable")[:K].tolist())
    B = np.argsort(np.asarray(dl, dtype=np.float64), kind="stable")[:K].tolist()
    return 1.0 - sum(1 for b in B if b in A) / K


# ---- acceptance of the resident loop ---------------------------------------------------
# k_loop_accept (gg_learn_search.hip): after the scorer's warm-up every step's
# acceptance is decided on the GPU.  The best candidate (connected first, then
# diameter, then distance total) goes through the rule of
# cppcode/annealing.cpp:406-412 (HillClimb.cpp:239 for hill climbing) with the
# uniform drawn from the counter stream (seed, step, 2^32 - 2), 53-bit mantissa.
ACCEPT_LANE = 0xFFFFFFFE


def accept_uniform(seed, step):
    return (devrng_ref.Stream(seed, step, ACCEPT_LANE).next() >> 11) * (1.0 / 9007199254740992.0)


def temperature(T0, T1, total_steps, step):
    """Geometric schedule T(i) = T0 (T1/T0)^(i/total_steps), step 1-based."""
    return T0 * (T1 / T0) ** (step / total_steps)


def accept(rule_hill, N, cur_total, cur_diam, ok, total, diam, step, seed, T0, T1, total_steps):
    """Acceptance of a candidate (ok: connected and within the level limit)
    against the current graph at `step` (1-based)."""
    import math
    NN1 = float(N) * (N - 1)
    fx = cur_total / NN1
    fy = total / NN1 if ok else 1e8
    dy = diam if ok else 100000000
    if rule_hill:
        return ok and (dy < cur_diam or (dy == cur_diam and fy < fx))
    acc = fy < fx or dy < cur_diam
    if not acc:
        T = temperature(T0, T1, total_steps, step)
        acc = accept_uniform(seed, step) < math.exp((fx - fy) * NN1 / (T * NN1))
    if dy > cur_diam or not ok:
        acc = False
    return acc
