// gg_search.cpp -- search drivers (host C++): simulated annealing and hill
// climbing over 2-opt edge swaps of a regular graph, on the exact evaluator.
//
// Reference drivers: cppcode/annealing.cpp:396-443 (acceptance + schedule),
// cppcode/HillClimb.cpp:222-249 (hill climbing), createNeighbour
// annealing.cpp:112-197 (2-opt).  Differences, fixed here on purpose:
//   * proposals follow the 2-opt of the configs: pick edges (a,b),(c,d)
//     uniformly (vertex then slot), rewire to (a,c),(b,d) or (a,d),(b,c) by a
//     coin, and redraw when the result is not simple (the reference always
//     takes (a,d),(b,c) and never rejects loops or multi-edges);
//   * every draw goes through gg::Rng (mt19937_64 + explicit bounded/real
//     mappings), so a GPU run and a CPU run of this driver draw the same moves;
//   * the temperature is geometric per step, T(i) = T0 (T1/T0)^(i/steps), on
//     the ASPL scale; the reference's exp((fx-fy) N(N-1)/temp) takes
//     temp = T(i) N(N-1).
//
// Evaluation modes:
//   PLAIN    one candidate per step, one exact launch (gg_eval_adj_swap)
//   BATCHED  K candidates per step in one launch (gg_batch_eval); the best by
//            (diameter, total) goes through the acceptance rule
//   LEARNED  P proposals scored by the MLP; the top K by score are evaluated
//            exactly as in BATCHED (see gg_learn.cpp)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>
#include "graphgolf_mi355x.h"
#include "gg_rng.hpp"
#include "graphgolf_search.h"
#include "gg_search_impl.hpp"

namespace gg {

// ---- evaluator backends ------------------------------------------------------------

struct PlainBackend : Backend {
    gg_eval *ev = nullptr;
    std::vector<gg_swap> last_sw;     // candidates of the last eval and their results:
    std::vector<gg_result> last_res;  // committing one of them needs no new evaluation
    // One exact evaluation per move as one BFS kernel launch that finalizes
    // in its last workgroup (swap read from and result written to pinned
    // memory, the result's sequence number polled: no copies, no events, no
    // stream wait), without the per-source eccentricities and level counts
    // the search never reads; WVC (and the level limit of WVCdiameterASPL)
    // under the WVC rules, which add a finalize kernel and a stream wait.
    // Commits are queued, not waited for.
    PlainBackend(int N, int d, int device, bool wvc, int *st) {
        *st = gg_create_adj(N, d, device, &ev);
        if (*st == GG_OK) *st = gg_eval_set_option(ev, GG_OPT_SWAP_GRAPH, 2);
        if (*st == GG_OK) *st = gg_eval_set_option(ev, GG_OPT_LEVELS, 0);
        if (*st == GG_OK && wvc) *st = gg_eval_set_option(ev, GG_OPT_WVC, 1);
        if (*st == GG_OK && wvc) *st = gg_eval_set_option(ev, GG_OPT_LEVEL_LIMIT, 100);
    }
    ~PlainBackend() override { gg_destroy(ev); }
    int set_base(const int32_t *adj, gg_result *out) override {
        last_sw.clear();
        return gg_eval_adj(ev, adj, out);
    }
    int eval(const gg_swap *sw, int K, gg_result *out) override {
        last_sw.assign(sw, sw + K);
        last_res.resize(K);
        for (int k = 0; k < K; k++) {
            int st = gg_eval_adj_swap(ev, &sw[k], &out[k]);
            if (st != GG_OK) { last_sw.clear(); return st; }
            last_res[k] = out[k];
        }
        return GG_OK;
    }
    int commit(const gg_swap *sw, gg_result *out) override {
        int st = gg_eval_adj_commit(ev, sw);
        if (st != GG_OK) return st;
        // the accepted graph is one just evaluated (annealing.cpp:426-428 keeps its
        // result rather than evaluating again)
        for (size_t k = 0; k < last_sw.size(); k++)
            if (memcmp(&last_sw[k], sw, sizeof(gg_swap)) == 0) {
                *out = last_res[k];
                last_sw.clear();
                return GG_OK;
            }
        last_sw.clear();
        gg_swap none;
        for (int j = 0; j < 4; j++) { none.v[j] = -1; none.slot[j] = 0; none.nbr[j] = 0; }
        return gg_eval_adj_swap(ev, &none, out);
    }
};

struct BatchBackend : Backend {
    gg_batch *b = nullptr;
    BatchBackend(int N, int d, int K, int device, int *st) {
        *st = gg_batch_create(N, d, K, device, &b);
    }
    ~BatchBackend() override { gg_batch_destroy(b); }
    gg_batch *batch() override { return b; }
    int set_base(const int32_t *adj, gg_result *out) override {
        return gg_batch_set_base(b, adj, out);
    }
    std::vector<gg_swap> last_sw;     // candidates of the last eval and their results
    std::vector<gg_result> last_res;
    int eval(const gg_swap *sw, int K, gg_result *out) override {
        int st = gg_batch_eval(b, sw, K, out);
        if (st == GG_OK) { last_sw.assign(sw, sw + K); last_res.assign(out, out + K); }
        else last_sw.clear();
        return st;
    }
    int commit(const gg_swap *sw, gg_result *out) override {
        // an evaluated candidate: its result is the new base's, and the
        // re-evaluation that refreshes the stored levels runs unwaited
        for (size_t k = 0; k < last_sw.size(); k++)
            if (memcmp(&last_sw[k], sw, sizeof(gg_swap)) == 0) {
                int st = gg_batch_commit_known(b, sw, &last_res[k]);
                if (st == GG_OK) *out = last_res[k];
                last_sw.clear();
                return st;
            }
        last_sw.clear();
        return gg_batch_commit(b, sw, out);
    }
    void forget() override { last_sw.clear(); }
};

struct CallbackBackend : Backend {
    gg_eval_callbacks cb;
    explicit CallbackBackend(const gg_eval_callbacks &c) : cb(c) {}
    int set_base(const int32_t *adj, gg_result *out) override {
        return cb.set_base(cb.ctx, adj, out);
    }
    int eval(const gg_swap *sw, int K, gg_result *out) override {
        return cb.eval(cb.ctx, sw, K, out);
    }
    int commit(const gg_swap *sw, gg_result *out) override { return cb.commit(cb.ctx, sw, out); }
};

// ---- moves ----------------------------------------------------------------------------

static bool row_has(const std::vector<int32_t> &adj, int d, int v, int x) {
    const int32_t *r = &adj[(size_t)v * d];
    for (int i = 0; i < d; i++)
        if (r[i] == x) return true;
    return false;
}

static int row_slot(const std::vector<int32_t> &adj, int d, int v, int x) {
    const int32_t *r = &adj[(size_t)v * d];
    for (int i = 0; i < d; i++)
        if (r[i] == x) return i;
    return -1;
}

// One proposal attempt: draws a, ia, c, ic, kind (5 draws).  Returns false
// when the rewired graph would not be simple.
bool propose_swap(const std::vector<int32_t> &adj, int N, int d, Rng &rng, gg_swap *sw,
                  MoveRec *mv) {
    int a = (int)rng.below(N), ia = (int)rng.below(d);
    int c = (int)rng.below(N), ic = (int)rng.below(d);
    int kind = (int)rng.below(2);
    int b = adj[(size_t)a * d + ia], dd = adj[(size_t)c * d + ic];
    if (a == b || a == c || a == dd || b == c || b == dd || c == dd) return false;
    // kind 1: (a,b),(c,dd) -> (a,dd),(b,c);  kind 0: -> (a,c),(b,dd)
    int xa = kind ? dd : c;   // new neighbour of a
    int xb = kind ? c : dd;   // new neighbour of b
    int xc = kind ? b : a;    // new neighbour of c
    int xd = kind ? a : b;    // new neighbour of dd
    if (row_has(adj, d, a, xa) || row_has(adj, d, b, xb)) return false;
    int ib = row_slot(adj, d, b, a), id = row_slot(adj, d, dd, c);
    if (ib < 0 || id < 0) return false;
    sw->v[0] = a;  sw->slot[0] = ia; sw->nbr[0] = xa;
    sw->v[1] = b;  sw->slot[1] = ib; sw->nbr[1] = xb;
    sw->v[2] = c;  sw->slot[2] = ic; sw->nbr[2] = xc;
    sw->v[3] = dd; sw->slot[3] = id; sw->nbr[3] = xd;
    if (mv) { mv->a = a; mv->b = b; mv->c = c; mv->d = dd; mv->kind = kind; }
    return true;
}

// `count` proposals, identical to `count` calls of the loop
//   while (!propose_swap(...)) {}
// An attempt's five draws do not depend on the graph, so the attempts of a
// round (as many as proposals still missing, never more than the loop would
// make) are drawn first; then their rows are prefetched and the attempts are
// checked in draw order.  Returns false after max_tries attempts without
// filling all proposals.
bool propose_swaps(const std::vector<int32_t> &adj, int N, int d, Rng &rng, int count,
                   gg_swap *sw, MoveRec *mv, int64_t max_tries) {
    struct Att {
        int a, ia, c, ic, kind;
    };
    std::vector<Att> att((size_t)count);
    const int32_t *A = adj.data();
    int filled = 0;
    int64_t tries = 0;
    while (filled < count) {
        const int need = count - filled;
        for (int j = 0; j < need; j++) {
            Att &t = att[j];
            t.a = (int)rng.below(N);
            t.ia = (int)rng.below(d);
            t.c = (int)rng.below(N);
            t.ic = (int)rng.below(d);
            t.kind = (int)rng.below(2);
            __builtin_prefetch(A + (size_t)t.a * d + t.ia);
            __builtin_prefetch(A + (size_t)t.c * d + t.ic);
        }
        for (int j = 0; j < need; j++) {  // rows of b and dd
            const Att &t = att[j];
            const int32_t *rb = A + (size_t)A[(size_t)t.a * d + t.ia] * d;
            const int32_t *rd = A + (size_t)A[(size_t)t.c * d + t.ic] * d;
            __builtin_prefetch(rb);
            __builtin_prefetch(rb + d - 1);
            __builtin_prefetch(rd);
            __builtin_prefetch(rd + d - 1);
            __builtin_prefetch(A + (size_t)t.a * d);
            __builtin_prefetch(A + (size_t)t.a * d + d - 1);
        }
        for (int j = 0; j < need; j++) {
            const Att &t = att[j];
            tries++;
            const int a = t.a, c = t.c, kind = t.kind;
            const int b = A[(size_t)a * d + t.ia], dd = A[(size_t)c * d + t.ic];
            if (a == b || a == c || a == dd || b == c || b == dd || c == dd) continue;
            const int xa = kind ? dd : c, xb = kind ? c : dd, xc = kind ? b : a, xd = kind ? a : b;
            if (row_has(adj, d, a, xa) || row_has(adj, d, b, xb)) continue;
            const int ib = row_slot(adj, d, b, a), id = row_slot(adj, d, dd, c);
            if (ib < 0 || id < 0) continue;
            gg_swap &o = sw[filled];
            o.v[0] = a;  o.slot[0] = t.ia; o.nbr[0] = xa;
            o.v[1] = b;  o.slot[1] = ib;   o.nbr[1] = xb;
            o.v[2] = c;  o.slot[2] = t.ic; o.nbr[2] = xc;
            o.v[3] = dd; o.slot[3] = id;   o.nbr[3] = xd;
            if (mv) { mv[filled].a = a; mv[filled].b = b; mv[filled].c = c; mv[filled].d = dd; mv[filled].kind = kind; }
            filled++;
        }
        if (filled < count && tries > max_tries) return false;
    }
    return true;
}

void apply_swap(std::vector<int32_t> &adj, int d, const gg_swap &sw) {
    for (int j = 0; j < 4; j++)
        if (sw.v[j] >= 0) adj[(size_t)sw.v[j] * d + sw.slot[j]] = sw.nbr[j];
}


}  // namespace gg

using namespace gg;

// ---- C ABI ------------------------------------------------------------------------------

static int search_init(gg_search *s, int N, int d, const int32_t *adj, const gg_search_cfg *cfg) {
    s->N = N; s->d = d; s->cfg = *cfg;
    s->adj.assign(adj, adj + (size_t)N * d);
    s->rng.reset(new Rng(cfg->seed));
    gg_result base;
    int st = s->be->set_base(s->adj.data(), &base);
    if (st != GG_OK) return st;
    s->cur = base;
    s->best = base;
    s->best_base = s->adj;
    s->trail.clear();
    s->best_len = 0;
    s->trail_live = true;
    s->sw.resize(std::max(1, std::max(cfg->K, cfg->P)));
    s->mv.resize(s->sw.size());
    s->res.resize(s->sw.size());
    return GG_OK;
}

static int validate_cfg(int N, int d, const int32_t *adj, const gg_search_cfg *cfg) {
    if (!cfg || !adj || N < 5 || d < 2) return GG_BAD_ARG;
    if (cfg->mode < GG_MODE_PLAIN || cfg->mode > GG_MODE_LEARNED) return GG_BAD_ARG;
    if (cfg->rule < GG_RULE_ANNEAL || cfg->rule > GG_RULE_HILL_WVC) return GG_BAD_ARG;
    // WVC is computed by the single-candidate evaluator only
    if (rule_uses_wvc(cfg->rule) && cfg->mode != GG_MODE_PLAIN) return GG_UNSUPPORTED;
    if (cfg->mode != GG_MODE_PLAIN && cfg->K < 1) return GG_BAD_ARG;
    if (cfg->mode == GG_MODE_LEARNED && cfg->P < cfg->K) return GG_BAD_ARG;
    if ((cfg->rule == GG_RULE_ANNEAL || cfg->rule == GG_RULE_ANNEAL_WVC) &&
        (!(cfg->T0 > 0) || !(cfg->T1 > 0)))
        return GG_BAD_ARG;
    if (cfg->schedule != GG_SCHED_GEOMETRIC && cfg->schedule != GG_SCHED_STEP) return GG_BAD_ARG;
    if (cfg->total_steps < 1) return GG_BAD_ARG;
    return GG_OK;
}

extern "C" int gg_search_create(int N, int d, const int32_t *adj, const gg_search_cfg *cfg,
                                int device, gg_search **out) {
    if (!out) return GG_BAD_ARG;
    *out = nullptr;
    int st = validate_cfg(N, d, adj, cfg);
    if (st != GG_OK) return st;
    gg_search *s = new gg_search();
    if (cfg->mode == GG_MODE_PLAIN)
        s->be.reset(new PlainBackend(N, d, device, rule_uses_wvc(cfg->rule), &st));
    else if (cfg->mode == GG_MODE_BATCHED)
        s->be.reset(new BatchBackend(N, d, std::max(cfg->K, 1), device, &st));
    else  // learned: room for audits that evaluate all P proposals
        s->be.reset(new BatchBackend(N, d, std::max(cfg->K, cfg->P), device, &st));
    if (st == GG_OK) st = search_init(s, N, d, adj, cfg);
    if (st == GG_OK && cfg->mode == GG_MODE_LEARNED) st = learn_create(s, device);
    if (st != GG_OK) { gg_search_destroy(s); return st; }
    *out = s;
    return GG_OK;
}

extern "C" int gg_search_create_cb(int N, int d, const int32_t *adj, const gg_search_cfg *cfg,
                                   const gg_eval_callbacks *cb, gg_search **out) {
    if (!out || !cb || !cb->set_base || !cb->eval || !cb->commit) return GG_BAD_ARG;
    *out = nullptr;
    int st = validate_cfg(N, d, adj, cfg);
    if (st != GG_OK) return st;
    if (cfg->mode == GG_MODE_LEARNED) return GG_UNSUPPORTED;
    gg_search *s = new gg_search();
    s->be.reset(new CallbackBackend(*cb));
    st = search_init(s, N, d, adj, cfg);
    if (st != GG_OK) { gg_search_destroy(s); return st; }
    *out = s;
    return GG_OK;
}

extern "C" void gg_search_destroy(gg_search *s) {
    if (!s) return;
    learn_destroy(s);
    delete s;
}

// Best-graph bookkeeping after an accepted swap: O(1) per step instead of a
// copy of the graph per improvement.  The trail is folded into best_base when
// it outgrows kTrailCap; when most of it lies past the best it is dropped, and
// the next improvement copies the graph.
void gg::track_best(gg_search *s, const gg_swap &sw) {
    constexpr size_t kTrailCap = 1 << 16;
    if (s->trail_live) s->trail.push_back(sw);
    if (better_by_rule(s->cfg.rule, s->cur, s->best)) {
        s->best = s->cur;
        if (s->trail_live) {
            s->best_len = s->trail.size();
        } else {
            s->best_base = s->adj;
            s->trail.clear();
            s->best_len = 0;
            s->trail_live = true;
        }
    }
    if (s->trail.size() > kTrailCap) {
        for (size_t i = 0; i < s->best_len; i++) apply_swap(s->best_base, s->d, s->trail[i]);
        if (s->trail.size() - s->best_len > kTrailCap / 2) {
            s->trail.clear();
            s->trail_live = false;
        } else {
            s->trail.erase(s->trail.begin(), s->trail.begin() + (ptrdiff_t)s->best_len);
        }
        s->best_len = 0;
    }
}

// One step: propose, evaluate exactly, accept or reject.
static int search_step(gg_search *s) {
    const gg_search_cfg &cfg = s->cfg;
    const int N = s->N, d = s->d;
    Rng &rng = *s->rng;
    int K = cfg.mode == GG_MODE_PLAIN ? 1 : cfg.K;
    int st;
    int nprop = cfg.mode == GG_MODE_LEARNED ? cfg.P : K;
    if (cfg.mode == GG_MODE_LEARNED && learn_on_device(s)) {
        // drawn, scored, selected and evaluated on the GPU in one graph; a
        // commit finds the candidates through the batch evaluator itself
        s->be->forget();
        st = learn_eval_device(s);
        if (st != GG_OK) return st;
        s->stats.proposals += nprop;
    } else {
        if (!propose_swaps(s->adj, N, d, rng, nprop, s->sw.data(), s->mv.data(),
                           1000000 + 4 * (int64_t)nprop))
            return GG_BAD_ARG;  // no simple swap exists
        s->stats.proposals += nprop;
        if (cfg.mode == GG_MODE_LEARNED) {
            st = learn_select(s, nprop, K);  // reorders sw/mv so the first K are the chosen ones
            if (st != GG_OK) return st;
        }
        st = s->be->eval(s->sw.data(), K, s->res.data());
        if (st != GG_OK) return st;
    }
    s->stats.exact_evals += K;
    if (cfg.mode == GG_MODE_LEARNED) {
        st = learn_observe(s, K);
        if (st != GG_OK) return st;
    }
    int bi = 0;
    for (int k = 1; k < K; k++)
        if (better_by_rule(cfg.rule, s->res[k], s->res[bi])) bi = k;
    const gg_result &y = s->res[bi];
    const bool ok = y.status == GG_OK;
    s->step++;
    const bool accept = accept_move(cfg, s->step, (double)N * (N - 1), s->cur, y, rng);
    if (s->log_cap > 0 && (int64_t)s->log.size() < s->log_cap * 10) {
        const MoveRec &m = s->mv[bi];
        // failed candidates log total = diameter = -1 (their partial values
        // depend on the evaluator)
        int64_t rec[10] = {s->step, m.a, m.b, m.c, m.d, m.kind, ok ? y.total_scaled : -1,
                           ok ? y.diameter : -1, y.status, accept ? 1 : 0};
        s->log.insert(s->log.end(), rec, rec + 10);
    }
    if (accept) {
        gg_result base;
        st = s->be->commit(&s->sw[bi], &base);
        if (st != GG_OK) return st;
        apply_swap(s->adj, d, s->sw[bi]);
        s->cur = base;
        s->stats.accepted++;
        track_best(s, s->sw[bi]);
    }
    return GG_OK;
}

extern "C" int gg_search_run(gg_search *s, int steps, gg_search_stats *out) {
    if (!s || steps < 0) return GG_BAD_ARG;
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < steps;) {
        if (learn_loop_ready(s)) {
            // after the scorer's warm-up: steps replayed on the GPU in chunks
            int done = 0;
            int st = learn_loop_run(s, steps - i, &done);
            if (st != GG_OK) return st;
            i += done;
            continue;
        }
        int st = search_step(s);
        if (st != GG_OK) return st;
        i++;
    }
    auto t1 = std::chrono::steady_clock::now();
    s->stats.seconds += std::chrono::duration<double>(t1 - t0).count();
    s->stats.steps = s->step;
    s->stats.total_cur = s->cur.total_scaled;
    s->stats.diam_cur = s->cur.diameter;
    s->stats.total_best = s->best.total_scaled;
    s->stats.diam_best = s->best.diameter;
    learn_fill_stats(s, &s->stats);
    if (out) *out = s->stats;
    return GG_OK;
}

extern "C" int gg_search_graph(gg_search *s, int32_t *adj_out, int best) {
    if (!s || !adj_out) return GG_BAD_ARG;
    if (!best) {
        memcpy(adj_out, s->adj.data(), sizeof(int32_t) * s->adj.size());
        return GG_OK;
    }
    memcpy(adj_out, s->best_base.data(), sizeof(int32_t) * s->best_base.size());
    for (size_t i = 0; i < s->best_len; i++) {
        const gg_swap &sw = s->trail[i];
        for (int j = 0; j < 4; j++)
            if (sw.v[j] >= 0) adj_out[(size_t)sw.v[j] * s->d + sw.slot[j]] = sw.nbr[j];
    }
    return GG_OK;
}

extern "C" int gg_search_set_log(gg_search *s, int capacity) {
    if (!s || capacity < 0) return GG_BAD_ARG;
    s->log_cap = capacity;
    s->log.clear();
    s->log.reserve((size_t)capacity * 10);
    return GG_OK;
}

extern "C" int gg_search_get_log(gg_search *s, int64_t *buf, int capacity, int *count) {
    if (!s || !count) return GG_BAD_ARG;
    int n = (int)(s->log.size() / 10);
    *count = n;
    if (buf) memcpy(buf, s->log.data(), sizeof(int64_t) * 10 * std::min(n, capacity));
    return GG_OK;
}
