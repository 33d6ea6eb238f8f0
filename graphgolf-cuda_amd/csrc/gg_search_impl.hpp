// gg_search_impl.hpp -- internal state of a search (host side).
#pragma once
#include <cmath>
#include <cstdint>
#include <memory>
#include <vector>
#include "graphgolf_search.h"
#include "gg_rng.hpp"

struct gg_batch;

// Internal to the library (gg_batch.hip): commit of a candidate whose exact
// result the caller holds; the base re-evaluation is queued, not waited for.
extern "C" int gg_batch_commit_known(gg_batch *b, const gg_swap *sw, const gg_result *known);

namespace gg {

struct Backend {
    virtual ~Backend() {}
    virtual gg_batch *batch() { return nullptr; }
    virtual int set_base(const int32_t *adj, gg_result *out) = 0;
    virtual int eval(const gg_swap *sw, int K, gg_result *out) = 0;
    virtual int commit(const gg_swap *sw, gg_result *out) = 0;
    // candidates evaluated outside eval() since: drop what eval() remembered
    virtual void forget() {}
};

struct MoveRec {
    int a, b, c, d, kind;
};

struct LearnState;  // gg_learn.cpp

bool propose_swap(const std::vector<int32_t> &adj, int N, int d, Rng &rng, gg_swap *sw,
                  MoveRec *mv);
// `count` proposals as `count` redraw loops over propose_swap (same draws,
// same results); false after max_tries attempts.  mv may be null.
bool propose_swaps(const std::vector<int32_t> &adj, int N, int d, Rng &rng, int count,
                   gg_swap *sw, MoveRec *mv, int64_t max_tries);

// Temperature of step i (1-based) on the ASPL scale.
inline double temperature(const gg_search_cfg &cfg, int64_t i) {
    if (cfg.schedule == GG_SCHED_STEP)
        return cfg.T0 * std::pow(0.995, (double)((i - 1) / 10000));
    return cfg.T0 * std::pow(cfg.T1 / cfg.T0, (double)i / cfg.total_steps);
}

inline bool rule_uses_wvc(int rule) { return rule == GG_RULE_ANNEAL_WVC || rule == GG_RULE_HILL_WVC; }

// Acceptance of candidate y against the current solution x at step `step`
// (1-based), cppcode/annealing.cpp:406-412, HillClimb.cpp:239,
// annealingWVC.cpp:412-417, HillClimbWVC.cpp:240.  A failed evaluation
// (unconnected, level limit) stands for the reference's sentinels: diameter
// 100000000 and ASPL 1e8, WVC 100000000.  The uniform is drawn only on the
// branch where the reference draws it.
inline bool accept_move(const gg_search_cfg &cfg, int64_t step, double NN1, const gg_result &x,
                        const gg_result &y, Rng &rng) {
    const double fx = (double)x.total_scaled / NN1;
    const int dx = x.diameter, wx = x.wvc;
    const bool ok = y.status == GG_OK;
    const double fy = ok ? (double)y.total_scaled / NN1 : 1e8;
    const int dy = ok ? y.diameter : 100000000, wy = ok ? y.wvc : 100000000;
    const bool wvc = rule_uses_wvc(cfg.rule);
    const bool lex = wvc ? (dy < dx || (dy == dx && wy < wx) || (dy == dx && wy == wx && fy < fx))
                         : (dy < dx || (dy == dx && fy < fx));
    if (cfg.rule == GG_RULE_HILL || cfg.rule == GG_RULE_HILL_WVC) return ok && lex;
    bool accept = false;
    if (wvc ? lex : (fy < fx || dy < dx)) accept = true;
    else if (rng.uniform() < std::exp((fx - fy) * NN1 / (temperature(cfg, step) * NN1)))
        accept = true;
    if (dy > dx || (wvc && dy == dx && wy > wx) || !ok) accept = false;
    return accept;
}

// Best-solution order: connected first, then (diameter, total), or
// (diameter, WVC, total) under the WVC rules.
inline bool better_by_rule(int rule, const gg_result &x, const gg_result &y) {
    const bool xo = x.status == GG_OK, yo = y.status == GG_OK;
    if (xo != yo) return xo;
    if (x.diameter != y.diameter) return x.diameter < y.diameter;
    if (rule_uses_wvc(rule) && x.wvc != y.wvc) return x.wvc < y.wvc;
    return x.total_scaled < y.total_scaled;
}

// Symmetric `part` tables (gg_orbit.cpp): one move attempt (edge-length change
// or orbit 2-opt) written to `out`; false when it must be redrawn.  lens:
// entries per row (nullptr = d in every row).
struct OrbitMove {
    int kind, a, ia, c, ic, k1, k2;
};
bool propose_orbit(const std::vector<int32_t> &cur, int N, int M, int d, Rng &rng,
                   std::vector<int32_t> &out, OrbitMove *mv, const int32_t *lens = nullptr);
void apply_swap(std::vector<int32_t> &adj, int d, const gg_swap &sw);

}  // namespace gg

struct gg_search {
    int N = 0, d = 0;
    gg_search_cfg cfg{};
    std::unique_ptr<gg::Backend> be;
    std::unique_ptr<gg::Rng> rng;
    std::vector<int32_t> adj;
    // best graph = best_base with trail[0, best_len) applied; while trail_live
    // the trail holds every accepted swap since best_base (so the current
    // graph is best_base with the whole trail applied)
    std::vector<int32_t> best_base;
    std::vector<gg_swap> trail;
    size_t best_len = 0;
    bool trail_live = true;
    gg_result cur{}, best{};
    std::vector<gg_swap> sw;
    std::vector<gg::MoveRec> mv;
    std::vector<gg_result> res;
    int64_t step = 0;
    gg_search_stats stats{};
    int log_cap = 0;
    std::vector<int64_t> log;
    gg::LearnState *learn = nullptr;
};

namespace gg {
int learn_create(gg_search *s, int device);
void learn_destroy(gg_search *s);
int learn_select(gg_search *s, int P, int K);
bool learn_on_device(const gg_search *s);
int learn_eval_device(gg_search *s);
int learn_observe(gg_search *s, int K);
void learn_fill_stats(gg_search *s, gg_search_stats *st);
// the learned search's resident loop (every decision of a step on the GPU)
bool learn_loop_ready(const gg_search *s);
int learn_loop_run(gg_search *s, int steps, int *done);
// best-graph bookkeeping after an accepted swap (s->cur already updated)
void track_best(gg_search *s, const gg_swap &sw);
}  // namespace gg
