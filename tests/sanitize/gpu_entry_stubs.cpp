// gpu_entry_stubs.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Stand-ins for the library's GPU entry points, so the host-side C++ (search
// drivers, move generators, instance generators) can be built on its own with
// g++ under AddressSanitizer / UndefinedBehaviorSanitizer.  The sanitized
// driver only takes the caller-evaluator paths (*_create_cb), which never
// reach these; any call that does gets GG_UNSUPPORTED.
#include <cstdint>
#include "graphgolf_mi355x.h"
#include "graphgolf_search.h"
#include "gg_search_impl.hpp"

extern "C" {
int gg_create_part(int, int, int, int, gg_eval **out) { if (out) *out = nullptr; return GG_UNSUPPORTED; }
int gg_create_piece(int, int, int, int, int, int, gg_eval **out) { if (out) *out = nullptr; return GG_UNSUPPORTED; }
int gg_create_box(int, int, int, int, int, int, int, int, gg_eval **out) { if (out) *out = nullptr; return GG_UNSUPPORTED; }
int gg_create_adj(int, int, int, gg_eval **out) { if (out) *out = nullptr; return GG_UNSUPPORTED; }
void gg_destroy(gg_eval *) {}
int gg_eval_part(gg_eval *, const int32_t *, gg_result *) { return GG_UNSUPPORTED; }
int gg_eval_piece(gg_eval *, const int32_t *, gg_result *) { return GG_UNSUPPORTED; }
int gg_eval_box(gg_eval *, const int32_t *, gg_result *) { return GG_UNSUPPORTED; }
int gg_eval_adj(gg_eval *, const int32_t *, gg_result *) { return GG_UNSUPPORTED; }
int gg_eval_adj_swap(gg_eval *, const gg_swap *, gg_result *) { return GG_UNSUPPORTED; }
int gg_eval_adj_commit(gg_eval *, const gg_swap *) { return GG_UNSUPPORTED; }
int gg_eval_set_option(gg_eval *, int, int) { return GG_UNSUPPORTED; }
int gg_batch_create(int, int, int, int, gg_batch **out) { if (out) *out = nullptr; return GG_UNSUPPORTED; }
void gg_batch_destroy(gg_batch *) {}
int gg_batch_set_base(gg_batch *, const int32_t *, gg_result *) { return GG_UNSUPPORTED; }
int gg_batch_eval(gg_batch *, const gg_swap *, int, gg_result *) { return GG_UNSUPPORTED; }
int gg_batch_commit(gg_batch *, const gg_swap *, gg_result *) { return GG_UNSUPPORTED; }
int gg_batch_commit_known(gg_batch *, const gg_swap *, const gg_result *) { return GG_UNSUPPORTED; }
}

namespace gg {
int learn_create(gg_search *, int) { return GG_UNSUPPORTED; }
void learn_destroy(gg_search *) {}
int learn_select(gg_search *, int, int) { return GG_UNSUPPORTED; }
bool learn_on_device(const gg_search *) { return false; }
int learn_eval_device(gg_search *) { return GG_UNSUPPORTED; }
int learn_observe(gg_search *, int) { return GG_UNSUPPORTED; }
void learn_fill_stats(gg_search *, gg_search_stats *) {}
bool learn_loop_ready(const gg_search *) { return false; }
int learn_loop_run(gg_search *, int, int *) { return GG_UNSUPPORTED; }
}  // namespace gg
