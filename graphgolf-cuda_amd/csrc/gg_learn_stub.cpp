// placeholder until the learned scorer lands
#include "gg_search_impl.hpp"
namespace gg {
int learn_create(gg_search *, int) { return GG_UNSUPPORTED; }
void learn_destroy(gg_search *) {}
int learn_select(gg_search *, int, int) { return GG_UNSUPPORTED; }
int learn_observe(gg_search *, int) { return GG_UNSUPPORTED; }
void learn_fill_stats(gg_search *, gg_search_stats *) {}
}
