// gg_inittemp.cpp -- initial-temperature estimate of the annealing drivers
// (cppcode/initTemp.cpp:399-430, pieceinitTemp.cpp:228-260): over `graphs`
// random graphs, the mean increase of the objective over the `neighbours`
// moves that make it worse (unconnected results skipped, as initTemp.cpp:419
// skips 1e9), and T = -mean / ln 0.4, the temperature at which such a move
// is accepted with probability 0.4.
//
// Explicit graphs: random regular graphs, 2-opt swaps; the neighbours of one
// graph are one batched evaluation (gg_batch_eval).  Symmetric `part` tables
// (N/M > 1): random orbit graphs, orbit moves, each neighbour a part
// evaluation.  Graph g uses generator seed seed + 1 + g; all moves draw from
// one gg::Rng(seed) stream in order.
#include <cmath>
#include <vector>
#include "graphgolf_mi355x.h"
#include "graphgolf_search.h"
#include "gg_host.h"
#include "gg_rng.hpp"
#include "gg_search_impl.hpp"

using namespace gg;

static void finish(double delta, int64_t cnt, int N, gg_init_temp_result *out) {
    out->count = cnt;
    out->mean_increase_total = cnt ? delta / (double)cnt : 0.0;
    out->temp_total = cnt ? -out->mean_increase_total / std::log(0.4) : 0.0;
    out->temp_aspl = out->temp_total / ((double)N * (N - 1));
}

extern "C" int gg_init_temp(int N, int degree, int graphs, int neighbours, uint64_t seed,
                            int device, gg_init_temp_result *out) {
    if (!out || N < 5 || degree < 2 || graphs < 1 || neighbours < 1 || (int64_t)N * degree % 2)
        return GG_BAD_ARG;
    gg_batch *b = nullptr;
    int st = gg_batch_create(N, degree, neighbours, device, &b);
    if (st != GG_OK) return st;
    Rng rng(seed);
    std::vector<int32_t> adj((size_t)N * degree);
    std::vector<gg_swap> sw(neighbours);
    std::vector<gg_result> res(neighbours);
    double delta = 0;
    int64_t cnt = 0;
    for (int g = 0; g < graphs && st == GG_OK; g++) {
        st = gg_gen_random_regular(N, degree, seed + 1 + (uint64_t)g, adj.data());
        if (st != GG_OK) break;
        gg_result base;
        st = gg_batch_set_base(b, adj.data(), &base);
        if (st != GG_OK) break;
        for (int k = 0; k < neighbours; k++) {
            int tries = 0;
            while (!propose_swap(adj, N, degree, rng, &sw[k], nullptr))
                if (++tries > 1000000) { st = GG_BAD_ARG; break; }
            if (st != GG_OK) break;
        }
        if (st != GG_OK) break;
        st = gg_batch_eval(b, sw.data(), neighbours, res.data());
        if (st != GG_OK || base.status != GG_OK) continue;
        for (int k = 0; k < neighbours; k++)
            if (res[k].status == GG_OK && res[k].total_scaled > base.total_scaled) {
                delta += (double)(res[k].total_scaled - base.total_scaled);
                cnt++;
            }
    }
    gg_batch_destroy(b);
    if (st != GG_OK) return st;
    finish(delta, cnt, N, out);
    return GG_OK;
}

extern "C" int gg_init_temp_part(int N, int M, int degree, int graphs, int neighbours,
                                 uint64_t seed, int device, gg_init_temp_result *out) {
    if (!out || M < 1 || N % M != 0 || N / M < 2 || degree < 2 || graphs < 1 || neighbours < 1 ||
        (int64_t)M * degree % 2)
        return GG_BAD_ARG;
    gg_eval *ev = nullptr;
    int st = gg_create_part(N, M, degree, device, &ev);
    if (st != GG_OK) return st;
    Rng rng(seed);
    std::vector<int32_t> tab((size_t)M * degree), cand;
    double delta = 0;
    int64_t cnt = 0;
    for (int g = 0; g < graphs && st == GG_OK; g++) {
        st = gg_gen_orbit_part(N / M, M, degree, seed + 1 + (uint64_t)g, tab.data());
        if (st != GG_OK) break;
        gg_result base, y;
        st = gg_eval_part(ev, tab.data(), &base);
        if (st != GG_OK) break;
        for (int k = 0; k < neighbours && st == GG_OK; k++) {
            OrbitMove mv{};
            int tries = 0;
            while (!propose_orbit(tab, N, M, degree, rng, cand, &mv))
                if (++tries > 1000000) { st = GG_BAD_ARG; break; }
            if (st != GG_OK) break;
            st = gg_eval_part(ev, cand.data(), &y);
            if (st != GG_OK) break;
            if (base.status == GG_OK && y.status == GG_OK && y.total_scaled > base.total_scaled) {
                delta += (double)(y.total_scaled - base.total_scaled);
                cnt++;
            }
        }
    }
    gg_destroy(ev);
    if (st != GG_OK) return st;
    finish(delta, cnt, N, out);
    return GG_OK;
}
