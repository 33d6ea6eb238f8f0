// host_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Runs the library's host-side C++ under the sanitizers: the seeded
// generators, the 2-opt search (plain and batched modes) and the orbit,
// piece and box searches, each with the oracle (oracle/gg_oracle.c) as the
// caller-supplied exact evaluator.  Prints "OK" when every call returns the
// expected status and the statistics add up.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "graphgolf_mi355x.h"
#include "graphgolf_search.h"

extern "C" {
int oracle_expand_part(int N, int M, int stride, const int32_t *lens, const int32_t *offs,
                       int dmax, int32_t *adj_out);
int oracle_expand_piece(int Nx, int Ny, int Mx, int My, int stride, const int32_t *lens,
                        const int32_t *offs, int dmax, int32_t *adj_out);
int oracle_expand_box(int Nx, int Ny, int Nz, int Mx, int My, int Mz, int stride,
                      const int32_t *lens, const int32_t *offs, int dmax, int32_t *adj_out);
int oracle_bfs(int N, int dmax, const int32_t *adj, int nsrc, const int32_t *srcs,
               int64_t *total, int32_t *diameter, int32_t *ecc_out, int64_t *reach_out,
               int reach_cap);
}

#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);       \
            std::exit(1);                                                 \
        }                                                                 \
    } while (0)

namespace {

// BFS from sources [0, M) of an explicit adjacency, total scaled by N/M.
void eval_adj(int N, int d, const int32_t *adj, int M, gg_result *out) {
    std::vector<int32_t> src(M);
    for (int i = 0; i < M; i++) src[i] = i;
    int64_t tot = 0;
    int32_t dia = 0;
    const int rc = oracle_bfs(N, d, adj, M, src.data(), &tot, &dia, nullptr, nullptr, 0);
    out->total_reps = tot;
    out->total_scaled = tot * (N / M);
    out->diameter = dia;
    out->wvc = 0;
    out->levels = dia + 1;
    out->status = rc == 0 ? GG_OK : GG_UNCONNECTED;
}

struct AdjCtx {
    int N, d;
    std::vector<int32_t> base;
};

void apply(std::vector<int32_t> &a, int d, const gg_swap &sw) {
    for (int j = 0; j < 4; j++)
        if (sw.v[j] >= 0) a[(size_t)sw.v[j] * d + sw.slot[j]] = sw.nbr[j];
}

int cb_set_base(void *ctx, const int32_t *adj, gg_result *out) {
    AdjCtx *c = static_cast<AdjCtx *>(ctx);
    c->base.assign(adj, adj + (size_t)c->N * c->d);
    eval_adj(c->N, c->d, c->base.data(), c->N, out);
    return GG_OK;
}

int cb_eval(void *ctx, const gg_swap *sw, int K, gg_result *out) {
    AdjCtx *c = static_cast<AdjCtx *>(ctx);
    for (int k = 0; k < K; k++) {
        std::vector<int32_t> a = c->base;
        apply(a, c->d, sw[k]);
        eval_adj(c->N, c->d, a.data(), c->N, &out[k]);
    }
    return GG_OK;
}

int cb_commit(void *ctx, const gg_swap *sw, gg_result *out) {
    AdjCtx *c = static_cast<AdjCtx *>(ctx);
    apply(c->base, c->d, *sw);
    eval_adj(c->N, c->d, c->base.data(), c->N, out);
    return GG_OK;
}

struct TabCtx {
    int dims, d;
    int Ns[3], Ms[3];
};

int cb_table(void *ctx, const int32_t *t, gg_result *out) {
    TabCtx *c = static_cast<TabCtx *>(ctx);
    int N = 1, M = 1;
    for (int k = 0; k < c->dims; k++) { N *= c->Ns[k]; M *= c->Ms[k]; }
    std::vector<int32_t> lens(M, c->d), adj((size_t)N * c->d);
    int rc;
    if (c->dims == 1) rc = oracle_expand_part(N, M, c->d, lens.data(), t, c->d, adj.data());
    else if (c->dims == 2)
        rc = oracle_expand_piece(c->Ns[0], c->Ns[1], c->Ms[0], c->Ms[1], c->d, lens.data(), t,
                                 c->d, adj.data());
    else
        rc = oracle_expand_box(c->Ns[0], c->Ns[1], c->Ns[2], c->Ms[0], c->Ms[1], c->Ms[2], c->d,
                               lens.data(), t, c->d, adj.data());
    if (rc != 0) return GG_BAD_ARG;
    // sources: the representatives x < Mx (, y < My, z < Mz)
    std::vector<int32_t> src;
    if (c->dims == 1) {
        for (int i = 0; i < M; i++) src.push_back(i);
    } else if (c->dims == 2) {
        for (int x = 0; x < c->Ms[0]; x++)
            for (int y = 0; y < c->Ms[1]; y++) src.push_back(x * c->Ns[1] + y);
    } else {
        for (int x = 0; x < c->Ms[0]; x++)
            for (int y = 0; y < c->Ms[1]; y++)
                for (int z = 0; z < c->Ms[2]; z++)
                    src.push_back((x * c->Ns[1] + y) * c->Ns[2] + z);
    }
    int64_t tot = 0;
    int32_t dia = 0;
    const int r = oracle_bfs(N, c->d, adj.data(), M, src.data(), &tot, &dia, nullptr, nullptr, 0);
    out->total_reps = tot;
    out->total_scaled = tot * (N / M);
    out->diameter = dia;
    out->wvc = 0;
    out->levels = dia + 1;
    out->status = r == 0 ? GG_OK : GG_UNCONNECTED;
    return GG_OK;
}

// A simple symmetric table: a d-regular graph on the M base rows (every edge
// {i, j} one orbit with lift vector k), offsets (j - i) + M*k per axis.
std::vector<int32_t> lattice_table(const TabCtx &c, uint64_t seed) {
    int M = 1;
    for (int k = 0; k < c.dims; k++) M *= c.Ms[k];
    std::vector<int32_t> base((size_t)M * c.d);
    CHECK(gg_gen_random_regular(M, c.d, seed, base.data()) == GG_OK);
    std::vector<int32_t> t((size_t)M * c.d * c.dims, 0);
    uint64_t x = seed * 0x9E3779B97F4A7C15ull + 1;
    auto rnd = [&](int n) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        return (int)(x % (uint64_t)n);
    };
    auto coords = [&](int r, int *cc) {
        for (int k = c.dims - 1; k >= 0; k--) { cc[k] = r % c.Ms[k]; r /= c.Ms[k]; }
    };
    for (int i = 0; i < M; i++)
        for (int s = 0; s < c.d; s++) {
            const int j = base[(size_t)i * c.d + s];
            if (j < i) continue;
            int js = -1;
            for (int q = 0; q < c.d; q++)
                if (base[(size_t)j * c.d + q] == i) js = q;
            CHECK(js >= 0);
            int ci[3] = {0, 0, 0}, cj[3] = {0, 0, 0};
            coords(i, ci);
            coords(j, cj);
            for (int k = 0; k < c.dims; k++) {
                const int off = (cj[k] - ci[k]) + c.Ms[k] * rnd(c.Ns[k] / c.Ms[k]);
                t[((size_t)i * c.d + s) * c.dims + k] = off;
                t[((size_t)j * c.d + js) * c.dims + k] = -off;
            }
        }
    return t;
}

void run_adj_search(int mode, int rule) {
    const int N = 200, d = 4;
    std::vector<int32_t> adj((size_t)N * d);
    CHECK(gg_gen_random_regular(N, d, 3, adj.data()) == GG_OK);
    AdjCtx ctx{N, d, {}};
    gg_eval_callbacks cb{&ctx, cb_set_base, cb_eval, cb_commit};
    gg_search_cfg cfg{};
    cfg.mode = mode;
    cfg.rule = rule;
    cfg.K = mode == GG_MODE_BATCHED ? 6 : 1;
    cfg.P = 1;
    cfg.total_steps = 60;
    cfg.T0 = 1e-2;
    cfg.T1 = 1e-4;
    cfg.seed = 9;
    cfg.schedule = GG_SCHED_GEOMETRIC;
    gg_search *s = nullptr;
    CHECK(gg_search_create_cb(N, d, adj.data(), &cfg, &cb, &s) == GG_OK);
    CHECK(gg_search_set_log(s, 60) == GG_OK);
    gg_search_stats st{};
    CHECK(gg_search_run(s, 60, &st) == GG_OK);
    CHECK(st.steps == 60 && st.exact_evals == 60 * cfg.K);
    std::vector<int32_t> g((size_t)N * d);
    CHECK(gg_search_graph(s, g.data(), 1) == GG_OK);
    std::vector<int64_t> log(60 * 10);
    int n = 0;
    CHECK(gg_search_get_log(s, log.data(), 60, &n) == GG_OK && n == 60);
    gg_result r;
    CHECK(gg_search_graph(s, g.data(), 0) == GG_OK);
    eval_adj(N, d, g.data(), N, &r);
    CHECK(r.status == GG_OK && r.total_scaled == st.total_cur);
    gg_search_destroy(s);
}

void run_table_search(TabCtx c, int rule) {
    int M = 1;
    for (int k = 0; k < c.dims; k++) M *= c.Ms[k];
    std::vector<int32_t> t = lattice_table(c, 5);
    gg_table_callbacks cb{&c, cb_table};
    gg_search_cfg cfg{};
    cfg.mode = GG_MODE_PLAIN;
    cfg.rule = rule;
    cfg.total_steps = 80;
    cfg.T0 = 1e-2;
    cfg.T1 = 1e-4;
    cfg.seed = 4;
    gg_orbit_search *s = nullptr;
    if (c.dims == 1) {
        CHECK(gg_orbit_search_create_cb(c.Ns[0], c.Ms[0], c.d, t.data(), &cfg, &cb, &s) == GG_OK);
    } else {
        CHECK(gg_lattice_search_create_cb(c.dims, c.Ns, c.Ms, c.d, t.data(), &cfg, &cb, &s) ==
              GG_OK);
    }
    CHECK(gg_orbit_search_set_log(s, 80) == GG_OK);
    gg_search_stats st{};
    CHECK(gg_orbit_search_run(s, 80, &st) == GG_OK);
    CHECK(st.steps == 80 && st.exact_evals == 80);
    std::vector<int32_t> out((size_t)M * c.d * c.dims);
    CHECK(gg_orbit_search_table(s, out.data(), 0) == GG_OK);
    gg_result r;
    CHECK(cb_table(&c, out.data(), &r) == GG_OK);
    CHECK(r.status == GG_OK && r.total_scaled == st.total_cur);
    gg_orbit_search_destroy(s);
}

}  // namespace

int main() {
    // generators: seeded, simple, regular
    std::vector<int32_t> a((size_t)300 * 6), b((size_t)300 * 6);
    CHECK(gg_gen_random_regular(300, 6, 7, a.data()) == GG_OK);
    CHECK(gg_gen_random_regular(300, 6, 7, b.data()) == GG_OK);
    CHECK(a == b);
    std::vector<int32_t> op((size_t)32 * 8);
    CHECK(gg_gen_orbit_part(4, 32, 8, 3, op.data()) == GG_OK);
    // explicit graphs: plain and batched 2-opt search, annealing and hill climbing
    run_adj_search(GG_MODE_PLAIN, GG_RULE_ANNEAL);
    run_adj_search(GG_MODE_PLAIN, GG_RULE_HILL);
    run_adj_search(GG_MODE_BATCHED, GG_RULE_ANNEAL);
    // the learned mode needs the GPU scorer: a clean error without it
    {
        AdjCtx ctx{300, 6, {}};
        gg_eval_callbacks cb{&ctx, cb_set_base, cb_eval, cb_commit};
        gg_search_cfg cfg{};
        cfg.mode = GG_MODE_LEARNED;
        cfg.K = 4;
        cfg.P = 16;
        cfg.total_steps = 10;
        cfg.T0 = 1e-2;
        cfg.T1 = 1e-4;
        gg_search *s = nullptr;
        CHECK(gg_search_create_cb(300, 6, a.data(), &cfg, &cb, &s) != GG_OK && s == nullptr);
    }
    // symmetric tables: part (1-D), piece (2-D), box (3-D)
    run_table_search(TabCtx{1, 6, {256, 1, 1}, {32, 1, 1}}, GG_RULE_ANNEAL);
    run_table_search(TabCtx{2, 4, {8, 12, 1}, {2, 3, 1}}, GG_RULE_HILL);
    run_table_search(TabCtx{2, 6, {16, 16, 1}, {4, 2, 1}}, GG_RULE_ANNEAL);
    run_table_search(TabCtx{3, 4, {4, 4, 4}, {2, 2, 2}}, GG_RULE_HILL);
    // argument checks reach their status codes
    gg_orbit_search *os = nullptr;
    int32_t Ns[2] = {8, 12}, Ms[2] = {3, 3};  // 8 % 3 != 0
    gg_search_cfg cfg{};
    cfg.total_steps = 1;
    gg_table_callbacks tcb{nullptr, cb_table};
    CHECK(gg_lattice_search_create_cb(2, Ns, Ms, 4, op.data(), &cfg, &tcb, &os) == GG_BAD_ARG);
    std::printf("OK\n");
    return 0;
}
