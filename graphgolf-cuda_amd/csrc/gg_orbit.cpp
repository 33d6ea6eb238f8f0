// gg_orbit.cpp -- search driver over cyclically symmetric ("orbit") graphs in
// the reference's `part` form: M representative rows of offsets, N = g*M,
// neighbour i of vertex v = (v + table[v mod M][i]) mod N (cppcode/part.hpp:6-44).
// Each candidate is a whole table, evaluated exactly from the M
// representatives with the total scaled by g (cudaASPLconv::diameterASPL,
// cudacode/cudaASPLconv.cu:197-243, through gg_eval_part).
//
// Moves follow createNeighbour for symmetric parts (cppcode/annealing.cpp:112-197):
//   coin 0 (only when g > 1): edge-length change -- edge orbit (from, slot)
//     with offset diff, partner row to = (from + diff) mod M and the slot
//     holding -diff; with from <= to the pair becomes
//     newdiff = (to - from) + M*k, -newdiff, k uniform in [0, g)
//     (annealing.cpp:114-139);
//   coin 1 (or g == 1): orbit 2-opt -- edge orbits (a->b) and (c->d) become
//     (a,d) and (b,c), each with a fresh lift k (annealing.cpp:140-197).
// Differences, fixed here on purpose (the same ones as the 2-opt driver):
// every draw goes through gg::Rng, so a GPU run and a CPU run of this driver
// draw the same moves; partner slots match offsets modulo N; and a move whose
// table is not a simple graph (a zero offset or a repeated offset in a row,
// i.e. a self-loop or a multi-edge in every copy) is redrawn, where the
// reference keeps it.  Acceptance and schedule are those of the 2-opt driver
// (annealing.cpp:406-412, HillClimb.cpp:239).
#include <chrono>
#include <climits>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>
#include "graphgolf_mi355x.h"
#include "graphgolf_search.h"
#include "gg_rng.hpp"
#include "gg_search_impl.hpp"

struct gg_orbit_search {
    int N = 0, M = 0, d = 0, g = 0;
    // lattice form (dims 2 = piece, 3 = box): per-axis orders and base
    // extents; dims 1 is the part form (Ns[0] = N, Ms[0] = M)
    int dims = 1;
    int Ns[3] = {0, 0, 0}, Ms[3] = {1, 1, 1};
    gg_search_cfg cfg{};
    gg_eval *ev = nullptr;
    gg_table_callbacks cb{};
    bool use_cb = false;
    std::unique_ptr<gg::Rng> rng;
    std::vector<int32_t> tab, cand, best_tab;
    // part form: entries per row (empty = every row has `d`); slots past a
    // row's length are padding, offset 0
    std::vector<int32_t> lens;
    gg_result cur{}, best{};
    int64_t step = 0;
    gg_search_stats stats{};
    int log_cap = 0;
    std::vector<int64_t> log;
};

namespace {

using gg::OrbitMove;

inline int64_t mod_n(int64_t x, int64_t N) {
    x %= N;
    return x < 0 ? x + N : x;
}

// Slot among the n entries of row `row` holding the reverse of offset `diff`
// seen from row `from`/slot `slot` (the same slot is skipped, as
// annealing.cpp:121 skips it for diff == 0); -1 when there is none.
int partner_slot(const std::vector<int32_t> &t, int N, int d, int row, int from, int slot,
                 int32_t diff, int n) {
    for (int i = 0; i < n; i++) {
        if (row == from && i == slot) continue;
        if (mod_n((int64_t)t[(size_t)row * d + i] + diff, N) == 0) return i;
    }
    return -1;
}

// The n entries of row `r` describe simple neighbourhoods: offsets nonzero
// and distinct mod N.
bool row_simple(const std::vector<int32_t> &t, int N, int d, int r, int n) {
    for (int i = 0; i < n; i++) {
        int64_t x = mod_n(t[(size_t)r * d + i], N);
        if (x == 0) return false;
        for (int j = 0; j < i; j++)
            if (mod_n(t[(size_t)r * d + j], N) == x) return false;
    }
    return true;
}

}  // namespace

// One proposal attempt on `cur`, written to `out` (a copy of cur).  Returns
// false when the move is degenerate or its table is not simple.  With row
// lengths (annealingNR, cppcode/annealingNR.cpp:117,146-147) a slot is drawn
// among its row's entries and partners are looked for among theirs.
bool gg::propose_orbit(const std::vector<int32_t> &cur, int N, int M, int d, gg::Rng &rng,
                       std::vector<int32_t> &out, OrbitMove *mv, const int32_t *lens) {
    const int g = N / M;
    auto len = [&](int r) { return lens ? (int)lens[r] : d; };
    const int coin = (int)rng.below(2);
    out = cur;
    if (g > 1 && coin == 0) {
        int from = (int)rng.below(M), ifrom = (int)rng.below(len(from));
        const int32_t diff = cur[(size_t)from * d + ifrom];
        int to = (int)(mod_n((int64_t)from + diff, N) % M);
        int ito = partner_slot(cur, N, d, to, from, ifrom, diff, len(to));
        if (ito < 0) return false;
        if (from > to) { std::swap(from, to); std::swap(ifrom, ito); }
        const int k = (int)rng.below(g);
        const int32_t nd = (to - from) + M * k;
        out[(size_t)from * d + ifrom] = nd;
        out[(size_t)to * d + ito] = -nd;
        *mv = OrbitMove{0, from, ifrom, to, ito, k, 0};
        return row_simple(out, N, d, from, len(from)) && row_simple(out, N, d, to, len(to));
    }
    int a = (int)rng.below(M), ia = (int)rng.below(len(a));
    int c = (int)rng.below(M), ic = (int)rng.below(len(c));
    const int32_t dab = cur[(size_t)a * d + ia], dcd = cur[(size_t)c * d + ic];
    int b = (int)(mod_n((int64_t)a + dab, N) % M);
    int dd = (int)(mod_n((int64_t)c + dcd, N) % M);
    int ib = partner_slot(cur, N, d, b, a, ia, dab, len(b));
    int id = partner_slot(cur, N, d, dd, c, ic, dcd, len(dd));
    if (ib < 0 || id < 0) return false;
    // the same edge orbit drawn twice (annealing.cpp:168-172)
    if ((c == a && ic == ia) || (c == b && ic == ib)) return false;
    const int ka = (int)rng.below(g), kb = (int)rng.below(g);
    *mv = OrbitMove{1, a, ia, c, ic, ka, kb};
    if (a > dd) { std::swap(a, dd); std::swap(ia, id); }
    const int32_t dad = (dd - a) + M * ka;
    out[(size_t)a * d + ia] = dad;
    out[(size_t)dd * d + id] = -dad;
    if (b > c) { std::swap(b, c); std::swap(ib, ic); }
    const int32_t dbc = (c - b) + M * kb;
    out[(size_t)b * d + ib] = dbc;
    out[(size_t)c * d + ic] = -dbc;
    return row_simple(out, N, d, a, len(a)) && row_simple(out, N, d, dd, len(dd)) &&
           row_simple(out, N, d, b, len(b)) && row_simple(out, N, d, c, len(c));
}

namespace {

// ---- lattice (piece / box) tables -------------------------------------------------------
// t[(row*d + slot)*D + k]: offset along axis k of slot `slot` of base row
// `row`; rows are mixed-radix over (M_0, .., M_{D-1}), axis 0 major
// (x*My + y for piece, (x*My + y)*Mz + z for box; cudaASPLpiece.cu:163-174).
struct Lattice {
    int D, d;
    int N[3], M[3], g[3];
    int Mtot;
    void coords(int row, int c[3]) const {
        for (int k = D - 1; k >= 0; k--) { c[k] = row % M[k]; row /= M[k]; }
    }
    int row_of(const int c[3]) const {
        int r = 0;
        for (int k = 0; k < D; k++) r = r * M[k] + c[k];
        return r;
    }
    // base row of the vertex reached from base row `row` by offset `off`
    int target(int row, const int32_t *off) const {
        int c[3];
        coords(row, c);
        for (int k = 0; k < D; k++) c[k] = (int)(mod_n((int64_t)c[k] + off[k], N[k]) % M[k]);
        return row_of(c);
    }
    const int32_t *at(const std::vector<int32_t> &t, int row, int slot) const {
        return t.data() + ((size_t)row * d + slot) * D;
    }
    int32_t *at(std::vector<int32_t> &t, int row, int slot) const {
        return t.data() + ((size_t)row * d + slot) * D;
    }
    bool zero(const int32_t *o) const {
        for (int k = 0; k < D; k++)
            if (mod_n(o[k], N[k]) != 0) return false;
        return true;
    }
    bool same(const int32_t *o, const int32_t *q) const {
        for (int k = 0; k < D; k++)
            if (mod_n((int64_t)o[k] - q[k], N[k]) != 0) return false;
        return true;
    }
    bool reverse(const int32_t *o, const int32_t *q) const {  // o == -q per axis
        for (int k = 0; k < D; k++)
            if (mod_n((int64_t)o[k] + q[k], N[k]) != 0) return false;
        return true;
    }
    // slot of `row` holding -diff seen from (from, slot), the same slot skipped
    // (pieceannealing.cpp:127-133); -1 when there is none
    int partner(const std::vector<int32_t> &t, int row, int from, int slot,
                const int32_t *diff) const {
        for (int i = 0; i < d; i++) {
            if (row == from && i == slot) continue;
            if (reverse(at(t, row, i), diff)) return i;
        }
        return -1;
    }
    // nonzero offsets, pairwise distinct (no self-loop or multi-edge in any copy)
    bool row_simple(const std::vector<int32_t> &t, int r) const {
        for (int i = 0; i < d; i++) {
            if (zero(at(t, r, i))) return false;
            for (int j = 0; j < i; j++)
                if (same(at(t, r, i), at(t, r, j))) return false;
        }
        return true;
    }
    // every offset of row r has its reverse in the row it leads to
    bool row_symmetric(const std::vector<int32_t> &t, int r) const {
        for (int i = 0; i < d; i++)
            if (partner(t, target(r, at(t, r, i)), r, i, at(t, r, i)) < 0) return false;
        return true;
    }
    // new offset (to - from) + M*k along every axis, k drawn per axis (x first)
    void lift(int from, int to, gg::Rng &rng, int32_t *out, int *kpack) const {
        int cf[3], ct[3];
        coords(from, cf);
        coords(to, ct);
        int kp = 0;
        for (int k = 0; k < D; k++) {
            const int kk = (int)rng.below(g[k]);
            out[k] = (ct[k] - cf[k]) + M[k] * kk;
            kp = kp * g[k] + kk;
        }
        *kpack = kp;
    }
};

// One proposal attempt on a piece / box table (pieceannealing.cpp:113-204,
// boxHillClimb.cpp:113-220): coin 0, when every axis has more than one copy,
// re-lifts one edge orbit; otherwise orbit 2-opt (a,b),(c,d) -> (a,d),(b,c)
// with fresh lifts.  The lifted pair keeps the reference's orientation (no
// from <= to reordering, unlike the 1-D form).  Redrawn unless the four
// touched rows stay simple and symmetric.
bool propose_lattice(const Lattice &L, const std::vector<int32_t> &cur, gg::Rng &rng,
                     std::vector<int32_t> &out, OrbitMove *mv) {
    const int d = L.d;
    bool lengths = true;
    for (int k = 0; k < L.D; k++) lengths = lengths && L.g[k] > 1;
    const int coin = (int)rng.below(2);
    out = cur;
    if (lengths && coin == 0) {
        const int from = (int)rng.below(L.Mtot), ifrom = (int)rng.below(d);
        const int32_t *diff = L.at(cur, from, ifrom);
        const int to = L.target(from, diff);
        const int ito = L.partner(cur, to, from, ifrom, diff);
        if (ito < 0) return false;
        int32_t nd[3];
        int kp = 0;
        L.lift(from, to, rng, nd, &kp);
        for (int k = 0; k < L.D; k++) {
            L.at(out, from, ifrom)[k] = nd[k];
            L.at(out, to, ito)[k] = -nd[k];
        }
        *mv = OrbitMove{0, from, ifrom, to, ito, kp, 0};
        return L.row_simple(out, from) && L.row_simple(out, to) && L.row_symmetric(out, from) &&
               L.row_symmetric(out, to);
    }
    const int a = (int)rng.below(L.Mtot), ia = (int)rng.below(d);
    const int c = (int)rng.below(L.Mtot), ic = (int)rng.below(d);
    const int32_t *dab = L.at(cur, a, ia), *dcd = L.at(cur, c, ic);
    const int b = L.target(a, dab), dd = L.target(c, dcd);
    const int ib = L.partner(cur, b, a, ia, dab);
    const int id = L.partner(cur, dd, c, ic, dcd);
    if (ib < 0 || id < 0) return false;
    if ((c == a && ic == ia) || (c == b && ic == ib)) return false;  // the same edge orbit
    int32_t dad[3], dbc[3];
    int ka = 0, kb = 0;
    L.lift(a, dd, rng, dad, &ka);
    L.lift(b, c, rng, dbc, &kb);
    *mv = OrbitMove{1, a, ia, c, ic, ka, kb};
    for (int k = 0; k < L.D; k++) {
        L.at(out, a, ia)[k] = dad[k];
        L.at(out, dd, id)[k] = -dad[k];
        L.at(out, b, ib)[k] = dbc[k];
        L.at(out, c, ic)[k] = -dbc[k];
    }
    const int rows[4] = {a, dd, b, c};
    for (int r : rows)
        if (!L.row_simple(out, r) || !L.row_symmetric(out, r)) return false;
    return true;
}

Lattice lattice_of(const gg_orbit_search *s) {
    Lattice L{};
    L.D = s->dims;
    L.d = s->d;
    L.Mtot = 1;
    for (int k = 0; k < 3; k++) {
        L.N[k] = k < s->dims ? s->Ns[k] : 1;
        L.M[k] = k < s->dims ? s->Ms[k] : 1;
        L.g[k] = L.N[k] / L.M[k];
        if (k < s->dims) L.Mtot *= L.M[k];
    }
    return L;
}

int eval_table(gg_orbit_search *s, const int32_t *t, gg_result *out) {
    if (s->use_cb) return s->cb.eval(s->cb.ctx, t, out);
    if (s->dims == 2) return gg_eval_piece(s->ev, t, out);
    if (s->dims == 3) return gg_eval_box(s->ev, t, out);
    return gg_eval_part(s->ev, t, out);
}

int validate(int N, int M, int d, const int32_t *table, const gg_search_cfg *cfg) {
    if (!cfg || !table || M < 1 || N < 5 || d < 2 || N % M != 0) return GG_BAD_ARG;
    if (cfg->mode != GG_MODE_PLAIN) return GG_UNSUPPORTED;
    if (cfg->rule < GG_RULE_ANNEAL || cfg->rule > GG_RULE_HILL_WVC) return GG_BAD_ARG;
    if ((cfg->rule == GG_RULE_ANNEAL || cfg->rule == GG_RULE_ANNEAL_WVC) &&
        (!(cfg->T0 > 0) || !(cfg->T1 > 0)))
        return GG_BAD_ARG;
    if (cfg->schedule != GG_SCHED_GEOMETRIC && cfg->schedule != GG_SCHED_STEP) return GG_BAD_ARG;
    if (cfg->total_steps < 1) return GG_BAD_ARG;
    return GG_OK;
}

int validate_lattice(int dims, const int32_t *Ns, const int32_t *Ms, int d,
                     const int32_t *table, const gg_search_cfg *cfg) {
    if (dims < 2 || dims > 3 || !Ns || !Ms) return GG_BAD_ARG;
    int64_t N = 1, M = 1;
    for (int k = 0; k < dims; k++) {
        if (Ms[k] < 1 || Ns[k] < 1 || Ns[k] % Ms[k] != 0) return GG_BAD_ARG;
        N *= Ns[k];
        M *= Ms[k];
    }
    if (N > INT32_MAX) return GG_BAD_ARG;
    // WVC exists for part tables only (cudaASPLconv::WVCdiameterASPL)
    if (cfg && gg::rule_uses_wvc(cfg->rule)) return GG_UNSUPPORTED;
    return validate((int)N, (int)M, d, table, cfg);
}

void set_lattice(gg_orbit_search *s, int dims, const int32_t *Ns, const int32_t *Ms) {
    s->dims = dims;
    for (int k = 0; k < dims; k++) { s->Ns[k] = Ns[k]; s->Ms[k] = Ms[k]; }
}

int init(gg_orbit_search *s, int N, int M, int d, const int32_t *table, const gg_search_cfg *cfg) {
    s->N = N; s->M = M; s->d = d; s->g = N / M; s->cfg = *cfg;
    if (s->dims == 1) { s->Ns[0] = N; s->Ms[0] = M; }
    s->tab.assign(table, table + (size_t)M * d * s->dims);
    s->rng.reset(new gg::Rng(cfg->seed));
    gg_result r;
    int st = eval_table(s, s->tab.data(), &r);
    if (st != GG_OK) return st;
    s->cur = s->best = r;
    s->best_tab = s->tab;
    return GG_OK;
}

int orbit_step(gg_orbit_search *s) {
    const gg_search_cfg &cfg = s->cfg;
    gg::Rng &rng = *s->rng;
    OrbitMove mv{};
    int tries = 0;
    if (s->dims == 1) {
        while (!gg::propose_orbit(s->tab, s->N, s->M, s->d, rng, s->cand, &mv,
                                  s->lens.empty() ? nullptr : s->lens.data()))
            if (++tries > 1000000) return GG_BAD_ARG;  // no simple move exists
    } else {
        const Lattice L = lattice_of(s);
        while (!propose_lattice(L, s->tab, rng, s->cand, &mv))
            if (++tries > 1000000) return GG_BAD_ARG;
    }
    s->stats.proposals++;
    gg_result y;
    int st = eval_table(s, s->cand.data(), &y);
    if (st != GG_OK) return st;
    s->stats.exact_evals++;
    s->step++;
    const bool ok = y.status == GG_OK;
    const bool accept =
        gg::accept_move(cfg, s->step, (double)s->N * (s->N - 1), s->cur, y, rng);
    if (s->log_cap > 0 && (int64_t)s->log.size() < (int64_t)s->log_cap * 10) {
        int64_t rec[10] = {s->step, mv.a * (int64_t)s->d + mv.ia, mv.c * (int64_t)s->d + mv.ic,
                           mv.k1, mv.k2, mv.kind, ok ? y.total_scaled : -1, ok ? y.diameter : -1,
                           y.status, accept ? 1 : 0};
        s->log.insert(s->log.end(), rec, rec + 10);
    }
    if (accept) {
        s->tab.swap(s->cand);
        s->cur = y;
        s->stats.accepted++;
        if (gg::better_by_rule(cfg.rule, s->cur, s->best)) {
            s->best = s->cur;
            s->best_tab = s->tab;
        }
    }
    return GG_OK;
}

}  // namespace

extern "C" int gg_orbit_search_create(int N, int M, int d, const int32_t *table,
                                      const gg_search_cfg *cfg, int device,
                                      gg_orbit_search **out) {
    if (!out) return GG_BAD_ARG;
    *out = nullptr;
    int st = validate(N, M, d, table, cfg);
    if (st != GG_OK) return st;
    gg_orbit_search *s = new gg_orbit_search();
    st = gg_create_part(N, M, d, device, &s->ev);
    // the WVC rules evaluate as WVCdiameterASPL does: levels up to 100
    // (cudaASPLconv.cu:264)
    if (st == GG_OK && gg::rule_uses_wvc(cfg->rule))
        st = gg_eval_set_option(s->ev, GG_OPT_LEVEL_LIMIT, 100);
    if (st == GG_OK) st = init(s, N, M, d, table, cfg);
    if (st != GG_OK) { gg_orbit_search_destroy(s); return st; }
    *out = s;
    return GG_OK;
}

extern "C" int gg_orbit_search_set_row_lengths(gg_orbit_search *s, const int32_t *lens) {
    if (!s || !lens || s->dims != 1) return GG_BAD_ARG;
    for (int r = 0; r < s->M; r++) {
        if (lens[r] < 1 || lens[r] > s->d) return GG_BAD_ARG;
        for (int i = lens[r]; i < s->d; i++)
            if (s->tab[(size_t)r * s->d + i] != 0) return GG_BAD_ARG;  // padding is offset 0
    }
    s->lens.assign(lens, lens + s->M);
    return GG_OK;
}

extern "C" int gg_orbit_search_create_cb(int N, int M, int d, const int32_t *table,
                                         const gg_search_cfg *cfg, const gg_table_callbacks *cb,
                                         gg_orbit_search **out) {
    if (!out || !cb || !cb->eval) return GG_BAD_ARG;
    *out = nullptr;
    int st = validate(N, M, d, table, cfg);
    if (st != GG_OK) return st;
    gg_orbit_search *s = new gg_orbit_search();
    s->cb = *cb;
    s->use_cb = true;
    st = init(s, N, M, d, table, cfg);
    if (st != GG_OK) { gg_orbit_search_destroy(s); return st; }
    *out = s;
    return GG_OK;
}

static int lattice_create(int dims, const int32_t *Ns, const int32_t *Ms, int d,
                          const int32_t *table, const gg_search_cfg *cfg, int device,
                          gg_orbit_search **out) {
    if (!out) return GG_BAD_ARG;
    *out = nullptr;
    int st = validate_lattice(dims, Ns, Ms, d, table, cfg);
    if (st != GG_OK) return st;
    gg_orbit_search *s = new gg_orbit_search();
    set_lattice(s, dims, Ns, Ms);
    const int N = Ns[0] * Ns[1] * (dims == 3 ? Ns[2] : 1);
    const int M = Ms[0] * Ms[1] * (dims == 3 ? Ms[2] : 1);
    st = dims == 2 ? gg_create_piece(Ns[0], Ns[1], Ms[0], Ms[1], d, device, &s->ev)
                   : gg_create_box(Ns[0], Ns[1], Ns[2], Ms[0], Ms[1], Ms[2], d, device, &s->ev);
    if (st == GG_OK) st = init(s, N, M, d, table, cfg);
    if (st != GG_OK) { gg_orbit_search_destroy(s); return st; }
    *out = s;
    return GG_OK;
}

extern "C" int gg_piece_search_create(int Nx, int Ny, int Mx, int My, int degree,
                                      const int32_t *table, const gg_search_cfg *cfg, int device,
                                      gg_orbit_search **out) {
    const int32_t Ns[2] = {Nx, Ny}, Ms[2] = {Mx, My};
    return lattice_create(2, Ns, Ms, degree, table, cfg, device, out);
}

extern "C" int gg_box_search_create(int Nx, int Ny, int Nz, int Mx, int My, int Mz, int degree,
                                    const int32_t *table, const gg_search_cfg *cfg, int device,
                                    gg_orbit_search **out) {
    const int32_t Ns[3] = {Nx, Ny, Nz}, Ms[3] = {Mx, My, Mz};
    return lattice_create(3, Ns, Ms, degree, table, cfg, device, out);
}

extern "C" int gg_lattice_search_create_cb(int dims, const int32_t *Ns, const int32_t *Ms,
                                           int degree, const int32_t *table,
                                           const gg_search_cfg *cfg,
                                           const gg_table_callbacks *cb, gg_orbit_search **out) {
    if (!out || !cb || !cb->eval) return GG_BAD_ARG;
    *out = nullptr;
    int st = validate_lattice(dims, Ns, Ms, degree, table, cfg);
    if (st != GG_OK) return st;
    gg_orbit_search *s = new gg_orbit_search();
    set_lattice(s, dims, Ns, Ms);
    s->cb = *cb;
    s->use_cb = true;
    int N = 1, M = 1;
    for (int k = 0; k < dims; k++) { N *= Ns[k]; M *= Ms[k]; }
    st = init(s, N, M, degree, table, cfg);
    if (st != GG_OK) { gg_orbit_search_destroy(s); return st; }
    *out = s;
    return GG_OK;
}

extern "C" void gg_orbit_search_destroy(gg_orbit_search *s) {
    if (!s) return;
    gg_destroy(s->ev);
    delete s;
}

extern "C" int gg_orbit_search_run(gg_orbit_search *s, int steps, gg_search_stats *out) {
    if (!s || steps < 0) return GG_BAD_ARG;
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < steps; i++) {
        int st = orbit_step(s);
        if (st != GG_OK) return st;
    }
    auto t1 = std::chrono::steady_clock::now();
    s->stats.seconds += std::chrono::duration<double>(t1 - t0).count();
    s->stats.steps = s->step;
    s->stats.total_cur = s->cur.total_scaled;
    s->stats.diam_cur = s->cur.diameter;
    s->stats.total_best = s->best.total_scaled;
    s->stats.diam_best = s->best.diameter;
    if (out) *out = s->stats;
    return GG_OK;
}

extern "C" int gg_orbit_search_table(gg_orbit_search *s, int32_t *table_out, int best) {
    if (!s || !table_out) return GG_BAD_ARG;
    const std::vector<int32_t> &t = best ? s->best_tab : s->tab;
    memcpy(table_out, t.data(), sizeof(int32_t) * t.size());
    return GG_OK;
}

extern "C" int gg_orbit_search_set_log(gg_orbit_search *s, int capacity) {
    if (!s || capacity < 0) return GG_BAD_ARG;
    s->log_cap = capacity;
    s->log.clear();
    s->log.reserve((size_t)capacity * 10);
    return GG_OK;
}

extern "C" int gg_orbit_search_get_log(gg_orbit_search *s, int64_t *buf, int capacity,
                                       int *count) {
    if (!s || !count) return GG_BAD_ARG;
    int n = (int)(s->log.size() / 10);
    *count = n;
    if (buf) memcpy(buf, s->log.data(), sizeof(int64_t) * 10 * (size_t)std::min(n, capacity));
    return GG_OK;
}
