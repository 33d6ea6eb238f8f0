// gg_gen.cpp -- seeded instance generators (host code).
//
// Random d-regular simple graphs: stub pairing (configuration model) as in
// cppcode/initTemp.cpp:399-408 and RNDpiecegen.cpp:89-127, but the reference
// never rejects loops or multi-edges; here small degrees resample the pairing
// until simple and larger degrees repair it with degree-preserving double-edge
// swaps (resampling would almost never succeed for d >= 12).
// Orbit ('part' with M < N) graphs: orbit edges (i, j, s) give row i the
// offset (j - i) + M*s and row j its negation (cppcode/annealing.cpp:132 form).
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <vector>
#include "graphgolf_mi355x.h"
#include "gg_rng.hpp"
#include "gg_host.h"

using gg::Rng;

namespace {

bool has_edge(const std::vector<std::vector<int>> &nb, int a, int b) {
    for (int x : nb[a])
        if (x == b) return true;
    return false;
}

void remove_one(std::vector<int> &row, int x) {
    for (size_t i = 0; i < row.size(); i++)
        if (row[i] == x) { row[i] = row.back(); row.pop_back(); return; }
}

void shuffle(std::vector<int> &v, Rng &rng) {
    for (size_t i = v.size(); i > 1; i--) {
        size_t j = rng.below(i);
        std::swap(v[i - 1], v[j]);
    }
}

}  // namespace

// Random simple d-regular graph, adjacency rows sorted.  Returns GG_OK or
// GG_BAD_ARG (n*d odd, d >= n, ...).
extern "C" int gg_gen_random_regular(int n, int d, uint64_t seed, int32_t *adj_out) {
    if (n < 1 || d < 0 || d >= n || ((long long)n * d) % 2 != 0 || !adj_out) return GG_BAD_ARG;
    Rng rng(seed);
    std::vector<int> stubs((size_t)n * d);
    std::vector<std::vector<int>> nb(n);
    const bool resample = d <= 6;
    for (int attempt = 0;; attempt++) {
        for (int v = 0; v < n; v++)
            for (int k = 0; k < d; k++) stubs[(size_t)v * d + k] = v;
        shuffle(stubs, rng);
        for (auto &r : nb) r.clear();
        std::vector<std::pair<int, int>> edges;
        edges.reserve(stubs.size() / 2);
        for (size_t i = 0; i + 1 < stubs.size(); i += 2) edges.emplace_back(stubs[i], stubs[i + 1]);
        for (auto &e : edges) { nb[e.first].push_back(e.second); nb[e.second].push_back(e.first); }
        auto bad = [&](int a, int b) {
            if (a == b) return true;
            int c = 0;
            for (int x : nb[a]) c += (x == b);
            return c > 1;
        };
        bool any_bad = false;
        for (auto &e : edges)
            if (bad(e.first, e.second)) { any_bad = true; break; }
        if (!any_bad) break;
        if (resample && attempt < 100000) continue;
        // repair: swap each bad edge with a random edge while that removes badness
        const size_t E = edges.size();
        for (long long iter = 0; iter < 1000000000LL; iter++) {
            std::vector<size_t> badidx;
            for (size_t i = 0; i < E; i++)
                if (bad(edges[i].first, edges[i].second)) badidx.push_back(i);
            if (badidx.empty()) break;
            for (size_t bi : badidx) {
                int a = edges[bi].first, b = edges[bi].second;
                if (!bad(a, b)) continue;
                for (int tries = 0; tries < 1000; tries++) {
                    size_t fj = rng.below(E);
                    if (fj == bi) continue;
                    int c = edges[fj].first, e2 = edges[fj].second;
                    if (rng.below(2)) std::swap(c, e2);
                    // (a,b),(c,e2) -> (a,c),(b,e2)
                    if (a == c || b == e2) continue;
                    if (has_edge(nb, a, c) || has_edge(nb, b, e2)) continue;
                    if (c == e2 && false) continue;
                    remove_one(nb[a], b); remove_one(nb[b], a);
                    remove_one(nb[c], e2); remove_one(nb[e2], c);
                    nb[a].push_back(c); nb[c].push_back(a);
                    nb[b].push_back(e2); nb[e2].push_back(b);
                    edges[bi] = {a, c};
                    edges[fj] = {b, e2};
                    break;
                }
            }
        }
        break;
    }
    for (int v = 0; v < n; v++) {
        std::sort(nb[v].begin(), nb[v].end());
        if ((int)nb[v].size() != d) return GG_BAD_ARG;
        for (int k = 0; k < d; k++) adj_out[(size_t)v * d + k] = nb[v][k];
    }
    return GG_OK;
}

// Orbit graph as a 'part' table (M = m rows of d offsets, N = g*m): orbit
// edges (i, j, s) with i, j < m and shift s in [0, g).  Self-loops (i = j,
// s = 0), half orbits (i = j, 2s = g) and duplicate orbits are rejected by
// resampling the whole pairing.
extern "C" int gg_gen_orbit_part(int g, int m, int d, uint64_t seed, int32_t *table_out) {
    if (g < 1 || m < 1 || d < 0 || ((long long)m * d) % 2 != 0 || !table_out) return GG_BAD_ARG;
    Rng rng(seed);
    const int N = g * m;
    std::vector<int> stubs((size_t)m * d);
    for (int attempt = 0; attempt < 1000000; attempt++) {
        for (int v = 0; v < m; v++)
            for (int k = 0; k < d; k++) stubs[(size_t)v * d + k] = v;
        shuffle(stubs, rng);
        std::vector<std::vector<int>> rows(m);
        std::vector<long long> keys;
        bool ok = true;
        for (size_t p = 0; p + 1 < stubs.size(); p += 2) {
            int i = stubs[p], j = stubs[p + 1];
            int s = (int)rng.below(g);
            if (i == j && (s == 0 || 2 * s == g)) { ok = false; }
            int ki = i, kj = j, ks = s;
            if (ki > kj) { std::swap(ki, kj); ks = (g - ks) % g; }
            if (ki == kj) ks = std::min(ks, (g - ks) % g);
            keys.push_back(((long long)ki * m + kj) * g + ks);
            int off = (j - i) + m * s;
            rows[i].push_back(off);
            rows[j].push_back(-off);
        }
        if (!ok) continue;
        std::sort(keys.begin(), keys.end());
        if (std::adjacent_find(keys.begin(), keys.end()) != keys.end()) continue;
        for (int r = 0; r < m; r++)
            for (int k = 0; k < d; k++) table_out[(size_t)r * d + k] = rows[r][k];
        (void)N;
        return GG_OK;
    }
    return GG_BAD_ARG;
}

// Symmetric tables of 1, 2 or 3 axes ('part', 'piece', 'box'; rows mixed-radix
// over Ms with axis 0 major, offsets [M][d][dims]).  The base rows' stubs are
// shuffled and paired as RNDpiecegen.cpp:89-127 / RNDboxgen.cpp:97-140 pair
// them, each pair an edge orbit with a lift drawn per axis.  Pairs that give a
// zero offset (a self-loop in every copy), or an offset equal to another of
// the same row (a multi-edge, half orbits included) are repaired: such a pair
// trades partners with a random other pair and both are lifted again, until
// no bad pair remains (the reference keeps them).  GG_BAD_ARG when the
// arguments are inconsistent or no simple table is reached.
extern "C" int gg_gen_lattice(int dims, const int32_t *Ns, const int32_t *Ms, int d,
                              uint64_t seed, int32_t *table_out) {
    if (dims < 1 || dims > 3 || !Ns || !Ms || d < 1 || !table_out) return GG_BAD_ARG;
    int64_t M = 1, N = 1;
    int g[3] = {1, 1, 1}, nk[3] = {1, 1, 1}, mk[3] = {1, 1, 1};
    for (int k = 0; k < dims; k++) {
        if (Ms[k] < 1 || Ns[k] < 1 || Ns[k] % Ms[k] != 0) return GG_BAD_ARG;
        nk[k] = Ns[k];
        mk[k] = Ms[k];
        g[k] = Ns[k] / Ms[k];
        M *= Ms[k];
        N *= Ns[k];
    }
    if (N > INT32_MAX || (M * d) % 2 != 0 || M * d > (int64_t)1 << 28) return GG_BAD_ARG;
    auto coords = [&](int r, int *c) {
        for (int k = dims - 1; k >= 0; k--) { c[k] = r % mk[k]; r /= mk[k]; }
    };
    // offset of pair (a -> b) with lifts kv, packed mod N per axis
    auto offset = [&](int a, int b, const int *kv, int *off) {
        int ca[3], cb[3];
        coords(a, ca);
        coords(b, cb);
        for (int k = 0; k < dims; k++) off[k] = (cb[k] - ca[k]) + mk[k] * kv[k];
    };
    auto key = [&](const int *off, int sign) {
        int64_t x = 0;
        for (int k = 0; k < dims; k++) {
            int64_t o = ((int64_t)sign * off[k]) % nk[k];
            if (o < 0) o += nk[k];
            x = x * nk[k] + o;
        }
        return x;
    };
    Rng rng(seed);
    const int64_t S = M * d, P = S / 2;
    std::vector<int> stubs((size_t)S);
    for (int64_t i = 0; i < S; i++) stubs[(size_t)i] = (int)(i / d);
    shuffle(stubs, rng);
    std::vector<int> pa((size_t)P), pb((size_t)P), lift((size_t)P * 3, 0);
    auto relift = [&](int64_t p) {
        for (int k = 0; k < dims; k++) lift[(size_t)p * 3 + k] = (int)rng.below(g[k]);
    };
    for (int64_t p = 0; p < P; p++) {
        pa[(size_t)p] = stubs[(size_t)(2 * p)];
        pb[(size_t)p] = stubs[(size_t)(2 * p + 1)];
        relift(p);
    }
    // entries (row, offset key, pair) of every pair's two ends
    std::vector<std::pair<std::pair<int, int64_t>, int64_t>> ent((size_t)S);
    std::vector<char> bad((size_t)P);
    for (int round = 0; round < 100000; round++) {
        for (int64_t p = 0; p < P; p++) {
            int off[3];
            offset(pa[(size_t)p], pb[(size_t)p], &lift[(size_t)p * 3], off);
            ent[(size_t)(2 * p)] = {{pa[(size_t)p], key(off, 1)}, p};
            ent[(size_t)(2 * p + 1)] = {{pb[(size_t)p], key(off, -1)}, p};
        }
        std::sort(ent.begin(), ent.end());
        std::fill(bad.begin(), bad.end(), 0);
        bool any = false;
        for (size_t i = 0; i < ent.size(); i++) {
            const bool zero = ent[i].first.second == 0;
            const bool dup = i + 1 < ent.size() && ent[i + 1].first == ent[i].first;
            if (zero || dup) {
                bad[(size_t)ent[i].second] = 1;
                if (dup) bad[(size_t)ent[i + 1].second] = 1;
                any = true;
            }
        }
        if (!any) {
            std::vector<int> fill((size_t)M, 0);
            for (int64_t p = 0; p < P; p++) {
                int off[3];
                const int a = pa[(size_t)p], b = pb[(size_t)p];
                offset(a, b, &lift[(size_t)p * 3], off);
                int32_t *ta = table_out + ((size_t)a * d + fill[(size_t)a]++) * dims;
                int32_t *tb = table_out + ((size_t)b * d + fill[(size_t)b]++) * dims;
                for (int k = 0; k < dims; k++) { ta[k] = off[k]; tb[k] = -off[k]; }
            }
            return GG_OK;
        }
        for (int64_t p = 0; p < P; p++) {
            if (!bad[(size_t)p]) continue;
            const int64_t q = (int64_t)rng.below((uint64_t)P);
            if (q != p) std::swap(pb[(size_t)p], pb[(size_t)q]);
            relift(p);
            relift(q);
        }
    }
    return GG_BAD_ARG;
}
