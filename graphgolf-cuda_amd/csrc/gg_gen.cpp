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
