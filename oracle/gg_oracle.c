/*
 * gg_oracle.c -- CPU oracle for the exact ASPL / diameter evaluator.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library never links,
 * loads or calls anything in this directory, and has no CPU fallback.
 *
 * It is a plain-C restatement of the reference's queue BFS:
 *   - cppcode/cpuASPLqueue.cpp:54-96   (part: expand offsets, BFS from the M
 *     representatives, total += level per dequeued vertex, diameter = last
 *     level whose next queue was empty, unconnected when a BFS misses a vertex,
 *     total *= N/M (integer division), ASPL = total / (N(N-1)))
 *   - cppcode/cpuASPLpiece.cpp:12-63   (piece: 2-D torus expansion, sources x<Mx,y<My)
 *   - cppcode/cpuASPLbox.cpp:12-66     (box: 3-D torus expansion, sources x<Mx,y<My,z<Mz)
 *   - cudacode/cudaASPLgraph.cu:141-199 (explicit adjacency, all N sources,
 *     unscaled total)
 * plus the reference GPU's WVC definition (cudacode/cudaASPLconv.cu:244-308):
 * bit j of the AND over all 32-bit source words merges sources j, j+32, ...;
 * WVC = number of those classes whose largest eccentricity equals the diameter.
 *
 * Unlike the reference (arrayQueue.hpp:4-34) there is no capacity template:
 * queues and the visited set are sized from N.
 *
 * Parity pins (see tests/test_oracle.py): closed forms (cycles, complete,
 * hypercube, Petersen, K_{d,d}), Floyd-Warshall brute force (oracle_apsp_fw),
 * symmetric-vs-expanded invariants, and the reference's own cpuASPL* templates
 * compiled from /root/reference into oracle/_ref (golden vectors under
 * tests/golden/).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_INT_MAX 2147483647

/* Expand a 1-D 'part' offset table into an explicit adjacency.
 * Row r of the table has lens[r] valid offsets (row stride = stride).
 * Vertex v gets neighbours (v + off + N) % N for off in row v % M
 * (cppcode/cpuASPLqueue.cpp:57-62).  adj_out is N*dmax, padded with -1. */
int oracle_expand_part(int N, int M, int stride, const int32_t *lens,
                       const int32_t *offs, int dmax, int32_t *adj_out)
{
    if (N <= 0 || M <= 0 || M > N) return -1;
    for (int v = 0; v < N; v++) {
        int r = v % M;
        int E = lens[r];
        if (E > dmax) return -1;
        for (int i = 0; i < dmax; i++) {
            if (i < E) {
                long long to = ((long long)v + offs[(size_t)r * stride + i] + N) % N;
                if (to < 0) to += N; /* the reference is undefined here; keep in range */
                adj_out[(size_t)v * dmax + i] = (int32_t)to;
            } else {
                adj_out[(size_t)v * dmax + i] = -1;
            }
        }
    }
    return 0;
}

/* 2-D 'piece' expansion (cppcode/cpuASPLpiece.cpp:16-28).  offs holds
 * (dx,dy) pairs, row (x*My+y) of stride 2*stride ints. Vertex id = x*Ny+y. */
int oracle_expand_piece(int Nx, int Ny, int Mx, int My, int stride,
                        const int32_t *lens, const int32_t *offs, int dmax,
                        int32_t *adj_out)
{
    if (Nx <= 0 || Ny <= 0 || Mx <= 0 || My <= 0) return -1;
    for (int x = 0; x < Nx; x++)
        for (int y = 0; y < Ny; y++) {
            int r = (x % Mx) * My + (y % My);
            int E = lens[r];
            if (E > dmax) return -1;
            int v = x * Ny + y;
            for (int i = 0; i < dmax; i++) {
                if (i < E) {
                    int dx = offs[((size_t)r * stride + i) * 2];
                    int dy = offs[((size_t)r * stride + i) * 2 + 1];
                    int tx = ((x + dx + Nx) % Nx + Nx) % Nx;
                    int ty = ((y + dy + Ny) % Ny + Ny) % Ny;
                    adj_out[(size_t)v * dmax + i] = tx * Ny + ty;
                } else {
                    adj_out[(size_t)v * dmax + i] = -1;
                }
            }
        }
    return 0;
}

/* 3-D 'box' expansion (cppcode/cpuASPLbox.cpp:16-31; id = x*Ny*Nz+y*Nz+z,
 * cppcode/box.hpp:29-36). offs holds (dx,dy,dz) triples. */
int oracle_expand_box(int Nx, int Ny, int Nz, int Mx, int My, int Mz, int stride,
                      const int32_t *lens, const int32_t *offs, int dmax,
                      int32_t *adj_out)
{
    if (Nx <= 0 || Ny <= 0 || Nz <= 0 || Mx <= 0 || My <= 0 || Mz <= 0) return -1;
    for (int x = 0; x < Nx; x++)
        for (int y = 0; y < Ny; y++)
            for (int z = 0; z < Nz; z++) {
                int r = ((x % Mx) * My + (y % My)) * Mz + (z % Mz);
                int E = lens[r];
                if (E > dmax) return -1;
                int v = (x * Ny + y) * Nz + z;
                for (int i = 0; i < dmax; i++) {
                    if (i < E) {
                        const int32_t *o = offs + ((size_t)r * stride + i) * 3;
                        int tx = ((x + o[0] + Nx) % Nx + Nx) % Nx;
                        int ty = ((y + o[1] + Ny) % Ny + Ny) % Ny;
                        int tz = ((z + o[2] + Nz) % Nz + Nz) % Nz;
                        adj_out[(size_t)v * dmax + i] = (tx * Ny + ty) * Nz + tz;
                    } else {
                        adj_out[(size_t)v * dmax + i] = -1;
                    }
                }
            }
    return 0;
}

/* Queue BFS from each listed source over an explicit adjacency (entries < 0
 * are padding and skipped).  Follows the per-source loop of
 * cppcode/cpuASPLqueue.cpp:65-91:
 *   total += level for every dequeued vertex,
 *   ecc(s) = the level after which the next queue is empty,
 *   a source that reaches fewer than N vertices marks the graph unconnected.
 * Outputs:
 *   *total        sum over sources of sum over reached v of d(s,v)  (unscaled)
 *   *diameter     max ecc, or ORACLE_INT_MAX when unconnected (cpuASPLqueue.cpp:92)
 *   ecc_out[k]    eccentricity of source k (may be NULL)
 *   reach_out[l]  number of (source, vertex) pairs at distance exactly l,
 *                 for l < reach_cap (may be NULL)
 * Returns 0, or 1 when unconnected. */
int oracle_bfs(int N, int dmax, const int32_t *adj, int nsrc, const int32_t *srcs,
               int64_t *total, int32_t *diameter, int32_t *ecc_out,
               int64_t *reach_out, int reach_cap)
{
    int32_t *cur = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    int32_t *nxt = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    uint64_t *vis = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)((N + 63) / 64));
    if (!cur || !nxt || !vis) { free(cur); free(nxt); free(vis); return -1; }
    if (reach_out) memset(reach_out, 0, sizeof(int64_t) * (size_t)reach_cap);
    int64_t tot = 0;
    int diam = 0, unconnected = 0;
    for (int k = 0; k < nsrc; k++) {
        int s = srcs[k];
        memset(vis, 0, sizeof(uint64_t) * (size_t)((N + 63) / 64));
        int nc = 0, nn = 0, cnt = 0, ecc = 0;
        vis[s >> 6] |= 1ull << (s & 63);
        cur[nc++] = s;
        for (int level = 0; nc > 0; level++) {
            nn = 0;
            for (int q = 0; q < nc; q++) {
                int v = cur[q];
                cnt++;
                tot += level;
                if (reach_out && level < reach_cap) reach_out[level]++;
                for (int i = 0; i < dmax; i++) {
                    int to = adj[(size_t)v * dmax + i];
                    if (to < 0) continue;
                    if (vis[to >> 6] & (1ull << (to & 63))) continue;
                    vis[to >> 6] |= 1ull << (to & 63);
                    nxt[nn++] = to;
                }
            }
            if (nn == 0) ecc = level;
            int32_t *t = cur; cur = nxt; nxt = t;
            nc = nn;
        }
        if (ecc_out) ecc_out[k] = ecc;
        if (ecc > diam) diam = ecc;
        if (cnt != N) unconnected = 1;
    }
    free(cur); free(nxt); free(vis);
    *total = tot;
    *diameter = unconnected ? ORACLE_INT_MAX : diam;
    return unconnected;
}

/* WVC as the reference GPU computes it (cudacode/cudaASPLconv.cu:256-285):
 * classes j in [0,32) merge sources j, j+32, ... (< M); count the classes whose
 * largest eccentricity equals the diameter. ecc[s] is indexed by source id. */
int oracle_wvc(int M, const int32_t *ecc, int diameter)
{
    int wvc = 0;
    for (int j = 0; j < 32 && j < M; j++) {
        int mx = -1;
        for (int s = j; s < M; s += 32)
            if (ecc[s] > mx) mx = ecc[s];
        if (mx == diameter) wvc++;
    }
    return wvc;
}

/* Brute-force all-pairs shortest paths (Floyd-Warshall, int distances,
 * INF = N) on an explicit adjacency.  Independent of the BFS above; used to
 * pin it on small graphs.  dist_out is N*N (may be NULL); returns the sum over
 * all ordered pairs from the first nsrc vertices (0..nsrc-1) and the max
 * finite distance; *unconnected set when any pair is unreachable. */
int oracle_apsp_fw(int N, int dmax, const int32_t *adj, int nsrc,
                   int64_t *total, int32_t *diameter, int32_t *unconnected,
                   int32_t *dist_out)
{
    int32_t *D = (int32_t *)malloc(sizeof(int32_t) * (size_t)N * N);
    if (!D) return -1;
    const int32_t INF = N + 1;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) D[(size_t)i * N + j] = (i == j) ? 0 : INF;
    for (int v = 0; v < N; v++)
        for (int i = 0; i < dmax; i++) {
            int u = adj[(size_t)v * dmax + i];
            if (u >= 0 && u != v) { D[(size_t)v * N + u] = 1; }
        }
    for (int k = 0; k < N; k++)
        for (int i = 0; i < N; i++) {
            int32_t dik = D[(size_t)i * N + k];
            if (dik >= INF) continue;
            for (int j = 0; j < N; j++) {
                int32_t c = dik + D[(size_t)k * N + j];
                if (c < D[(size_t)i * N + j]) D[(size_t)i * N + j] = c;
            }
        }
    int64_t tot = 0;
    int32_t diam = 0, unc = 0;
    for (int i = 0; i < nsrc; i++)
        for (int j = 0; j < N; j++) {
            int32_t d = D[(size_t)i * N + j];
            if (d >= INF) { unc = 1; continue; }
            tot += d;
            if (d > diam) diam = d;
        }
    if (dist_out) memcpy(dist_out, D, sizeof(int32_t) * (size_t)N * N);
    free(D);
    *total = tot;
    *diameter = diam;
    *unconnected = unc;
    return 0;
}

/* OpenMP variant of oracle_bfs over sources (CPU baseline on all host cores;
 * the reference's own OpenMP pragma is commented out, cpuASPLqueue.cpp:64).
 * Same outputs except reach_out, which is not produced. */
#include <omp.h>
int oracle_bfs_omp(int N, int dmax, const int32_t *adj, int nsrc, const int32_t *srcs,
                   int nthreads, int64_t *total, int32_t *diameter)
{
    int64_t tot = 0;
    int diam = 0, unconnected = 0, err = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel reduction(+ : tot) reduction(max : diam) reduction(| : unconnected, err)
    {
        int32_t *cur = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
        int32_t *nxt = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
        uint64_t *vis = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)((N + 63) / 64));
        if (!cur || !nxt || !vis) err = 1;
#pragma omp for schedule(dynamic, 4)
        for (int k = 0; k < nsrc; k++) {
            if (err) continue;
            int s = srcs[k];
            memset(vis, 0, sizeof(uint64_t) * (size_t)((N + 63) / 64));
            int nc = 0, cnt = 0, ecc = 0;
            vis[s >> 6] |= 1ull << (s & 63);
            cur[nc++] = s;
            for (int level = 0; nc > 0; level++) {
                int nn = 0;
                for (int q = 0; q < nc; q++) {
                    int v = cur[q];
                    cnt++;
                    tot += level;
                    for (int i = 0; i < dmax; i++) {
                        int to = adj[(size_t)v * dmax + i];
                        if (to < 0) continue;
                        if (vis[to >> 6] & (1ull << (to & 63))) continue;
                        vis[to >> 6] |= 1ull << (to & 63);
                        nxt[nn++] = to;
                    }
                }
                if (nn == 0) ecc = level;
                int32_t *t = cur; cur = nxt; nxt = t;
                nc = nn;
            }
            if (ecc > diam) diam = ecc;
            if (cnt != N) unconnected = 1;
        }
        free(cur); free(nxt); free(vis);
    }
    if (err) return -1;
    *total = tot;
    *diameter = unconnected ? ORACLE_INT_MAX : diam;
    return unconnected;
}
