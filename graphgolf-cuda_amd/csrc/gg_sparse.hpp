// gg_sparse.hpp -- incremental exact evaluation of swap candidates against a
// base whose per-level BFS state is stored ("difference propagation").
//
// For column w (64 sources) let B_l[v] be the base visited word of v after
// level l and V_l[v] the candidate's.  Since the candidate graph G' differs
// from the base G only in the rows of its swap endpoints E,
//     V_l[v] = V_{l-1}[v] | OR_{u in adj'(v)} V_{l-1}[u]
// equals B_l[v] unless v is in E or v or one of its neighbours differed at
// level l-1.  So with Delta_l = { v : V_l[v] != B_l[v] } (Delta_0 empty), only
//     C_l = E  u  Delta_{l-1}  u  adj'(Delta_{l-1})
// needs recomputing, and the column's reached-pair count after level l is the
// base count plus sum_{v in Delta_l} popc(V_l[v]) - popc(B_l[v]).  The distance
// total and eccentricity of the column follow from those counts.  Iteration
// stops at a fixed point (level >= base column eccentricity with Delta_l empty,
// or level > base eccentricity with Delta_l == Delta_{l-1}).
//
// One workgroup per column keeps B_{l-1}[w][*] in LDS and walks the levels
// for all candidates of the launch.  All its waves draw the level's items
// (C_l of every active candidate, flattened) from one pool, so a candidate
// with a large Delta does not idle the others.  Delta_{l-1} membership is a
// per-task bitmap and its words a per-task dense override array in HBM
// (written only at Delta positions), so the D neighbour lookups of an item
// are independent loads.  Tasks whose Delta outgrows the list capacity, and
// the columns that contain a swap endpoint as a source, are queued for the
// dense kernel (k_dense_lds_queue) or run densely in the kernel's own tail.
//
// A column that holds a swap endpoint x as a source differs from the base in
// x's bit nearly everywhere (x's own BFS tree changes), so its Delta would
// overflow.  Bits are independent BFSes, though: in fixed-degree kernels
// with a tail (TAIL = 1) such a column runs incrementally with the endpoint
// sources' bits masked out of every word compared, counted and logged (the
// base's counts of those bits are taken from the stored levels as they are
// staged), and the tail runs a single-source BFS from each endpoint on the
// candidate graph (top-down while the frontier is small,
// bottom-up after it, as Beamer et al. 2012 "Direction-optimizing
// breadth-first search"; bottom-up only where the candidate graph may be
// directed, since the evaluators pull from listed neighbours).  The column's
// reached count after level l is then
//     base count - base counts of the masked sources + masked delta
//                + candidate counts of the masked sources,
// which k_sparse_finalize combines.  Without a tail (TAIL = 0) those columns
// run densely in the queue kernel.
//
// Every Delta_l (l <= S) is also logged as (vertex, V_l word) pairs, and the
// single-source BFSes log the candidate's visited bitmap after each level:
// committing a candidate then rewrites the stored base levels in place
// (B_l := V_l on Delta_l, the endpoint sources' bits from their bitmaps)
// instead of running a new BFS.
#pragma once
#include "gg_dense.hpp"

#define GG_TASK_ACTIVE 0
#define GG_TASK_DONE 1
#define GG_TASK_DENSE 2
#define GG_SPARSE_MAX_K 64  // candidates per column workgroup (LDS bookkeeping)

struct SparseArgs {
    int N, D, W, K, lmax;
    int k0;                          // first candidate of this launch (swaps/results offset)
    int kg;                          // candidates per column workgroup (<= GG_SPARSE_MAX_K):
                                     // ceil(K / kg) groups of W column workgroups
    int grid;                        // workgroups of the dense-tail instances (>= groups * W)
    int block;                       // threads per workgroup of the tail instances (512, 1024)
    const int *adjR;                 // [N][D] base neighbours, row-major
    const gg_swap *swaps;            // [k0 + K]
    const unsigned long long *snap;  // [S][W][N] base words after levels 1..S
    int S;
    const ColPart *base_parts;       // [W] base column parts (ecc, levels)
    const long long *base_reach;     // [W][lmax+1] base per-level new pairs
    int cap;                         // Delta list capacity per task (list stride)
    int climit;                      // Delta entries kept before a task overflows (<= cap)
    int skcap;                       // Delta keys of the previous level staged in LDS (max)
    // per task t = k*W + w (k local to the launch), parity p = level & 1
    int *dl_key;                     // [T][2][cap] Delta vertices
    unsigned *dl_bits;               // [T][2][ceil(N/32)] Delta membership
    unsigned long long *dl_ovr;      // [T][2][N] Delta words (valid where the bit is set)
    long long *dcnt;                 // [k0+K][W][lmax+1] count deltas per level
    int *task_state;                 // [k0+K][W]
    int *task_end;                   // [k0+K][W] level of the fixed point
    int *task_flags;                 // [k0+K][W]
    int *task_ecc;                   // [k0+K][W] the candidate's column eccentricity (finalize)
    int2 *queue;                     // dense fallback tasks (global k, w)
    int *queue_count;
    int *queue_next;
    int *ep_next;                    // endpoint tasks claimed (kernels with a dense tail)
    ColPart *dense_parts;            // [k0+K][W] written by the dense queue kernel
    ResultRec *res;                  // [k0+K]
    // What committing one of the launch's candidates needs without a new BFS
    // (all null: not kept).  Per task and level l <= S the Delta_l entries,
    // and per endpoint task 4k+j of the dense tail the column's words after
    // each level and its per-level new pairs.
    int *lg_key;                     // [K*W][S][cap] Delta_l vertices
    unsigned long long *lg_word;     // [K*W][S][cap] their candidate words V_l
    int *lg_n;                       // [K*W][S] |Delta_l|
    unsigned long long *ep_snap;     // [4K][S][N]
    long long *ep_reach;             // [4K][lmax+1]
    // single-source BFS of endpoint j of candidate g on the candidate graph
    // (task 4g + j, g global)
    int *ss_cnt;                     // [4(k0+K)][lmax+1] new vertices per level
    int *ss_ecc;                     // [4(k0+K)] eccentricity of the source
    int *ss_flags;                   // [4(k0+K)] GG_FLAG_LIMIT
    // per task (g, w) with masked endpoint sources: their base reached pairs
    // after levels 0..S (popcount of the masked bits of B_l)
    int *mb_cnt;                     // [(k0+K)][W][S+1]
    int *old_ecc;                    // [W] the base columns' eccentricities as this launch
                                     // read them (the commit reads them there, so the
                                     // new parts can be written beside it)
    unsigned *ss_vis;                // [4K][S][ceil(N/32)] candidate visited after level l
    int undirected;                  // every candidate graph is undirected (top-down allowed)
};

// The endpoint sources of swap `sw` among column w's 64 sources (bits of the
// column word); rows may repeat in a swap, vertices count once.
// (Inlined with its loops unrolled, so a swap held in registers stays there:
// an out-of-line call would force the caller's copy into memory.)
__host__ __device__ __forceinline__ unsigned long long endpoint_mask(const gg_swap &sw, int w) {
    unsigned long long m = 0;
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (sw.v[j] >= 0 && (sw.v[j] >> 6) == w) m |= 1ull << (sw.v[j] & 63);
    return m;
}

// Task index of endpoint j of a swap, or -1 when it is unused or repeats an
// earlier endpoint's vertex.
__host__ __device__ __forceinline__ bool endpoint_first(const gg_swap &sw, int j) {
    if (sw.v[j] < 0) return false;
#pragma unroll
    for (int jj = 0; jj < j; jj++)
        if (sw.v[jj] == sw.v[j]) return false;
    return true;
}

// Row of candidate vertex v with the swap's replacements applied: DT > 0
// loads the whole row as 16-byte vectors (one memory latency), DT = 0 loads
// slot by slot.
// Whether v is one of the swap's endpoints (rows other than those keep the
// base neighbours).
__device__ __forceinline__ bool swap_touches(const SwapRegs &sw, int v) {
    return v == sw.v[0] || v == sw.v[1] || v == sw.v[2] || v == sw.v[3];
}

template <int DT>
struct CandRow {
    int u[DT > 0 ? DT : 1];
    __device__ __forceinline__ void load(const int *adjR, int v, const SwapRegs &sw) {
        const int4 *r = reinterpret_cast<const int4 *>(adjR + (size_t)v * DT);
#pragma unroll
        for (int q = 0; q < DT / 4; q++) {
            int4 x = r[q];
            u[4 * q] = x.x; u[4 * q + 1] = x.y; u[4 * q + 2] = x.z; u[4 * q + 3] = x.w;
        }
        // (unconditional: a branch around these compare-selects costs the
        // column code more in spilled scalar registers than it saves)
#pragma unroll
        for (int i = 0; i < DT; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (sw.v[j] == v && sw.s[j] == i) u[i] = sw.n[j];
    }
};

// Candidate neighbour i of v (base row with the swap's slot replacements).
__device__ __forceinline__ int cand_nbr(const int *adjR, int D, int v, int i, const SwapRegs &sw) {
    int u = adjR[(size_t)v * D + i];
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (sw.v[j] == v && sw.s[j] == i) u = sw.n[j];
    return u;
}

// Word of x at the previous level: the task's override where Delta_{l-1}
// holds x, else the base word in LDS.  `ubits` (LDS) is the union of
// Delta_{l-1} over the column's tasks: outside it the base word is the
// answer without touching the task's bitmap in global memory.
template <class BW>
__device__ __forceinline__ unsigned long long prev_word(const unsigned *ubits, const unsigned *bits,
                                                        const unsigned long long *ovr,
                                                        const BW &Bprev, int x) {
    if (!((ubits[x >> 5] >> (x & 31)) & 1u)) return Bprev[x];
    // the membership word and the override are loaded together (one memory
    // latency instead of two); ovr[x] is in bounds and simply discarded
    // where the task's bit is clear
    const unsigned bw = bits[x >> 5];
    const unsigned long long o = ovr[x];
    return ((bw >> (x & 31)) & 1u) ? o : Bprev[x];
}

// Single-source BFS from x, by the whole workgroup, in `lds` (at least
// 12 * ceil(N/32) + 8 + 8 * lc bytes): visited / frontier / next bitmaps and
// two frontier lists of lc entries.  A level runs top-down from the frontier
// list (each listed vertex ORs its neighbours' bits in; allowed only on an
// undirected graph) while the list holds the frontier and the frontier's
// edges are fewer than twice the unvisited vertices, else bottom-up (each
// unvisited vertex looks for a neighbour in the frontier bitmap: the pull
// semantics of the evaluators).  Writes the new vertices per level to
// cnt[0..lmax] (cnt[0] = 1), the visited bitmap after levels 1..S to
// vis_log (when set; levels after the last are the final set) and returns
// the eccentricity of x; *limit = 1 when level lmax still found vertices.
template <int BS, int DT>
__device__ int ss_bfs(const int *adjR, int N, int D, int lmax, int S, int x,
                      const SwapRegs &sw, bool topdown_ok, unsigned *lds, int lc, int *cnt,
                      unsigned *vis_log, int *limit) {
    __shared__ int s_new, s_len[2], s_un;
    const int tid = threadIdx.x;
    const int NB = (N + 31) >> 5, NBE = (NB + 1) & ~1;
    unsigned *vis = lds, *front = lds + NBE, *nxt = lds + 2 * NBE;
    int *lst0 = reinterpret_cast<int *>(lds + 3 * NBE), *lst1 = lst0 + lc;
    for (int i = tid; i < NBE; i += BS) { vis[i] = 0u; front[i] = 0u; nxt[i] = 0u; }
    __syncthreads();
    if (tid == 0) {
        vis[x >> 5] = 1u << (x & 31);
        front[x >> 5] = 1u << (x & 31);
        lst0[0] = x;
        s_len[0] = 1;
        cnt[0] = 1;
    }
    __syncthreads();
    int reached = 1, fsize = 1, in = 0, ecc = 0, l = 1;
    bool list_ok = true;
    *limit = 0;
    for (; l <= lmax; l++) {
        const bool td = topdown_ok && list_ok && (long long)fsize * D <= 2LL * (N - reached);
        int *lin = in ? lst1 : lst0, *lout = in ? lst0 : lst1;
        if (tid == 0) { s_new = 0; s_len[in ^ 1] = 0; }
        __syncthreads();
        // a new vertex: counted, appended to the next list while it fits
        auto append = [&](bool isnew, int y, bool list) {
            if (!isnew) return;
            atomicAdd(&s_new, 1);
            if (!list) return;
            const int pos = atomicAdd(&s_len[in ^ 1], 1);
            if (pos < lc) lout[pos] = y;
        };
        auto claim = [&](int y) {
            const unsigned bit = 1u << (y & 31);
            if (vis[y >> 5] & bit) return false;
            const unsigned old = atomicOr(&vis[y >> 5], bit);
            if (old & bit) return false;
            atomicOr(&nxt[y >> 5], bit);
            return true;
        };
        // the lanes of a wave take consecutive entries
        const int wbase = tid & ~63, lane = tid & 63;
        if (td) {
            // a frontier whose edges fit one pass of the workgroup: a lane
            // per (vertex, slot) item; a larger one (fixed degrees): a lane
            // per four consecutive slots of a frontier vertex's row (one
            // 16-byte load)
            if (DT > 0 && fsize * D > BS) {
                // two quads per lane, both row loads in flight together
                constexpr int Q = DT > 0 ? DT / 4 : 1;
                const int items = fsize * Q;
                for (int i0 = 2 * wbase; i0 < items; i0 += 2 * BS) {
                    int y4[2][4];
                    bool ok[2];
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const int it = i0 + 64 * h + lane;
                        ok[h] = it < items;
                        y4[h][0] = y4[h][1] = y4[h][2] = y4[h][3] = 0;
                        if (ok[h]) {
                            const int e = it / Q, q = it - e * Q;
                            const int v = lin[e];
                            const int4 x = reinterpret_cast<const int4 *>(adjR + (size_t)v * D)[q];
                            y4[h][0] = x.x; y4[h][1] = x.y; y4[h][2] = x.z; y4[h][3] = x.w;
                            if (swap_touches(sw, v)) {
#pragma unroll
                                for (int i = 0; i < 4; i++)
#pragma unroll
                                    for (int j = 0; j < 4; j++)
                                        if (sw.v[j] == v && sw.s[j] == 4 * q + i) y4[h][i] = sw.n[j];
                            }
                        }
                    }
#pragma unroll
                    for (int h = 0; h < 2; h++)
#pragma unroll
                        for (int i = 0; i < 4; i++)
                            append(ok[h] && claim(y4[h][i]), y4[h][i], true);
                }
            } else {
                const int items = fsize * D;
                for (int i0 = wbase; i0 < items; i0 += BS) {
                    const int it = i0 + lane;
                    const bool ok = it < items;
                    int y = 0;
                    if (ok) {
                        const int e = it / D, i = it - e * D;
                        y = cand_nbr(adjR, D, lin[e], i, sw);
                    }
                    append(ok && claim(y), y, true);
                }
            }
        } else {
            // the unvisited vertices listed first (a lane per vertex then
            // waits for one row load, not one per vertex it scans)
            if (tid == 0) s_un = 0;
            __syncthreads();
            for (int i = tid; i < NB; i += BS) {
                unsigned u = ~vis[i];
                if (i == NB - 1 && (N & 31)) u &= (1u << (N & 31)) - 1u;
                if (!u) continue;
                int pos = atomicAdd(&s_un, __popc(u));
                while (u) {
                    const int b = __builtin_ctz(u);
                    u &= u - 1u;
                    if (pos < lc) lout[pos] = 32 * i + b;
                    pos++;
                }
            }
            __syncthreads();
            const int nun = s_un;
            const bool listed = nun <= lc;
            const int span = listed ? nun : N;
            if constexpr (DT > 0) {
                // two vertices per lane: the first eight slots of both rows
                // loaded together (a bottom-up frontier is large, so a vertex
                // nearly always meets it there), the rest only where not
                auto quad_hit = [&](int v, int q, int4 x) {
                    int y[4] = {x.x, x.y, x.z, x.w};
                    if (swap_touches(sw, v)) {
#pragma unroll
                        for (int i = 0; i < 4; i++)
#pragma unroll
                            for (int j = 0; j < 4; j++)
                                if (sw.v[j] == v && sw.s[j] == 4 * q + i) y[i] = sw.n[j];
                    }
                    bool hit = false;
#pragma unroll
                    for (int i = 0; i < 4; i++) hit = hit || ((front[y[i] >> 5] >> (y[i] & 31)) & 1u);
                    return hit;
                };
                constexpr int Q = DT / 4;
                for (int v0 = 2 * wbase; v0 < span; v0 += 2 * BS) {
                    int vv[2];
                    bool un[2];
                    int4 x0[2], x1[2];
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const int e = v0 + 64 * h + lane;
                        vv[h] = listed ? (e < nun ? lout[e] : N) : e;
                        un[h] = vv[h] < N && !((vis[vv[h] >> 5] >> (vv[h] & 31)) & 1u);
                        x0[h] = x1[h] = make_int4(0, 0, 0, 0);
                        if (un[h]) {
                            const int4 *r4 = reinterpret_cast<const int4 *>(adjR + (size_t)vv[h] * D);
                            x0[h] = r4[0];
                            if (Q > 1) x1[h] = r4[1];
                        }
                    }
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        bool hit = false;
                        if (un[h]) {
                            const int v = vv[h];
                            hit = quad_hit(v, 0, x0[h]) || (Q > 1 && quad_hit(v, 1, x1[h]));
                            const int4 *r4 = reinterpret_cast<const int4 *>(adjR + (size_t)v * D);
                            for (int q = 2; q < Q && !hit; q++) hit = quad_hit(v, q, r4[q]);
                            if (hit) atomicOr(&nxt[v >> 5], 1u << (v & 31));
                        }
                        append(hit, vv[h], false);
                    }
                }
            } else {
                for (int v0 = wbase; v0 < span; v0 += BS) {
                    const int e = v0 + lane;
                    const int v = listed ? (e < nun ? lout[e] : N) : e;
                    bool hit = false;
                    if (v < N && !((vis[v >> 5] >> (v & 31)) & 1u)) {
                        for (int i = 0; i < D && !hit; i++) {
                            const int y = cand_nbr(adjR, D, v, i, sw);
                            hit = (front[y >> 5] >> (y & 31)) & 1u;
                        }
                        if (hit) atomicOr(&nxt[v >> 5], 1u << (v & 31));
                    }
                    append(hit, v, false);
                }
            }
        }
        __syncthreads();
        const int nnew = s_new;
        const bool lok = td && s_len[in ^ 1] <= lc;
        for (int i = tid; i < NBE; i += BS) {
            const unsigned n = nxt[i];
            const unsigned vv = td ? vis[i] : (vis[i] | n);
            vis[i] = vv;
            front[i] = n;
            nxt[i] = 0u;
            if (vis_log && l <= S && i < NB) vis_log[(size_t)(l - 1) * NB + i] = vv;
        }
        if (tid == 0) cnt[l] = nnew;
        __syncthreads();
        if (nnew == 0) break;
        ecc = l;
        reached += nnew;
        fsize = nnew;
        list_ok = lok;
        in ^= 1;
    }
    if (l > lmax) *limit = 1;
    // levels after the last: the final visited set
    if (vis_log)
        for (int ll = (l > lmax ? lmax : l) + 1; ll <= S; ll++)
            for (int i = tid; i < NB; i += BS) vis_log[(size_t)(ll - 1) * NB + i] = vis[i];
    __syncthreads();
    return ecc;
}

// Dynamic LDS of k_sparse: B_{l-1}[N] words, the Delta union bitmap, the
// staged Delta keys.
__host__ __device__ inline size_t sparse_lds_bytes(int N, int skcap) {
    const size_t nb = ((size_t)(N + 31) / 32 + 1) & ~(size_t)1;  // even: keeps 8-byte alignment
    return (size_t)N * 8 + nb * 4 + (size_t)skcap * 4;
}

// TAIL = 1 (fixed-degree instances): after its column, each workgroup runs
// the single-source BFS tasks of the swap endpoints (candidate k, endpoint
// j -> task 4k+j, claimed from a.ep_next; the list follows from the swaps
// alone) in the same LDS; workgroups beyond the W columns (launched on spare
// CUs) start on those tasks at once.  TAIL = 0 queues the columns holding an
// endpoint for k_dense_lds_queue instead.
template <int BS, int DT, int TAIL>
__global__ void __launch_bounds__(BS) k_sparse(SparseArgs a) {
    extern __shared__ unsigned long long Bprev[];
    __shared__ gg_swap s_sw[GG_SPARSE_MAX_K];     // the group's swaps
    __shared__ int s_state[GG_SPARSE_MAX_K];      // task state of this column
    __shared__ int s_cnt[2][GG_SPARSE_MAX_K];     // |Delta| per parity
    __shared__ int s_koff[GG_SPARSE_MAX_K + 1];   // staged Delta keys: offset per candidate
    __shared__ int s_ooff[GG_SPARSE_MAX_K + 1];   // stale Delta_{l-2} entries: offset per candidate
    __shared__ int s_nold[GG_SPARSE_MAX_K];
    __shared__ int s_staged;
    __shared__ int s_items[GG_SPARSE_MAX_K + 1];  // prefix sums of items per candidate
    __shared__ int s_kc[GG_SPARSE_MAX_K];         // keys that spawn items (compacted levels)
    __shared__ int s_nprev[GG_SPARSE_MAX_K];
    __shared__ int s_ncur[GG_SPARSE_MAX_K];
    __shared__ int s_active[GG_SPARSE_MAX_K];     // 1 if the task runs this level
    __shared__ int s_over[GG_SPARSE_MAX_K];
    __shared__ int s_diff[GG_SPARSE_MAX_K];
    __shared__ unsigned long long s_dc[GG_SPARSE_MAX_K];
    __shared__ unsigned long long s_lost[GG_SPARSE_MAX_K];  // bits lost on Delta_{l-1}
    __shared__ unsigned long long s_endmask;      // tasks that ended last level: clear their bitmaps
    __shared__ unsigned long long s_maskset;      // tasks with masked endpoint sources
    __shared__ unsigned long long s_mask[GG_SPARSE_MAX_K];  // endpoint sources' bits (masked)
    __shared__ int s_mcnt[GG_SPARSE_MAX_K];       // base pairs of the masked bits in B_{l-1}
    __shared__ int s_nmasked;
    __shared__ int s_next;
    __shared__ int s_any;
    const int tid = threadIdx.x, lane = tid & 63;
    const int N = a.N, D = a.D, W = a.W;
    // workgroup (group gi, column w) runs the group's candidates kb..kb+K-1
    // of the launch on column w; workgroups beyond the groups' columns
    // (dense-tail instances launched with spare CUs) go straight to the
    // endpoint tasks
    const int groups = (a.K + a.kg - 1) / a.kg;
    const bool has_col = (int)blockIdx.x < groups * W;
    const int gi = has_col ? (int)blockIdx.x / W : 0;
    const int w = has_col ? (int)blockIdx.x - gi * W : 0;
    const int kb = gi * a.kg;                                     // launch-local
    const int K = has_col ? (a.K - kb < a.kg ? a.K - kb : a.kg) : 0;
    const int kc = a.k0 + kb;                                     // global
    const int NB = (N + 31) >> 5;
    const int ecc_w = a.base_parts[w].ecc;
    if (has_col && gi == 0 && threadIdx.x == 0 && a.old_ecc) a.old_ecc[w] = ecc_w;
    const unsigned long long pad = pad_mask(w, N);
    const int NBE = (NB + 1) & ~1;
    unsigned *s_ubits = reinterpret_cast<unsigned *>(Bprev + N);  // [NBE]
    int *s_keys = reinterpret_cast<int *>(s_ubits + NBE);          // [skcap]
    for (int k = tid; k < K; k += BS) s_sw[k] = a.swaps[kc + k];
    auto gt = [&](int k) { return (size_t)(kc + k) * W + w; };       // global (cand, col)
    auto lt = [&](int k) { return (size_t)(kb + k) * W + w; };       // task slot of this launch
    if (tid == 0) { s_nmasked = 0; s_endmask = 0; s_maskset = 0; }
    __syncthreads();
    if (has_col) {

    // task init: a column holding a swap endpoint as a source runs with the
    // endpoint bits masked (the tail's single-source BFSes supply them), or
    // densely in the queue kernel when there is no tail
    for (int k = tid; k < K; k += BS) {
        const gg_swap &s = a.swaps[kc + k];
        const unsigned long long em = endpoint_mask(s, w);
        const bool dense = TAIL == 0 && em != 0;
        s_mask[k] = TAIL ? em : 0ull;
        if (TAIL && em) {
            atomicAdd(&s_nmasked, 1);
            atomicOr(&s_maskset, 1ull << k);
        }
        s_cnt[0][k] = 0;
        s_cnt[1][k] = 0;
        a.task_flags[gt(k)] = 0;
        a.task_end[gt(k)] = 0;
        s_state[k] = dense ? GG_TASK_DENSE : GG_TASK_ACTIVE;
        a.task_state[gt(k)] = s_state[k];
        if (dense) {
            int q = atomicAdd(a.queue_count, 1);
            a.queue[q] = make_int2(kc + k, w);
        }
    }
    __threadfence_block();
    __syncthreads();
    // Bitmaps are all zero between launches: a task clears both of its
    // bitmaps at the level after it ends (and after the last level).
    auto clear_ended = [&]() {
        for (unsigned long long m = s_endmask; m; m &= m - 1) {
            unsigned *b0 = a.dl_bits + lt(__builtin_ctzll(m)) * 2 * NB;
            for (int i = tid; i < 2 * NB; i += BS) b0[i] = 0u;
        }
    };

    for (int l = 1; l <= a.lmax; l++) {
        const int par = (l - 1) & 1, cur = l & 1;
        const int lp = l - 1;
        // B_{l-1} of this column into LDS (stored levels: 16-byte loads,
        // eight per thread in flight before the LDS stores)
        if (lp >= 1 && lp < ecc_w && (N & 1) == 0) {
            const uint4 *src =
                reinterpret_cast<const uint4 *>(a.snap + ((size_t)(lp - 1) * W + w) * N);
            uint4 *dst = reinterpret_cast<uint4 *>(Bprev);
            const int n2 = N >> 1;
#pragma unroll 8
            for (int i = tid; i < n2; i += BS) dst[i] = src[i];
        } else {
            for (int v = tid; v < N; v += BS) {
                unsigned long long b;
                if (lp == 0) b = pad | (((v >> 6) == w) ? (1ull << (v & 63)) : 0ull);
                else if (lp >= ecc_w) b = ~0ull;
                else b = a.snap[((size_t)(lp - 1) * W + w) * N + v];
                Bprev[v] = b;
            }
        }
        for (int i = tid; i < NBE; i += BS) s_ubits[i] = 0u;
        // per-candidate bookkeeping
        for (int k = tid; k < K; k += BS) {
            const bool act = s_state[k] == GG_TASK_ACTIVE;
            s_active[k] = act;
            s_nprev[k] = act ? s_cnt[par][k] : 0;
            s_nold[k] = act ? s_cnt[cur][k] : 0;
            s_ncur[k] = 0;
            s_over[k] = 0;
            s_diff[k] = 0;
            s_dc[k] = 0;
            s_lost[k] = 0;
            s_mcnt[k] = 0;
            s_kc[k] = 0;
        }
        if (tid == 0) { s_next = 0; s_any = 0; }
        clear_ended();
        __syncthreads();
        if (tid == 0) s_endmask = 0;
        if (s_nmasked) {
            // base pairs of the masked endpoint bits after level l-1: what the
            // finalize replaces with the single-source counts
            for (unsigned long long mm = s_maskset; mm; mm &= mm - 1) {
                const int k = __builtin_ctzll(mm);
                if (!s_active[k]) continue;
                const unsigned long long m = s_mask[k];
                int c = 0;
                for (int v = tid; v < N; v += BS) c += __popcll(Bprev[v] & m);
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
                if (lane == 0 && c) atomicAdd(&s_mcnt[k], c);
            }
        }
        if (tid < 64) {  // prefix sums over the candidates (K <= 64), one wave
            const bool act = tid < K && s_active[tid];
            int it = act ? 4 + s_nprev[tid] * (D + 1) : 0;
            int kk = act ? s_nprev[tid] : 0;
            int oo = act ? s_nold[tid] : 0;
            const int it0 = it, kk0 = kk, oo0 = oo;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t1 = __shfl_up(it, o, 64), t2 = __shfl_up(kk, o, 64),
                          t3 = __shfl_up(oo, o, 64);
                if (lane >= o) { it += t1; kk += t2; oo += t3; }
            }
            if (tid < K) { s_items[tid] = it - it0; s_koff[tid] = kk - kk0; s_ooff[tid] = oo - oo0; }
            if (tid == 63) {
                s_items[K] = it;
                s_koff[K] = kk;
                s_ooff[K] = oo;
                s_staged = kk <= a.skcap;
            }
        }
        __syncthreads();
        // last k with off[k] <= i (empty candidates share their successor's start)
        auto owner = [&](const int *off, int i) {
            int k = 0, hi = K;
            while (hi - k > 1) {
                const int mid = (k + hi) >> 1;
                if (off[mid] <= i) k = mid;
                else hi = mid;
            }
            return k;
        };
        // flat over (candidate, entry): clear the stale bits of the cur parity
        // (Delta_{l-2}) so Delta_l starts empty; build the union of the
        // Delta_{l-1} (LDS bitmap) and stage their keys
        // At and above the column's eccentricity only the keys that lost bits
        // spawn items: with staged keys those are packed to the front of their
        // candidate's key range and the item counts recomputed, so the pool
        // deals out no items that would be skipped.
        const bool compact = s_staged && l >= ecc_w;
        {
            const int nold_all = s_ooff[K], nprev_all = s_koff[K];
            const bool stg = s_staged;
            for (int i = tid; i < nold_all; i += BS) {
                const int k = owner(s_ooff, i);
                const int x = a.dl_key[(lt(k) * 2 + cur) * a.cap + (i - s_ooff[k])];
                atomicAnd(&a.dl_bits[(lt(k) * 2 + cur) * NB + (x >> 5)], ~(1u << (x & 31)));
            }
            for (int i = tid; i < nprev_all; i += BS) {
                const int k = owner(s_koff, i);
                const int x = a.dl_key[(lt(k) * 2 + par) * a.cap + (i - s_koff[k])];
                atomicOr(&s_ubits[x >> 5], 1u << (x & 31));
                int xs = x;
                if (l >= ecc_w - 1) {
                    // Where B_l[v] is full, a bit can be missing from V_l[v] (v
                    // not a swap endpoint) only if every closed neighbour of v
                    // that had it in B_{l-1} is in Delta_{l-1} and lost it: the
                    // lost bits are collected.  At and above the column's
                    // eccentricity every B_l word is full, and keys that lost
                    // nothing spawn no items; one level below most are.
                    const unsigned long long lost =
                        Bprev[x] & ~a.dl_ovr[(lt(k) * 2 + par) * (size_t)N + x] & ~s_mask[k];
                    if (lost) atomicOr(&s_lost[k], lost);
                    else if (l >= ecc_w) xs = -1;
                }
                if (compact) {
                    if (xs >= 0) s_keys[s_koff[k] + atomicAdd(&s_kc[k], 1)] = xs;
                } else if (stg) {
                    s_keys[i] = xs;
                }
            }
        }
        __threadfence_block();
        __syncthreads();
        if (compact) {
            if (tid < 64) {  // item prefix sums over the packed keys, one wave
                const bool act = tid < K && s_active[tid];
                int it = act ? 4 + s_kc[tid] * (D + 1) : 0;
                const int it0 = it;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int t1 = __shfl_up(it, o, 64);
                    if (lane >= o) it += t1;
                }
                if (tid < K) s_items[tid] = it - it0;
                if (tid == 63) s_items[K] = it;
            }
            __syncthreads();
        }
        const int total = s_items[K];
        const bool staged = s_staged;
        // shared pool of items over all active candidates
        for (;;) {
            int base_it = 0;
            if (lane == 0) base_it = atomicAdd(&s_next, 64);
            base_it = __shfl(base_it, 0, 64);
            if (base_it >= total) break;
            const int it = base_it + lane;
            if (it >= total) continue;
            const int k = owner(s_items, it);
            const int loc = it - s_items[k];
            const size_t tl = lt(k);
            SwapRegs sw;
            bool has_sw;
            load_swap(s_sw, k, sw, has_sw);
            const unsigned *pbits = a.dl_bits + (tl * 2 + par) * NB;
            const unsigned long long *povr = a.dl_ovr + (tl * 2 + par) * (size_t)N;
            int v;
            if (loc < 4) {
                v = loc == 0 ? sw.v[0] : loc == 1 ? sw.v[1] : loc == 2 ? sw.v[2] : sw.v[3];
                if (v < 0) continue;
            } else {
                const int e = (loc - 4) / (D + 1), i = (loc - 4) - e * (D + 1);
                const int u = staged ? s_keys[s_koff[k] + e] : a.dl_key[(tl * 2 + par) * a.cap + e];
                if (u < 0) continue;  // lost no bit (levels where B_l is full)
                v = i == 0 ? u : cand_nbr(a.adjR, D, u, i - 1, sw);
            }
            // masked endpoint sources: their bits take no part in any word
            // compared, counted or logged
            const unsigned long long keep = ~s_mask[k];
            unsigned long long basew = ~0ull;  // issued before the gather
            if (l < ecc_w) basew = a.snap[((size_t)(l - 1) * W + w) * N + v];
            basew &= keep;
            unsigned long long acc;
            const bool none = s_nprev[k] == 0;
            if constexpr (DT > 0) {
                const bool inE = v == sw.v[0] || v == sw.v[1] || v == sw.v[2] || v == sw.v[3];
                unsigned long long need = ~0ull;
                if (l >= ecc_w && !none) {
                    // B_l is full here, so the item only has to show V_l[v]
                    // full.  Off the swap's endpoints only the task's lost bits
                    // can be missing (none lost: nothing to show).
                    need = inE ? keep : s_lost[k];
                    if (!need) continue;
                }
                CandRow<DT> row;
                row.load(a.adjR, v, sw);
                // one level below the column's eccentricity: the same for the
                // vertices whose B_l word is full (the base word and the row
                // are loaded together)
                const bool bfull = !none && (l >= ecc_w || (l == ecc_w - 1 && basew == keep));
                if (bfull && l < ecc_w) {
                    need = inE ? keep : s_lost[k];
                    if (!need) continue;
                }
                if (bfull) {
                    // words of neighbours outside the Delta union are B_{l-1}
                    // (LDS), a lower bound of the OR; groups of 4 until the
                    // needed bits are covered.  Otherwise the exact OR below.
                    auto outside = [&](int x) {
                        return ((s_ubits[x >> 5] >> (x & 31)) & 1u) ? 0ull : Bprev[x];
                    };
                    acc = outside(v);
                    bool full = (acc & need) == need;
#pragma unroll
                    for (int g = 0; g < DT && !full; g += 4) {
#pragma unroll
                        for (int i = g; i < (g + 4 < DT ? g + 4 : DT); i++) acc |= outside(row.u[i]);
                        full = (acc & need) == need;
                    }
                    if (full) continue;
                }
                if (none) {
                    acc = Bprev[v];
#pragma unroll
                    for (int i = 0; i < DT; i++) acc |= Bprev[row.u[i]];
                } else {
                    acc = prev_word(s_ubits, pbits, povr, Bprev, v);
#pragma unroll
                    for (int i = 0; i < DT; i++)
                        acc |= prev_word(s_ubits, pbits, povr, Bprev, row.u[i]);
                }
            } else {
                if (none) {
                    acc = Bprev[v];
                    for (int i = 0; i < D; i++) acc |= Bprev[cand_nbr(a.adjR, D, v, i, sw)];
                } else {
                    acc = prev_word(s_ubits, pbits, povr, Bprev, v);
                    for (int i = 0; i < D; i++)
                        acc |= prev_word(s_ubits, pbits, povr, Bprev, cand_nbr(a.adjR, D, v, i, sw));
                }
            }
            acc &= keep;
            if (acc == basew) continue;
            unsigned *cbits = a.dl_bits + (tl * 2 + cur) * NB;
            const unsigned m = 1u << (v & 31);
            const unsigned old = atomicOr(&cbits[v >> 5], m);
            if (old & m) continue;  // another item of this task recorded v
            const int idx = atomicAdd(&s_ncur[k], 1);
            if (idx >= a.climit) { s_over[k] = 1; continue; }
            a.dl_key[(tl * 2 + cur) * a.cap + idx] = v;
            a.dl_ovr[(tl * 2 + cur) * (size_t)N + v] = acc;
            if (a.lg_key && l <= a.S) {
                const size_t li = (tl * a.S + (l - 1)) * a.cap + idx;
                a.lg_key[li] = v;
                a.lg_word[li] = acc;
            }
            atomicAdd(&s_dc[k], (unsigned long long)((long long)__popcll(acc) -
                                                      (long long)__popcll(basew)));
            // whether Delta_l differs from Delta_{l-1} matters only above the
            // column's eccentricity (the fixed-point test)
            if (l > ecc_w && (none || !((pbits[v >> 5] >> (v & 31)) & 1u) || povr[v] != acc))
                s_diff[k] = 1;
        }
        __threadfence_block();
        __syncthreads();
        for (int k = tid; k < K; k += BS) {
            if (!s_active[k]) continue;
            const size_t g = gt(k);
            if (s_over[k]) {
                // its levels are not kept: committing this candidate needs a new BFS
                s_state[k] = GG_TASK_DENSE;
                a.task_state[g] = GG_TASK_DENSE;
                a.task_flags[g] = GG_FLAG_QUEUED;
                int q = atomicAdd(a.queue_count, 1);
                a.queue[q] = make_int2(kc + k, w);
                atomicOr(&s_endmask, 1ull << k);
                continue;
            }
            const int n_cur = s_ncur[k];
            a.dcnt[g * (a.lmax + 1) + l] = (long long)s_dc[k];
            if (s_mask[k] && l - 1 <= a.S) a.mb_cnt[g * (a.S + 1) + (l - 1)] = s_mcnt[k];
            if (a.lg_n && l <= a.S) a.lg_n[lt(k) * a.S + (l - 1)] = n_cur;
            s_cnt[cur][k] = n_cur;
            const bool fixed = (l >= ecc_w && n_cur == 0) ||
                               (l > ecc_w && !s_diff[k] && n_cur == s_nprev[k]);
            if (fixed || l >= a.lmax) {
                s_state[k] = GG_TASK_DONE;
                a.task_state[g] = GG_TASK_DONE;
                a.task_end[g] = l;
                if (!fixed) a.task_flags[g] = GG_FLAG_LIMIT;
                atomicOr(&s_endmask, 1ull << k);
            } else {
                atomicOr(&s_any, 1);
            }
        }
        __threadfence_block();
        __syncthreads();
        if (!s_any) break;
    }
    clear_ended();
    }  // has_col
    if constexpr (TAIL > 0) {
        // tail: a single-source BFS from every endpoint of every candidate
        // of the launch on the candidate graph (task 4k + j)
        __shared__ int s_task;
        const int ntask = 4 * a.K;
        const gg_swap *lsw = a.swaps + a.k0;  // the launch's swaps
        const int NB = (N + 31) >> 5;
        // frontier lists in the LDS left after the three bitmaps
        const int lc = (int)(((size_t)N * 8 - 12 * (size_t)((NB + 1) & ~1)) / 8);
        unsigned *scratch = reinterpret_cast<unsigned *>(Bprev);
        for (;;) {
            __syncthreads();  // the LDS is free again
            if (tid == 0) {
                int t;
                for (;;) {
                    t = atomicAdd(a.ep_next, 1);
                    if (t >= ntask) break;
                    if (endpoint_first(lsw[t >> 2], t & 3)) break;
                }
                s_task = t;
            }
            __syncthreads();
            const int t = __builtin_amdgcn_readfirstlane(s_task);
            if (t >= ntask) break;
            const int k = t >> 2, x = lsw[k].v[t & 3];
            const size_t gt4 = (size_t)4 * a.k0 + t;
            const int L1 = a.lmax + 1;
            SwapRegs sw;
            bool has;
            load_swap(lsw, k, sw, has);
            int lim_c = 0;
            const int ec = ss_bfs<BS, DT>(a.adjR, N, D, a.lmax, a.S, x, sw, a.undirected,
                                          scratch, lc, a.ss_cnt + gt4 * L1,
                                          a.ss_vis ? a.ss_vis + (size_t)t * a.S * NB : nullptr,
                                          &lim_c);
            if (tid == 0) {
                a.ss_ecc[gt4] = ec;
                a.ss_flags[gt4] = lim_c ? GG_FLAG_LIMIT : 0;
            }
        }
    }
}

// Candidate k's result, by one workgroup: merges sparse column results (base
// counts + deltas) and dense fallback parts into its ResultRec (AGENT: stored
// for another workgroup of the same launch, complete on return).
template <int BS, bool AGENT = false>
__device__ __forceinline__ void sparse_finalize_body(const SparseArgs &a, int nvalid_last, int k) {
    __shared__ long long s_tot[BS / 64], s_rch[BS / 64];
    __shared__ int s_diam[BS / 64], s_lev[BS / 64], s_fl[BS / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int W = a.W, L1 = a.lmax + 1;
    long long tot = 0, rch = 0;
    int diam = 0, lev = 0, fl = 0;
    // the candidate's swap and its endpoints' single-source results (the
    // same for every thread): loaded once, up front
    constexpr int kLU = 8;
    static_assert(GG_LEVEL_LIMIT >= kLU, "rows hold at least kLU + 1 levels");
    const gg_swap sk = a.swaps[k];
    int sse[4], ssf[4], ssc[4][kLU];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const size_t t4 = (size_t)4 * k + j;
        sse[j] = a.ss_ecc[t4];
        ssf[j] = a.ss_flags[t4];
#pragma unroll
        for (int i = 0; i < kLU; i++) ssc[j][i] = a.ss_cnt[t4 * L1 + i + 1];
    }
    for (int w = tid; w < W; w += BS) {
        const size_t t = (size_t)k * W + w;
        const long long *br = a.base_reach + (size_t)w * L1;
        const long long *dd = a.dcnt + t * L1;
        // every load of the first kLU levels issued with the task's state (the
        // entries past the column's levels are masked below, not skipped)
        const int ts = a.task_state[t];
        const int end = a.task_end[t];
        const ColPart bp = a.base_parts[w];
        const int tflags = a.task_flags[t];
        long long pbr[kLU], pdd[kLU];
        int pmb[kLU];  // base pairs of the masked bits (used by endpoint columns)
        const int *mb = a.mb_cnt + t * (a.S + 1);
#pragma unroll
        for (int i = 0; i < kLU; i++) {
            pbr[i] = br[i + 1];
            pdd[i] = dd[i + 1];
            pmb[i] = mb[i + 1 <= a.S ? i + 1 : a.S];
        }
        if (ts == GG_TASK_DENSE) {
            const ColPart &p = a.dense_parts[t];
            a.task_ecc[t] = p.ecc;
            tot += p.total; rch += p.reached;
            diam = p.ecc > diam ? p.ecc : diam;
            lev = p.levels > lev ? p.levels : lev;
            fl |= p.flags | (tflags & GG_FLAG_QUEUED);
            continue;
        }
        // endpoint sources of this column (masked in the sparse pass): their
        // single-source counts replace the base's
        const unsigned long long em = endpoint_mask(sk, w);
        const long long nm = __popcll(em);
        int top = end, sfl = 0;
        bool mine[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            mine[j] = em && endpoint_first(sk, j) && (sk.v[j] >> 6) == w;
            if (!mine[j]) continue;
            top = sse[j] > top ? sse[j] : top;
            sfl |= ssf[j];
        }
        if (top > a.lmax) top = a.lmax;
        long long cb = (w == W - 1) ? nvalid_last : 64;  // base reached after level 0
        long long cprev = cb, ctot = 0, cx = nm, dend = 0;
        int ecc = 0;
        long long pcx[kLU], pbx[kLU];
#pragma unroll
        for (int i = 0; i < kLU; i++) {
            const int l = i + 1;
            if (!(l <= bp.levels && l < L1)) pbr[i] = 0;
            if (!(l <= end)) pdd[i] = 0;
            pcx[i] = 0;
            pbx[i] = 0;
            if (em && l <= top) {
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (mine[j] && l <= sse[j]) pcx[i] += ssc[j][i];
                pbx[i] = l < bp.ecc ? (long long)pmb[i] : (long long)a.N * nm;
            }
            if (l == end) dend = pdd[i];
        }
        if (end > kLU) dend = dd[end];
#pragma unroll
        for (int i = 0; i < kLU; i++) {
            const int l = i + 1;
            if (l > top) break;
            cb += pbr[i];
            cx += pcx[i];
            // after the masked pass's fixed point its count delta stays
            const long long c = cb + (l <= end ? pdd[i] : dend) + (em ? cx - pbx[i] : 0);
            if (c > cprev) { ecc = l; ctot += (long long)l * (c - cprev); }
            cprev = c;
        }
        for (int l = kLU + 1; l <= top; l++) {
            if (l <= bp.levels) cb += br[l];
            long long bx = 0;
            if (em) {
                for (int j = 0; j < 4; j++) {
                    if (!mine[j]) continue;
                    const size_t t4 = (size_t)4 * k + j;
                    if (l <= sse[j]) cx += a.ss_cnt[t4 * L1 + l];
                }
                bx = l < bp.ecc ? a.mb_cnt[t * (a.S + 1) + l] : (long long)a.N * nm;
            }
            const long long c = cb + dd[l <= end ? l : end] + (em ? cx - bx : 0);
            if (c > cprev) { ecc = l; ctot += (long long)l * (c - cprev); }
            cprev = c;
        }
        a.task_ecc[t] = ecc;
        tot += ctot;
        rch += cprev;
        diam = ecc > diam ? ecc : diam;
        const int tf = tflags | sfl;
        int lv = tf ? top : ecc + 1;
        lev = lv > lev ? lv : lev;
        fl |= tf;
    }
    tot = wave_sum64(tot);
    rch = wave_sum64(rch);
    diam = wave_max32(diam);
    lev = wave_max32(lev);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) fl |= __shfl_xor(fl, o, 64);
    if (lane == 0) { s_tot[wid] = tot; s_rch[wid] = rch; s_diam[wid] = diam; s_lev[wid] = lev; s_fl[wid] = fl; }
    __syncthreads();
    if (tid == 0) {
        for (int i = 1; i < BS / 64; i++) {
            s_tot[0] += s_tot[i]; s_rch[0] += s_rch[i];
            s_diam[0] = s_diam[i] > s_diam[0] ? s_diam[i] : s_diam[0];
            s_lev[0] = s_lev[i] > s_lev[0] ? s_lev[i] : s_lev[0];
            s_fl[0] |= s_fl[i];
        }
        ResultRec r;
        r.total = s_tot[0];
        r.reached = s_rch[0];
        r.diameter = s_diam[0];
        r.levels = s_lev[0];
        r.flags = s_fl[0];
        r.wvc = 0;
        r.seq = 0;
        r.pad = 0;
        if (AGENT) {
            ResultRec *d = a.res + k;
            st_agent(&d->total, r.total);
            st_agent(&d->reached, r.reached);
            st_agent(&d->diameter, r.diameter);
            st_agent(&d->levels, r.levels);
            st_agent(&d->wvc, r.wvc);
            st_agent(&d->flags, r.flags);
            __builtin_amdgcn_s_waitcnt(0);
        } else {
            a.res[k] = r;
        }
    }
}

// One workgroup per candidate (sparse_finalize_body).
template <int BS>
__global__ void __launch_bounds__(BS) k_sparse_finalize(SparseArgs a, int nvalid_last) {
    sparse_finalize_body<BS>(a, nvalid_last, blockIdx.x);
}

// Resets the dense queue counters and the endpoint-task counters (after the
// dense kernel and the finalize).
static __global__ void k_queue_reset(int *ctl, int n) {
    if (blockIdx.x == 0)
        for (int i = threadIdx.x; i < n; i += blockDim.x) ctl[i] = 0;
}

// Launch of the sparse kernel of degree DT (fixed-degree instances live in
// gg_sparse_d*.hip); tail = 1 picks the instance with the single-source BFS
// tail (fixed degrees only).
template <int DT>
hipError_t launch_sparse_deg(const SparseArgs &s, size_t lds, int tail, hipStream_t st) {
    auto go = [&](auto fn, gg::LdsAttr &attr, int grid, int bs) -> hipError_t {
        hipError_t e = gg::ensure_dynamic_lds((const void *)fn, lds, attr);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(fn, dim3(grid), dim3(bs), lds, st, s);
        return hipGetLastError();
    };
    static gg::LdsAttr c0, c1, c2;
    if constexpr (DT > 0) {
        // tail instances may take extra workgroups; one per (group, column)
        // otherwise
        if (tail && s.block == 512) return go(&k_sparse<512, DT, 1>, c2, s.grid, 512);
        if (tail) return go(&k_sparse<1024, DT, 1>, c1, s.grid, 1024);
    }
    return go(&k_sparse<1024, DT, 0>, c0, (s.K + s.kg - 1) / s.kg * s.W, 1024);
}

namespace gg {
hipError_t launch_sparse_d4(const SparseArgs &, size_t, int, hipStream_t);
hipError_t launch_sparse_d8(const SparseArgs &, size_t, int, hipStream_t);
hipError_t launch_sparse_d12(const SparseArgs &, size_t, int, hipStream_t);
hipError_t launch_sparse_d16(const SparseArgs &, size_t, int, hipStream_t);
hipError_t launch_sparse_d24(const SparseArgs &, size_t, int, hipStream_t);
}  // namespace gg
