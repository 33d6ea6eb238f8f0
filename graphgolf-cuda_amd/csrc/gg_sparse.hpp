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
// One workgroup per column keeps B_{l-1}[w][*] in LDS and walks the levels;
// each wave handles some candidates of the column, holding Delta_{l-1} in a
// small LDS hash (Delta lists persist across levels in global scratch).
// Tasks whose Delta outgrows the capacity, and the columns that contain a swap
// endpoint as a source, are queued for the dense kernel (k_dense_lds_queue).
#pragma once
#include "gg_dense.hpp"

#define GG_TASK_ACTIVE 0
#define GG_TASK_DONE 1
#define GG_TASK_DENSE 2

struct SparseArgs {
    int N, D, W, K, lmax;
    const int *adjR;                 // [N][D] base neighbours, row-major
    const gg_swap *swaps;            // [K]
    const unsigned long long *snap;  // [S][W][N] base words after levels 1..S
    int S;
    const ColPart *base_parts;       // [W] base column parts (ecc, levels)
    const long long *base_reach;     // [W][lmax+1] base per-level new pairs
    int cap;                         // Delta list capacity per task
    int *dl_key;                     // [K*W][2][cap]
    unsigned long long *dl_val;      // [K*W][2][cap]
    int *dl_cnt;                     // [K*W][2]
    long long *dcnt;                 // [K*W][lmax+1] count deltas per level
    int *task_state;                 // [K*W]
    int *task_end;                   // [K*W] level of the fixed point
    int *task_flags;                 // [K*W]
    int2 *queue;                     // [K*W] dense fallback tasks
    int *queue_count;
    int *queue_next;
    ColPart *dense_parts;            // [K*W] written by the dense queue kernel
    ResultRec *res;                  // [K]
};

template <int HS>
__device__ __forceinline__ bool hash_find(const int *hk, const unsigned long long *hv, int key,
                                          unsigned long long &val) {
    unsigned h = ((unsigned)key * 2654435761u) & (HS - 1);
    for (int p = 0; p < HS; p++) {
        int k = hk[h];
        if (k == key) { val = hv[h]; return true; }
        if (k < 0) return false;
        h = (h + 1) & (HS - 1);
    }
    return false;
}

// Inserts key if absent.  Returns 1 if inserted, 0 if present, -1 if full.
template <int HS>
__device__ __forceinline__ int hash_insert(int *hk, unsigned long long *hv, int key,
                                           unsigned long long val) {
    unsigned h = ((unsigned)key * 2654435761u) & (HS - 1);
    for (int p = 0; p < HS; p++) {
        int prev = atomicCAS(&hk[h], -1, key);
        if (prev == -1) { hv[h] = val; return 1; }
        if (prev == key) return 0;
        h = (h + 1) & (HS - 1);
    }
    return -1;
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Candidate neighbour i of v (base row with the swap's slot replacements).
__device__ __forceinline__ int cand_nbr(const int *adjR, int D, int v, int i, const SwapRegs &sw) {
    int u = adjR[(size_t)v * D + i];
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (sw.v[j] == v && sw.s[j] == i) u = sw.n[j];
    return u;
}

template <int NW, int HS>
__global__ void __launch_bounds__(NW * 64) k_sparse(SparseArgs a) {
    extern __shared__ unsigned long long Bprev[];
    __shared__ int hk[NW][2][HS];
    __shared__ unsigned long long hv[NW][2][HS];
    __shared__ int s_ncur[NW];
    __shared__ int s_active[2];  // parity-indexed: reset two levels ahead of its readers
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int w = blockIdx.x;
    const int N = a.N, D = a.D, W = a.W, K = a.K;
    const int ecc_w = a.base_parts[w].ecc;
    const unsigned long long pad = pad_mask(w, N);

    // task init: columns holding a swap endpoint as a source go dense at once
    for (int k = tid; k < K; k += NW * 64) {
        const size_t t = (size_t)k * W + w;
        const gg_swap &s = a.swaps[k];
        bool dense = false;
        for (int j = 0; j < 4; j++)
            if (s.v[j] >= 0 && (s.v[j] >> 6) == w) dense = true;
        a.dl_cnt[t * 2] = 0;
        a.task_flags[t] = 0;
        a.task_end[t] = 0;
        if (dense) {
            a.task_state[t] = GG_TASK_DENSE;
            int q = atomicAdd(a.queue_count, 1);
            a.queue[q] = make_int2(k, w);
        } else {
            a.task_state[t] = GG_TASK_ACTIVE;
        }
    }
    __threadfence_block();
    __syncthreads();

    for (int l = 1; l <= a.lmax; l++) {
        // B_{l-1} of this column into LDS
        const int lp = l - 1;
        for (int v = tid; v < N; v += NW * 64) {
            unsigned long long b;
            if (lp == 0) b = pad | (((v >> 6) == w) ? (1ull << (v & 63)) : 0ull);
            else if (lp >= ecc_w) b = ~0ull;
            else b = a.snap[((size_t)(lp - 1) * W + w) * N + v];
            Bprev[v] = b;
        }
        if (tid == 0) s_active[l & 1] = 0;
        __syncthreads();
        int *hkp = hk[wave][0], *hkc = hk[wave][1];
        unsigned long long *hvp = hv[wave][0], *hvc = hv[wave][1];
        for (int k = wave; k < K; k += NW) {
            const size_t t = (size_t)k * W + w;
            if (a.task_state[t] != GG_TASK_ACTIVE) continue;
            SwapRegs sw;
            bool has_sw;
            load_swap(a.swaps, k, sw, has_sw);
            const int par = (l - 1) & 1, cur = l & 1;
            const int n_prev = a.dl_cnt[t * 2 + par];
            const int *pkey = a.dl_key + (t * 2 + par) * a.cap;
            const unsigned long long *pval = a.dl_val + (t * 2 + par) * a.cap;
            int *ckey = a.dl_key + (t * 2 + cur) * a.cap;
            unsigned long long *cval = a.dl_val + (t * 2 + cur) * a.cap;
            for (int i = lane; i < HS; i += 64) { hkp[i] = -1; hkc[i] = -1; }
            if (lane == 0) s_ncur[wave] = 0;
            wave_sync_lds();
            for (int e = lane; e < n_prev; e += 64) hash_insert<HS>(hkp, hvp, pkey[e], pval[e]);
            wave_sync_lds();
            const int items = 4 + n_prev * (D + 1);
            long long dc = 0;
            int overflow = 0, differs = 0;
            for (int it = lane; it < items; it += 64) {
                int v;
                if (it < 4) {
                    v = it == 0 ? sw.v[0] : it == 1 ? sw.v[1] : it == 2 ? sw.v[2] : sw.v[3];
                    if (v < 0) continue;
                } else {
                    int e = (it - 4) / (D + 1), i = (it - 4) - e * (D + 1);
                    int u = pkey[e];
                    v = i == 0 ? u : cand_nbr(a.adjR, D, u, i - 1, sw);
                }
                unsigned long long own, acc;
                if (n_prev == 0) {
                    acc = Bprev[v];
                    for (int i = 0; i < D; i++) acc |= Bprev[cand_nbr(a.adjR, D, v, i, sw)];
                } else {
                    if (!hash_find<HS>(hkp, hvp, v, own)) own = Bprev[v];
                    acc = own;
                    for (int i = 0; i < D; i++) {
                        int x = cand_nbr(a.adjR, D, v, i, sw);
                        unsigned long long xv;
                        if (!hash_find<HS>(hkp, hvp, x, xv)) xv = Bprev[x];
                        acc |= xv;
                    }
                }
                unsigned long long base;
                if (l >= ecc_w) base = ~0ull;
                else base = a.snap[((size_t)(l - 1) * W + w) * N + v];
                if (acc != base) {
                    int ins = hash_insert<HS>(hkc, hvc, v, acc);
                    if (ins < 0) { overflow = 1; continue; }
                    if (ins == 1) {
                        int idx = atomicAdd(&s_ncur[wave], 1);
                        if (idx >= a.cap) { overflow = 1; continue; }
                        ckey[idx] = v;
                        cval[idx] = acc;
                        dc += (long long)__popcll(acc) - (long long)__popcll(base);
                        unsigned long long pv;
                        if (!hash_find<HS>(hkp, hvp, v, pv) || pv != acc) differs = 1;
                    }
                }
            }
            wave_sync_lds();
            dc = wave_sum64(dc);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                overflow |= __shfl_xor(overflow, o, 64);
                differs |= __shfl_xor(differs, o, 64);
            }
            const int n_cur = s_ncur[wave];
            if (lane == 0) {
                if (overflow) {
                    a.task_state[t] = GG_TASK_DENSE;
                    int q = atomicAdd(a.queue_count, 1);
                    a.queue[q] = make_int2(k, w);
                } else {
                    a.dcnt[t * (a.lmax + 1) + l] = dc;
                    a.dl_cnt[t * 2 + cur] = n_cur;
                    bool fixed = (l >= ecc_w && n_cur == 0) ||
                                 (l > ecc_w && !differs && n_cur == n_prev);
                    if (fixed || l >= a.lmax) {
                        a.task_state[t] = GG_TASK_DONE;
                        a.task_end[t] = l;
                        if (!fixed) a.task_flags[t] = GG_FLAG_LIMIT;
                    } else {
                        atomicAdd(&s_active[l & 1], 1);
                    }
                }
            }
            wave_sync_lds();
        }
        __syncthreads();
        if (s_active[l & 1] == 0) break;
    }
}

// One workgroup per candidate: merges sparse column results (base counts +
// deltas) and dense fallback parts into the candidate's ResultRec.
template <int BS>
__global__ void __launch_bounds__(BS) k_sparse_finalize(SparseArgs a, int nvalid_last) {
    __shared__ long long s_tot[BS / 64], s_rch[BS / 64];
    __shared__ int s_diam[BS / 64], s_lev[BS / 64], s_fl[BS / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int k = blockIdx.x;
    const int W = a.W, L1 = a.lmax + 1;
    long long tot = 0, rch = 0;
    int diam = 0, lev = 0, fl = 0;
    for (int w = tid; w < W; w += BS) {
        const size_t t = (size_t)k * W + w;
        if (a.task_state[t] == GG_TASK_DENSE) {
            const ColPart &p = a.dense_parts[t];
            tot += p.total; rch += p.reached;
            diam = p.ecc > diam ? p.ecc : diam;
            lev = p.levels > lev ? p.levels : lev;
            fl |= p.flags;
            continue;
        }
        const int end = a.task_end[t];
        const ColPart &bp = a.base_parts[w];
        const long long *br = a.base_reach + (size_t)w * L1;
        const long long *dd = a.dcnt + t * L1;
        long long cb = (w == W - 1) ? nvalid_last : 64;  // base reached after level 0
        long long cprev = cb, ctot = 0;
        int ecc = 0;
        for (int l = 1; l <= end; l++) {
            if (l <= bp.levels) cb += br[l];
            long long c = cb + dd[l];
            if (c > cprev) { ecc = l; ctot += (long long)l * (c - cprev); }
            cprev = c;
        }
        tot += ctot;
        rch += cprev;
        diam = ecc > diam ? ecc : diam;
        int lv = a.task_flags[t] ? end : ecc + 1;
        lev = lv > lev ? lv : lev;
        fl |= a.task_flags[t];
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
        a.res[k] = r;
    }
}

// Resets the dense queue counters (after the dense kernel and the finalize).
__global__ void k_queue_reset(int *queue_count, int *queue_next) {
    if (threadIdx.x == 0 && blockIdx.x == 0) { *queue_count = 0; *queue_next = 0; }
}
