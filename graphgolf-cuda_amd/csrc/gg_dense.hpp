// gg_dense.hpp -- dense column BFS kernels (bit-parallel, 64 sources / column).
//
// Algorithm (semantics of cudacode/cudaASPLconv.cu:197-266, re-designed):
// every 64 consecutive sources form an independent "column"; one workgroup
// runs the whole level-synchronous BFS of one column.  The column's visited
// bitset (one uint64 per vertex) lives in LDS, so every level is an LDS
// neighbour-OR gather and the level barrier is a workgroup barrier -- no host
// round trip per level (the reference synchronises and copies to the host
// after every level, cudaASPLconv.cu:208-217).
//
//   new[v]  = (OR_{u in adj(v)} vis[u]) & ~vis[v]     (pull over visited)
//   vis[v] |= new[v];  total += level * popc(new[v])
//
// Pulling over the full visited set is exact: a bit of vis[u] that was set
// before the previous level reached v one level earlier, so it is already in
// vis[v].  A vertex whose word is full is skipped, and a gather stops as soon
// as (acc | vis[v]) is full.  Bits of sources >= M are padding and start set
// (cudaASPLconv.cu:38-44).  Per-source eccentricity = last level in which the
// source's bit appeared anywhere in the column.
//
// Each (candidate, column) workgroup writes a ColPart; the last workgroup of a
// candidate (agent-scope counter, reset by that workgroup) reduces the parts
// into the candidate's ResultRec, so one launch yields the final answer.
#pragma once
#include <type_traits>
#include "gg_common.hpp"

namespace gg {
// Dynamic-LDS size a kernel has been allowed on each device (the attribute
// only grows).  ensure_dynamic_lds raises it to at least `lds` bytes on the
// current device, under a process-wide lock: launchers on different host
// threads share these per-kernel records.
constexpr int kMaxDevices = 64;
struct LdsAttr {
    int bytes[kMaxDevices];
    LdsAttr() { for (int &b : bytes) b = -1; }
};
hipError_t ensure_dynamic_lds(const void *fn, size_t lds, LdsAttr &attr);
}  // namespace gg

// ---- topologies ---------------------------------------------------------------
// Offset form, 1-D ring: neighbour i of v is v + tab[i*R + v % R] (mod N) with
// offsets stored normalised to [0, N).  R = M for 'part' (cppcode/part.hpp),
// R = N for an explicit adjacency (offset = (u - v) mod N).
template <class IdxT>
struct TopoRing {
    const IdxT *tab;
    unsigned N, R;
    struct Ctx { unsigned v, row; };
    __device__ __forceinline__ Ctx ctx(unsigned v) const {
        return Ctx{v, v < R ? v : v % R};
    }
    __device__ __forceinline__ unsigned nbr(const Ctx &c, int i) const {
        unsigned u = c.v + (unsigned)tab[(size_t)i * R + c.row];
        return u >= N ? u - N : u;
    }
};

// 16-bit entries packed four per 64-bit word (N <= 65535): word (q, r) holds
// slots 4q..4q+3 of row r, slot 4q+j in bits 16j..16j+15.  A vertex's D
// neighbours are ceil(D/4) coalesced 8-byte loads per wave.  When every
// vertex is its own row (R == N) the entries are the neighbour ids
// themselves; otherwise (symmetric tables, R < N) they are offsets mod N.
struct TopoRingQ {
    const uint64_t *tab;  // [ceil(D/4)][R]
    unsigned N, R;
    struct Ctx { unsigned v, row; };
    __device__ __forceinline__ Ctx ctx(unsigned v) const {
        return Ctx{v, v < R ? v : v % R};
    }
    __device__ __forceinline__ unsigned nbr(const Ctx &c, int i) const {
        uint64_t q = tab[(size_t)(i >> 2) * R + c.row];
        unsigned e = (unsigned)(q >> (16 * (i & 3))) & 0xFFFFu;
        if (R == N) return e;
        unsigned u = c.v + e;
        return u >= N ? u - N : u;
    }
};

// 2-D torus 'piece' (cppcode/piece.hpp, cudacode/cudaASPLpiece.cu:35-54):
// v = x*Ny + y; offsets (dx,dy) normalised, packed dx | dy << 16.
struct TopoPiece {
    const uint32_t *tab;
    unsigned Nx, Ny, Mx, My, R;
    struct Ctx { unsigned x, y, row; };
    __device__ __forceinline__ Ctx ctx(unsigned v) const {
        unsigned x = v / Ny, y = v - x * Ny;
        return Ctx{x, y, (x % Mx) * My + (y % My)};
    }
    __device__ __forceinline__ unsigned nbr(const Ctx &c, int i) const {
        uint32_t o = tab[(size_t)i * R + c.row];
        unsigned tx = c.x + (o & 0xFFFFu); tx = tx >= Nx ? tx - Nx : tx;
        unsigned ty = c.y + (o >> 16);     ty = ty >= Ny ? ty - Ny : ty;
        return tx * Ny + ty;
    }
};

// 3-D torus 'box' (cppcode/box.hpp:29-36, cudacode/cudaASPLbox.cu:38-61):
// v = (x*Ny + y)*Nz + z; offsets packed dx | dy << 21 | dz << 42.
struct TopoBox {
    const uint64_t *tab;
    unsigned Nx, Ny, Nz, Mx, My, Mz, R;
    struct Ctx { unsigned x, y, z, row; };
    __device__ __forceinline__ Ctx ctx(unsigned v) const {
        unsigned yz = Ny * Nz;
        unsigned x = v / yz, r = v - x * yz;
        unsigned y = r / Nz, z = r - y * Nz;
        return Ctx{x, y, z, ((x % Mx) * My + (y % My)) * Mz + (z % Mz)};
    }
    __device__ __forceinline__ unsigned nbr(const Ctx &c, int i) const {
        uint64_t o = tab[(size_t)i * R + c.row];
        unsigned tx = c.x + (unsigned)(o & 0x1FFFFFu);         tx = tx >= Nx ? tx - Nx : tx;
        unsigned ty = c.y + (unsigned)((o >> 21) & 0x1FFFFFu); ty = ty >= Ny ? ty - Ny : ty;
        unsigned tz = c.z + (unsigned)((o >> 42) & 0x1FFFFFu); tz = tz >= Nz ? tz - Nz : tz;
        return (tx * Ny + ty) * Nz + tz;
    }
};

// ---- kernel arguments ---------------------------------------------------------
// Running totals of a fused single-candidate launch, updated with
// device-scope atomics by each column workgroup as it finishes: the
// workgroups spread over kFuseSlots slots of their own 64-byte line (fewer
// atomics queue on one line), the arrivals on a line of their own.
constexpr int kFuseSlots = 16;
struct FuseSlot {
    unsigned long long total, reached;
    int ecc_levels;  // max of ecc * 256 + levels (levels follows ecc)
    int flags;
    int pad[10];
};
struct FuseAcc {
    FuseSlot slot[kFuseSlots];
    int arrived;
    int pad[15];
};

struct DenseArgs {
    int N, M, D;            // vertices, sources, slots per row
    int W;                  // 64-source columns per candidate
    int K;                  // candidates in this launch
    int lmax;               // highest level evaluated (GG_LEVEL_LIMIT)
    const int *src_vertex;  // [M] vertex of source s
    const gg_swap *swaps;   // [K] slot replacements, or nullptr
    ColPart *parts;         // [K*W]
    ResultRec *res;         // [K]
    int *ecc_out;           // [M] per-source eccentricity, or nullptr
    long long *reach_col;   // [W*(lmax+1)] per-column level counts, or nullptr
    long long *reach_out;   // [lmax+1] summed level counts, or nullptr
    int want_wvc;           // finalize computes WVC from ecc_out
    unsigned long long *gscratch;  // global variant: 2*N words per workgroup
    // base snapshots: after level l (1 <= l <= snap_levels) the column's words
    // are stored at snap[((l-1)*W + w)*N + v]
    unsigned long long *snap;
    int snap_levels;
    // queue mode: tasks (k, w) = queue[i] for i < *queue_count, claimed via
    // *queue_next; parts are written but no candidate is finalised
    const int2 *queue;
    const int *queue_count;
    int *queue_next;
    // push levels: capacity of the level-1 frontier list that follows vis[N]
    // in dynamic LDS (0 = pull from level 1)
    int push_cap;
    // queue mode: the words after each level of an endpoint task (candidate
    // k >= ep_k0, column holding endpoint j of its swap) go to slot
    // 4(k - ep_k0) + j of ep_snap [.][ep_S][N] / ep_reach [.][lmax+1]
    unsigned long long *ep_snap;
    long long *ep_reach;
    int ep_k0, ep_S;
    // fused finalize (k_dense_lds, single candidate, no eccentricities or
    // level counts asked for): the candidate's accumulators, zero between
    // launches; nullptr = separate finalize kernel
    FuseAcc *fuse;
    int seq;  // fused launches: written to res->seq after the result
};

struct SwapRegs {
    int v[4], s[4], n[4];
};

__device__ __forceinline__ void load_swap(const gg_swap *sw, int k, SwapRegs &r, bool &has) {
    has = false;
    if (!sw) { for (int j = 0; j < 4; j++) { r.v[j] = -1; r.s[j] = 0; r.n[j] = 0; } return; }
    const gg_swap &s = sw[k];
#pragma unroll
    for (int j = 0; j < 4; j++) { r.v[j] = s.v[j]; r.s[j] = s.slot[j]; r.n[j] = s.nbr[j]; }
    has = true;
}

// Neighbour-OR gather for vertex v with early exit once (acc | own) is full.
template <class Topo, class Vis>
__device__ __forceinline__ unsigned long long gather_plain(const Topo &topo, Vis vis, unsigned v,
                                                           unsigned long long own, int D) {
    unsigned long long acc = 0;
    typename Topo::Ctx c = topo.ctx(v);
    int i = 0;
    for (; i + 4 <= D; i += 4) {
        unsigned u0 = topo.nbr(c, i), u1 = topo.nbr(c, i + 1);
        unsigned u2 = topo.nbr(c, i + 2), u3 = topo.nbr(c, i + 3);
        acc |= vis[u0] | vis[u1] | vis[u2] | vis[u3];
        if ((acc | own) == ~0ull) return acc;
    }
    for (; i < D; i++) acc |= vis[topo.nbr(c, i)];
    return acc;
}

// Gather of a vertex whose row a candidate swap patches (slot by slot).
template <class Topo, class Vis>
__device__ __forceinline__ unsigned long long gather_patched(const Topo &topo, Vis vis, unsigned v,
                                                             int D, const SwapRegs &sw) {
    unsigned long long acc = 0;
    typename Topo::Ctx c = topo.ctx(v);
    for (int i = 0; i < D; i++) {
        unsigned u = topo.nbr(c, i);
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (sw.v[j] == (int)v && sw.s[j] == i) u = (unsigned)sw.n[j];
        acc |= vis[u];
    }
    return acc;
}

// Generic entry used by the global-memory kernel.
template <class Topo, class Vis>
__device__ __forceinline__ unsigned long long gather_or(const Topo &topo, Vis vis, unsigned v,
                                                        unsigned long long own, int D,
                                                        const SwapRegs &sw, bool has_sw) {
    bool hit = has_sw && ((int)v == sw.v[0] || (int)v == sw.v[1] || (int)v == sw.v[2] ||
                          (int)v == sw.v[3]);
    return hit ? gather_patched(topo, vis, v, D, sw) : gather_plain(topo, vis, v, own, D);
}

// Fixed-degree gather on the quad-packed ring topology with R == N (entries
// are neighbour ids): the ceil(DT/4) words of a vertex are loaded up front (one
// memory latency per vertex), then the neighbours are OR-ed from LDS in
// groups of four with the early exit.
template <int DT>
struct FixedRow {
    static constexpr int Q = (DT + 3) / 4;
    unsigned long long q[Q];
    // voff = 8 * vertex: 32-bit byte offsets from the table base, so each
    // word's address is one add of a uniform stride
    __device__ __forceinline__ void load(const char *base, unsigned stride8, unsigned voff) {
#pragma unroll
        for (int j = 0; j < Q; j++)
            q[j] = *reinterpret_cast<const unsigned long long *>(base + (j * stride8 + voff));
    }
    // a patched row (the candidate's swap applied) staged in LDS
    __device__ __forceinline__ void load_lds(const unsigned long long *p) {
#pragma unroll
        for (int j = 0; j < Q; j++) q[j] = p[j];
    }
    __device__ __forceinline__ unsigned u(int i) const {
        return (unsigned)(q[i >> 2] >> (16 * (i & 3))) & 0xFFFFu;
    }
};

// Reads go out in groups of GG_GATHER_GROUP (all in flight together); the
// word is tested after each group when `exits` (a wave-uniform flag: false on
// levels where few words can fill up, which then skip the tests' round trips).
#ifndef GG_GATHER_GROUP
#define GG_GATHER_GROUP 8
#endif
template <int DT, class Vis>
__device__ __forceinline__ unsigned long long gather_fixed(const FixedRow<DT> &row, Vis vis,
                                                           unsigned long long own, bool exits) {
    constexpr int G = GG_GATHER_GROUP;
    unsigned long long acc = 0;
#pragma unroll
    for (int i = 0; i < DT; i += G) {
        unsigned long long g = vis[row.u(i)];
#pragma unroll
        for (int t = 1; t < G; t++)
            if (i + t < DT) g |= vis[row.u(i + t)];
        acc |= g;
        if (i + G < DT && exits && (acc | own) == ~0ull) break;
        // compiler-only barrier: keeps the next group's LDS loads below the
        // early exit (hoisting all DT loads costs registers and spills)
        __asm__ volatile("" ::: "memory");
    }
    return acc;
}

// Neighbour i of u in the candidate graph (the swap's slot replacements applied).
template <class Topo>
__device__ __forceinline__ unsigned nbr_patched(const Topo &topo, unsigned u, int i,
                                                const SwapRegs &sw, bool has_sw) {
    unsigned x = topo.nbr(topo.ctx(u), i);
    if (has_sw) {
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (sw.v[j] == (int)u && sw.s[j] == i) x = (unsigned)sw.n[j];
    }
    return x;
}

// Padding bits of column w: sources 64w+j >= M are treated as already reached.
__device__ __forceinline__ unsigned long long pad_mask(int w, int M) {
    int base = w * 64;
    if (base + 64 <= M) return 0ull;
    if (base >= M) return ~0ull;
    return ~0ull << (M - base);
}

// Reduces the column parts of candidate k into its ResultRec (and the WVC
// and per-level reach of a single-candidate launch), by the whole workgroup:
// k_dense_finalize (one workgroup per candidate, after the BFS launch), or
// the last column workgroup of a fused single-candidate BFS launch.
template <int BS>
__device__ void finalize_candidate(const DenseArgs &a, int k) {
    __shared__ long long s_tot[BS / 64], s_rch[BS / 64];
    __shared__ int s_diam[BS / 64], s_lev[BS / 64], s_fl[BS / 64];
    __shared__ int s_cls[32];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid < 32) s_cls[tid] = -1;
    __syncthreads();
    long long tot = 0, rch = 0;
    int diam = 0, lev = 0, fl = 0;
    for (int i = tid; i < a.W; i += BS) {
        const ColPart *q = a.parts + (size_t)k * a.W + i;
        tot += q->total;
        rch += q->reached;
        int e = q->ecc, l = q->levels;
        diam = e > diam ? e : diam;
        lev = l > lev ? l : lev;
        fl |= q->flags;
    }
    tot = wave_sum64(tot);
    rch = wave_sum64(rch);
    diam = wave_max32(diam);
    lev = wave_max32(lev);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) fl |= __shfl_xor(fl, o, 64);
    if (lane == 0) { s_tot[wid] = tot; s_rch[wid] = rch; s_diam[wid] = diam; s_lev[wid] = lev; s_fl[wid] = fl; }
    // WVC: classes j in [0,32) merge sources j, j+32, ... (cudaASPLconv.cu:258-285)
    if (a.want_wvc && a.ecc_out) {
        for (int s = tid; s < a.M; s += BS) atomicMax(&s_cls[s & 31], a.ecc_out[s]);
    }
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
        int wvc = 0;
        if (a.want_wvc && a.ecc_out)
            for (int j = 0; j < 32; j++) wvc += (s_cls[j] == r.diameter) ? 1 : 0;
        r.wvc = wvc;
        a.res[k] = r;
    }
    if (a.reach_out && a.reach_col) {
        // per-level histogram: one block-wide pass over the columns per level
        // up to the deepest level executed, zeros above it
        __shared__ long long s_red[BS / 64];
        __syncthreads();  // s_lev[0] is the deepest level
        const int L = a.lmax + 1;
        const int top = s_lev[0] < L - 1 ? s_lev[0] : L - 1;
        for (int l = top + 1 + tid; l < L; l += BS) a.reach_out[l] = 0;
        for (int l = 0; l <= top; l++) {
            long long s = 0;
            for (int i = tid; i < a.W; i += BS) {
                const ColPart *q = a.parts + (size_t)k * a.W + i;
                if (l <= q->levels) s += a.reach_col[(size_t)i * L + l];
            }
            s = wave_sum64(s);
            if (lane == 0) s_red[wid] = s;
            __syncthreads();
            if (tid == 0) {
                long long t = 0;
                for (int x = 0; x < BS / 64; x++) t += s_red[x];
                a.reach_out[l] = t;
            }
            __syncthreads();
        }
    }
}

// The launch boundary orders the BFS kernel's parts before these loads.
template <int BS>
__global__ void __launch_bounds__(BS) k_dense_finalize(DenseArgs a) {
    finalize_candidate<BS>(a, blockIdx.x);
}

// Where a column BFS stores its per-level words and counts: the words after
// level l (1 <= l <= levels) at snap + (l-1)*lstride, the new pairs of level l
// at reach[l] (either pointer may be null).
struct ColOut {
    unsigned long long *snap;
    size_t lstride;
    int levels;
    long long *reach;
};

// The evaluator's own layout: snap[(l-1)][w][v], reach_col[w][l].
__device__ __forceinline__ ColOut col_out(const DenseArgs &a, int w) {
    ColOut o;
    o.snap = a.snap ? a.snap + (size_t)w * a.N : nullptr;
    o.lstride = (size_t)a.W * a.N;
    o.levels = a.snap_levels;
    o.reach = a.reach_col ? a.reach_col + (size_t)w * (a.lmax + 1) : nullptr;
    return o;
}

// ---- LDS-resident column BFS ---------------------------------------------------
// One workgroup runs one (candidate, column) task with vis[N] in dynamic LDS.
// Each thread owns VPT vertices, so new words stay in registers across the
// level barrier (no second LDS buffer).  Returns the column part in
// thread 0.
template <class Topo, int BS, int VPT, int DT>
__device__ __forceinline__ ColPart dense_column(const Topo &topo, const DenseArgs &a, int k, int w,
                                                unsigned long long *vis, const ColOut &out) {
    __shared__ unsigned long long s_or[2];
    __shared__ unsigned s_cnt[2];
    __shared__ int s_fn;
    const int tid = threadIdx.x, lane = tid & 63;
    const unsigned N = (unsigned)a.N;
    const int D = a.D;
    SwapRegs sw;
    bool has_sw;
    load_swap(a.swaps, k, sw, has_sw);

    const unsigned long long pad = pad_mask(w, a.M);
    for (unsigned v = tid; v < N; v += BS) vis[v] = pad;
    if (tid < 2) { s_or[tid] = 0; s_cnt[tid] = 0; }
    // fixed-degree kernels: the rows of the swapped vertices, patched, so every
    // slot runs the same gather (row t of swap entry t, all entries on it applied)
    constexpr bool kFixed = DT > 0 && std::is_same<Topo, TopoRingQ>::value;
    constexpr int kQ = (DT + 3) / 4;
    __shared__ unsigned long long s_prow[4 * (kQ > 0 ? kQ : 1)];
    if constexpr (kFixed) {
        if (has_sw && tid < 4 * kQ) {
            const int t = tid / kQ, qq = tid - t * kQ;
            const int vt = t == 0 ? sw.v[0] : t == 1 ? sw.v[1] : t == 2 ? sw.v[2] : sw.v[3];
            if (vt >= 0) {
                unsigned long long word = topo.tab[(size_t)qq * topo.R + (unsigned)vt];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    if (sw.v[e] == vt && (sw.s[e] >> 2) == qq) {
                        const int sh = 16 * (sw.s[e] & 3);
                        word = (word & ~(0xFFFFull << sh)) |
                               ((unsigned long long)(unsigned)sw.n[e] << sh);
                    }
                }
                s_prow[t * kQ + qq] = word;
            }
        }
    }
    __syncthreads();
    const int nvalid = (a.M - w * 64) < 64 ? (a.M - w * 64) : 64;
    if (tid < nvalid) {
        unsigned sv = (unsigned)a.src_vertex[w * 64 + tid];
        atomicOr(&vis[sv], 1ull << tid);
    }
    __syncthreads();

    long long total = 0, reached = nvalid;
    int my_ecc = 0, last_new = 0, level = 0, flags = 0;
    long long *reach = out.reach;
    if (reach && tid == 0) reach[0] = nvalid;
    // wave-blocked ownership: wave q owns vertices [64*VPT*q, 64*VPT*(q+1)),
    // slot j of lane l is vb + 64*j (coalesced per slot, small LDS offsets)
    const unsigned vb = (unsigned)(tid >> 6) * 64u * VPT + (unsigned)lane;
    // slots of this lane whose row the candidate's swap patches (one VGPR),
    // and for each the first swap entry on that vertex (2 bits per slot)
    unsigned hitmask = 0, hitsel = 0;
    if (has_sw) {
#pragma unroll
        for (int j = 0; j < VPT; j++) {
            const int v = (int)(vb + j * 64);
            const int t = v == sw.v[0] ? 0 : v == sw.v[1] ? 1 : v == sw.v[2] ? 2 : v == sw.v[3] ? 3 : -1;
            if (t >= 0) { hitmask |= 1u << j; hitsel |= (unsigned)t << (2 * j); }
        }
    }
    auto snap_of = [&](int l) -> unsigned long long * {
        return (out.snap && l <= out.levels) ? out.snap + (size_t)(l - 1) * out.lstride : nullptr;
    };
    // level bookkeeping shared by push and pull levels: folds this thread's
    // new-pair count and new-bit mask into the block, returns true when the
    // BFS of the column is finished
    unsigned long long cover = (unsigned)nvalid;
    auto end_level = [&](int l, unsigned cnt, unsigned long long lor) -> bool {
        lor = wave_or64(lor);
        cnt = wave_sum32(cnt);
        if (lane == 0 && cnt) {
            atomicOr(&s_or[l & 1], lor);
            atomicAdd(&s_cnt[l & 1], cnt);
        }
        __syncthreads();
        const unsigned long long blk = s_or[l & 1];
        cover += s_cnt[l & 1];  // every thread: pairs reached so far
        if (tid == 0) {
            unsigned c = s_cnt[l & 1];
            total += (long long)l * c;
            reached += c;
            if (reach) reach[l] = c;
            s_or[(l + 1) & 1] = 0;
            s_cnt[(l + 1) & 1] = 0;
        }
        if (tid < 64 && ((blk >> tid) & 1ull)) my_ecc = l;
        if (blk == 0) return true;
        last_new = l;
        if (l >= a.lmax) { flags |= GG_FLAG_LIMIT; return true; }
        return false;
    };
    bool done = false;
    level = 1;
    // push follows edges from the frontier, pull reads each vertex's listed
    // neighbours: the same BFS only on undirected graphs, which the host
    // guarantees before it sets push_cap
    if (a.push_cap > 0) {
        // levels 1 and 2 push from the (small) frontier with LDS atomics;
        // the old word tells whether a bit is new, so counts stay exact
        unsigned *flist = reinterpret_cast<unsigned *>(vis + N);
        if (tid == 0) s_fn = 0;
        __syncthreads();
        unsigned cnt = 0;
        unsigned long long lor = 0;
        for (int it = tid; it < nvalid * D; it += BS) {
            const int j = it / D, i = it - j * D;
            const unsigned sv = (unsigned)a.src_vertex[w * 64 + j];
            const unsigned u = nbr_patched(topo, sv, i, sw, has_sw);
            const unsigned long long bit = 1ull << j;
            const unsigned long long old = atomicOr(&vis[u], bit);
            if (!(old & bit)) {
                cnt++;
                lor |= bit;
                const int idx = atomicAdd(&s_fn, 1);
                if (idx < a.push_cap) flist[idx] = u | ((unsigned)j << 17);
            }
        }
        __syncthreads();
        if (unsigned long long *sn = snap_of(1))
            for (unsigned v = tid; v < N; v += BS) sn[v] = vis[v];
        done = end_level(1, cnt, lor);
        const int nf = s_fn;
        if (!done && nf <= a.push_cap) {
            level = 2;
            cnt = 0;
            lor = 0;
            if constexpr (kFixed) {
                // one item per (frontier entry, table word): four neighbours
                // from one 8-byte load; swapped rows come from s_prow
                for (int it = tid; it < nf * kQ; it += BS) {
                    const int e = it / kQ, q = it - e * kQ;
                    const unsigned f = flist[e];
                    const unsigned u = f & 0x1FFFFu, j = f >> 17;
                    unsigned long long word;
                    if (has_sw && ((int)u == sw.v[0] || (int)u == sw.v[1] || (int)u == sw.v[2] ||
                                   (int)u == sw.v[3])) {
                        const int t = (int)u == sw.v[0] ? 0 : (int)u == sw.v[1] ? 1
                                    : (int)u == sw.v[2] ? 2 : 3;
                        word = s_prow[t * kQ + q];
                    } else {
                        word = topo.tab[(size_t)q * topo.R + u];
                    }
                    const unsigned long long bit = 1ull << j;
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        if (4 * q + t >= DT) break;
                        const unsigned x = (unsigned)(word >> (16 * t)) & 0xFFFFu;
                        const unsigned long long old = atomicOr(&vis[x], bit);
                        if (!(old & bit)) { cnt++; lor |= bit; }
                    }
                }
            } else {
                for (int it = tid; it < nf * D; it += BS) {
                    const int e = it / D, i = it - e * D;
                    const unsigned f = flist[e];
                    const unsigned u = f & 0x1FFFFu, j = f >> 17;
                    const unsigned x = nbr_patched(topo, u, i, sw, has_sw);
                    const unsigned long long bit = 1ull << j;
                    const unsigned long long old = atomicOr(&vis[x], bit);
                    if (!(old & bit)) { cnt++; lor |= bit; }
                }
            }
            __syncthreads();
            if (unsigned long long *sn = snap_of(2))
                for (unsigned v = tid; v < N; v += BS) sn[v] = vis[v];
            done = end_level(2, cnt, lor);
        }
    } else {
        level = 0;
    }
    // `level` is the last level executed; pull continues from the next one
    if (!done) for (level = level + 1;; level++) {
        unsigned long long nw[VPT];
        if constexpr (DT > 0 && std::is_same<Topo, TopoRingQ>::value) {
            // fixed degree: the next vertex's words load while this one gathers
            const char *tbase = reinterpret_cast<const char *>(topo.tab);
            const unsigned stride8 = topo.R * 8u;
            // opaque per level: the row addresses are recomputed inside the
            // level loop rather than kept live across it for every slot
            unsigned voff = vb * 8u;
            __asm__ volatile("" : "+v"(voff));
            FixedRow<DT> row[2];
            auto load_row = [&](FixedRow<DT> &r, int j) {
                if ((hitmask >> j) & 1u) r.load_lds(s_prow + ((hitsel >> (2 * j)) & 3u) * kQ);
                else r.load(tbase, stride8, voff + (unsigned)j * 512u);
            };
            // under half of the column's pairs reached: words rarely fill up
            // during this level, so skip the early-exit tests
            const bool exits = 2 * cover >= (unsigned long long)N * (unsigned)nvalid;
            // slots whose word is not full yet: only their rows are loaded
            // (late levels touch few rows; the table reads are L2 traffic)
            unsigned need = 0;
#pragma unroll
            for (int j = 0; j < VPT; j++) {
                const unsigned v = vb + j * 64;
                if (v < N && ~vis[v]) need |= 1u << j;
            }
            if (need & 1u) load_row(row[0], 0);
#pragma unroll
            for (int j = 0; j < VPT; j++) {
                unsigned v = vb + j * 64;
                if (j + 1 < VPT && ((need >> (j + 1)) & 1u)) load_row(row[(j + 1) & 1], j + 1);
                nw[j] = 0;
                if ((need >> j) & 1u) {
                    unsigned long long own = vis[v];
                    nw[j] = gather_fixed<DT>(row[j & 1], vis, own, exits) & ~own;
                }
                __asm__ volatile("" ::: "memory");  // one row in flight ahead, no more
            }
        } else {
#pragma unroll
            for (int j = 0; j < VPT; j++) {
                unsigned v = vb + j * 64;
                nw[j] = 0;
                if (v < N) {
                    unsigned long long own = vis[v];
                    if (~own) {
                        unsigned long long acc = ((hitmask >> j) & 1u)
                                                     ? gather_patched(topo, vis, v, D, sw)
                                                     : gather_plain(topo, vis, v, own, D);
                        nw[j] = acc & ~own;
                    }
                }
            }
        }
        __syncthreads();
        unsigned long long lor = 0;
        unsigned cnt = 0;
        unsigned long long *snap = snap_of(level);
#pragma unroll
        for (int j = 0; j < VPT; j++) {
            unsigned v = vb + j * 64;
            if (nw[j]) {
                vis[v] |= nw[j];  // only this thread writes vertex v
                lor |= nw[j];
                cnt += __popcll(nw[j]);
            }
            if (snap && v < N) snap[v] = vis[v];
        }
        if (end_level(level, cnt, lor)) break;
    }
    if (a.ecc_out && tid < nvalid) a.ecc_out[w * 64 + tid] = my_ecc;
    ColPart p;
    p.total = total;
    p.reached = reached;
    p.ecc = last_new;
    p.levels = level;
    p.flags = flags;
    p.pad = 0;
    return p;
}

// One workgroup per (candidate, column); k_dense_finalize then reduces the
// parts.  With a.fuse (one candidate) each workgroup instead adds its part
// into its slot of the accumulators with device-scope atomics, waits until
// they have completed, and only then counts itself in; the last one to
// arrive takes the slots' sums with device-scope atomic exchanges (leaving
// them zero for the next launch) and writes the result.  No release fence
// (an L2 writeback per workgroup) is needed: every value crosses workgroups
// through atomics.
template <class Topo, int BS, int VPT, int DT>
__global__ void __launch_bounds__(BS) k_dense_lds(Topo topo, DenseArgs a) {
    extern __shared__ unsigned long long vis[];
    const int t = blockIdx.x;
    const int k = t / a.W, w = t - k * a.W;
    ColPart p = dense_column<Topo, BS, VPT, DT>(topo, a, k, w, vis, col_out(a, w));
    if (threadIdx.x == 0) a.parts[(size_t)k * a.W + w] = p;
    if (a.fuse) {
        __shared__ int s_last;
        FuseAcc *f = a.fuse;
        if (threadIdx.x == 0) {
            FuseSlot *q = &f->slot[t % kFuseSlots];
            atomicAdd(&q->total, (unsigned long long)p.total);
            atomicAdd(&q->reached, (unsigned long long)p.reached);
            atomicMax(&q->ecc_levels, p.ecc * 256 + p.levels);
            if (p.flags) atomicOr(&q->flags, p.flags);
            __builtin_amdgcn_s_waitcnt(0);  // the atomics above have completed
            s_last = atomicAdd(&f->arrived, 1) == a.W - 1;
        }
        __syncthreads();
        if (s_last && threadIdx.x < 64) {
            const int i = threadIdx.x;
            unsigned long long tot = 0, rch = 0;
            int el = 0, fl = 0;
            if (i < kFuseSlots) {
                FuseSlot *q = &f->slot[i];
                tot = atomicExch(&q->total, 0ull);
                rch = atomicExch(&q->reached, 0ull);
                el = atomicExch(&q->ecc_levels, 0);
                fl = atomicExch(&q->flags, 0);
            }
            tot = (unsigned long long)wave_sum64((long long)tot);
            rch = (unsigned long long)wave_sum64((long long)rch);
            el = wave_max32(el);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) fl |= __shfl_xor(fl, o, 64);
            if (i == 0) {
                atomicExch(&f->arrived, 0);
                // the fields first, then the sequence number a host may be
                // polling (res can be pinned host memory)
                ResultRec *r = a.res;
                r->total = (long long)tot;
                r->reached = (long long)rch;
                r->diameter = el >> 8;
                r->levels = el & 255;
                r->flags = fl;
                r->wvc = 0;
                __threadfence_system();
                __hip_atomic_store(&r->seq, a.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
    }
}

// Persistent workgroups draining a device-side task queue; parts only.
template <class Topo, int BS, int VPT, int DT>
__global__ void __launch_bounds__(BS) k_dense_lds_queue(Topo topo, DenseArgs a) {
    extern __shared__ unsigned long long vis[];
    __shared__ int s_task;
    // the queue is complete when this launch starts: an empty one (the usual
    // case) costs one load, no claim
    const int qn = *a.queue_count;
    if (qn == 0) return;
    for (;;) {
        if (threadIdx.x == 0) s_task = atomicAdd(a.queue_next, 1);
        __syncthreads();
        const int t = s_task;
        if (t >= qn) break;
        const int2 kw = a.queue[t];
        ColOut o = col_out(a, kw.y);
        if (a.ep_snap && a.swaps && kw.x >= a.ep_k0) {
            const gg_swap &sw = a.swaps[kw.x];
            int j = -1;
            for (int e = 3; e >= 0; e--)
                if (sw.v[e] >= 0 && (sw.v[e] >> 6) == kw.y) j = e;
            if (j >= 0) {
                const size_t slot = 4 * (size_t)(kw.x - a.ep_k0) + j;
                o.snap = a.ep_snap + slot * a.ep_S * a.N;
                o.lstride = (size_t)a.N;
                o.levels = a.ep_S;
                o.reach = a.ep_reach + slot * (a.lmax + 1);
            }
        }
        ColPart p = dense_column<Topo, BS, VPT, DT>(topo, a, kw.x, kw.y, vis, o);
        if (threadIdx.x == 0) a.parts[(size_t)kw.x * a.W + kw.y] = p;
        __syncthreads();
    }
}

// ---- global-memory column BFS (N too large for LDS) -------------------------------
// Persistent grid; workgroup b owns scratch cur/nxt of N words each.
template <class Topo>
__global__ void __launch_bounds__(1024) k_dense_global(Topo topo, DenseArgs a, int tasks) {
    constexpr int BS = 1024;
    // three buffers: with one barrier per level, the buffer reset at the end
    // of level l is the one used at level l+2 (read by everyone before the
    // barrier of level l), never the one of level l+1.
    __shared__ unsigned long long s_or[3];
    __shared__ unsigned s_cnt[3];
    const int tid = threadIdx.x, lane = tid & 63;
    const unsigned N = (unsigned)a.N;
    const int D = a.D;
    unsigned long long *bufA = a.gscratch + (size_t)blockIdx.x * 2 * N;
    unsigned long long *bufB = bufA + N;
    for (int t = blockIdx.x; t < tasks; t += gridDim.x) {
        const int k = t / a.W, w = t - k * a.W;
        SwapRegs sw;
        bool has_sw;
        load_swap(a.swaps, k, sw, has_sw);
        unsigned long long *cur = bufA, *nxt = bufB;
        const unsigned long long pad = pad_mask(w, a.M);
        for (unsigned v = tid; v < N; v += BS) cur[v] = pad;
        if (tid < 3) { s_or[tid] = 0; s_cnt[tid] = 0; }
        __syncthreads();
        const int nvalid = (a.M - w * 64) < 64 ? (a.M - w * 64) : 64;
        if (tid < nvalid) {
            unsigned sv = (unsigned)a.src_vertex[w * 64 + tid];
            cur[sv] |= 1ull << tid;
        }
        __syncthreads();
        long long total = 0, reached = nvalid;
        int my_ecc = 0, last_new = 0, level = 0, flags = 0;
        long long *reach = a.reach_col ? a.reach_col + (size_t)w * (a.lmax + 1) : nullptr;
        if (reach && tid == 0) reach[0] = nvalid;
        for (level = 1;; level++) {
            unsigned long long lor = 0;
            unsigned cnt = 0;
            for (unsigned v = tid; v < N; v += BS) {
                unsigned long long own = cur[v], n = 0;
                if (~own) n = gather_or(topo, cur, v, own, D, sw, has_sw) & ~own;
                nxt[v] = own | n;
                lor |= n;
                cnt += __popcll(n);
            }
            lor = wave_or64(lor);
            cnt = wave_sum32(cnt);
            if (lane == 0 && cnt) {
                atomicOr(&s_or[level % 3], lor);
                atomicAdd(&s_cnt[level % 3], cnt);
            }
            __syncthreads();
            const unsigned long long blk = s_or[level % 3];
            if (tid == 0) {
                unsigned c = s_cnt[level % 3];
                total += (long long)level * c;
                reached += c;
                if (reach) reach[level] = c;
                s_or[(level + 2) % 3] = 0;
                s_cnt[(level + 2) % 3] = 0;
            }
            if (tid < 64 && ((blk >> tid) & 1ull)) my_ecc = level;
            unsigned long long *tmp = cur; cur = nxt; nxt = tmp;
            if (blk == 0) break;
            last_new = level;
            if (level >= a.lmax) { flags |= GG_FLAG_LIMIT; break; }
        }
        if (a.ecc_out && tid < nvalid) a.ecc_out[w * 64 + tid] = my_ecc;
        ColPart p;
        p.total = total;
        p.reached = reached;
        p.ecc = last_new;
        p.levels = level;
        p.flags = flags;
        p.pad = 0;
        if (tid == 0) a.parts[(size_t)k * a.W + w] = p;
        __syncthreads();
    }
}

