// gg_propose.hpp -- 2-opt proposals drawn on the GPU (the learned search's
// draw, also carried by the commit launch of the step before).
//
// Proposal p of step t draws from stream (seed, t, p) until it finds a simple
// 2-opt swap, with the host generator's attempt (propose_swap): a, ia, c, ic,
// kind, edges (a, adj[a][ia]) and (c, adj[c][ic]) rewired to (a,d),(b,c)
// (kind 1) or (a,c),(b,d) (kind 0), redrawn unless the four vertices are
// distinct and neither new edge exists.
#pragma once
#include "gg_common.hpp"
#include "gg_devrng.hpp"

#define GG_PROPOSE_MAX_ATTEMPTS 4096
#define GG_PROPOSE_ROW 32

namespace gg {

// Row of v (d neighbours) in registers: 16-byte loads issued together when
// d is a multiple of 4 (one memory latency), else one by one.
template <int DM>
struct PRow {
    int u[DM];
    __device__ __forceinline__ void load(const int *adjR, int v, int d) {
        const int *r = adjR + (size_t)v * d;
        // every index is a compile-time constant (predicated, no dynamic
        // indexing), so the row stays in registers
        if ((d & 3) == 0) {
            const int4 *r4 = reinterpret_cast<const int4 *>(r);
#pragma unroll
            for (int q = 0; q < DM / 4; q++) {
                int4 x = make_int4(-1, -1, -1, -1);
                if (4 * q < d) x = r4[q];
                u[4 * q] = x.x; u[4 * q + 1] = x.y; u[4 * q + 2] = x.z; u[4 * q + 3] = x.w;
            }
        } else {
#pragma unroll
            for (int i = 0; i < DM; i++) u[i] = i < d ? r[i] : -1;
        }
    }
    // the row as DM ints in 16-byte loads with no branch (d a multiple of 4,
    // DM <= 32: the array's GG_ADJ_ROW_PAD ints of padding keep the last
    // row's read in bounds; entries past d are ignored by slot_of)
    __device__ __forceinline__ void load_all(const int *adjR, int v, int d) {
        const int4 *r4 = reinterpret_cast<const int4 *>(adjR + (size_t)v * d);
#pragma unroll
        for (int q = 0; q < DM / 4; q++) {
            const int4 x = r4[q];
            u[4 * q] = x.x; u[4 * q + 1] = x.y; u[4 * q + 2] = x.z; u[4 * q + 3] = x.w;
        }
    }
    __device__ __forceinline__ int slot_of(int d, int x) const {
        int s = -1;
#pragma unroll
        for (int i = DM - 1; i >= 0; i--)
            if (i < d && u[i] == x) s = i;
        return s;
    }
};

// Scalar fallback for degrees above the register row.
__device__ __forceinline__ int row_slot_of(const int *row, int d, int x) {
    for (int i = 0; i < d; i++)
        if (row[i] == x) return i;
    return -1;
}

// ctl: [0] step, [1] warm, [2] diameter; workgroup 0 copies them to dctl
// (read by the later kernels of the step).  A proposal with no simple swap
// after GG_PROPOSE_MAX_ATTEMPTS ORs dctl[3] and writes an empty swap.
struct ProposeArgs {
    const int *adjR;
    int N, d;
    uint64_t seed;
    const int *ctl;
    int P;
    int *moves;      // [P][5]: a, b, c, d, kind
    gg_swap *out;    // [P]
    int *dctl;
};

// The proposals of workgroup bid (64 threads, one proposal each).
__device__ __forceinline__ void propose_body(const ProposeArgs &pa, int bid) {
    const int *adjR = pa.adjR;
    const int N = pa.N, d = pa.d;
    // the header, read by every thread (no barrier); workgroup 0 copies it
    const int step = pa.ctl[0];
    if (bid == 0 && threadIdx.x == 0) {
        const int h1 = pa.ctl[1], h2 = pa.ctl[2];
        pa.dctl[0] = step;
        pa.dctl[1] = h1;
        pa.dctl[2] = h2;
    }
    const int p = bid * 64 + threadIdx.x;
    if (p >= pa.P) return;
    uint64_t st = sm64_stream(pa.seed, (uint64_t)step, (uint64_t)p);
    const bool reg = d <= GG_PROPOSE_ROW;
    // every row of an attempt in two rounds of branch-free loads: b, d and
    // a's row, then the rows of b and d
    const bool vec = reg && (d & 3) == 0;
    for (int att = 0; att < GG_PROPOSE_MAX_ATTEMPTS; att++) {
        const int a = (int)sm64_below(st, N), ia = (int)sm64_below(st, d);
        const int c = (int)sm64_below(st, N), ic = (int)sm64_below(st, d);
        const int kind = (int)sm64_below(st, 2);
        const int b = adjR[(size_t)a * d + ia], dd = adjR[(size_t)c * d + ic];
        PRow<GG_PROPOSE_ROW> ra, rb, rd;
        if (vec) {
            ra.load_all(adjR, a, d);
            rb.load_all(adjR, b, d);
            rd.load_all(adjR, dd, d);
        } else if (reg) {
            ra.load(adjR, a, d);
        }
        if (a == b || a == c || a == dd || b == c || b == dd || c == dd) continue;
        const int xa = kind ? dd : c, xb = kind ? c : dd, xc = kind ? b : a, xd = kind ? a : b;
        int ib, id;
        if (reg) {
            if (!vec) {
                rb.load(adjR, b, d);
                rd.load(adjR, dd, d);
            }
            if (ra.slot_of(d, xa) >= 0 || rb.slot_of(d, xb) >= 0) continue;
            ib = rb.slot_of(d, a);
            id = rd.slot_of(d, c);
        } else {
            const int *ra = adjR + (size_t)a * d, *rb = adjR + (size_t)b * d;
            if (row_slot_of(ra, d, xa) >= 0 || row_slot_of(rb, d, xb) >= 0) continue;
            ib = row_slot_of(rb, d, a);
            id = row_slot_of(adjR + (size_t)dd * d, d, c);
        }
        if (ib < 0 || id < 0) continue;
        gg_swap o;
        o.v[0] = a;  o.slot[0] = ia; o.nbr[0] = xa;
        o.v[1] = b;  o.slot[1] = ib; o.nbr[1] = xb;
        o.v[2] = c;  o.slot[2] = ic; o.nbr[2] = xc;
        o.v[3] = dd; o.slot[3] = id; o.nbr[3] = xd;
        pa.out[p] = o;
        int *m = pa.moves + (size_t)p * 5;
        m[0] = a; m[1] = b; m[2] = c; m[3] = dd; m[4] = kind;
        return;
    }
    atomicOr(pa.dctl + 3, 1);  // no simple swap found: the host reports it
    gg_swap none;
    for (int j = 0; j < 4; j++) { none.v[j] = -1; none.slot[j] = 0; none.nbr[j] = 0; }
    pa.out[p] = none;
    int *m = pa.moves + (size_t)p * 5;
    m[0] = m[1] = m[2] = m[3] = 0; m[4] = 0;
}

// Proposal p drawn by a group of four consecutive lanes (lane q = 0..3 of
// the group holds quads q and q + 4 of the three rows an attempt reads; d a
// multiple of 4 up to 32), on the adjacency with swap `patch` applied (null:
// as stored): the swap's four entries are replaced wherever they are read, so
// the draw may run beside the launch that rewrites them.  The same draws,
// checks and result as propose_one on the patched adjacency.
__device__ __forceinline__ void propose_split(const ProposeArgs &pa, int step, int p, int q,
                                              const gg_swap *patch) {
    const int *adjR = pa.adjR;
    const int N = pa.N, d = pa.d;
    int pv[4] = {-1, -1, -1, -1}, ps[4] = {0, 0, 0, 0}, pn[4] = {0, 0, 0, 0};
    if (patch) {
#pragma unroll
        for (int j = 0; j < 4; j++) { pv[j] = patch->v[j]; ps[j] = patch->slot[j]; pn[j] = patch->nbr[j]; }
    }
    auto fix = [&](int v, int slot, int u) {
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (pv[j] == v && ps[j] == slot) u = pn[j];
        return u;
    };
    auto comp = [](const int4 &x, int k) { return k == 0 ? x.x : k == 1 ? x.y : k == 2 ? x.z : x.w; };
    uint64_t st = sm64_stream(pa.seed, (uint64_t)step, (uint64_t)p);
    for (int att = 0; att < GG_PROPOSE_MAX_ATTEMPTS; att++) {
        const int a = (int)sm64_below(st, N), ia = (int)sm64_below(st, d);
        const int c = (int)sm64_below(st, N), ic = (int)sm64_below(st, d);
        const int kind = (int)sm64_below(st, 2);
        const int b = fix(a, ia, adjR[(size_t)a * d + ia]);
        const int dd = fix(c, ic, adjR[(size_t)c * d + ic]);
        // the lane's quads of the three rows (the array's GG_ADJ_ROW_PAD ints
        // of padding keep quads past the row in bounds; their entries are
        // ignored)
        int4 xa4[2], xb4[2], xd4[2];
        const int4 *ra = reinterpret_cast<const int4 *>(adjR + (size_t)a * d);
        const int4 *rb = reinterpret_cast<const int4 *>(adjR + (size_t)b * d);
        const int4 *rd = reinterpret_cast<const int4 *>(adjR + (size_t)dd * d);
#pragma unroll
        for (int h = 0; h < 2; h++) {
            xa4[h] = ra[q + 4 * h];
            xb4[h] = rb[q + 4 * h];
            xd4[h] = rd[q + 4 * h];
        }
        if (a == b || a == c || a == dd || b == c || b == dd || c == dd) continue;
        const int xa = kind ? dd : c, xb = kind ? c : dd, xc = kind ? b : a, xd = kind ? a : b;
        bool hxa = false, hxb = false;
        int sb = 1 << 30, sd = 1 << 30;
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int slot = 4 * (q + 4 * h) + k;
                if (slot >= d) continue;
                const int ya = fix(a, slot, comp(xa4[h], k));
                const int yb = fix(b, slot, comp(xb4[h], k));
                const int yd = fix(dd, slot, comp(xd4[h], k));
                hxa = hxa || ya == xa;
                hxb = hxb || yb == xb;
                if (yb == a && slot < sb) sb = slot;
                if (yd == c && slot < sd) sd = slot;
            }
        // the group's four lanes combined (the group's lanes always run
        // together: they draw the same attempts)
#pragma unroll
        for (int o = 1; o <= 2; o <<= 1) {
            // every lane shuffles (no short-circuit: a lane left out of a
            // shuffle would hand its partner a zero)
            const int ta = __shfl_xor((int)hxa, o, 64), tb = __shfl_xor((int)hxb, o, 64);
            const int ob = __shfl_xor(sb, o, 64), od = __shfl_xor(sd, o, 64);
            hxa = hxa || ta != 0;
            hxb = hxb || tb != 0;
            sb = ob < sb ? ob : sb;
            sd = od < sd ? od : sd;
        }
        if (hxa || hxb) continue;
        if (sb == (1 << 30) || sd == (1 << 30)) continue;
        if (q == 0) {
            gg_swap o;
            o.v[0] = a;  o.slot[0] = ia; o.nbr[0] = xa;
            o.v[1] = b;  o.slot[1] = sb; o.nbr[1] = xb;
            o.v[2] = c;  o.slot[2] = ic; o.nbr[2] = xc;
            o.v[3] = dd; o.slot[3] = sd; o.nbr[3] = xd;
            pa.out[p] = o;
            int *m = pa.moves + (size_t)p * 5;
            m[0] = a; m[1] = b; m[2] = c; m[3] = dd; m[4] = kind;
        }
        return;
    }
    if (q != 0) return;
    atomicOr(pa.dctl + 3, 1);  // no simple swap found: the host reports it
    gg_swap none;
    for (int j = 0; j < 4; j++) { none.v[j] = -1; none.slot[j] = 0; none.nbr[j] = 0; }
    pa.out[p] = none;
    int *m = pa.moves + (size_t)p * 5;
    m[0] = m[1] = m[2] = m[3] = 0; m[4] = 0;
}

// The draw of workgroup bid by groups of four lanes (propose_split), with
// workgroup 0 copying the header.
__device__ __forceinline__ void propose_split_body(const ProposeArgs &pa, const gg_swap *patch,
                                                   int bid) {
    const int step = pa.ctl[0];
    if (bid == 0 && threadIdx.x == 0) {
        const int h1 = pa.ctl[1], h2 = pa.ctl[2];
        pa.dctl[0] = step;
        pa.dctl[1] = h1;
        pa.dctl[2] = h2;
    }
    const int g = bid * (int)blockDim.x + (int)threadIdx.x;
    const int p = g >> 2;
    if (p >= pa.P) return;
    propose_split(pa, step, p, g & 3, patch);
}

// Degrees the four-lane draw takes.
__host__ __device__ inline bool propose_split_ok(int d) { return d <= 32 && (d & 3) == 0; }

}  // namespace gg
