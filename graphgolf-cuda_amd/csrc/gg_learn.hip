// gg_learn.hip -- learned proposal scorer: feature kernel and a 32-128-1 MLP
// (BF16 weights, FP32 accumulation) with hand-written forward, backward and
// Adam kernels.  No reference counterpart: the reference evaluates every
// proposal exactly (cppcode/annealing.cpp:399-403).
//
// Features of a swap (a,b),(c,d) -> (p1,q1),(p2,q2) come from the base BFS
// level words the batch evaluator stores: bit s of B_l[w][x] (s = 64w + j) is
// [d(s, x) <= l], so the N-bit row R_l(x) = (B_l[w][x])_w is the radius-l ball
// of x, and L_l(x) = R_l(x) & ~R_{l-1}(x) its l-th layer.
#include <hip/hip_bf16.h>
#include <cmath>
#include <cstring>
#include <vector>
#include "gg_common.hpp"
#include "gg_internal.hpp"
#include "gg_batch.hpp"
#include "gg_learn.hpp"

using namespace gg;

// MLP parameter layout (fp32 masters and bf16 copies): W1[H][F], b1[H], W2[H], b2.
#define OFF_W1 0
#define OFF_B1 (GG_HIDDEN * GG_FEATS)
#define OFF_W2 (OFF_B1 + GG_HIDDEN)
#define OFF_B2 (OFF_W2 + GG_HIDDEN)

__device__ __forceinline__ float bf(const __hip_bfloat16 v) { return __bfloat162float(v); }

// ---- features ----------------------------------------------------------------------------
// Per proposal (added edges (p1,q1),(p2,q2), removed (a,b),(c,d)), 32 features:
//  added edge t, X[8t..8t+8): pre-swap d(p,q)/diam, |L3(p)&L3(q)|/n,
//    |~R3(p)&~R3(q)|/n, |L2(p)&L2(q)|/d^2, |L2(q)&~R3(p)|/d^2,
//    |L2(p)&~R3(q)|/d^2, (|L1(q)&~R3(p)| + |L1(p)&~R3(q)|)/d,
//    (|L1(q)&~R2(p)| + |L1(p)&~R2(q)|)/d -- vertices the new edge brings
//    closer to p or q (first-order gain of the swap);
//  removed edge t = (x,y), X[16+6t..16+6t+6): U(x,y)/d, U(y,x)/d,
//    |L2(x)&L2(y)|/d^2, |R2(x)&R2(y)|/d^2, (|L1(x)&L2(y)| + |L2(x)&L1(y)|)/d^2,
//    (ecc(x)+ecc(y)-2 diam)/2, where U(u,v) = #{z in N(v), z not in B1(u) :
//    |N(z) & N(u)| = 1}, the second-layer vertices of u reached only through
//    v (first-order loss of the swap);
//  endpoint e of (a,b,c,d), X[28+e]: |~R3(x_e)|/n.
// Bits of sources >= N are set in every stored level (they start reached);
// level 0 sets them too, so no layer L_l (l >= 1) and no complement ~R_l
// holds them; R2&R2 is masked explicitly.
struct FeatArgs {
    int N, W, S, diam, d;
    const int *diam_dev;             // diameter on the device (overrides diam)
    const unsigned long long *snap;  // [S][W][N]
    const unsigned long long *rows;  // [RL][N][W] levels 1..RL row-major (full words stored)
    int RL;
    const ColPart *base_parts;       // [W]
    const int *adjR;                 // [N][d] row-major neighbours of the base
    const int *moves;                // [P][5]: a, b, c, d, kind
    int P;
    float *X;                        // [P][GG_FEATS]
    const __hip_bfloat16 *wq;        // MLP weights: with `score`, also the scores
    const __hip_bfloat16 *wq2;       // with wpar: the copy used when *wpar is odd
    const int *wpar;
    float *score;                    // [P] (null: features only)
};

__device__ __forceinline__ unsigned long long pad_bits(const FeatArgs &f, int w) {
    const int r = f.N & 63;
    return (w == f.W - 1 && r) ? (~0ull << r) : 0ull;
}

// B_l[w][x] with levels >= the column's eccentricity full and level 0 = {x}
// plus the padding sources.
__device__ __forceinline__ unsigned long long ball_word(const FeatArgs &f, int l, int w, int x,
                                                       int ecc_w) {
    if (l <= 0) return (((x >> 6) == w) ? (1ull << (x & 63)) : 0ull) | pad_bits(f, w);
    if (l <= f.RL) return f.rows[((size_t)(l - 1) * f.N + x) * f.W + w];  // lanes over w: coalesced
    if (l >= ecc_w) return ~0ull;
    return f.snap[((size_t)(l - 1) * f.W + w) * f.N + x];
}

// Bit z of the radius-1 ball row of u: z in B1(u).
__device__ __forceinline__ bool in_ball1(const FeatArgs &f, int u, int z) {
    const int w = z >> 6;
    return (ball_word(f, 1, w, u, f.base_parts[w].ecc) >> (z & 63)) & 1ull;
}

// Radius-1 ball rows of the four endpoints staged in LDS per wave when the
// row fits (W <= GG_FEAT_LDS_W words), else read from the stored levels.
#define GG_FEAT_LDS_W 256

// Value 0..3 selected by a wave-uniform or per-lane index without an indexed
// register array.
__device__ __forceinline__ int pick4(int v0, int v1, int v2, int v3, int i) {
    return i == 0 ? v0 : i == 1 ? v1 : i == 2 ? v2 : v3;
}

// U(u,v) for the four directions (u,v) = (a,b), (b,a), (c,d), (d,c)
// (u = endpoint e, v = endpoint e^1): lanes over the 4d (direction, entry z
// of v's row) items; a lane reads z's row and tests its entries against u's
// radius-1 ball row.
__device__ void uniq_second_layer4(const FeatArgs &f, const int ends[4],
                                   const unsigned long long (*r1)[GG_FEAT_LDS_W], int lane,
                                   int out[4]) {
    auto ball1 = [&](int e, int z) -> bool {
        if (r1) return (r1[e][z >> 6] >> (z & 63)) & 1ull;
        return in_ball1(f, pick4(ends[0], ends[1], ends[2], ends[3], e), z);
    };
    int part[4] = {0, 0, 0, 0};
    const int items = 4 * f.d;
    for (int k = lane; k < items; k += 64) {
        const int e = k / f.d, i = k - e * f.d;
        const int u = pick4(ends[0], ends[1], ends[2], ends[3], e),
                  v = pick4(ends[1], ends[0], ends[3], ends[2], e);
        const int z = f.adjR[(size_t)v * f.d + i];
        if (z == u || ball1(e, z)) continue;
        const int *rz = f.adjR + (size_t)z * f.d;
        int common = 0;
#pragma unroll 8
        for (int j = 0; j < f.d; j++) common += ball1(e, rz[j]) ? 1 : 0;
        const int hit = common == 1 ? 1 : 0;
        part[0] += e == 0 ? hit : 0;
        part[1] += e == 1 ? hit : 0;
        part[2] += e == 2 ? hit : 0;
        part[3] += e == 3 ? hit : 0;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        int x = part[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        out[q] = x;
    }
}

// Diameter the features normalise by (the device copy when given).
__device__ __forceinline__ int feat_diam(const FeatArgs &f) {
    return f.diam_dev ? (*f.diam_dev > 0 ? *f.diam_dev : 1) : f.diam;
}

// d(x, y) before the swap given the words of x's ball rows at y's column for
// levels 1..3 (b3): the first level whose word holds y, levels past 3 read one
// by one (levels at or above the column's eccentricity are full).
__device__ __forceinline__ int pair_dist(const FeatArgs &f, int x, int y, int ew,
                                         const unsigned long long b3[3]) {
    const int w = y >> 6, bit = y & 63;
#pragma unroll
    for (int l = 1; l <= 3; l++)
        if ((b3[l - 1] >> bit) & 1ull) return l;
    for (int l = 4; l < 64; l++)
        if ((ball_word(f, l, w, x, ew) >> bit) & 1ull) return l;
    return 64;
}

__device__ __forceinline__ int pair_dist(const FeatArgs &f, int x, int y) {
    const int w = y >> 6, ew = f.base_parts[w].ecc;
    unsigned long long b3[3];
#pragma unroll
    for (int l = 1; l <= 3; l++) b3[l - 1] = ball_word(f, l, w, x, ew);
    return pair_dist(f, x, y, ew, b3);
}

// The 32 features of a proposal into X (lane 0; the counts are wave totals,
// dist[t] the pre-swap distance of added edge t).
__device__ __forceinline__ void write_features(const FeatArgs &f, int diam, const int dist[2],
                                               const float *cnt, const int *eccl, const int *uq,
                                               float *X) {
    const float d1 = (float)f.d, d2 = d1 * d1, n = (float)f.N, inv_diam = 1.f / (float)diam;
#pragma unroll
    for (int t = 0; t < 2; t++) {
        X[t * 8 + 0] = dist[t] * inv_diam;
        X[t * 8 + 1] = cnt[t * 7 + 0] / n;
        X[t * 8 + 2] = cnt[t * 7 + 1] / n;
        X[t * 8 + 3] = cnt[t * 7 + 2] / d2;
        X[t * 8 + 4] = cnt[t * 7 + 3] / d2;
        X[t * 8 + 5] = cnt[t * 7 + 4] / d2;
        X[t * 8 + 6] = cnt[t * 7 + 5] / d1;
        X[t * 8 + 7] = cnt[t * 7 + 6] / d1;
    }
#pragma unroll
    for (int t = 0; t < 2; t++) {
        X[16 + t * 6 + 0] = (float)uq[2 * t] / d1;
        X[16 + t * 6 + 1] = (float)uq[2 * t + 1] / d1;
        X[16 + t * 6 + 2] = cnt[14 + t * 3 + 0] / d2;
        X[16 + t * 6 + 3] = cnt[14 + t * 3 + 1] / d2;
        X[16 + t * 6 + 4] = cnt[14 + t * 3 + 2] / d2;
        X[16 + t * 6 + 5] = (float)(eccl[2 * t] + eccl[2 * t + 1] - 2 * diam) * 0.5f;
    }
#pragma unroll
    for (int e = 0; e < 4; e++) X[28 + e] = cnt[20 + e] / n;
}

// Counts of one column w given the endpoints' ball words of levels 0..3
// (R[endpoint][level]) and the column's eccentricity ew: the per-lane partial
// counts and the endpoints' full levels.
template <class T>
__device__ __forceinline__ void column_counts(const FeatArgs &f, int w, int ew, const int ends[4],
                                              int kind, const unsigned long long (&R)[4][4],
                                              T *cnt, int *eccl) {
    const unsigned long long valid = ~pad_bits(f, w);
#pragma unroll
    for (int e = 0; e < 4; e++) {
        int fl = 0;
        for (int l = 0; l <= ew; l++) {
            const unsigned long long x = l < 4 ? R[e][l] : ball_word(f, l, w, ends[e], ew);
            if (x == ~0ull) { fl = l; break; }
        }
        eccl[e] = fl > eccl[e] ? fl : eccl[e];
    }
    auto L = [&](int e, int l) { return R[e][l] & ~R[e][l - 1]; };
    // index of p1,q1,p2,q2 in ends: a=0,b=1,c=2,dd=3
    const int ip1 = 0, iq1 = kind ? 3 : 2, ip2 = 1, iq2 = kind ? 2 : 3;
    const int pairs[2][2] = {{ip1, iq1}, {ip2, iq2}};
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const int x = pairs[t][0], y = pairs[t][1];
        cnt[t * 7 + 0] += __popcll(L(x, 3) & L(y, 3));
        cnt[t * 7 + 1] += __popcll(~R[x][3] & ~R[y][3]);
        cnt[t * 7 + 2] += __popcll(L(x, 2) & L(y, 2));
        cnt[t * 7 + 3] += __popcll(L(y, 2) & ~R[x][3]);
        cnt[t * 7 + 4] += __popcll(L(x, 2) & ~R[y][3]);
        cnt[t * 7 + 5] += __popcll(L(y, 1) & ~R[x][3]) + __popcll(L(x, 1) & ~R[y][3]);
        cnt[t * 7 + 6] += __popcll(L(y, 1) & ~R[x][2]) + __popcll(L(x, 1) & ~R[y][2]);
    }
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const int x = 2 * t, y = 2 * t + 1;
        cnt[14 + t * 3 + 0] += __popcll(L(x, 2) & L(y, 2));
        cnt[14 + t * 3 + 1] += __popcll(R[x][2] & R[y][2] & valid);
        cnt[14 + t * 3 + 2] += __popcll(L(x, 1) & L(y, 2)) + __popcll(L(x, 2) & L(y, 1));
    }
#pragma unroll
    for (int e = 0; e < 4; e++) cnt[20 + e] += __popcll(~R[e][3]);
}

// bf16 pair of a 32-bit word as floats (element 0 in the low half).
__device__ __forceinline__ float bf_lo(unsigned u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf_hi(unsigned u) { return __uint_as_float(u & 0xffff0000u); }

// The MLP score of a proposal from its features xs (every lane), with the
// arithmetic of k_mlp_forward (same products, fma order and sums; products
// rounded before they are summed, as k_mlp_forward's pass through LDS does):
// bit-identical.  w1 holds lane's hidden units lane and lane + 64 (rows of
// W1 as bf16 pairs), b1/w2 their bias and output weight.
__device__ __forceinline__ float mlp_score(const float xs[GG_FEATS],
                                           const uint4 (&w1)[GG_HIDDEN / 64][GG_FEATS / 8],
                                           const float (&b1)[GG_HIDDEN / 64],
                                           const float (&w2)[GG_HIDDEN / 64], float b2) {
#pragma clang fp contract(off)
    float o = 0.f;
#pragma unroll
    for (int q = 0; q < GG_HIDDEN / 64; q++) {
        float acc = b1[q];
#pragma unroll
        for (int i = 0; i < GG_FEATS / 8; i++) {
            const uint4 u = w1[q][i];
            acc = fmaf(bf_lo(u.x), xs[8 * i + 0], acc);
            acc = fmaf(bf_hi(u.x), xs[8 * i + 1], acc);
            acc = fmaf(bf_lo(u.y), xs[8 * i + 2], acc);
            acc = fmaf(bf_hi(u.y), xs[8 * i + 3], acc);
            acc = fmaf(bf_lo(u.z), xs[8 * i + 4], acc);
            acc = fmaf(bf_hi(u.z), xs[8 * i + 5], acc);
            acc = fmaf(bf_lo(u.w), xs[8 * i + 6], acc);
            acc = fmaf(bf_hi(u.w), xs[8 * i + 7], acc);
        }
        const float r = w2[q] * (acc > 0.f ? acc : 0.f);
        o = o + r;
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) o = o + __shfl_xor(o, d, 64);
    return o + b2;
}

// Lane's MLP weights (hidden units lane, lane + 64) from the bf16 copy.
struct LaneWeights {
    uint4 w1[GG_HIDDEN / 64][GG_FEATS / 8];
    float b1[GG_HIDDEN / 64], w2[GG_HIDDEN / 64], b2;
};
__device__ __forceinline__ void load_lane_weights(const __hip_bfloat16 *wq, int lane,
                                                  LaneWeights &lw) {
#pragma unroll
    for (int q = 0; q < GG_HIDDEN / 64; q++) {
        const int h = lane + 64 * q;
        const uint4 *r = reinterpret_cast<const uint4 *>(wq + OFF_W1 + h * GG_FEATS);
#pragma unroll
        for (int i = 0; i < GG_FEATS / 8; i++) lw.w1[q][i] = r[i];
        lw.b1[q] = bf(wq[OFF_B1 + h]);
        lw.w2[q] = bf(wq[OFF_W2 + h]);
    }
    lw.b2 = bf(wq[OFF_B2]);
}

// Shape of the feature pass with one workgroup per proposal: row-major
// levels 1..3 kept, at most 256 columns (one per thread, the radius-1 rows
// staged in LDS), degree a multiple of 4 up to 32 (the 4d second-layer
// entries one per thread, their rows as 16-byte loads).
#define GG_FEAT_FAST_D 32
__host__ __device__ inline bool features_fast_ok(const FeatArgs &f) {
    return f.rows && f.RL >= 3 && f.W <= GG_FEAT_LDS_W && f.d <= GG_FEAT_FAST_D &&
           (f.d & 3) == 0;
}

// Workgroups of the feature pass for P proposals.
__host__ inline int features_groups(const FeatArgs &f) {
    return features_fast_ok(f) ? f.P : (f.P + 3) / 4;
}

// Proposal p by one workgroup of 256 threads: thread t takes column t and
// second-layer entry t, threads 0 and 1 the added edges' distances, threads
// below 128 the hidden units of the score.  Every global load is issued
// before any is used (three dependent rounds: the move, the level words and
// entries, the entries' rows).  Counts are summed per wave, then over the
// four waves (integers held exactly in fp32: any order gives the same sum).
// Wave sums of 12 words of two 16-bit counts each, reduce-scattered: on
// return lane L holds the wave total of word (L&32 ? 6 : 0) + (L&16 ? 3 : 0) +
// (L&8 ? 2 : 0) + (L&4 ? 1 : 0) (no word where L&8 and L&4 are both set).
// Halves cannot carry: every lane's half is at most 128, so a total is at
// most 8192.
__device__ __forceinline__ unsigned wave_sum12_scatter(unsigned (&pk)[12], int lane) {
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8, b2 = lane & 4;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const unsigned mine = b5 ? pk[j + 6] : pk[j], other = b5 ? pk[j] : pk[j + 6];
        pk[j] = mine + (unsigned)__shfl_xor((int)other, 32, 64);
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const unsigned mine = b4 ? pk[j + 3] : pk[j], other = b4 ? pk[j] : pk[j + 3];
        pk[j] = mine + (unsigned)__shfl_xor((int)other, 16, 64);
    }
    unsigned q[2];
    {
        const unsigned hi0 = pk[2], hi1 = 0u;
        const unsigned m0 = b3 ? hi0 : pk[0], o0 = b3 ? pk[0] : hi0;
        const unsigned m1 = b3 ? hi1 : pk[1], o1 = b3 ? pk[1] : hi1;
        q[0] = m0 + (unsigned)__shfl_xor((int)o0, 8, 64);
        q[1] = m1 + (unsigned)__shfl_xor((int)o1, 8, 64);
    }
    unsigned x;
    {
        const unsigned mine = b2 ? q[1] : q[0], other = b2 ? q[0] : q[1];
        x = mine + (unsigned)__shfl_xor((int)other, 4, 64);
    }
    x += (unsigned)__shfl_xor((int)x, 2, 64);
    x += (unsigned)__shfl_xor((int)x, 1, 64);
    return x;
}

struct FeatWgLds {
    unsigned long long r1[4][GG_FEAT_LDS_W];  // R_1 of the endpoints
    unsigned cnt[4][12];                       // per wave: counts 2j, 2j+1 as 16-bit halves
    int ecc[4][4];
    unsigned uq[4];                            // per wave: U counts as bytes
    int dist[2];
    float x[GG_FEATS];
    float rh[GG_HIDDEN];
};

// Proposal p by one workgroup of 256 threads: thread t takes column t and
// second-layer entry t, threads 0 and 1 the added edges' distances, threads
// below 128 the hidden units of the score.  Every global load is issued
// before any is used (three dependent rounds: the move, the level words and
// entries, the entries' rows).  Counts are integers summed per wave, then
// over the four waves; lane i of wave 0 writes feature i.
__device__ __forceinline__ void features_wg_body(const FeatArgs &f, int p, FeatWgLds &L) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int W = f.W, N = f.N, d = f.d, items = 4 * d;
    const int *m = f.moves + (size_t)p * 5;
    const int a = m[0], b = m[1], c = m[2], dd = m[3], kind = m[4];
    const bool wodd = f.score && f.wpar && (*f.wpar & 1);
    const int p1 = a, q1 = kind ? dd : c, p2 = b, q2 = kind ? c : dd;
    const int ends[4] = {a, b, c, dd};
    const int diam = feat_diam(f);
    // (1) second-layer entry z (item k = tid: endpoint e = k / d, entry k % d
    // of the row of e's partner e ^ 1); index clamped so every thread loads
    const int k = tid < items ? tid : items - 1;
    const int ue = k / d;
    const int uz = f.adjR[(size_t)pick4(b, a, dd, c, ue) * d + (k - ue * d)];
    // (2) the endpoints' ball words of levels 1..3 at column tid
    const int w = tid < W ? tid : W - 1;
    const int ew = f.base_parts[w].ecc;
    unsigned long long R[4][4];
#pragma unroll
    for (int e = 0; e < 4; e++)
#pragma unroll
        for (int l = 1; l <= 3; l++) R[e][l] = f.rows[((size_t)(l - 1) * N + ends[e]) * W + w];
    // (3) added edge t = tid & 1: x's words at y's column
    const int tx = (tid & 1) ? p2 : p1, ty = (tid & 1) ? q2 : q1;
    const int ewd = f.base_parts[ty >> 6].ecc;
    unsigned long long b3[3];
#pragma unroll
    for (int l = 1; l <= 3; l++) b3[l - 1] = f.rows[((size_t)(l - 1) * N + tx) * W + (ty >> 6)];
    const __hip_bfloat16 *wq = wodd ? f.wq2 : f.wq;
    // (4) the row of the second-layer entry
    int4 ur[GG_FEAT_FAST_D / 4];
    {
        const int4 *r = reinterpret_cast<const int4 *>(f.adjR + (size_t)uz * d);
#pragma unroll
        for (int j = 0; j < GG_FEAT_FAST_D / 4; j++)
            ur[j] = 4 * j < d ? r[j] : make_int4(-1, -1, -1, -1);
    }
    // counts of column tid (each at most 128)
    int cnt[24];
#pragma unroll
    for (int i = 0; i < 24; i++) cnt[i] = 0;
    int eccl[4] = {0, 0, 0, 0};
    if (tid < W) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            R[e][0] = ball_word(f, 0, tid, ends[e], ew);
            L.r1[e][tid] = R[e][1];
        }
        column_counts(f, tid, ew, ends, kind, R, cnt, eccl);
    }
    {
        unsigned pk[12];
#pragma unroll
        for (int j = 0; j < 12; j++) pk[j] = (unsigned)cnt[2 * j] | ((unsigned)cnt[2 * j + 1] << 16);
        const unsigned tot = wave_sum12_scatter(pk, lane);
        const int word = ((lane & 32) ? 6 : 0) + ((lane & 16) ? 3 : 0) + ((lane & 8) ? 2 : 0) +
                         ((lane & 4) ? 1 : 0);
        if ((lane & 3) == 0 && !((lane & 8) && (lane & 4))) L.cnt[wv][word] = tot;
    }
#pragma unroll
    for (int e = 0; e < 4; e++) eccl[e] = wave_max32(eccl[e]);
    if (lane == 0) {
#pragma unroll
        for (int e = 0; e < 4; e++) L.ecc[wv][e] = eccl[e];
    }
    if (tid < 2) L.dist[tid] = pair_dist(f, tx, ty, ewd, b3);
    __syncthreads();
    // U(u, v) of the four directions: entry z of v's row not in B1(u) with
    // exactly one neighbour in B1(u); per-direction counts (at most d <= 32)
    // as the bytes of one word
    {
        auto ball1 = [&](int e, int z) -> bool {
            return (L.r1[e][z >> 6] >> (z & 63)) & 1ull;
        };
        int hit = 0;
        if (tid < items && !(uz == pick4(a, b, c, dd, ue) || ball1(ue, uz))) {
            int common = 0;
#pragma unroll
            for (int j = 0; j < GG_FEAT_FAST_D / 4; j++) {
                if (4 * j >= d) break;
                const int4 q = ur[j];
                common += (ball1(ue, q.x) ? 1 : 0) + (ball1(ue, q.y) ? 1 : 0) +
                          (ball1(ue, q.z) ? 1 : 0) + (ball1(ue, q.w) ? 1 : 0);
            }
            hit = common == 1 ? 1 : 0;
        }
        unsigned x = (unsigned)hit << (8 * ue);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += (unsigned)__shfl_xor((int)x, o, 64);
        if (lane == 0) L.uq[wv] = x;
    }
    __syncthreads();
    if (wv == 0 && lane < GG_FEATS) {
        // feature `lane`, with the expression of write_features
        const int i = lane;
        auto count = [&](int j) {
            const int h = (j & 1) ? 16 : 0, wj = j >> 1;
            return (int)(((L.cnt[0][wj] >> h) & 0xffffu) + ((L.cnt[1][wj] >> h) & 0xffffu) +
                         ((L.cnt[2][wj] >> h) & 0xffffu) + ((L.cnt[3][wj] >> h) & 0xffffu));
        };
        const float d1 = (float)d, d2 = d1 * d1, n = (float)N, inv_diam = 1.f / (float)diam;
        float x;
        if (i < 16) {
            const int t = i >> 3, r = i & 7;
            if (r == 0) {
                x = L.dist[t] * inv_diam;
            } else {
                const float den = r <= 2 ? n : r <= 5 ? d2 : d1;
                x = (float)count(t * 7 + r - 1) / den;
            }
        } else if (i < 28) {
            const int t = (i - 16) / 6, r = (i - 16) - 6 * t;
            if (r < 2) {
                const int e = 2 * t + r;
                const unsigned u = L.uq[0] + L.uq[1] + L.uq[2] + L.uq[3];
                x = (float)(int)((u >> (8 * e)) & 0xffu) / d1;
            } else if (r < 5) {
                x = (float)count(14 + t * 3 + r - 2) / d2;
            } else {
                int e0 = 0, e1 = 0;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    e0 = L.ecc[q][2 * t] > e0 ? L.ecc[q][2 * t] : e0;
                    e1 = L.ecc[q][2 * t + 1] > e1 ? L.ecc[q][2 * t + 1] : e1;
                }
                x = (float)(e0 + e1 - 2 * diam) * 0.5f;
            }
        } else {
            x = (float)count(20 + i - 28) / n;
        }
        f.X[(size_t)p * GG_FEATS + i] = x;
        L.x[i] = x;
    }
    if (!f.score) return;
    // hidden unit tid of the score (its weights loaded while wave 0 writes
    // the features)
    uint4 w1[GG_FEATS / 8];
    float b1 = 0.f, w2 = 0.f;
    if (tid < GG_HIDDEN) {
        const uint4 *r = reinterpret_cast<const uint4 *>(wq + OFF_W1 + tid * GG_FEATS);
#pragma unroll
        for (int i = 0; i < GG_FEATS / 8; i++) w1[i] = r[i];
        b1 = bf(wq[OFF_B1 + tid]);
        w2 = bf(wq[OFF_W2 + tid]);
    }
    __syncthreads();
    // the MLP score with the arithmetic of k_mlp_forward (same products, fma
    // order and sums; products rounded before they are summed, as
    // k_mlp_forward's pass through LDS does): bit-identical
    {
#pragma clang fp contract(off)
        if (tid < GG_HIDDEN) {
            float acc = b1;
#pragma unroll
            for (int i = 0; i < GG_FEATS / 8; i++) {
                const uint4 u = w1[i];
                acc = fmaf(bf_lo(u.x), L.x[8 * i + 0], acc);
                acc = fmaf(bf_hi(u.x), L.x[8 * i + 1], acc);
                acc = fmaf(bf_lo(u.y), L.x[8 * i + 2], acc);
                acc = fmaf(bf_hi(u.y), L.x[8 * i + 3], acc);
                acc = fmaf(bf_lo(u.z), L.x[8 * i + 4], acc);
                acc = fmaf(bf_hi(u.z), L.x[8 * i + 5], acc);
                acc = fmaf(bf_lo(u.w), L.x[8 * i + 6], acc);
                acc = fmaf(bf_hi(u.w), L.x[8 * i + 7], acc);
            }
            L.rh[tid] = w2 * (acc > 0.f ? acc : 0.f);
        }
        __syncthreads();
        if (wv == 0) {
            float o = 0.f;
#pragma unroll
            for (int q = 0; q < GG_HIDDEN / 64; q++) o = o + L.rh[lane + 64 * q];
#pragma unroll
            for (int x = 32; x > 0; x >>= 1) o = o + __shfl_xor(o, x, 64);
            if (lane == 0) f.score[p] = o + bf(wq[OFF_B2]);
        }
    }
}

// One wave per proposal (workgroup bid: proposals 4 bid ..) where the
// workgroup pass does not apply.  Counts are wave-reduced popcounts over the
// W words.
__device__ __forceinline__ void features_body(const FeatArgs &f, int bid) {
    __shared__ unsigned long long s_r1[4][4][GG_FEAT_LDS_W];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // a wave past the last proposal computes proposal 0 and writes nothing,
    // so every wave reaches the workgroup barrier
    const int p0 = bid * 4 + wv;
    const bool active = p0 < f.P;
    const int p = active ? p0 : 0;
    const bool stage = f.W <= GG_FEAT_LDS_W;
    const int *m = f.moves + (size_t)p * 5;
    const int a = m[0], b = m[1], c = m[2], dd = m[3], kind = m[4];
    const bool wodd = f.score && f.wpar && (*f.wpar & 1);
    // added edges (p1,q1),(p2,q2); removed (a,b),(c,dd)
    const int p1 = a, q1 = kind ? dd : c, p2 = b, q2 = kind ? c : dd;
    const int ends[4] = {a, b, c, dd};
    // per-lane partial counts
    float cnt[24];
#pragma unroll
    for (int i = 0; i < 24; i++) cnt[i] = 0.f;
    // eccentricity of each endpoint x: the first level whose ball row R_l(x)
    // is full in every column (per lane over its columns, then the wave max)
    int eccl[4] = {0, 0, 0, 0};
    int uq[4];
    int dist[2];
    float xf[GG_FEATS];
#pragma unroll
    for (int i = 0; i < GG_FEATS; i++) xf[i] = 0.f;
    const int diam = feat_diam(f);
    for (int w = lane; w < f.W; w += 64) {
        const int ew = f.base_parts[w].ecc;
        unsigned long long R[4][4];  // R[endpoint][level 0..3]
#pragma unroll
        for (int e = 0; e < 4; e++)
#pragma unroll
            for (int l = 0; l < 4; l++) R[e][l] = ball_word(f, l, w, ends[e], ew);
        if (stage) {
#pragma unroll
            for (int e = 0; e < 4; e++) s_r1[wv][e][w] = R[e][1];
        }
        column_counts(f, w, ew, ends, kind, R, cnt, eccl);
    }
#pragma unroll
    for (int i = 0; i < 24; i++) {
        float v = cnt[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        cnt[i] = v;
    }
#pragma unroll
    for (int e = 0; e < 4; e++) eccl[e] = wave_max32(eccl[e]);
    __syncthreads();
    uniq_second_layer4(f, ends, stage ? s_r1[wv] : nullptr, lane, uq);
    if (lane == 0) {
        dist[0] = pair_dist(f, p1, q1);
        dist[1] = pair_dist(f, p2, q2);
        write_features(f, diam, dist, cnt, eccl, uq, xf);
    }
    if (lane == 0 && active) {
        float *X = f.X + (size_t)p * GG_FEATS;
#pragma unroll
        for (int i = 0; i < GG_FEATS; i++) X[i] = xf[i];
    }
    if (!f.score) return;
    float xs[GG_FEATS];
#pragma unroll
    for (int i = 0; i < GG_FEATS; i++) xs[i] = __shfl(xf[i], 0, 64);
    LaneWeights lw;
    load_lane_weights(wodd ? f.wq2 : f.wq, lane, lw);
    const float o = mlp_score(xs, lw.w1, lw.b1, lw.w2, lw.b2);
    if (lane == 0 && active) f.score[p] = o;
}

__global__ void __launch_bounds__(256) k_features(FeatArgs f) { features_body(f, blockIdx.x); }
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_features_wg(FeatArgs f) {
    __shared__ FeatWgLds L;
    features_wg_body(f, blockIdx.x, L);
}

// ---- MLP ---------------------------------------------------------------------------------

// Forward of P samples: workgroup g takes samples [16g, 16g+16), threads h and
// h+128 compute hidden unit h for half of them each, one wave per four
// samples sums the output layer.  bf16 weights staged in LDS, fp32
// accumulation.
#define GG_FWD_SAMPLES 16
__global__ void __launch_bounds__(256) k_mlp_forward(const float *X, int P,
                                                     const __hip_bfloat16 *wq, float *out) {
    __shared__ float sw[GG_NPARAMS];
    __shared__ float sx[GG_FWD_SAMPLES][GG_FEATS];
    __shared__ float sr[GG_FWD_SAMPLES][GG_HIDDEN];  // w2[h] * relu(hidden h)
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int s0 = blockIdx.x * GG_FWD_SAMPLES;
    const int ns = min(GG_FWD_SAMPLES, P - s0);
    for (int i = tid; i < GG_NPARAMS; i += 256) sw[i] = bf(wq[i]);
    for (int i = tid; i < GG_FWD_SAMPLES * GG_FEATS; i += 256) {
        const int s = i / GG_FEATS, f = i - s * GG_FEATS;
        sx[s][f] = s < ns ? X[(size_t)(s0 + s) * GG_FEATS + f] : 0.f;
    }
    __syncthreads();
    const int h = tid & (GG_HIDDEN - 1), half = tid >> 7;
    {
        float w1[GG_FEATS];
#pragma unroll
        for (int i = 0; i < GG_FEATS; i++) w1[i] = sw[OFF_W1 + h * GG_FEATS + i];
        const float b1 = sw[OFF_B1 + h], w2 = sw[OFF_W2 + h];
        for (int q = 0; q < GG_FWD_SAMPLES / 2; q++) {
            const int s = half * (GG_FWD_SAMPLES / 2) + q;
            float acc = b1;
#pragma unroll
            for (int i = 0; i < GG_FEATS; i++) acc = fmaf(w1[i], sx[s][i], acc);
            sr[s][h] = w2 * (acc > 0.f ? acc : 0.f);
        }
    }
    __syncthreads();
    for (int s = wv; s < GG_FWD_SAMPLES; s += 4) {
        float o = 0.f;
        for (int k = lane; k < GG_HIDDEN; k += 64) o += sr[s][k];
        for (int d = 32; d > 0; d >>= 1) o += __shfl_xor(o, d, 64);
        if (lane == 0 && s < ns) out[s0 + s] = o + sw[OFF_B2];
    }
}

// Gradient of the mean squared error over a minibatch of B samples (rows
// idx[0..B) of X/Y), forward with the bf16 weights, fp32 accumulation.
// Group g takes samples [16g, 16g+16) and writes its partial gradient (and
// partial squared error at index GG_NPARAMS) to part[g][]; the reduction
// sums the partials in group order, so the result is deterministic.
struct GradLds {
    float sx[GG_GRAD_SAMPLES][GG_FEATS];
    float sa[GG_GRAD_SAMPLES][GG_HIDDEN];  // hidden pre-activations
    float sg[GG_GRAD_SAMPLES];             // dLoss/dOutput (0 for padding)
    float se[GG_GRAD_SAMPLES];             // squared error
    int si[GG_GRAD_SAMPLES];               // sample rows (idx may be host memory)
};
__device__ __forceinline__ void grad_part_body(const float *X, const float *Y, const int *idx,
                                               int B, const __hip_bfloat16 *wq, float *part,
                                               int g, GradLds &L) {
    float(&sx)[GG_GRAD_SAMPLES][GG_FEATS] = L.sx;
    float(&sa)[GG_GRAD_SAMPLES][GG_HIDDEN] = L.sa;
    float(&sg)[GG_GRAD_SAMPLES] = L.sg;
    float(&se)[GG_GRAD_SAMPLES] = L.se;
    int(&si)[GG_GRAD_SAMPLES] = L.si;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int s0 = g * GG_GRAD_SAMPLES;
    const int ns = min(GG_GRAD_SAMPLES, B - s0);
    // threads h and h+128 each take half of the samples for hidden unit h
    const int h = tid & (GG_HIDDEN - 1), half = tid >> 7;
    if (tid < ns) si[tid] = idx ? idx[s0 + tid] : s0 + tid;
    // the weights this thread uses, read from the bf16 copy as they are
    // needed (row h of W1, b1[h], w2[h]; the output weights of hidden units
    // lane and lane + 64, b2)
    uint4 w1q[GG_FEATS / 8];
    {
        const uint4 *r = reinterpret_cast<const uint4 *>(wq + OFF_W1 + h * GG_FEATS);
#pragma unroll
        for (int i = 0; i < GG_FEATS / 8; i++) w1q[i] = r[i];
    }
    const float b1 = bf(wq[OFF_B1 + h]), w2h = bf(wq[OFF_W2 + h]);
    const float w2o[2] = {bf(wq[OFF_W2 + lane]), bf(wq[OFF_W2 + lane + 64])};
    const float b2 = bf(wq[OFF_B2]);
    __syncthreads();
    for (int i = tid; i < GG_GRAD_SAMPLES * GG_FEATS; i += 256) {
        const int s = i / GG_FEATS, f = i - s * GG_FEATS;
        float v = 0.f;
        if (s < ns) v = X[(size_t)si[s] * GG_FEATS + f];
        sx[s][f] = v;
    }
    __syncthreads();
    {
        float w1[GG_FEATS];
#pragma unroll
        for (int i = 0; i < GG_FEATS / 8; i++) {
            w1[8 * i + 0] = bf_lo(w1q[i].x);
            w1[8 * i + 1] = bf_hi(w1q[i].x);
            w1[8 * i + 2] = bf_lo(w1q[i].y);
            w1[8 * i + 3] = bf_hi(w1q[i].y);
            w1[8 * i + 4] = bf_lo(w1q[i].z);
            w1[8 * i + 5] = bf_hi(w1q[i].z);
            w1[8 * i + 6] = bf_lo(w1q[i].w);
            w1[8 * i + 7] = bf_hi(w1q[i].w);
        }
        for (int q = 0; q < GG_GRAD_SAMPLES / 2; q++) {
            const int s = half * (GG_GRAD_SAMPLES / 2) + q;
            float acc = b1;
#pragma unroll
            for (int i = 0; i < GG_FEATS; i++) acc = fmaf(w1[i], sx[s][i], acc);
            sa[s][h] = acc;
        }
    }
    __syncthreads();
    for (int s = wv; s < GG_GRAD_SAMPLES; s += 4) {
        float p = 0.f;
#pragma unroll
        for (int q = 0; q < GG_HIDDEN / 64; q++) {
            const float a = sa[s][lane + 64 * q];
            p = fmaf(w2o[q], a > 0.f ? a : 0.f, p);
        }
        for (int o = 32; o > 0; o >>= 1) p += __shfl_xor(p, o, 64);
        if (lane == 0) {
            float e = 0.f;
            if (s < ns) e = (p + b2) - Y[si[s]];
            sg[s] = 2.f * e / (float)B;
            se[s] = e * e;
        }
    }
    __syncthreads();
    const float w2 = w2h;
    float gw1[GG_FEATS / 2];
#pragma unroll
    for (int i = 0; i < GG_FEATS / 2; i++) gw1[i] = 0.f;
    float gb1 = 0.f, gw2 = 0.f;
    for (int s = 0; s < GG_GRAD_SAMPLES; s++) {
        const float a = sa[s][h];
        if (a <= 0.f) continue;
        const float g = sg[s];
        gw2 = fmaf(g, a, gw2);
        const float gh = g * w2;
        gb1 += gh;
#pragma unroll
        for (int i = 0; i < GG_FEATS / 2; i++)
            gw1[i] = fmaf(gh, sx[s][half * (GG_FEATS / 2) + i], gw1[i]);
    }
    float *out = part + (size_t)g * GG_PART_STRIDE;
#pragma unroll
    for (int i = 0; i < GG_FEATS / 2; i++)
        out[OFF_W1 + h * GG_FEATS + half * (GG_FEATS / 2) + i] = gw1[i];
    if (half == 0) {
        out[OFF_B1 + h] = gb1;
        out[OFF_W2 + h] = gw2;
    }
    if (tid == 0) {
        float gb2 = 0.f, l = 0.f;
        for (int s = 0; s < GG_GRAD_SAMPLES; s++) { gb2 += sg[s]; l += se[s]; }
        out[OFF_B2] = gb2;
        out[GG_NPARAMS] = l;
    }
}

__global__ void __launch_bounds__(256) k_mlp_grad_part(const float *X, const float *Y,
                                                       const int *idx, int B,
                                                       const __hip_bfloat16 *wq, float *part,
                                                       const TrainGate *gate,
                                                       const __hip_bfloat16 *wq2,
                                                       const int *par) {
    if (gate && !gate->on) return;
    __shared__ GradLds L;
    grad_part_body(X, Y, idx, B, (par && (*par & 1)) ? wq2 : wq, part, blockIdx.x, L);
}

// Features and scores of the P proposals (workgroups < nfeat) beside the
// gradient of a gated training step (the workgroups after them): a kernel that
// leaves most CUs idle carries the learned search's training on them.
__global__ void __launch_bounds__(256) k_features_grad(FeatArgs f, GradArgs g, int nfeat) {
    if ((int)blockIdx.x < nfeat) {
        features_body(f, blockIdx.x);
        return;
    }
    if (!g.gate->on) return;
    __shared__ GradLds L;
    grad_part_body(g.X, g.Y, g.idx, g.B, (*g.par & 1) ? g.wq2 : g.wq, g.part,
                   (int)blockIdx.x - nfeat, L);
}

// The same with one workgroup per proposal (features_fast_ok), the two
// bodies sharing one LDS block.  The ngrad gradient workgroups come first, so
// they are dispatched with the first proposals rather than after the last.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_features_wg_grad(FeatArgs f, GradArgs g, int ngrad) {
    __shared__ union {
        FeatWgLds f;
        GradLds g;
    } L;
    if ((int)blockIdx.x >= ngrad) {
        features_wg_body(f, (int)blockIdx.x - ngrad, L.f);
        return;
    }
    if (!g.gate->on) return;
    grad_part_body(g.X, g.Y, g.idx, g.B, (*g.par & 1) ? g.wq2 : g.wq, g.part, blockIdx.x, L.g);
}

// Gradient reduction and Adam step in one launch, one thread per parameter
// (reduce_adam_body).  Same results as the two kernels.
__global__ void __launch_bounds__(256) k_mlp_reduce_adam(AdamArgs a) {
    reduce_adam_body(a, blockIdx.x * 256 + threadIdx.x);
}

// Sums the G partial gradients (fixed order) into grad[] and the mean loss.
__global__ void __launch_bounds__(256) k_mlp_reduce(const float *part, int G, int B, float *grad,
                                                    float *loss, const TrainGate *gate) {
    if (gate && !gate->on) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i > GG_NPARAMS) return;
    float acc = 0.f;
    for (int g = 0; g < G; g++) acc += part[(size_t)g * GG_PART_STRIDE + i];
    if (i < GG_NPARAMS) grad[i] = acc;
    else if (loss) *loss = acc / (float)B;
}

// Adam step on the fp32 masters, then refresh the bf16 copies.  With a gate
// the bias corrections beta^t come from it (the step count lives on the GPU).
__global__ void __launch_bounds__(256) k_mlp_adam(float *w, float *m, float *v,
                                                  const float *grad, __hip_bfloat16 *wq, float lr,
                                                  float b1t, float b2t, const TrainGate *gate) {
    if (gate) {
        if (!gate->on) return;
        b1t = gate->b1t;
        b2t = gate->b2t;
    }
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= GG_NPARAMS) return;
    adam_update(w, m, v, wq, i, grad[i], lr, b1t, b2t);
}

__global__ void k_to_bf16(const float *w, __hip_bfloat16 *wq, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) wq[i] = __float2bfloat16(w[i]);
}

// ---- host objects ------------------------------------------------------------------------
static int mlp_alloc(gg_mlp *m) {
    GG_HIP(hipSetDevice(m->device));
    GG_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    const size_t P = GG_NPARAMS;
    if (hipMalloc((void **)&m->d_w, 4 * P) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&m->d_m, 4 * P) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&m->d_v, 4 * P) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&m->d_g, 4 * P) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&m->d_wq, 2 * P) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&m->d_wq2, 2 * P) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&m->d_pre,
                  4 * (size_t)(GG_MAX_BATCH / GG_GRAD_SAMPLES) * GG_PART_STRIDE) != hipSuccess)
        return GG_OOM;
    if (hipMalloc((void **)&m->d_loss, 4) != hipSuccess) return GG_OOM;
    m->xcap = 0;
    return GG_OK;
}

static int ensure_io(gg_mlp *m, int n) {
    if (n <= m->xcap) return GG_OK;
    if (m->d_x) (void)hipFree(m->d_x);
    if (m->d_y) (void)hipFree(m->d_y);
    if (m->d_out) (void)hipFree(m->d_out);
    m->d_x = nullptr; m->d_y = nullptr; m->d_out = nullptr;
    if (hipMalloc((void **)&m->d_x, 4 * (size_t)n * GG_FEATS) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&m->d_y, 4 * (size_t)n) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&m->d_out, 4 * (size_t)n) != hipSuccess) return GG_OOM;
    m->xcap = n;
    return GG_OK;
}

// He-style init from the shared explicit RNG (deterministic per seed).
int gg::mlp_init_params(std::vector<float> &w, uint64_t seed) {
    gg::Rng rng(seed);
    w.assign(GG_NPARAMS, 0.f);
    const double s1 = std::sqrt(2.0 / GG_FEATS), s2 = std::sqrt(1.0 / GG_HIDDEN);
    auto normal = [&]() {  // Box-Muller on two uniforms
        double u1 = rng.uniform(), u2 = rng.uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int i = 0; i < GG_HIDDEN * GG_FEATS; i++) w[OFF_W1 + i] = (float)(normal() * s1);
    for (int i = 0; i < GG_HIDDEN; i++) w[OFF_W2 + i] = (float)(normal() * s2);
    return GG_OK;
}

extern "C" int gg_mlp_create(int device, uint64_t seed, gg_mlp **out) {
    if (!out) return GG_BAD_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return GG_NO_DEVICE;
    if (device < 0 || device >= n) return GG_BAD_ARG;
    gg_mlp *m = new gg_mlp();
    m->device = device;
    int st = mlp_alloc(m);
    if (st == GG_OK) {
        std::vector<float> w;
        gg::mlp_init_params(w, seed);
        st = gg_mlp_set_params(m, w.data());
    }
    if (st != GG_OK) { gg_mlp_destroy(m); return st; }
    *out = m;
    return GG_OK;
}

extern "C" void gg_mlp_destroy(gg_mlp *m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void *bufs[] = {m->d_w, m->d_m, m->d_v, m->d_g, m->d_wq, m->d_wq2, m->d_pre, m->d_loss, m->d_x,
                    m->d_y, m->d_out};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

extern "C" int gg_mlp_set_params(gg_mlp *m, const float *params) {
    if (!m || !params) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(m->device));
    GG_HIP(hipMemcpyAsync(m->d_w, params, 4 * GG_NPARAMS, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_to_bf16, dim3((GG_NPARAMS + 255) / 256), dim3(256), 0, m->stream, m->d_w,
                       m->d_wq, GG_NPARAMS);
    GG_HIP(hipGetLastError());
    std::vector<float> z(GG_NPARAMS, 0.f);
    GG_HIP(hipMemcpyAsync(m->d_m, z.data(), 4 * GG_NPARAMS, hipMemcpyHostToDevice, m->stream));
    GG_HIP(hipMemcpyAsync(m->d_v, z.data(), 4 * GG_NPARAMS, hipMemcpyHostToDevice, m->stream));
    GG_HIP(hipStreamSynchronize(m->stream));
    m->t = 0;
    return GG_OK;
}

extern "C" int gg_mlp_get_params(gg_mlp *m, float *params) {
    if (!m || !params) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(m->device));
    GG_HIP(hipMemcpyAsync(params, m->d_w, 4 * GG_NPARAMS, hipMemcpyDeviceToHost, m->stream));
    GG_HIP(hipStreamSynchronize(m->stream));
    return GG_OK;
}

int gg::mlp_forward_dev(gg_mlp *m, const float *d_x, int P, float *d_out, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_forward, dim3((P + GG_FWD_SAMPLES - 1) / GG_FWD_SAMPLES), dim3(256),
                       0, st, d_x, P, m->d_wq, d_out);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

int gg::mlp_grad_dev(gg_mlp *m, const float *d_x, const float *d_y, const int *d_idx, int B,
                     hipStream_t st) {
    if (B < 1 || B > GG_MAX_BATCH) return GG_BAD_ARG;
    const int G = (B + GG_GRAD_SAMPLES - 1) / GG_GRAD_SAMPLES;
    hipLaunchKernelGGL(k_mlp_grad_part, dim3(G), dim3(256), 0, st, d_x, d_y, d_idx, B, m->d_wq,
                       m->d_pre, (const TrainGate *)nullptr, (const __hip_bfloat16 *)nullptr,
                       (const int *)nullptr);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(GG_NPARAMS / 256 + 1), dim3(256), 0, st, m->d_pre, G,
                       B, m->d_g, m->d_loss, (const TrainGate *)nullptr);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

int gg::mlp_train_gated_dev(gg_mlp *m, const float *d_x, const float *d_y, const int *d_idx,
                            int B, float lr, const TrainGate *gate, hipStream_t st,
                            const int *par) {
    if (B < 1 || B > GG_MAX_BATCH || !gate) return GG_BAD_ARG;
    const int G = (B + GG_GRAD_SAMPLES - 1) / GG_GRAD_SAMPLES;
    hipLaunchKernelGGL(k_mlp_grad_part, dim3(G), dim3(256), 0, st, d_x, d_y, d_idx, B, m->d_wq,
                       m->d_pre, gate, (const __hip_bfloat16 *)m->d_wq2, par);
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3(GG_ADAM_GROUPS), dim3(256), 0, st,
                       mlp_adam_args(m, B, lr, gate, par));
    GG_HIP(hipGetLastError());
    return GG_OK;
}

GradArgs gg::mlp_grad_args(gg_mlp *m, const float *d_x, const float *d_y, const int *d_idx,
                           int B, const TrainGate *gate, const int *par) {
    GradArgs g{};
    g.X = d_x; g.Y = d_y; g.idx = d_idx; g.B = B;
    g.wq = m->d_wq; g.wq2 = m->d_wq2; g.par = par; g.part = m->d_pre; g.gate = gate;
    return g;
}

AdamArgs gg::mlp_adam_args(gg_mlp *m, int B, float lr, const TrainGate *gate, const int *par) {
    AdamArgs a{};
    a.part = m->d_pre; a.G = (B + GG_GRAD_SAMPLES - 1) / GG_GRAD_SAMPLES; a.B = B;
    a.grad = m->d_g; a.loss = m->d_loss; a.w = m->d_w; a.m = m->d_m; a.v = m->d_v;
    a.wq = m->d_wq; a.wq2 = m->d_wq2; a.lr = lr; a.gate = gate; a.par = par;
    return a;
}

int gg::mlp_refresh_dev(gg_mlp *m, hipStream_t st) {
    hipLaunchKernelGGL(k_to_bf16, dim3((GG_NPARAMS + 255) / 256), dim3(256), 0, st, m->d_w,
                       m->d_wq, GG_NPARAMS);
    hipLaunchKernelGGL(k_to_bf16, dim3((GG_NPARAMS + 255) / 256), dim3(256), 0, st, m->d_w,
                       m->d_wq2, GG_NPARAMS);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

int gg::mlp_adam_dev(gg_mlp *m, float lr, hipStream_t st) {
    m->t++;
    const float b1t = std::pow(0.9f, (float)m->t), b2t = std::pow(0.999f, (float)m->t);
    hipLaunchKernelGGL(k_mlp_adam, dim3((GG_NPARAMS + 255) / 256), dim3(256), 0, st, m->d_w,
                       m->d_m, m->d_v, m->d_g, m->d_wq, lr, b1t, b2t, (const TrainGate *)nullptr);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

extern "C" int gg_mlp_forward(gg_mlp *m, const float *X, int P, float *out) {
    if (!m || !X || !out || P < 1) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(m->device));
    int st = ensure_io(m, P);
    if (st != GG_OK) return st;
    GG_HIP(hipMemcpyAsync(m->d_x, X, 4 * (size_t)P * GG_FEATS, hipMemcpyHostToDevice, m->stream));
    st = mlp_forward_dev(m, m->d_x, P, m->d_out, m->stream);
    if (st != GG_OK) return st;
    GG_HIP(hipMemcpyAsync(out, m->d_out, 4 * (size_t)P, hipMemcpyDeviceToHost, m->stream));
    GG_HIP(hipStreamSynchronize(m->stream));
    return GG_OK;
}

extern "C" int gg_mlp_grad(gg_mlp *m, const float *X, const float *Y, int B, float *loss,
                           float *grad) {
    if (!m || !X || !Y || !grad || B < 1 || B > GG_MAX_BATCH) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(m->device));
    int st = ensure_io(m, B);
    if (st != GG_OK) return st;
    GG_HIP(hipMemcpyAsync(m->d_x, X, 4 * (size_t)B * GG_FEATS, hipMemcpyHostToDevice, m->stream));
    GG_HIP(hipMemcpyAsync(m->d_y, Y, 4 * (size_t)B, hipMemcpyHostToDevice, m->stream));
    st = mlp_grad_dev(m, m->d_x, m->d_y, nullptr, B, m->stream);
    if (st != GG_OK) return st;
    GG_HIP(hipMemcpyAsync(grad, m->d_g, 4 * GG_NPARAMS, hipMemcpyDeviceToHost, m->stream));
    if (loss) GG_HIP(hipMemcpyAsync(loss, m->d_loss, 4, hipMemcpyDeviceToHost, m->stream));
    GG_HIP(hipStreamSynchronize(m->stream));
    return GG_OK;
}

extern "C" int gg_mlp_train(gg_mlp *m, const float *X, const float *Y, int B, float lr,
                            float *loss) {
    if (!m || !X || !Y || B < 1 || B > GG_MAX_BATCH || !(lr > 0)) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(m->device));
    int st = ensure_io(m, B);
    if (st != GG_OK) return st;
    GG_HIP(hipMemcpyAsync(m->d_x, X, 4 * (size_t)B * GG_FEATS, hipMemcpyHostToDevice, m->stream));
    GG_HIP(hipMemcpyAsync(m->d_y, Y, 4 * (size_t)B, hipMemcpyHostToDevice, m->stream));
    st = mlp_grad_dev(m, m->d_x, m->d_y, nullptr, B, m->stream);
    if (st != GG_OK) return st;
    st = mlp_adam_dev(m, lr, m->stream);
    if (st != GG_OK) return st;
    if (loss) GG_HIP(hipMemcpyAsync(loss, m->d_loss, 4, hipMemcpyDeviceToHost, m->stream));
    GG_HIP(hipStreamSynchronize(m->stream));
    return GG_OK;
}

// Features of P moves (a, b, c, d, kind) on the batch evaluator's base.
int gg::features_dev(gg_batch *b, const int *d_moves, int P, float *d_x, hipStream_t st,
                     const int *diam_dev) {
    return features_score_dev(b, nullptr, d_moves, P, d_x, nullptr, st, diam_dev);
}

// Features and, with m and d_score, the MLP scores of the P moves in one launch.
int gg::features_score_dev(gg_batch *b, gg_mlp *m, const int *d_moves, int P, float *d_x,
                           float *d_score, hipStream_t st, const int *diam_dev,
                           const int *wpar, const GradArgs *grad) {
    if (!b->sparse_ok) return GG_UNSUPPORTED;
    FeatArgs f{};
    f.diam_dev = diam_dev;
    f.N = b->N; f.W = b->W; f.S = b->snap_cap;
    f.diam = b->base_res.diameter > 0 ? b->base_res.diameter : 1;
    f.d = b->degree;
    f.snap = b->d_snap;
    f.rows = b->d_rows;
    f.RL = b->d_rows ? b->rows_levels : 0;
    f.base_parts = b->d_parts + (size_t)b->max_k * b->W;
    f.adjR = b->d_adjR;
    f.moves = d_moves;
    f.P = P;
    f.X = d_x;
    f.wq = m && d_score ? m->d_wq : nullptr;
    f.wq2 = m && d_score ? m->d_wq2 : nullptr;
    f.wpar = m && d_score ? wpar : nullptr;
    f.score = m ? d_score : nullptr;
    const int nfeat = features_groups(f);
    const bool wg = features_fast_ok(f);
    if (grad) {
        const int G = (grad->B + GG_GRAD_SAMPLES - 1) / GG_GRAD_SAMPLES;
        hipLaunchKernelGGL(wg ? k_features_wg_grad : k_features_grad, dim3(nfeat + G), dim3(256),
                           0, st, f, *grad, wg ? G : nfeat);
    } else {
        hipLaunchKernelGGL(wg ? k_features_wg : k_features, dim3(nfeat), dim3(256), 0, st, f);
    }
    GG_HIP(hipGetLastError());
    return GG_OK;
}

extern "C" int gg_batch_features(gg_batch *b, const int32_t *moves, int P, float *X) {
    if (!b || !moves || !X || P < 1 || !b->have_base) return GG_BAD_ARG;
    for (int p = 0; p < P; p++)
        for (int j = 0; j < 4; j++)
            if (moves[p * 5 + j] < 0 || moves[p * 5 + j] >= b->N) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(b->device));
    int *d_m = nullptr;
    float *d_x = nullptr;
    if (hipMalloc((void **)&d_m, 4 * (size_t)P * 5) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&d_x, 4 * (size_t)P * GG_FEATS) != hipSuccess) {
        (void)hipFree(d_m);
        return GG_OOM;
    }
    int st = GG_OK;
    if (hipMemcpyAsync(d_m, moves, 4 * (size_t)P * 5, hipMemcpyHostToDevice, b->stream) !=
        hipSuccess)
        st = GG_HIP_ERROR;
    if (st == GG_OK) st = features_dev(b, d_m, P, d_x, b->stream);
    if (st == GG_OK &&
        hipMemcpyAsync(X, d_x, 4 * (size_t)P * GG_FEATS, hipMemcpyDeviceToHost, b->stream) !=
            hipSuccess)
        st = GG_HIP_ERROR;
    if (hipStreamSynchronize(b->stream) != hipSuccess) st = GG_HIP_ERROR;
    (void)hipFree(d_m);
    (void)hipFree(d_x);
    return st;
}

extern "C" int gg_batch_score(gg_batch *b, gg_mlp *m, const int32_t *moves, int P, float *X,
                              float *scores) {
    if (!b || !m || !moves || !X || !scores || P < 1 || !b->have_base) return GG_BAD_ARG;
    if (m->device != b->device) return GG_BAD_ARG;
    for (int p = 0; p < P; p++)
        for (int j = 0; j < 4; j++)
            if (moves[p * 5 + j] < 0 || moves[p * 5 + j] >= b->N) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(b->device));
    int *d_m = nullptr;
    float *d_x = nullptr, *d_s = nullptr;
    int st = GG_OK;
    if (hipMalloc((void **)&d_m, 4 * (size_t)P * 5) != hipSuccess ||
        hipMalloc((void **)&d_x, 4 * (size_t)P * GG_FEATS) != hipSuccess ||
        hipMalloc((void **)&d_s, 4 * (size_t)P) != hipSuccess)
        st = GG_OOM;
    // the weights are written on the MLP's stream: order after them
    if (st == GG_OK && hipStreamSynchronize(m->stream) != hipSuccess) st = GG_HIP_ERROR;
    if (st == GG_OK &&
        hipMemcpyAsync(d_m, moves, 4 * (size_t)P * 5, hipMemcpyHostToDevice, b->stream) !=
            hipSuccess)
        st = GG_HIP_ERROR;
    if (st == GG_OK) st = features_score_dev(b, m, d_m, P, d_x, d_s, b->stream);
    if (st == GG_OK &&
        (hipMemcpyAsync(X, d_x, 4 * (size_t)P * GG_FEATS, hipMemcpyDeviceToHost, b->stream) !=
             hipSuccess ||
         hipMemcpyAsync(scores, d_s, 4 * (size_t)P, hipMemcpyDeviceToHost, b->stream) !=
             hipSuccess))
        st = GG_HIP_ERROR;
    if (hipStreamSynchronize(b->stream) != hipSuccess && st == GG_OK) st = GG_HIP_ERROR;
    if (d_m) (void)hipFree(d_m);
    if (d_x) (void)hipFree(d_x);
    if (d_s) (void)hipFree(d_s);
    return st;
}
