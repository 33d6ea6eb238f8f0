// gg_learn_search.hip -- the learned mode of the search driver.
//
// Per step: the P proposals' features are computed on the GPU from the base
// BFS levels, the MLP scores them, and the K lowest predicted distance-total
// changes are evaluated exactly (a quarter of K is always drawn uniformly so
// the training data stays unbiased; all K are uniform until warmup_pairs exact
// pairs exist).  Every exactly evaluated candidate becomes a training pair
// (features, normalised exact delta) in a device replay buffer, and the MLP
// takes Adam steps on random minibatches of that buffer.  Acceptance is always
// decided on the exact results.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>
#include "gg_common.hpp"
#include "gg_batch.hpp"
#include "gg_learn.hpp"
#include "gg_search_impl.hpp"
#include "gg_devrng.hpp"
#include "gg_sparse.hpp"
#include "gg_propose.hpp"

namespace gg {

// ---- the resident loop's device state --------------------------------------------------
// After the scorer's warm-up a learned step replays as one captured graph that
// also decides acceptance, writes the training targets, draws the minibatch,
// takes the Adam step and commits in place: the host launches steps back to
// back and reads one StepRec per step when a chunk of steps ends.
#define GG_LOOP_HOST_COMMIT 1  // the accepted candidate has no kept logs: host commit
#define GG_LOOP_NO_SWAP 2      // a proposal found no simple swap
#define GG_LOOP_REC_CAP 512    // steps per chunk (records in device memory, read after it)
#define GG_LOOP_INFLIGHT 4     // graph launches enqueued ahead of the GPU
#define GG_LOOP_UNROLL 8       // steps per graph launch
#define GG_ACCEPT_LANE 0xFFFFFFFEull
#define GG_BATCH_SALT 0x5851F42D4C957F2Dull

struct LoopCtl {
    int hdr[4];            // step before this one (low 32 bits), warm-up flag, diameter (>= 1), -
    long long step;        // search step before this one
    long long cur_total;   // current graph: distance total, diameter, levels, status, WVC
    int cur_diam, cur_levels, cur_status, cur_wvc;
    int head, count;       // replay ring position and fill
    int halted;            // a step of this chunk stopped it: the later steps are void
    int reason;            // GG_LOOP_*
    int nrec;              // steps recorded in this chunk
    int trained;           // Adam steps taken in this chunk
    int adam_t;            // Adam step count (bias corrections)
    int pad;
};

struct StepRec {
    long long step;        // the step's number, 1-based
    long long total;       // the best candidate's exact result
    int diameter, levels, status, wvc;
    int bi, accept, inplace, pad;
    int mv[5];
    int pad2;
    gg_swap sw;
};

// ---- proposals and selection on the GPU ------------------------------------------------
#define GG_SELECT_MAX_P 4096
#define GG_EXPLORE_LANE 0xFFFFFFFFull

__global__ void __launch_bounds__(64) k_propose(ProposeArgs pa) { propose_body(pa, blockIdx.x); }

// Rank of every score among the P (NaN as +inf, ties by index).  A workgroup
// of 256 threads ranks 16 proposals: all P scores staged in LDS, thread t
// counts the predecessors of proposal 16g + (t mod 16) within the (t / 16)-th
// sixteenth of the scores, and the sixteen partial counts are summed in LDS.
#define GG_RANK_PER 16
// Where the K chosen proposals go (the rows k_select writes).
struct SelArgs {
    int K;
    uint64_t seed;
    const int *ctl;          // [0] step, [1] warm-up flag
    const gg_swap *psw;      // [P] proposals
    const int *pmoves;       // [P][5]
    gg_swap *out_sw;         // [K]
    int *out_moves;          // [K][5]
    int *out_sel;            // [K]
};

__device__ __forceinline__ void sel_write(const SelArgs &s, int slot, int p) {
    s.out_sw[slot] = s.psw[p];
#pragma unroll
    for (int j = 0; j < 5; j++) s.out_moves[slot * 5 + j] = s.pmoves[(size_t)p * 5 + j];
    s.out_sel[slot] = p;
}

// Ranks of proposals [16 bid, 16 bid + 16): rank[p] = #{q : s_q < s_p or
// (s_q = s_p and q < p)} with NaN scores last.  With sel, a proposal the
// selection takes writes its own row: ranks below G = K - K/4 go to slot
// rank (greedy part), and during the warm-up proposal p < K to slot p.
__device__ __forceinline__ void rank_body(const float *score, int P, int *rank, int bid,
                                          float *s_sc, const SelArgs *sel = nullptr) {
    __shared__ int s_part[16][GG_RANK_PER];
    const bool warm = sel && sel->ctl[1] != 0;  // loaded with the scores
    for (int q = threadIdx.x; q < P; q += 256) {
        const float x = score[q];
        s_sc[q] = x == x ? x : INFINITY;
    }
    __syncthreads();
    const int j = threadIdx.x & (GG_RANK_PER - 1), seg = threadIdx.x >> 4;
    const int p = bid * GG_RANK_PER + j;
    int r = 0;
    if (p < P) {
        // segment seg counts the scores q = seg (mod 16): the four segments
        // of a wave read four banks
        const float x = s_sc[p];
#pragma unroll 8
        for (int q = seg; q < P; q += 16) {
            const float y = s_sc[q];
            r += (y < x || (y == x && q < p)) ? 1 : 0;
        }
    }
    s_part[seg][j] = r;
    __syncthreads();
    if (seg == 0 && p < P) {
        int t = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) t += s_part[i][j];
        rank[p] = t;
        if (sel) {
            const int G = sel->K - sel->K / 4;
            if (warm ? p < sel->K : t < G) sel_write(*sel, warm ? p : t, p);
        }
    }
}

// The exploration part of a selection after the warm-up, by one workgroup:
// slots G..K-1 take the first draws of stream (seed, step, GG_EXPLORE_LANE)
// that fall on a proposal of rank >= G not drawn before (k_select's rule).
// Draws are made 32 at a time: draw i of a batch from stream state s0 is
// mix(s0 + (i+1) * 0x9E3779B97F4A7C15) whenever no draw of the batch needs
// Lemire's rejection step (x * P mod 2^64 >= P for every one; otherwise the
// batch is redrawn one at a time).  Draw i is taken iff its rank is >= G and
// no earlier draw fell on the same proposal (an earlier draw of that
// proposal was taken or had rank < G); picks fill the slots in draw order.
#define GG_EXPLORE_BATCH 32
#define GG_EXPLORE_MAX_K 64  // the resident loop's K bound
__device__ void explore_body(const float *score, int P, const SelArgs &s, float *s_sc) {
    __shared__ int s_x[GG_EXPLORE_BATCH], s_r[GG_EXPLORE_BATCH], s_pick[GG_EXPLORE_MAX_K];
    __shared__ int s_n;
    const int tid = threadIdx.x, lane = tid & 63;
    const int K = s.K, G = K - K / 4;
    // the header and the scores loaded together (the scores are not needed
    // during the warm-up)
    const int step = s.ctl[0], warm = s.ctl[1];
    for (int q = tid; q < P; q += 256) {
        const float x = score[q];
        s_sc[q] = x == x ? x : INFINITY;
    }
    if (warm != 0 || G >= K) return;
    // wave 0 draws; every thread keeps the stream state
    uint64_t st = sm64_stream(s.seed, (uint64_t)step, GG_EXPLORE_LANE);
    int n = G;
    for (;;) {
        if (tid < 64) {
            const bool in = lane < GG_EXPLORE_BATCH;
            const uint64_t x = sm64_mix(st + (uint64_t)(lane + 1) * 0x9E3779B97F4A7C15ull);
            const bool dirty = in && x * (uint64_t)P < (uint64_t)P;
            if (__ballot(dirty) == 0ull) {
                if (in) s_x[lane] = (int)mul_hi64(x, (uint64_t)P);
            } else if (lane == 0) {
                uint64_t t = st;
                for (int i = 0; i < GG_EXPLORE_BATCH; i++) s_x[i] = (int)sm64_below(t, (uint64_t)P);
            }
        }
        __syncthreads();
        // rank of draw tid / 8 over the scores q = tid % 8 (mod 8)
        {
            const int i = tid >> 3, part = tid & 7;
            const int xi = s_x[i];
            const float xv = s_sc[xi];
            int r = 0;
#pragma unroll 8
            for (int q = part; q < P; q += 8) {
                const float y = s_sc[q];
                r += (y < xv || (y == xv && q < xi)) ? 1 : 0;
            }
            r += __shfl_xor(r, 1, 64);
            r += __shfl_xor(r, 2, 64);
            r += __shfl_xor(r, 4, 64);
            if (part == 0) s_r[i] = r;
        }
        __syncthreads();
        if (tid < 64) {
            const bool in = lane < GG_EXPLORE_BATCH;
            const int x = in ? s_x[lane] : -1;
            bool first = true;
#pragma unroll
            for (int j = 0; j < GG_EXPLORE_BATCH; j++) {
                const int xj = __shfl(x, j, 64);
                first = first && !(j < lane && xj == x);
            }
            for (int j = G; j < n; j++) first = first && s_pick[j] != x;  // earlier batches
            const bool take = in && first && s_r[in ? lane : 0] >= G;
            const unsigned long long m = __ballot(take);
            const int slot = n + __popcll(m & ((1ull << lane) - 1ull));
            if (take && slot < K) {
                s_pick[slot] = x;
                sel_write(s, slot, x);
            }
            if (lane == 0) s_n = min(K, n + __popcll(m));
        }
        __syncthreads();
        n = s_n;
        if (n >= K) break;
        // a batch held too few picks: the stream continues after its draws
        // (every thread advances its copy of the state the same way)
        for (int i = 0; i < GG_EXPLORE_BATCH; i++) (void)sm64_below(st, (uint64_t)P);
    }
}

__global__ void __launch_bounds__(256) k_rank(const float *score, int P, int *rank) {
    __shared__ float s_sc[GG_SELECT_MAX_P];
    rank_body(score, P, rank, blockIdx.x, s_sc);
}

// The ranks and the selection (workgroup 0 the exploration draws, workgroups
// 1..nrank the ranks, each writing its greedy picks), beside the reduction
// and Adam update of a gated training step whose gradient the last feature
// launch carried.  Writes what k_rank followed by k_select writes.
__global__ void __launch_bounds__(256) k_rank_adam(const float *score, int P, int *rank,
                                                   int nrank, AdamArgs a, SelArgs sel) {
    __shared__ float s_sc[GG_SELECT_MAX_P];
    const int b = blockIdx.x;
    if (b == 0) {
        explore_body(score, P, sel, s_sc);
        return;
    }
    if (b <= nrank) {
        rank_body(score, P, rank, b - 1, s_sc, &sel);
        return;
    }
    reduce_adam_body(a, (b - 1 - nrank) * 256 + threadIdx.x);
}

__global__ void __launch_bounds__(1024) k_select(const int *rank, int P, int K, uint64_t seed,
                                                 const int *ctl, const gg_swap *psw,
                                                 const int *pmoves, gg_swap *out_sw,
                                                 int *out_moves, int *out_sel) {
    __shared__ unsigned char s_used[GG_SELECT_MAX_P];
    __shared__ int s_sel[GG_SELECT_MAX_P];
    const int tid = threadIdx.x;
    const bool warm = ctl[1] != 0;
    if (warm) {
        for (int k = tid; k < K; k += 1024) s_sel[k] = k;
    } else {
        const int G = K - K / 4;
        for (int p = tid; p < P; p += 1024) {
            const int r = rank[p];
            const bool g = r < G;
            s_used[p] = g ? 1 : 0;
            if (g) s_sel[r] = p;
        }
        __syncthreads();
        if (tid == 0) {
            uint64_t st = sm64_stream(seed, (uint64_t)ctl[0], GG_EXPLORE_LANE);
            for (int n = G; n < K;) {
                const int x = (int)sm64_below(st, P);
                if (!s_used[x]) { s_used[x] = 1; s_sel[n++] = x; }
            }
        }
    }
    __syncthreads();
    for (int k = tid; k < K; k += 1024) {
        const int p = s_sel[k];
        out_sw[k] = psw[p];
        for (int j = 0; j < 5; j++) out_moves[k * 5 + j] = pmoves[(size_t)p * 5 + j];
        out_sel[k] = p;
    }
}

// End of a captured step: the K results, swaps and move rows and the control
// words go straight to pinned host memory (no copy nodes), and the batch's
// queue counters are reset for the next step.
__global__ void __launch_bounds__(256) k_pack_out(int K, const ResultRec *res, ResultRec *h_res,
                                                  const gg_swap *sw, gg_swap *h_sw,
                                                  const int *msel, int *h_msel, int *dctl,
                                                  int *h_ctl, int *qctl, int nqctl) {
    const int t = threadIdx.x;
    for (int k = t; k < K; k += 256) {
        h_res[k] = res[k];
        h_sw[k] = sw[k];
    }
    for (int i = t; i < 5 * K; i += 256) h_msel[i] = msel[i];
    if (t < 4) h_ctl[t] = dctl[t];
    if (t == 3) dctl[3] = 0;  // k_propose only ORs its error word: clear it for the next replay
    for (int i = t; i < nqctl; i += 256) qctl[i] = 0;
}

// Rows idx[r] of src into ring slots head + r of dst; with ysrc, the targets
// ysrc[(head + r) % cap] (pinned host memory) into ydst at the same slots.
__global__ void k_gather_rows(const float *src, const int *idx, int n, float *dst, int dst_cap,
                              int head, const float *ysrc, float *ydst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * GG_FEATS) return;
    const int r = i / GG_FEATS, c = i - r * GG_FEATS;
    const int slot = (head + r) % dst_cap;
    dst[(size_t)slot * GG_FEATS + c] = src[(size_t)idx[r] * GG_FEATS + c];
    if (ysrc && c == 0) ydst[slot] = ysrc[slot];
}

struct AcceptArgs {
    int N, K, rule, sched, total_steps, S, warmup_pairs, every, B, cap, nqctl;
    double T0, T1, mu, sigma;
    uint64_t seed;
    const ResultRec *res;    // the launch's K exact results
    const gg_swap *swaps;    // their swaps
    const int *msel;         // their moves (a, b, c, d, kind)
    const float *x;          // the P proposals' features
    const int *psel;         // the K chosen proposals (indices into x)
    LoopCtl *ctl;
    int *dctl;               // k_propose's copy of the header: [3] its error word
    CommitSel *sel;
    TrainGate *gate;
    float *bx, *by;          // replay ring: features, targets
    int *idx;                // minibatch rows
    StepRec *rec;            // GG_LOOP_REC_CAP entries in device memory
    int *qctl;               // the incremental launch's queue counters
    unsigned *done;          // finalize workgroups done (zero between launches)
};

__device__ __forceinline__ int rec_status(const ResultRec &r, long long pairs) {
    if (r.flags & GG_FLAG_LIMIT) return GG_DIAMETER_LIMIT;
    return r.reached != pairs ? GG_UNCONNECTED : GG_OK;
}

// Acceptance of one resident-loop step, by one wave: the K exact results become
// training pairs (feature rows and normalised targets into the replay ring);
// the best candidate by (connected, diameter, total, index) goes through the
// annealing or hill-climbing rule (cppcode/annealing.cpp:406-412,
// HillClimb.cpp:239; the uniform from the counter stream (seed, step,
// GG_ACCEPT_LANE)); an accepted candidate is committed in place by the next
// kernels, or, when its levels were not kept, halts the chunk for a commit by
// the host.  Then the Adam step's gate and minibatch (rows drawn from
// (seed ^ GG_BATCH_SALT, t, row)), the next step's header and the step's
// record.  A halted chunk's later steps are void: only the queue counters of
// their launches are reset.
__device__ __forceinline__ void loop_accept_body(const AcceptArgs &a) {
    LoopCtl *__restrict__ c = a.ctl;
    const int lane = threadIdx.x, K = a.K;  // K <= 64: lane k holds candidate k
    // every load first (stores below could alias them)
    const int halted = c->halted, err = a.dctl[3];
    const long long cur_total = c->cur_total, step = c->step;
    const int cur_diam = c->cur_diam, head = c->head, count0 = c->count, adam_t = c->adam_t,
              nrec = c->nrec, trained = c->trained;
    ResultRec r{};
    int mvk[5] = {0, 0, 0, 0, 0}, swk[12];  // lane k: candidate k's move and swap words
#pragma unroll
    for (int j = 0; j < 12; j++) swk[j] = 0;
    if (lane < K) {
        r = a.res[lane];
#pragma unroll
        for (int j = 0; j < 5; j++) mvk[j] = a.msel[lane * 5 + j];
        const int *sp = reinterpret_cast<const int *>(a.swaps + lane);
#pragma unroll
        for (int j = 0; j < 12; j++) swk[j] = sp[j];
    }
    // feature rows of the chosen proposals: lane = (row, half), 16 floats
    // (row index clamped: every lane loads, lanes past 2K store nothing)
    const int hr0 = lane < 2 * K ? lane : 0, hr1 = lane + 64 < 2 * K ? lane + 64 : 0;
    const float4 *src0 = reinterpret_cast<const float4 *>(
        a.x + (size_t)a.psel[hr0 >> 1] * GG_FEATS + (hr0 & 1) * (GG_FEATS / 2));
    const float4 *src1 = reinterpret_cast<const float4 *>(
        a.x + (size_t)a.psel[hr1 >> 1] * GG_FEATS + (hr1 & 1) * (GG_FEATS / 2));
    const float4 f00 = src0[0], f01 = src0[1], f02 = src0[2], f03 = src0[3];
    const float4 f10 = src1[0], f11 = src1[1], f12 = src1[2], f13 = src1[3];
    if (halted || err) {
        for (int i = lane; i < a.nqctl; i += 64) a.qctl[i] = 0;
        if (lane == 0) {
            if (!halted) { c->halted = 1; c->reason = GG_LOOP_NO_SWAP; }
            a.sel->k = -1;
            a.gate->on = 0;
            a.dctl[3] = 0;
        }
        return;
    }
    // training rows: the feature rows go to the ring now (their slots follow
    // from the head alone), so they are not held through the decision
    if (lane < 2 * K) {
        float4 *dst = reinterpret_cast<float4 *>(
            a.bx + (size_t)((head + (lane >> 1)) % a.cap) * GG_FEATS + (lane & 1) * (GG_FEATS / 2));
        dst[0] = f00; dst[1] = f01; dst[2] = f02; dst[3] = f03;
    }
    if (lane + 64 < 2 * K) {
        const int hr = lane + 64;
        float4 *dst = reinterpret_cast<float4 *>(
            a.bx + (size_t)((head + (hr >> 1)) % a.cap) * GG_FEATS + (hr & 1) * (GG_FEATS / 2));
        dst[0] = f10; dst[1] = f11; dst[2] = f12; dst[3] = f13;
    }
    const long long pairs = (long long)a.N * a.N;
    // the best candidate by (connected, diameter, total, index)
    const bool rok = lane < K && rec_status(r, pairs) == GG_OK;
    int key_bad = lane < K ? (rok ? 0 : 1) : 2, key_d = r.diameter, key_k = lane;
    long long key_t = r.total;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int ob = __shfl_xor(key_bad, o, 64), od = __shfl_xor(key_d, o, 64),
                  okk = __shfl_xor(key_k, o, 64);
        const long long ot = __shfl_xor(key_t, o, 64);
        const bool better = ob != key_bad ? ob < key_bad
                            : od != key_d ? od < key_d
                            : ot != key_t ? ot < key_t
                                          : okk < key_k;
        if (better) { key_bad = ob; key_d = od; key_t = ot; key_k = okk; }
    }
    const int bi = key_k;
    int mv = 0, swv = 0;  // lane j: word j of the best candidate's move / swap
#pragma unroll
    for (int j = 0; j < 5; j++) {
        const int x = __shfl(mvk[j], bi, 64);
        if (lane == j) mv = x;
    }
#pragma unroll
    for (int j = 0; j < 12; j++) {
        const int x = __shfl(swk[j], bi, 64);
        if (lane == j) swv = x;
    }
    ResultRec y;
    y.total = __shfl(r.total, bi, 64);
    y.reached = __shfl(r.reached, bi, 64);
    y.diameter = __shfl(r.diameter, bi, 64);
    y.levels = __shfl(r.levels, bi, 64);
    y.wvc = __shfl(r.wvc, bi, 64);
    y.flags = __shfl(r.flags, bi, 64);
    const int ys = rec_status(y, pairs);
    const long long step1 = step + 1;
    const int count = count0 + K < a.cap ? count0 + K : a.cap;
    const bool train = step % a.every == 0 && count >= a.B;
    const int t = adam_t + (train ? 1 : 0);
    // acceptance (every lane computes the same)
    const double NN1 = (double)a.N * (a.N - 1);
    const double fxv = (double)cur_total / NN1;
    const bool ok = ys == GG_OK;
    const double fy = ok ? (double)y.total / NN1 : 1e8;
    const int dy = ok ? y.diameter : 100000000;
    bool accept;
    if (a.rule == GG_RULE_HILL) {
        accept = ok && (dy < cur_diam || (dy == cur_diam && fy < fxv));
    } else {
        accept = fy < fxv || dy < cur_diam;
        if (!accept) {
            const double T = a.sched == GG_SCHED_STEP
                                 ? a.T0 * pow(0.995, (double)((step1 - 1) / 10000))
                                 : a.T0 * pow(a.T1 / a.T0, (double)step1 / a.total_steps);
            uint64_t st = sm64_stream(a.seed, (uint64_t)step1, GG_ACCEPT_LANE);
            const double u = (double)(sm64_next(st) >> 11) * (1.0 / 9007199254740992.0);
            accept = u < exp((fxv - fy) * NN1 / (T * NN1));
        }
        if (dy > cur_diam || !ok) accept = false;
    }
    const bool inplace = accept && ok && y.diameter - 1 <= a.S &&
                         !(y.flags & (GG_FLAG_QUEUED | GG_FLAG_LIMIT));
    // ---- stores: normalised targets of the training pairs
    if (lane < K) {
        const double raw = rok ? (double)(r.total - cur_total) +
                                     (double)(r.diameter - cur_diam) * (double)a.N * a.N
                               : 1e30;
        double yv = (raw - a.mu) / a.sigma;
        yv = yv < 8.0 ? yv : 8.0;
        yv = -8.0 < yv ? yv : -8.0;
        a.by[(head + lane) % a.cap] = (float)yv;
    }
    // the step's record
    StepRec *rec = a.rec + nrec % GG_LOOP_REC_CAP;
    if (lane < 5) rec->mv[lane] = mv;
    if (lane < 12) reinterpret_cast<int *>(&rec->sw)[lane] = swv;
    if (lane == 0) {
        rec->step = step1;
        rec->total = y.total;
        rec->diameter = y.diameter;
        rec->levels = y.levels;
        rec->status = ys;
        rec->wvc = y.wvc;
        rec->bi = bi;
        rec->accept = accept;
        rec->inplace = inplace;
        a.sel->k = inplace ? bi : -1;
        const int diam = accept ? y.diameter : cur_diam;
        if (accept) {
            c->cur_total = y.total;
            c->cur_diam = y.diameter;
            c->cur_levels = y.levels;
            c->cur_status = ys;
            c->cur_wvc = y.wvc;
            if (!inplace) { c->halted = 1; c->reason = GG_LOOP_HOST_COMMIT; }
        }
        c->head = (head + K) % a.cap;
        c->count = count;
        if (train) {
            c->adam_t = t;
            c->trained = trained + 1;
            a.gate->b1t = powf(0.9f, (float)t);
            a.gate->b2t = powf(0.999f, (float)t);
        }
        a.gate->on = train ? 1 : 0;
        c->step = step1;
        c->hdr[0] = (int)step1;
        c->hdr[1] = count < a.warmup_pairs ? 1 : 0;
        c->hdr[2] = diam > 0 ? diam : 1;
        c->nrec = nrec + 1;
    }
    // the Adam step's minibatch rows
    if (train)
        for (int i = lane; i < a.B; i += 64) {
            uint64_t st = sm64_stream(a.seed ^ GG_BATCH_SALT, (uint64_t)t, (uint64_t)i);
            a.idx[i] = (int)sm64_below(st, (uint64_t)count);
        }
    for (int i = lane; i < a.nqctl; i += 64) a.qctl[i] = 0;
}

// The resident loop's finalize: workgroup k merges candidate k's result
// (sparse_finalize_body); the last workgroup to finish (counter) runs the
// step's acceptance with its first wave once every result is visible.
__global__ void __launch_bounds__(256) k_finalize_accept(SparseArgs sa, int nvalid_last,
                                                         AcceptArgs a) {
    __shared__ int s_last;
    // the results cross workgroups as stores through to the device-coherent
    // level (no L2 writeback fence per workgroup)
    sparse_finalize_body<256, true>(sa, nvalid_last, blockIdx.x);
    __syncthreads();
    if (threadIdx.x == 0) s_last = atomicAdd(a.done, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!s_last) return;
    // the other workgroups' results were stored through to the coherent
    // level: drop stale lines of this XCD's caches before reading them
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (threadIdx.x == 0) *a.done = 0u;
    if (threadIdx.x < 64) loop_accept_body(a);
}

int gather_rows_dev(const float *src, const int *idx, int n, float *dst, int dst_cap, int head,
                    hipStream_t st, const float *ysrc, float *ydst) {
    hipLaunchKernelGGL(k_gather_rows, dim3((n * GG_FEATS + 255) / 256), dim3(256), 0, st, src, idx,
                       n, dst, dst_cap, head, ysrc, ydst);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

hipStream_t batch_stream(gg_batch *b) { return b->stream; }

struct LearnState {
    gg_batch *batch = nullptr;
    gg_mlp *mlp = nullptr;
    hipStream_t st = nullptr;                    // the batch evaluator's stream (all work)
    int P = 0, K = 0;
    int *d_moves = nullptr, *h_moves = nullptr;  // [P][5]
    float *d_x = nullptr;                        // [P][F]
    float *d_score = nullptr, *h_score = nullptr;
    int *d_sel = nullptr, *h_sel = nullptr;      // chosen proposal indices [K]
    int cap = 0, count = 0, head = 0;            // replay buffer
    float *d_bx = nullptr, *d_by = nullptr;
    int *d_idx = nullptr, *h_idx = nullptr;      // minibatch indices
    double mu = 0, sigma = 1;                    // target normalisation (frozen after warmup)
    bool norm_frozen = false;
    int64_t train_steps = 0;
    double last_loss = 0;
    bool loss_pending = false;                   // a training step ran since the loss was read
    double spearman = 0, miss = 0;
    std::vector<int> order;
    std::unique_ptr<Rng> rng;                    // minibatch sampling / exploration
    std::vector<int> last_sel;                   // original indices of the chosen K
    std::vector<double> raw;                     // raw deltas of the replay buffer (host mirror)
    float *h_ynorm = nullptr;                    // normalised targets (pinned host mirror)
    // device-drawn proposals: the whole step (draw, features, scores,
    // selection, incremental evaluation) replays as one captured graph
    bool dev = false;                            // this search draws on the GPU
    bool sel_on_dev = false;                     // d_sel holds the last step's choice
    gg_swap *d_psw = nullptr;                    // [P] proposals
    int *d_rank = nullptr;                       // [P] score ranks
    int *d_msel = nullptr, *h_msel = nullptr;    // [K][5] chosen move rows
    int *d_ctl = nullptr, *h_ctl = nullptr;      // [4] step, warm, diameter, error; h_ctl[4..8) read back
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int g_push = -1;
    // resident loop (after the warm-up): every decision of a step on the GPU
    LoopCtl *d_loop = nullptr, *h_loop = nullptr;  // control block, pinned staging copy
    CommitSel *d_csel = nullptr;                   // the step's in-place commit
    TrainGate *d_gate = nullptr;                   // the step's Adam step
    int *d_bidx = nullptr;                         // minibatch rows drawn on the GPU
    unsigned *d_done = nullptr;                    // finalize workgroups done
    StepRec *h_rec = nullptr;                      // one record per step (pinned host copy)
    StepRec *d_rec = nullptr;                      // the device writes them here
    // the resident loop's captured launches, [unrolled][first][draws]: one or
    // GG_LOOP_UNROLL steps; the first step draws its own proposals (a chunk's
    // first launch) or the step before drew them; the last step's commit
    // draws the next step's (every launch but a chunk's last)
    hipGraph_t lgraph[2][2][2] = {};
    hipGraphExec_t lexec[2][2][2] = {};
    int l_push = -1;
    hipEvent_t lev[GG_LOOP_INFLIGHT] = {};          // steps in flight: one event each
};

static int alloc_learn(LearnState *L) {
    const int P = L->P, K = L->K;
    if (hipMalloc((void **)&L->d_moves, 4 * (size_t)P * 5) != hipSuccess) return GG_OOM;
    if (hipHostMalloc((void **)&L->h_moves, 4 * (size_t)P * 5, hipHostMallocDefault) != hipSuccess)
        return GG_OOM;
    if (hipMalloc((void **)&L->d_x, 4 * (size_t)P * GG_FEATS) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&L->d_score, 4 * (size_t)P) != hipSuccess) return GG_OOM;
    if (hipHostMalloc((void **)&L->h_score, 4 * (size_t)P, hipHostMallocDefault) != hipSuccess)
        return GG_OOM;
    if (hipMalloc((void **)&L->d_sel, 4 * (size_t)P) != hipSuccess) return GG_OOM;
    if (hipHostMalloc((void **)&L->h_sel, 4 * (size_t)P, hipHostMallocDefault) != hipSuccess)
        return GG_OOM;
    if (hipMalloc((void **)&L->d_bx, 4 * (size_t)L->cap * GG_FEATS) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&L->d_by, 4 * (size_t)L->cap) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&L->d_idx, 4 * GG_MAX_BATCH) != hipSuccess) return GG_OOM;
    if (hipHostMalloc((void **)&L->h_idx, 4 * GG_MAX_BATCH, hipHostMallocDefault) != hipSuccess)
        return GG_OOM;
    if (hipHostMalloc((void **)&L->h_ynorm, 4 * (size_t)L->cap, hipHostMallocDefault) != hipSuccess)
        return GG_OOM;
    if (hipMalloc((void **)&L->d_psw, sizeof(gg_swap) * (size_t)P) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&L->d_rank, 4 * (size_t)P) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&L->d_msel, 4 * (size_t)K * 5) != hipSuccess) return GG_OOM;
    if (hipHostMalloc((void **)&L->h_msel, 4 * (size_t)K * 5, hipHostMallocDefault) != hipSuccess)
        return GG_OOM;
    if (hipMalloc((void **)&L->d_ctl, 4 * 4) != hipSuccess) return GG_OOM;
    // the error word d_ctl[3] is only ever ORed by k_propose and cleared by
    // k_pack_out, so it must start at zero (a fresh allocation is not)
    if (hipMemset(L->d_ctl, 0, 4 * 4) != hipSuccess) return GG_HIP_ERROR;
    if (hipHostMalloc((void **)&L->h_ctl, 4 * 8, hipHostMallocDefault) != hipSuccess)
        return GG_OOM;
    if (hipMalloc((void **)&L->d_loop, sizeof(LoopCtl)) != hipSuccess) return GG_OOM;
    if (hipHostMalloc((void **)&L->h_loop, sizeof(LoopCtl), hipHostMallocDefault) != hipSuccess)
        return GG_OOM;
    if (hipMalloc((void **)&L->d_csel, sizeof(CommitSel)) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&L->d_gate, sizeof(TrainGate)) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&L->d_bidx, 4 * GG_MAX_BATCH) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&L->d_done, 4) != hipSuccess) return GG_OOM;
    if (hipMemset(L->d_done, 0, 4) != hipSuccess) return GG_HIP_ERROR;
    if (hipHostMalloc((void **)&L->h_rec, sizeof(StepRec) * GG_LOOP_REC_CAP,
                      hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void **)&L->d_rec, sizeof(StepRec) * GG_LOOP_REC_CAP) != hipSuccess)
        return GG_OOM;
    for (hipEvent_t &e : L->lev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return GG_HIP_ERROR;
    return GG_OK;
}

int learn_create(gg_search *s, int device) {
    gg_batch *b = s->be->batch();
    if (!b) return GG_UNSUPPORTED;
    LearnState *L = new LearnState();
    s->learn = L;
    L->batch = b;
    L->st = b->stream;
    L->P = s->cfg.P;
    L->K = s->cfg.K;
    L->cap = 1 << 17;  // 131072 pairs
    L->rng.reset(new Rng(s->cfg.seed ^ 0x9E3779B97F4A7C15ull));
    L->dev = L->P <= GG_SELECT_MAX_P;
    int st = gg_mlp_create(device, s->cfg.seed + 1, &L->mlp);
    if (st != GG_OK) return st;
    return alloc_learn(L);
}

void learn_destroy(gg_search *s) {
    LearnState *L = s->learn;
    if (!L) return;
    if (L->st) (void)hipStreamSynchronize(L->st);
    if (L->gexec) (void)hipGraphExecDestroy(L->gexec);
    if (L->graph) (void)hipGraphDestroy(L->graph);
    for (int i = 0; i < 8; i++) {
        hipGraphExec_t &x = L->lexec[i >> 2][(i >> 1) & 1][i & 1];
        hipGraph_t &g = L->lgraph[i >> 2][(i >> 1) & 1][i & 1];
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
    }
    for (hipEvent_t e : L->lev)
        if (e) (void)hipEventDestroy(e);
    gg_mlp_destroy(L->mlp);
    void *dev[] = {L->d_moves, L->d_x, L->d_score, L->d_sel, L->d_bx,   L->d_by,   L->d_idx,
                   L->d_psw,   L->d_msel, L->d_ctl, L->d_rank, L->d_loop, L->d_csel, L->d_gate,
                   L->d_bidx,  L->d_done, L->d_rec};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    void *host[] = {L->h_moves, L->h_score, L->h_sel,  L->h_idx,
                    L->h_ynorm, L->h_msel,  L->h_ctl,  L->h_loop, L->h_rec};
    for (void *p : host)
        if (p) (void)hipHostFree(p);
    delete L;
    s->learn = nullptr;
}

// Scores the P proposals in s->sw / s->mv (features + MLP forward on the GPU).
static int score_proposals(gg_search *s, int P) {
    LearnState *L = s->learn;
    for (int p = 0; p < P; p++) {
        const MoveRec &m = s->mv[p];
        int *r = L->h_moves + (size_t)p * 5;
        r[0] = m.a; r[1] = m.b; r[2] = m.c; r[3] = m.d; r[4] = m.kind;
    }
    GG_HIP(hipMemcpyAsync(L->d_moves, L->h_moves, 4 * (size_t)P * 5, hipMemcpyHostToDevice, L->st));
    int st = features_score_dev(L->batch, L->mlp, L->d_moves, P, L->d_x, L->d_score, L->st);
    if (st != GG_OK) return st;
    GG_HIP(hipMemcpyAsync(L->h_score, L->d_score, 4 * (size_t)P, hipMemcpyDeviceToHost, L->st));
    GG_HIP(hipStreamSynchronize(L->st));
    return GG_OK;
}

int learn_select(gg_search *s, int P, int K) {
    LearnState *L = s->learn;
    int st = score_proposals(s, P);
    if (st != GG_OK) return st;
    std::vector<int> &ord = L->order;
    ord.resize(P);
    std::iota(ord.begin(), ord.end(), 0);
    const bool warm = L->count < s->cfg.warmup_pairs;
    std::vector<int> sel;
    sel.reserve(K);
    if (warm) {
        for (int k = 0; k < K; k++) sel.push_back(k);  // proposals are already uniform draws
    } else {
        const int explore = K / 4;
        const int greedy = K - explore;
        std::partial_sort(ord.begin(), ord.begin() + greedy, ord.end(), [&](int x, int y) {
            return L->h_score[x] < L->h_score[y] || (L->h_score[x] == L->h_score[y] && x < y);
        });
        std::vector<char> used(P, 0);
        for (int k = 0; k < greedy; k++) { sel.push_back(ord[k]); used[ord[k]] = 1; }
        while ((int)sel.size() < K) {
            int x = (int)L->rng->below(P);
            if (!used[x]) { used[x] = 1; sel.push_back(x); }
        }
    }
    // move the chosen proposals to the front, keeping their original indices
    std::vector<gg_swap> sw(K);
    std::vector<MoveRec> mv(K);
    for (int k = 0; k < K; k++) { sw[k] = s->sw[sel[k]]; mv[k] = s->mv[sel[k]]; }
    for (int k = 0; k < K; k++) { s->sw[k] = sw[k]; s->mv[k] = mv[k]; }
    L->last_sel = sel;
    return GG_OK;
}

bool learn_on_device(const gg_search *s) { return s->learn && s->learn->dev; }

// One learned step's candidates, drawn, scored, selected and evaluated on the
// GPU (one captured graph, one wait): fills s->sw, s->mv and s->res for the K
// chosen proposals.
int learn_eval_device(gg_search *s) {
    LearnState *L = s->learn;
    gg_batch *b = L->batch;
    const int P = L->P, K = L->K;
    L->h_ctl[0] = (int)s->step;
    L->h_ctl[1] = L->count < s->cfg.warmup_pairs ? 1 : 0;
    L->h_ctl[2] = s->cur.diameter > 0 ? s->cur.diameter : 1;
    L->h_ctl[3] = 0;
    const int push = batch_push_state(b);
    if (!L->gexec || L->g_push != push) {
        if (L->gexec) { (void)hipGraphExecDestroy(L->gexec); L->gexec = nullptr; }
        if (L->graph) { (void)hipGraphDestroy(L->graph); L->graph = nullptr; }
        GG_HIP(hipStreamBeginCapture(L->st, hipStreamCaptureModeThreadLocal));
        int st = GG_OK;
        auto fail = [&](int code) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(L->st, &g);
            if (g) (void)hipGraphDestroy(g);
            return code;
        };
        // the control words are read from pinned host memory by k_propose
        hipLaunchKernelGGL(k_propose, dim3((P + 63) / 64), dim3(64), 0, L->st,
                           ProposeArgs{b_adjR(b), s->N, s->d, s->cfg.seed, L->h_ctl, P,
                                       L->d_moves, L->d_psw, L->d_ctl});
        if (hipGetLastError() != hipSuccess) return fail(GG_HIP_ERROR);
        if ((st = features_score_dev(b, L->mlp, L->d_moves, P, L->d_x, L->d_score, L->st,
                                     L->d_ctl + 2)) != GG_OK)
            return fail(st);
        hipLaunchKernelGGL(k_rank, dim3((P + GG_RANK_PER - 1) / GG_RANK_PER), dim3(256), 0, L->st,
                           L->d_score, P, L->d_rank);
        hipLaunchKernelGGL(k_select, dim3(1), dim3(1024), 0, L->st, L->d_rank, P, K,
                           s->cfg.seed, L->d_ctl, L->d_psw, L->d_moves, b_swaps(b), L->d_msel,
                           L->d_sel);
        if (hipGetLastError() != hipSuccess) return fail(GG_HIP_ERROR);
        if ((st = batch_enqueue_dev(b, K)) != GG_OK) return fail(st);
        BatchOut bo = batch_out(b);
        hipLaunchKernelGGL(k_pack_out, dim3(1), dim3(256), 0, L->st, K, bo.d_res, bo.h_res,
                           b_swaps(b), bo.h_swaps, L->d_msel, L->h_msel, L->d_ctl, L->h_ctl + 4,
                           bo.qctl, bo.nqctl);
        if (hipGetLastError() != hipSuccess) return fail(GG_HIP_ERROR);
        hipGraph_t g = nullptr;
        GG_HIP(hipStreamEndCapture(L->st, &g));
        L->graph = g;
        GG_HIP(hipGraphInstantiate(&L->gexec, g, nullptr, nullptr, 0));
        L->g_push = push;
    }
    GG_HIP(hipGraphLaunch(L->gexec, L->st));
    GG_HIP(hipStreamSynchronize(L->st));
    if (L->h_ctl[4 + 3]) return GG_BAD_ARG;  // some proposal found no simple swap
    const gg_swap *hs = b_hswaps(b);
    for (int k = 0; k < K; k++) {
        s->sw[k] = hs[k];
        const int *m = L->h_msel + (size_t)k * 5;
        s->mv[k] = MoveRec{m[0], m[1], m[2], m[3], m[4]};
    }
    L->sel_on_dev = true;
    return batch_finish_dev(b, K, s->res.data());
}

// The resident loop runs once the targets' normalisation is frozen, on the
// plain annealing / hill-climbing rules, while the base allows the
// incremental path with one launch whose logs are kept (in-place commits).
bool learn_loop_ready(const gg_search *s) {
    const LearnState *L = s->learn;
    if (!L || !L->dev || !L->norm_frozen) return false;
    if (s->cfg.rule != GG_RULE_ANNEAL && s->cfg.rule != GG_RULE_HILL) return false;
    if (s->cfg.schedule != GG_SCHED_GEOMETRIC && s->cfg.schedule != GG_SCHED_STEP) return false;
    gg_batch *b = L->batch;
    // the acceptance wave holds one candidate per lane
    return L->K <= 64 && batch_push_state(b) == 1 && batch_sparse_ok(b) &&
           batch_single_launch(b, L->K);
}

// The resident loop's draw: the step number, warm-up flag and diameter are
// read from the control block and copied to d_ctl, which the later kernels of
// the step read.
static ProposeArgs loop_propose_args(gg_search *s) {
    LearnState *L = s->learn;
    return ProposeArgs{b_adjR(L->batch), s->N, s->d, s->cfg.seed, (const int *)L->d_loop->hdr,
                       L->P, L->d_moves, L->d_psw, L->d_ctl};
}

// One resident-loop step's launches on the learned search's stream.  first:
// the step draws its own proposals (else the last step's commit launch drew
// them); next: the commit launch also draws the next step's.
static int loop_enqueue(gg_search *s, bool first, bool next) {
    LearnState *L = s->learn;
    gg_batch *b = L->batch;
    const int P = L->P, K = L->K;
    int nq = 0;
    int *qctl = batch_qctl(b, &nq);
    AcceptArgs a{};
    a.N = s->N; a.K = K; a.rule = s->cfg.rule; a.sched = s->cfg.schedule;
    a.total_steps = s->cfg.total_steps; a.S = batch_snap_cap(b);
    a.warmup_pairs = s->cfg.warmup_pairs; a.every = std::max(1, s->cfg.train_every);
    a.B = std::min(std::max(s->cfg.train_batch, 1), GG_MAX_BATCH); a.cap = L->cap;
    a.nqctl = nq;
    a.T0 = s->cfg.T0; a.T1 = s->cfg.T1; a.mu = L->mu; a.sigma = L->sigma; a.seed = s->cfg.seed;
    a.res = batch_dres(b); a.swaps = b_swaps(b); a.msel = L->d_msel; a.x = L->d_x;
    a.psel = L->d_sel; a.ctl = L->d_loop; a.dctl = L->d_ctl; a.sel = L->d_csel;
    a.gate = L->d_gate; a.bx = L->d_bx; a.by = L->d_by; a.idx = L->d_bidx; a.rec = L->d_rec;
    a.qctl = qctl;
    a.done = L->d_done;
    const float lr = (float)(s->cfg.lr > 0 ? s->cfg.lr : 1e-3);
    int st = GG_OK;
    if (first) {
        hipLaunchKernelGGL(k_propose, dim3((P + 63) / 64), dim3(64), 0, L->st, loop_propose_args(s));
        GG_HIP(hipGetLastError());
    }
    // The last step's training step runs in this step's feature and rank
    // launches, on the CUs their own workgroups leave idle: its gradient
    // beside the features, its update beside the ranks.  The scores use the
    // bf16 copy of step parity, trained through two steps back, so the
    // update writing the other copy does not race them.
    const int *par = (const int *)L->d_loop->hdr;
    const GradArgs ga = mlp_grad_args(L->mlp, L->d_bx, L->d_by, L->d_bidx, a.B, L->d_gate, par);
    if ((st = features_score_dev(b, L->mlp, L->d_moves, P, L->d_x, L->d_score, L->st,
                                 L->d_ctl + 2, par, &ga)) != GG_OK)
        return st;
    const int nrank = (P + GG_RANK_PER - 1) / GG_RANK_PER;
    const SelArgs sel{K, s->cfg.seed, L->d_ctl, L->d_psw, L->d_moves, b_swaps(b), L->d_msel,
                      L->d_sel};
    hipLaunchKernelGGL(k_rank_adam, dim3(1 + nrank + GG_ADAM_GROUPS), dim3(256), 0, L->st,
                       L->d_score, P, L->d_rank, nrank,
                       mlp_adam_args(L->mlp, a.B, lr, L->d_gate, par), sel);
    GG_HIP(hipGetLastError());
    SparseArgs fin{};
    int nvalid = 0;
    if ((st = batch_enqueue_dev(b, K, &fin, &nvalid)) != GG_OK) return st;
    hipLaunchKernelGGL(k_finalize_accept, dim3(K), dim3(256), 0, L->st, fin, nvalid, a);
    GG_HIP(hipGetLastError());
    const ProposeArgs pn = loop_propose_args(s);
    return batch_enqueue_commit_dev(b, L->d_csel, next ? &pn : nullptr);
}

// Captures `steps` consecutive steps into one executable graph (first: the
// first step draws its own proposals; draws: the last step's commit draws
// the next step's).
static int loop_capture_n(gg_search *s, int steps, bool first, bool draws, hipGraph_t *gr,
                          hipGraphExec_t *ex) {
    LearnState *L = s->learn;
    if (*ex) { (void)hipGraphExecDestroy(*ex); *ex = nullptr; }
    if (*gr) { (void)hipGraphDestroy(*gr); *gr = nullptr; }
    GG_HIP(hipStreamBeginCapture(L->st, hipStreamCaptureModeThreadLocal));
    int st = GG_OK;
    for (int i = 0; i < steps && st == GG_OK; i++)
        st = loop_enqueue(s, first && i == 0, i + 1 < steps || draws);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(L->st, &g);
    if (st != GG_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    GG_HIP(e);
    *gr = g;
    GG_HIP(hipGraphInstantiate(ex, g, nullptr, nullptr, 0));
    return GG_OK;
}

// The launch variant [unrolled][first][draws], captured the first time it is
// needed after the batch evaluator's state last changed (which drops them all).
static int loop_graph(gg_search *s, int u, bool first, bool draws, hipGraphExec_t *out) {
    LearnState *L = s->learn;
    const int push = batch_push_state(L->batch);
    if (L->l_push != push) {
        for (int i = 0; i < 8; i++) {
            hipGraphExec_t &x = L->lexec[i >> 2][(i >> 1) & 1][i & 1];
            hipGraph_t &g = L->lgraph[i >> 2][(i >> 1) & 1][i & 1];
            if (x) { (void)hipGraphExecDestroy(x); x = nullptr; }
            if (g) { (void)hipGraphDestroy(g); g = nullptr; }
        }
        L->l_push = push;
    }
    const int ui = u > 1 ? 1 : 0;
    hipGraphExec_t &x = L->lexec[ui][first][draws];
    if (!x) {
        int st = loop_capture_n(s, u, first, draws, &L->lgraph[ui][first][draws], &x);
        if (st != GG_OK) return st;
    }
    *out = x;
    return GG_OK;
}

// Runs up to `steps` steps of the resident loop (one chunk: the graph
// launched back to back, one wait), then replays the records on the host:
// statistics, log, current and best graph, the batch evaluator's mirror of
// every commit the device made.  A step whose accepted candidate had no kept
// logs ends the chunk and is committed here by a new BFS.  *done = steps
// completed (>= 1 unless an error is returned).
int learn_loop_run(gg_search *s, int steps, int *done) {
    LearnState *L = s->learn;
    gg_batch *b = L->batch;
    *done = 0;
    s->be->forget();
    // GG_LOOP_GRAPH=0: the step's kernels launched one by one instead of a
    // graph replay (for tools that do not follow graph launches)
    const char *lg = getenv("GG_LOOP_GRAPH");
    const bool use_graph = !(lg && lg[0] == '0');
    const int n = std::min(steps, GG_LOOP_REC_CAP);
    // the chunk's launch variants, captured (on the idle stream) before any
    // work of the chunk is queued
    if (use_graph)
        for (int i = 0; i < n;) {
            const int u = n - i >= GG_LOOP_UNROLL ? GG_LOOP_UNROLL : 1;
            hipGraphExec_t x = nullptr;
            int st = loop_graph(s, u, i == 0, i + u < n, &x);
            if (st != GG_OK) return st;
            i += u;
        }
    LoopCtl &h = *L->h_loop;
    memset(&h, 0, sizeof(h));
    h.step = s->step;
    h.hdr[0] = (int)s->step;
    h.hdr[1] = L->count < s->cfg.warmup_pairs ? 1 : 0;
    h.hdr[2] = s->cur.diameter > 0 ? s->cur.diameter : 1;
    h.cur_total = s->cur.total_scaled;
    h.cur_diam = s->cur.diameter;
    h.cur_levels = s->cur.levels;
    h.cur_status = s->cur.status;
    h.cur_wvc = s->cur.wvc;
    h.head = L->head;
    h.count = L->count;
    h.adam_t = L->mlp->t;
    GG_HIP(hipMemcpyAsync(L->d_loop, L->h_loop, sizeof(LoopCtl), hipMemcpyHostToDevice, L->st));
    // both bf16 copies of the weights from the masters (the loop alternates
    // them), and no training step pending for the first step to carry
    int rst = mlp_refresh_dev(L->mlp, L->st);
    if (rst != GG_OK) return rst;
    GG_HIP(hipMemsetAsync(L->d_gate, 0, sizeof(TrainGate), L->st));
    // at most GG_LOOP_INFLIGHT launches queued ahead of the one running
    for (int i = 0, j = 0; i < n; j++) {
        hipEvent_t &e = L->lev[j % GG_LOOP_INFLIGHT];
        if (j >= GG_LOOP_INFLIGHT) GG_HIP(hipEventSynchronize(e));
        // every step's commit draws the next step's proposals, but the
        // chunk's last (its draws, scores and choice stay readable)
        const int u = n - i >= GG_LOOP_UNROLL ? GG_LOOP_UNROLL : 1;
        const bool first = i == 0, draws = i + u < n;
        if (use_graph) {
            hipGraphExec_t x = nullptr;
            int st = loop_graph(s, u, first, draws, &x);
            if (st != GG_OK) return st;
            GG_HIP(hipGraphLaunch(x, L->st));
        } else {
            for (int q = 0; q < u; q++) {
                int st = loop_enqueue(s, first && q == 0, q + 1 < u || draws);
                if (st != GG_OK) return st;
            }
        }
        GG_HIP(hipEventRecord(e, L->st));
        i += u;
    }
    // the last step's training step (no later step carries it), then the
    // host path's single copy holds the latest weights again
    const float lr = (float)(s->cfg.lr > 0 ? s->cfg.lr : 1e-3);
    rst = mlp_train_gated_dev(L->mlp, L->d_bx, L->d_by, L->d_bidx,
                              std::min(std::max(s->cfg.train_batch, 1), GG_MAX_BATCH), lr,
                              L->d_gate, L->st, (const int *)L->d_loop->hdr);
    if (rst != GG_OK) return rst;
    rst = mlp_refresh_dev(L->mlp, L->st);
    if (rst != GG_OK) return rst;
    GG_HIP(hipMemcpyAsync(L->h_loop, L->d_loop, sizeof(LoopCtl), hipMemcpyDeviceToHost, L->st));
    GG_HIP(hipMemcpyAsync(L->h_rec, L->d_rec, sizeof(StepRec) * (size_t)n, hipMemcpyDeviceToHost,
                          L->st));
    GG_HIP(hipStreamSynchronize(L->st));
    const int P = L->P, K = L->K;
    for (int i = 0; i < h.nrec; i++) {
        const StepRec &r = L->h_rec[i];
        s->step = r.step;
        s->stats.proposals += P;
        s->stats.exact_evals += K;
        s->stats.train_pairs += K;
        const bool ok = r.status == GG_OK;
        if (s->log_cap > 0 && (int64_t)s->log.size() < s->log_cap * 10) {
            int64_t rec[10] = {r.step, r.mv[0], r.mv[1], r.mv[2], r.mv[3], r.mv[4],
                               ok ? r.total : -1, ok ? r.diameter : -1, r.status, r.accept};
            s->log.insert(s->log.end(), rec, rec + 10);
        }
        if (!r.accept) continue;
        gg_result y{};
        y.total_reps = r.total;
        y.total_scaled = r.total;
        y.diameter = r.diameter;
        y.wvc = r.wvc;
        y.levels = r.levels;
        y.status = r.status;
        if (r.inplace) {
            batch_note_device_commit(b, &r.sw, &y);
        } else {
            int st = batch_commit_rebuild(b, &r.sw, &y);
            if (st != GG_OK) return st;
        }
        apply_swap(s->adj, s->d, r.sw);
        s->cur = y;
        s->stats.accepted++;
        track_best(s, r.sw);
    }
    L->head = h.head;
    L->count = h.count;
    L->mlp->t = h.adam_t;
    L->train_steps += h.trained;
    if (h.trained) L->loss_pending = true;
    L->sel_on_dev = false;
    *done = h.nrec;
    if (h.halted && h.reason == GG_LOOP_NO_SWAP) return GG_BAD_ARG;
    return GG_OK;
}

static double raw_delta(const gg_search *s, const gg_result &r) {
    if (r.status != GG_OK) return 1e300;
    return (double)(r.total_scaled - s->cur.total_scaled) +
           (double)(r.diameter - s->cur.diameter) * (double)s->N * s->N;
}

static int train_steps(gg_search *s, int steps) {
    LearnState *L = s->learn;
    const int B = std::min(std::max(s->cfg.train_batch, 1), GG_MAX_BATCH);
    if (L->count < B) return GG_OK;
    for (int t = 0; t < steps; t++) {
        // h_idx (pinned) is read by the gradient kernel itself; the step's
        // wait for its candidates orders the next rewrite after that read
        if (t > 0) GG_HIP(hipStreamSynchronize(L->st));
        for (int i = 0; i < B; i++) L->h_idx[i] = (int)L->rng->below(L->count);
        int st = mlp_grad_dev(L->mlp, L->d_bx, L->d_by, L->h_idx, B, L->st);
        if (st != GG_OK) return st;
        st = mlp_adam_dev(L->mlp, (float)(s->cfg.lr > 0 ? s->cfg.lr : 1e-3), L->st);
        if (st != GG_OK) return st;
        L->train_steps++;
    }
    L->loss_pending = true;  // read when the stats are (no wait per step)
    return GG_OK;
}

// Normalised targets of ring positions [pos, pos + n) into the pinned mirror.
static void normalise_targets(LearnState *L, int pos, int n) {
    for (int i = 0; i < n; i++) {
        double y = (L->raw[(pos + i) % L->cap] - L->mu) / L->sigma;
        L->h_ynorm[(pos + i) % L->cap] = (float)std::max(-8.0, std::min(8.0, y));
    }
}

// Uploads normalised targets for ring positions [pos, pos + n) (n <= cap).
static int upload_targets(LearnState *L, int pos, int n) {
    int first = std::min(n, L->cap - pos);
    normalise_targets(L, pos, n);
    GG_HIP(hipMemcpyAsync(L->d_by + pos, L->h_ynorm + pos, 4 * (size_t)first,
                          hipMemcpyHostToDevice, L->st));
    if (n > first)
        GG_HIP(hipMemcpyAsync(L->d_by, L->h_ynorm, 4 * (size_t)(n - first),
                              hipMemcpyHostToDevice, L->st));
    return GG_OK;
}

int learn_observe(gg_search *s, int K) {
    LearnState *L = s->learn;
    if (L->raw.empty()) L->raw.assign(L->cap, 0.0);
    const int pos = L->head;
    for (int k = 0; k < K; k++) {
        double dlt = raw_delta(s, s->res[k]);
        L->raw[(pos + k) % L->cap] = dlt < 1e299 ? dlt : 1e30;
        if (!L->sel_on_dev) L->h_sel[k] = L->last_sel[k];
    }
    if (!L->sel_on_dev)
        GG_HIP(hipMemcpyAsync(L->d_sel, L->h_sel, 4 * (size_t)K, hipMemcpyHostToDevice, L->st));
    L->sel_on_dev = false;
    // once the normalisation is frozen the K new targets travel with the
    // gathered feature rows (read from pinned memory by the gather kernel)
    const bool frozen = L->norm_frozen;
    if (frozen) normalise_targets(L, pos, K);
    int st = gather_rows_dev(L->d_x, L->d_sel, K, L->d_bx, L->cap, pos, L->st,
                             frozen ? L->h_ynorm : nullptr, frozen ? L->d_by : nullptr);
    if (st != GG_OK) return st;
    L->head = (L->head + K) % L->cap;
    L->count = std::min(L->count + K, L->cap);
    s->stats.train_pairs += K;
    if (!frozen) {
        const int need = std::max(64, std::min(s->cfg.warmup_pairs, 4096));
        if (L->count < need) return GG_OK;
        // Median and 1.4826 x MAD of the finite raw deltas seen so far, then
        // normalise all.  A diameter change adds (+-1) N^2 to a delta (about
        // 1e6 times the spread of the ordinary ones): a single such move in the
        // sample would set a mean / std scale that flattens every other target
        // to ~0.  With the robust scale those moves land in the +-8 clip.
        std::vector<double> fin;
        fin.reserve(L->count);
        for (int i = 0; i < L->count; i++)
            if (L->raw[i] < 1e29) fin.push_back(L->raw[i]);
        double med = 0, mad = 0;
        if (!fin.empty()) {
            auto mid = fin.begin() + fin.size() / 2;
            std::nth_element(fin.begin(), mid, fin.end());
            med = *mid;
            for (double &x : fin) x = std::fabs(x - med);
            std::nth_element(fin.begin(), mid, fin.end());
            mad = *mid;
        }
        L->mu = med;
        L->sigma = mad > 0 ? 1.4826 * mad : 1.0;
        L->norm_frozen = true;
        st = upload_targets(L, 0, L->count);
        if (st != GG_OK) return st;
    }
    const int every = std::max(1, s->cfg.train_every);
    if (s->step % every == 0) {
        st = train_steps(s, 1);
        if (st != GG_OK) return st;
    }
    return GG_OK;
}

void learn_fill_stats(gg_search *s, gg_search_stats *st) {
    LearnState *L = s->learn;
    if (!L) return;
    st->train_steps = L->train_steps;
    if (L->loss_pending) {
        float loss = 0;
        if (hipMemcpyAsync(&loss, L->mlp->d_loss, 4, hipMemcpyDeviceToHost, L->st) == hipSuccess &&
            hipStreamSynchronize(L->st) == hipSuccess)
            L->last_loss = loss;
        L->loss_pending = false;
    }
    st->train_loss = L->last_loss;
    st->spearman = L->spearman;
    st->miss_rate = L->miss;
}

static std::vector<double> ranks(const std::vector<double> &x) {
    std::vector<int> o(x.size());
    std::iota(o.begin(), o.end(), 0);
    std::sort(o.begin(), o.end(), [&](int a, int b) { return x[a] < x[b]; });
    std::vector<double> r(x.size());
    for (size_t i = 0; i < o.size();) {
        size_t j = i;
        while (j + 1 < o.size() && x[o[j + 1]] == x[o[i]]) j++;
        double avg = 0.5 * (double)(i + j);
        for (size_t t = i; t <= j; t++) r[o[t]] = avg;
        i = j + 1;
    }
    return r;
}

static double pearson(const std::vector<double> &a, const std::vector<double> &b) {
    const size_t n = a.size();
    double ma = 0, mb = 0;
    for (size_t i = 0; i < n; i++) { ma += a[i]; mb += b[i]; }
    ma /= n; mb /= n;
    double sab = 0, saa = 0, sbb = 0;
    for (size_t i = 0; i < n; i++) {
        sab += (a[i] - ma) * (b[i] - mb);
        saa += (a[i] - ma) * (a[i] - ma);
        sbb += (b[i] - mb) * (b[i] - mb);
    }
    return (saa > 0 && sbb > 0) ? sab / std::sqrt(saa * sbb) : 0.0;
}

}  // namespace gg

using namespace gg;

extern "C" int gg_batch_propose(gg_batch *b, uint64_t seed, int64_t step, int P,
                                int32_t *moves_out) {
    if (!b || !moves_out || P < 1 || step < 0 || step > INT32_MAX) return GG_BAD_ARG;
    int N = 0, d = 0;
    if (batch_shape(b, &N, &d) != GG_OK) return GG_BAD_ARG;
    hipStream_t st = batch_stream(b);
    // d_c[0..4): control words in (step, warm, diameter, -); d_c[4..8): the
    // kernel's copy, [7] its error word
    int *d_m = nullptr, *d_c = nullptr;
    gg_swap *d_s = nullptr;
    int h_c[8] = {(int)step, 0, 0, 0, 0, 0, 0, 0};
    int rc = GG_OK;
    if (hipMalloc((void **)&d_m, 4 * (size_t)P * 5) != hipSuccess ||
        hipMalloc((void **)&d_s, sizeof(gg_swap) * (size_t)P) != hipSuccess ||
        hipMalloc((void **)&d_c, sizeof(h_c)) != hipSuccess)
        rc = GG_OOM;
    if (rc == GG_OK && hipMemcpyAsync(d_c, h_c, sizeof(h_c), hipMemcpyHostToDevice, st) != hipSuccess)
        rc = GG_HIP_ERROR;
    if (rc == GG_OK) {
        hipLaunchKernelGGL(k_propose, dim3((P + 63) / 64), dim3(64), 0, st,
                           ProposeArgs{b_adjR(b), N, d, seed, d_c, P, d_m, d_s, d_c + 4});
        if (hipGetLastError() != hipSuccess) rc = GG_HIP_ERROR;
    }
    if (rc == GG_OK &&
        (hipMemcpyAsync(moves_out, d_m, 4 * (size_t)P * 5, hipMemcpyDeviceToHost, st) != hipSuccess ||
         hipMemcpyAsync(h_c, d_c, sizeof(h_c), hipMemcpyDeviceToHost, st) != hipSuccess))
        rc = GG_HIP_ERROR;
    if (hipStreamSynchronize(st) != hipSuccess && rc == GG_OK) rc = GG_HIP_ERROR;
    if (rc == GG_OK && h_c[7]) rc = GG_BAD_ARG;
    if (d_m) (void)hipFree(d_m);
    if (d_s) (void)hipFree(d_s);
    if (d_c) (void)hipFree(d_c);
    return rc;
}

// Audit: draw P fresh proposals (own RNG, the search state is unchanged),
// score them, evaluate all P exactly, and report the scorer's Spearman rank
// correlation with the exact delta and the fraction of the exact top-K it
// misses.  Optionally returns the P (score, exact delta) pairs.
extern "C" int gg_search_audit(gg_search *s, uint64_t seed, double *spearman, double *miss_rate,
                               float *scores_out, double *delta_out) {
    if (!s || !s->learn) return GG_BAD_ARG;
    LearnState *L = s->learn;
    const int P = s->cfg.P, K = s->cfg.K;
    Rng rng(seed);
    for (int p = 0; p < P; p++) {
        int tries = 0;
        while (!propose_swap(s->adj, s->N, s->d, rng, &s->sw[p], &s->mv[p]))
            if (++tries > 1000000) return GG_BAD_ARG;
    }
    int st = score_proposals(s, P);
    if (st != GG_OK) return st;
    std::vector<gg_result> res(P);
    st = s->be->eval(s->sw.data(), P, res.data());
    if (st != GG_OK) return st;
    std::vector<double> sc(P), dl(P);
    for (int p = 0; p < P; p++) {
        sc[p] = L->h_score[p];
        dl[p] = raw_delta(s, res[p]);
        if (dl[p] > 1e299) dl[p] = 1e30;
        if (scores_out) scores_out[p] = L->h_score[p];
        if (delta_out) delta_out[p] = dl[p];
    }
    const double rho = pearson(ranks(sc), ranks(dl));
    std::vector<int> a(P), b(P);
    std::iota(a.begin(), a.end(), 0);
    std::iota(b.begin(), b.end(), 0);
    // ties broken by proposal index, so the sets are a function of (sc, dl)
    // alone (a stable argsort of either array gives the same K)
    std::partial_sort(a.begin(), a.begin() + K, a.end(), [&](int x, int y) {
        return sc[x] < sc[y] || (sc[x] == sc[y] && x < y);
    });
    std::partial_sort(b.begin(), b.begin() + K, b.end(), [&](int x, int y) {
        return dl[x] < dl[y] || (dl[x] == dl[y] && x < y);
    });
    std::vector<char> inA(P, 0);
    for (int k = 0; k < K; k++) inA[a[k]] = 1;
    int hit = 0;
    for (int k = 0; k < K; k++) hit += inA[b[k]];
    L->spearman = rho;
    L->miss = 1.0 - (double)hit / K;
    if (spearman) *spearman = rho;
    if (miss_rate) *miss_rate = L->miss;
    return GG_OK;
}

extern "C" int gg_search_last_selection(gg_search *s, int64_t *step, int32_t *warm,
                                        float *scores_out, int32_t *moves_out, int32_t *sel_out) {
    if (!s || !s->learn) return GG_BAD_ARG;
    LearnState *L = s->learn;
    if (!L->dev) return GG_UNSUPPORTED;
    if (s->step < 1) return GG_BAD_ARG;  // no step has run
    const int P = L->P, K = L->K;
    // the device buffers hold the last captured step's draws, scores and
    // choice until the next step: nothing after the step rewrites them
    GG_HIP(hipStreamSynchronize(L->st));
    if (scores_out)
        GG_HIP(hipMemcpy(scores_out, L->d_score, 4 * (size_t)P, hipMemcpyDeviceToHost));
    if (moves_out)
        GG_HIP(hipMemcpy(moves_out, L->d_moves, 4 * (size_t)P * 5, hipMemcpyDeviceToHost));
    if (sel_out) GG_HIP(hipMemcpy(sel_out, L->d_sel, 4 * (size_t)K, hipMemcpyDeviceToHost));
    int hdr[4];
    GG_HIP(hipMemcpy(hdr, L->d_ctl, sizeof(hdr), hipMemcpyDeviceToHost));
    if (step) *step = hdr[0];
    if (warm) *warm = hdr[1];
    return GG_OK;
}
