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
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>
#include "gg_common.hpp"
#include "gg_batch.hpp"
#include "gg_learn.hpp"
#include "gg_search_impl.hpp"

namespace gg {

__global__ void k_gather_rows(const float *src, const int *idx, int n, float *dst, int dst_cap,
                              int head) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * GG_FEATS) return;
    const int r = i / GG_FEATS, c = i - r * GG_FEATS;
    dst[(size_t)((head + r) % dst_cap) * GG_FEATS + c] = src[(size_t)idx[r] * GG_FEATS + c];
}

int gather_rows_dev(const float *src, const int *idx, int n, float *dst, int dst_cap, int head,
                    hipStream_t st) {
    hipLaunchKernelGGL(k_gather_rows, dim3((n * GG_FEATS + 255) / 256), dim3(256), 0, st, src, idx,
                       n, dst, dst_cap, head);
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
    int st = gg_mlp_create(device, s->cfg.seed + 1, &L->mlp);
    if (st != GG_OK) return st;
    return alloc_learn(L);
}

void learn_destroy(gg_search *s) {
    LearnState *L = s->learn;
    if (!L) return;
    if (L->st) (void)hipStreamSynchronize(L->st);
    gg_mlp_destroy(L->mlp);
    void *dev[] = {L->d_moves, L->d_x, L->d_score, L->d_sel, L->d_bx, L->d_by, L->d_idx};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    void *host[] = {L->h_moves, L->h_score, L->h_sel, L->h_idx, L->h_ynorm};
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
    int st = features_dev(L->batch, L->d_moves, P, L->d_x, L->st);
    if (st != GG_OK) return st;
    st = mlp_forward_dev(L->mlp, L->d_x, P, L->d_score, L->st);
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
        if (t > 0) GG_HIP(hipStreamSynchronize(L->st));  // h_idx is reused
        for (int i = 0; i < B; i++) L->h_idx[i] = (int)L->rng->below(L->count);
        GG_HIP(hipMemcpyAsync(L->d_idx, L->h_idx, 4 * B, hipMemcpyHostToDevice, L->st));
        int st = mlp_grad_dev(L->mlp, L->d_bx, L->d_by, L->d_idx, B, L->st);
        if (st != GG_OK) return st;
        st = mlp_adam_dev(L->mlp, (float)(s->cfg.lr > 0 ? s->cfg.lr : 1e-3), L->st);
        if (st != GG_OK) return st;
        L->train_steps++;
    }
    L->loss_pending = true;  // read when the stats are (no wait per step)
    return GG_OK;
}

// Uploads normalised targets for ring positions [pos, pos + n) (n <= cap).
static int upload_targets(LearnState *L, int pos, int n) {
    int first = std::min(n, L->cap - pos);
    for (int i = 0; i < n; i++) {
        double y = (L->raw[(pos + i) % L->cap] - L->mu) / L->sigma;
        L->h_ynorm[(pos + i) % L->cap] = (float)std::max(-8.0, std::min(8.0, y));
    }
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
        L->h_sel[k] = L->last_sel[k];
    }
    GG_HIP(hipMemcpyAsync(L->d_sel, L->h_sel, 4 * (size_t)K, hipMemcpyHostToDevice, L->st));
    int st = gather_rows_dev(L->d_x, L->d_sel, K, L->d_bx, L->cap, pos, L->st);
    if (st != GG_OK) return st;
    L->head = (L->head + K) % L->cap;
    L->count = std::min(L->count + K, L->cap);
    s->stats.train_pairs += K;
    if (!L->norm_frozen) {
        const int need = std::max(64, std::min(s->cfg.warmup_pairs, 4096));
        if (L->count < need) return GG_OK;
        // mean / std of the finite raw deltas seen so far, then normalise all
        double m = 0, v = 0;
        int n = 0;
        for (int i = 0; i < L->count; i++)
            if (L->raw[i] < 1e29) { m += L->raw[i]; n++; }
        m = n ? m / n : 0.0;
        for (int i = 0; i < L->count; i++)
            if (L->raw[i] < 1e29) v += (L->raw[i] - m) * (L->raw[i] - m);
        v = n > 1 ? v / n : 1.0;
        L->mu = m;
        L->sigma = v > 0 ? std::sqrt(v) : 1.0;
        L->norm_frozen = true;
        st = upload_targets(L, 0, L->count);
    } else {
        st = upload_targets(L, pos, K);
    }
    if (st != GG_OK) return st;
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
    std::partial_sort(a.begin(), a.begin() + K, a.end(), [&](int x, int y) { return sc[x] < sc[y]; });
    std::partial_sort(b.begin(), b.begin() + K, b.end(), [&](int x, int y) { return dl[x] < dl[y]; });
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
