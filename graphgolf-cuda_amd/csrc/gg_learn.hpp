// gg_learn.hpp -- learned proposal scorer (host side declarations).
#pragma once
#include <vector>
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include "gg_rng.hpp"
#include "graphgolf_learn.h"
#include "gg_common.hpp"

#define GG_FEATS 32
#define GG_HIDDEN 128
#define GG_NPARAMS (GG_HIDDEN * GG_FEATS + GG_HIDDEN + GG_HIDDEN + 1)
#define GG_MAX_BATCH 512
#define GG_GRAD_SAMPLES 16                          // minibatch samples per gradient group
#define GG_PART_STRIDE (GG_NPARAMS + 1)             // partial gradient + squared error
#define GG_ADAM_GROUPS (GG_NPARAMS / 256 + 1)       // 256-thread groups of the update

struct gg_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float *d_w = nullptr, *d_m = nullptr, *d_v = nullptr, *d_g = nullptr;  // fp32 masters, Adam, grad
    __hip_bfloat16 *d_wq = nullptr;  // bf16 weights used by every forward
    __hip_bfloat16 *d_wq2 = nullptr; // second bf16 copy (the learned search's resident loop
                                     // alternates the two by step parity)
    float *d_pre = nullptr, *d_loss = nullptr;  // d_pre: per-workgroup partial gradients
    float *d_x = nullptr, *d_y = nullptr, *d_out = nullptr;
    int xcap = 0;
    int t = 0;  // Adam step
};

struct gg_batch;
struct SparseArgs;

// Device-side training decision of a step (the learned search's resident
// loop): on = take one Adam step, b1t / b2t = 0.9^t and 0.999^t.
struct TrainGate {
    int on;
    float b1t, b2t;
    int pad;
};

// A gated training step's gradient (groups of GG_GRAD_SAMPLES samples): rows
// idx[0..B) of X / Y, the bf16 copy *par & 1 (wq, wq2), partials into part.
struct GradArgs {
    const float *X, *Y;
    const int *idx;
    int B;
    const __hip_bfloat16 *wq, *wq2;
    const int *par;
    float *part;
    const TrainGate *gate;
};

// Its reduction and Adam update: the G partials summed in group order, the
// fp32 masters updated and the bf16 copy (*par - 1) & 1 written (from the
// unchanged masters when the gate is off); without par, copy wq.
struct AdamArgs {
    const float *part;
    int G, B;
    float *grad, *loss, *w, *m, *v;
    __hip_bfloat16 *wq, *wq2;
    float lr;
    const TrainGate *gate;
    const int *par;
};

// One Adam step of a parameter from its moments m0, v0 and value w0.
__device__ __forceinline__ void adam_math(float m0, float v0, float w0, float gi, float lr,
                                          float b1t, float b2t, float &mi, float &vi, float &wi) {
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    mi = beta1 * m0 + (1.f - beta1) * gi;
    vi = beta2 * v0 + (1.f - beta2) * gi * gi;
    const float mh = mi / (1.f - b1t), vh = vi / (1.f - b2t);
    wi = w0 - lr * mh / (sqrtf(vh) + eps);
}

__device__ __forceinline__ void adam_update(float *w, float *m, float *v, __hip_bfloat16 *wq,
                                            int i, float gi, float lr, float b1t, float b2t) {
    float mi, vi, wi;
    adam_math(m[i], v[i], w[i], gi, lr, b1t, b2t, mi, vi, wi);
    m[i] = mi;
    v[i] = vi;
    w[i] = wi;
    wq[i] = __float2bfloat16(wi);
}

// Parameter i (i == GG_NPARAMS: the loss) of an Adam step.  Every load the
// step needs (gate, moments, master value, the partial gradients) is issued
// before any store.
__device__ __forceinline__ void reduce_adam_body(const AdamArgs &a, int i) {
    const int par = a.par ? *a.par : 0;
    const int on = a.gate->on;
    const float b1t = a.gate->b1t, b2t = a.gate->b2t;
    const int ic = i < GG_NPARAMS ? i : 0;
    const float m0 = a.m[ic], v0 = a.v[ic], w0 = a.w[ic];
    __hip_bfloat16 *wq = (a.par && ((par - 1) & 1)) ? a.wq2 : a.wq;
    if (!on) {
        if (a.par && i < GG_NPARAMS) wq[i] = __float2bfloat16(w0);
        return;
    }
    if (i > GG_NPARAMS) return;
    float acc = 0.f;
#pragma unroll 16
    for (int g = 0; g < a.G; g++) acc += a.part[(size_t)g * GG_PART_STRIDE + i];
    if (i == GG_NPARAMS) {
        *a.loss = acc / (float)a.B;
        return;
    }
    float mi, vi, wi;
    adam_math(m0, v0, w0, acc, a.lr, b1t, b2t, mi, vi, wi);
    a.grad[i] = acc;
    a.m[i] = mi;
    a.v[i] = vi;
    a.w[i] = wi;
    wq[i] = __float2bfloat16(wi);
}

namespace gg {
int mlp_init_params(std::vector<float> &w, uint64_t seed);
int mlp_forward_dev(gg_mlp *m, const float *d_x, int P, float *d_out, hipStream_t st);
int mlp_grad_dev(gg_mlp *m, const float *d_x, const float *d_y, const int *d_idx, int B,
                 hipStream_t st);
int mlp_adam_dev(gg_mlp *m, float lr, hipStream_t st);
// gradient + Adam step gated on the device (*gate read by the kernels).  With
// par (a device step counter p, read at launch): the gradient's forward uses
// the bf16 copy p & 1 and the update writes copy (p - 1) & 1 (also when the
// gate is off, refreshed from the fp32 masters).
int mlp_train_gated_dev(gg_mlp *m, const float *d_x, const float *d_y, const int *d_idx, int B,
                        float lr, const TrainGate *gate, hipStream_t st,
                        const int *par = nullptr);
// both bf16 copies from the fp32 masters
int mlp_refresh_dev(gg_mlp *m, hipStream_t st);
// arguments of a gated step's gradient / update kernels carried by other launches
GradArgs mlp_grad_args(gg_mlp *m, const float *d_x, const float *d_y, const int *d_idx, int B,
                       const TrainGate *gate, const int *par);
AdamArgs mlp_adam_args(gg_mlp *m, int B, float lr, const TrainGate *gate, const int *par);
// diam_dev: the base diameter read on the device (null: the host's value)
int features_dev(gg_batch *b, const int *d_moves, int P, float *d_x, hipStream_t st,
                 const int *diam_dev = nullptr);
// wpar: device step counter p; the scores use bf16 copy p & 1 (null: d_wq).
// grad: a gated training step's gradient carried by the same launch.
int features_score_dev(gg_batch *b, gg_mlp *m, const int *d_moves, int P, float *d_x,
                       float *d_score, hipStream_t st, const int *diam_dev = nullptr,
                       const int *wpar = nullptr, const GradArgs *grad = nullptr);
// fin: null launches the finalize; else its arguments are returned there (and
// *fin_nvalid its second argument) for the caller's own finalize launch
int batch_enqueue_dev(gg_batch *b, int K, SparseArgs *fin = nullptr, int *fin_nvalid = nullptr);
struct BatchOut {
    const ResultRec *d_res;
    ResultRec *h_res;      // pinned
    gg_swap *h_swaps;      // pinned
    int *qctl;
    int nqctl;
};
BatchOut batch_out(gg_batch *b);
int batch_finish_dev(gg_batch *b, int K, gg_result *out);
int batch_push_state(gg_batch *b);
const int *b_adjR(gg_batch *b);
int batch_shape(gg_batch *b, int *N, int *d);
gg_swap *b_swaps(gg_batch *b);
const gg_swap *b_hswaps(gg_batch *b);
int gather_rows_dev(const float *src, const int *idx, int n, float *dst, int dst_cap, int head,
                    hipStream_t st, const float *ysrc = nullptr, float *ydst = nullptr);
hipStream_t batch_stream(gg_batch *b);
struct ProposeArgs;
// next: the commit's parts launch also draws those proposals (extra workgroups)
int batch_enqueue_commit_dev(gg_batch *b, const CommitSel *sel,
                             const ProposeArgs *next = nullptr);
void batch_note_device_commit(gg_batch *b, const gg_swap *sw, const gg_result *known);
int batch_commit_rebuild(gg_batch *b, const gg_swap *sw, const gg_result *known);
bool batch_single_launch(gg_batch *b, int K);
int batch_snap_cap(gg_batch *b);
bool batch_sparse_ok(gg_batch *b);
int *batch_qctl(gg_batch *b, int *n);
const ResultRec *batch_dres(gg_batch *b);
}  // namespace gg
