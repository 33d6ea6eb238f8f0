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

struct gg_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float *d_w = nullptr, *d_m = nullptr, *d_v = nullptr, *d_g = nullptr;  // fp32 masters, Adam, grad
    __hip_bfloat16 *d_wq = nullptr;  // bf16 weights used by every forward
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

namespace gg {
int mlp_init_params(std::vector<float> &w, uint64_t seed);
int mlp_forward_dev(gg_mlp *m, const float *d_x, int P, float *d_out, hipStream_t st);
int mlp_grad_dev(gg_mlp *m, const float *d_x, const float *d_y, const int *d_idx, int B,
                 hipStream_t st);
int mlp_adam_dev(gg_mlp *m, float lr, hipStream_t st);
// gradient + Adam step gated on the device (*gate read by the kernels)
int mlp_train_gated_dev(gg_mlp *m, const float *d_x, const float *d_y, const int *d_idx, int B,
                        float lr, const TrainGate *gate, hipStream_t st);
// diam_dev: the base diameter read on the device (null: the host's value)
int features_dev(gg_batch *b, const int *d_moves, int P, float *d_x, hipStream_t st,
                 const int *diam_dev = nullptr);
int features_score_dev(gg_batch *b, gg_mlp *m, const int *d_moves, int P, float *d_x,
                       float *d_score, hipStream_t st, const int *diam_dev = nullptr);
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
int batch_enqueue_commit_dev(gg_batch *b, const CommitSel *sel);
void batch_note_device_commit(gg_batch *b, const gg_swap *sw, const gg_result *known);
int batch_commit_rebuild(gg_batch *b, const gg_swap *sw, const gg_result *known);
bool batch_single_launch(gg_batch *b, int K);
int batch_snap_cap(gg_batch *b);
bool batch_sparse_ok(gg_batch *b);
int *batch_qctl(gg_batch *b, int *n);
const ResultRec *batch_dres(gg_batch *b);
}  // namespace gg
