/*
 * graphgolf_learn.h -- C ABI of the learned proposal scorer: swap features
 * from the batch evaluator's stored base BFS levels, and a 32-128-1 MLP
 * (BF16 weights, FP32 accumulation) trained on the GPU by its own forward,
 * backward and Adam kernels.  No reference counterpart: the reference
 * evaluates every proposal exactly (cppcode/annealing.cpp:399-403).
 */
#ifndef GRAPHGOLF_LEARN_H
#define GRAPHGOLF_LEARN_H

#include <stdint.h>
#include "graphgolf_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GG_MLP_FEATURES 32
#define GG_MLP_HIDDEN 128
/* W1[128][32], b1[128], W2[128], b2 */
#define GG_MLP_PARAMS (128 * 32 + 128 + 128 + 1)
#define GG_MLP_MAX_BATCH 512

typedef struct gg_mlp gg_mlp;

int gg_mlp_create(int device, uint64_t seed, gg_mlp **out);
void gg_mlp_destroy(gg_mlp *m);
int gg_mlp_set_params(gg_mlp *m, const float *params);  /* fp32 masters; resets Adam */
int gg_mlp_get_params(gg_mlp *m, float *params);
/* scores for P feature rows (P x 32) */
int gg_mlp_forward(gg_mlp *m, const float *X, int P, float *out);
/* mean squared error and its gradient w.r.t. the parameters, forward with the
 * bf16 weights (B <= 512) */
int gg_mlp_grad(gg_mlp *m, const float *X, const float *Y, int B, float *loss, float *grad);
/* one Adam step (beta 0.9/0.999, eps 1e-8) on the minibatch */
int gg_mlp_train(gg_mlp *m, const float *X, const float *Y, int B, float lr, float *loss);

/* Features (P x 32) of P swap moves {a, b, c, d, kind} (kind 1: (a,d),(b,c);
 * kind 0: (a,c),(b,d)) on the batch evaluator's current base. */
int gg_batch_features(gg_batch *b, const int32_t *moves, int P, float *X);
/* The same features and the MLP's scores of the P moves, computed in one
 * launch as the learned search computes them; scores equal gg_mlp_forward on
 * X bit for bit.  No reference counterpart. */
int gg_batch_score(gg_batch *b, gg_mlp *m, const int32_t *moves, int P, float *X, float *scores);

/* P uniform 2-opt proposals on the batch evaluator's current base, drawn on
 * the GPU as the learned search draws them: proposal p of step `step` takes
 * splitmix64 stream (seed, step, p) (start state
 * mix(mix(seed ^ 0xA0761D6478BD642F) ^ step * 0xE7037ED1A0B428DB)
 * ^ p * 0x8EBC6AF09C88C6E3) and redraws a, ia, c, ic, kind (Lemire bounded
 * integers) until the rewired graph is simple, as createNeighbour's 2-opt
 * does (cppcode/annealing.cpp:140-197, here with both rewirings and the
 * simplicity test).  Writes P rows {a, b, c, d, kind}; GG_BAD_ARG when some
 * proposal found no simple swap in 4096 attempts. */
int gg_batch_propose(gg_batch *b, uint64_t seed, int64_t step, int P, int32_t *moves_out);

#ifdef __cplusplus
}
#endif
#endif
