/*
 * graphgolf_search.h -- C ABI of the search drivers built on the exact
 * evaluator: simulated annealing / hill climbing over 2-opt swaps of a
 * regular graph, with plain, batched and learned-proposal evaluation.
 *
 * Replaces the driver loops of cppcode/annealing.cpp:396-443 (acceptance and
 * schedule), cppcode/HillClimb.cpp:222-249 and createNeighbour
 * (annealing.cpp:112-197), which the reference compiles into each CLI.
 */
#ifndef GRAPHGOLF_SEARCH_H
#define GRAPHGOLF_SEARCH_H

#include <stdint.h>
#include "graphgolf_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gg_search gg_search;

#define GG_MODE_PLAIN   0  /* one exact evaluation per step                  */
#define GG_MODE_BATCHED 1  /* K candidates per step, one launch               */
#define GG_MODE_LEARNED 2  /* P proposals scored by the MLP, top K exact     */

#define GG_RULE_ANNEAL  0  /* cppcode/annealing.cpp:406-412                   */
#define GG_RULE_HILL    1  /* cppcode/HillClimb.cpp:239                        */
/* The same two rules on (diameter, WVC, ASPL) compared lexicographically, WVC
 * as WVCdiameterASPL computes it (cudacode/cudaASPLconv.cu:256-285):
 * annealing accepts a lexicographic improvement, else draws the uniform; a
 * larger diameter, or an equal diameter with a larger WVC, is always rejected
 * (cppcode/annealingWVC.cpp:412-417); hill climbing accepts only a
 * lexicographic improvement (cppcode/HillClimbWVC.cpp:240).  The best
 * solution is kept in the same order (annealingWVC.cpp:433).  Plain mode of
 * gg_search (explicit graphs, all N vertices as sources) and part tables of
 * gg_orbit_search only: the reference has WVC for cudaASPLconv alone.  The
 * level limit of these searches is WVCdiameterASPL's, 100 (:264). */
#define GG_RULE_ANNEAL_WVC 2
#define GG_RULE_HILL_WVC   3

#define GG_SCHED_GEOMETRIC 0  /* T(i) = T0 (T1/T0)^(i/total_steps)              */
#define GG_SCHED_STEP      1  /* T(i) = T0 0.995^floor((i-1)/10^4), updated after
                                 every 10^4 iterations (annealing.cpp:442)     */

typedef struct gg_search_cfg {
    int32_t mode;
    int32_t rule;
    int32_t K;            /* exact candidates per step (batched, learned)    */
    int32_t P;            /* proposals scored per step (learned)             */
    int32_t total_steps;  /* length of the temperature schedule              */
    int32_t warmup_pairs; /* learned: uniform selection until this many exact pairs */
    int32_t train_every;  /* learned: optimiser steps after this many steps  */
    int32_t train_batch;  /* learned: minibatch size                         */
    double T0, T1;        /* temperature on the ASPL scale, geometric T0 -> T1 */
    double lr;            /* learned: Adam learning rate                     */
    uint64_t seed;
    int32_t schedule;     /* GG_SCHED_*                                       */
} gg_search_cfg;

typedef struct gg_search_stats {
    int64_t steps, exact_evals, proposals, accepted;
    int64_t total_cur, total_best;   /* scaled distance totals              */
    int32_t diam_cur, diam_best;
    double seconds;                  /* wall time inside gg_search_run       */
    /* learned mode */
    int64_t train_pairs, train_steps;
    double train_loss;               /* last minibatch MSE (scaled target)   */
    double spearman;                 /* rank correlation score vs exact delta (last audit) */
    double miss_rate;                /* fraction of the exact top-K the scorer missed (last audit) */
} gg_search_stats;

/* Exact evaluator supplied by the caller (the reference's compile-time
 * evaluator substitution, annealing.cpp:296 vs annealing_cpu.cpp:292). */
typedef struct gg_eval_callbacks {
    void *ctx;
    int (*set_base)(void *ctx, const int32_t *adj, gg_result *out);
    int (*eval)(void *ctx, const gg_swap *swaps, int K, gg_result *out);
    int (*commit)(void *ctx, const gg_swap *sw, gg_result *out);
} gg_eval_callbacks;

int gg_search_create(int N, int d, const int32_t *adj, const gg_search_cfg *cfg, int device,
                     gg_search **out);
int gg_search_create_cb(int N, int d, const int32_t *adj, const gg_search_cfg *cfg,
                        const gg_eval_callbacks *cb, gg_search **out);
void gg_search_destroy(gg_search *s);
int gg_search_run(gg_search *s, int steps, gg_search_stats *out);
/* Current (best = 0) or best (best = 1) graph as N x d adjacency. */
int gg_search_graph(gg_search *s, int32_t *adj_out, int best);
/* Per-step log: 10 int64 per step {step, a, b, c, d, kind, total, diameter,
 * status, accepted} of the candidate that went through the acceptance rule
 * (total = diameter = -1 when its status is not GG_OK). */
int gg_search_set_log(gg_search *s, int capacity);
int gg_search_get_log(gg_search *s, int64_t *buf, int capacity, int *count);

/* Learned mode only: draw P fresh proposals with an RNG seeded by `seed`
 * (the search state is unchanged), score them, evaluate all P exactly, and
 * report the scorer's Spearman rank correlation with the exact change of the
 * objective and the fraction of the exact top-K it misses.  scores_out /
 * delta_out (P entries each) may be NULL. */
int gg_search_audit(gg_search *s, uint64_t seed, double *spearman, double *miss_rate,
                    float *scores_out, double *delta_out);

/* Learned mode with proposals drawn on the GPU (P <= 4096) only: what the
 * last step drew, scored and chose, for checking the selection rule from
 * outside.  step = the step index the draws used (stream (seed, step, p) of
 * proposal p, stream (seed, step, 2^32 - 1) of the exploration draws),
 * warm = 1 when all K were taken uniformly (proposals 0..K-1, before
 * warmup_pairs exact pairs existed); scores_out P floats, moves_out P rows
 * (a, b, c, d, kind), sel_out the K chosen proposal indices in evaluation
 * order.  Any output pointer may be NULL.  Valid until the next step or
 * audit (an audit scores its own proposals in the same buffers).  Counterpart of the reference's
 * one draw per move (cppcode/annealing.cpp:399-403), which has no selection. */
int gg_search_last_selection(gg_search *s, int64_t *step, int32_t *warm, float *scores_out,
                             int32_t *moves_out, int32_t *sel_out);

/* ---- search over cyclically symmetric graphs ------------------------------
 * The reference's `part` form (cppcode/part.hpp:6-44): M representative rows
 * of `degree` offsets, N = g*M, neighbour i of vertex v is
 * (v + table[v mod M][i]) mod N.  Moves are createNeighbour's for symmetric
 * parts (cppcode/annealing.cpp:112-197: edge-length change and orbit 2-opt,
 * with fresh lifts), redrawn when the table is not simple; each candidate
 * table is evaluated exactly from the M representatives (gg_eval_part), and
 * acceptance is the rule of gg_search (annealing.cpp:406-412 /
 * HillClimb.cpp:239; pieceannealing.cpp:252-260 for the piece form of the
 * same graphs).  Mode must be GG_MODE_PLAIN. */
typedef struct gg_orbit_search gg_orbit_search;

/* Exact table evaluator supplied by the caller (CPU twin of the driver). */
typedef struct gg_table_callbacks {
    void *ctx;
    int (*eval)(void *ctx, const int32_t *table, gg_result *out);
} gg_table_callbacks;

int gg_orbit_search_create(int N, int M, int degree, const int32_t *table,
                           const gg_search_cfg *cfg, int device, gg_orbit_search **out);
int gg_orbit_search_create_cb(int N, int M, int degree, const int32_t *table,
                              const gg_search_cfg *cfg, const gg_table_callbacks *cb,
                              gg_orbit_search **out);
/* Rows of different lengths, as annealingNR draws them (a row, then a slot
 * among that row's entries; partners among the partner row's entries:
 * cppcode/annealingNR.cpp:117,146-147): lens[r] in [1, degree] entries of row
 * r are edges, the slots after them padding that must hold offset 0 (a
 * self-loop, as cudaASPLconv.cu:199-201 pads short rows).  Row lengths never
 * change.  Part form only; call before the first run. */
int gg_orbit_search_set_row_lengths(gg_orbit_search *s, const int32_t *lens);
void gg_orbit_search_destroy(gg_orbit_search *s);
int gg_orbit_search_run(gg_orbit_search *s, int steps, gg_search_stats *out);
/* Current (best = 0) or best (best = 1) table, M x degree offsets. */
int gg_orbit_search_table(gg_orbit_search *s, int32_t *table_out, int best);
/* Per-step log: 10 int64 per step {step, a*degree+slot_a, c*degree+slot_c,
 * k1, k2, kind (0 = length change, 1 = 2-opt), total, diameter, status,
 * accepted} (total = diameter = -1 when the status is not GG_OK).  Lattice
 * searches pack a lift vector as k = (kx*gy + ky)*gz + kz. */
int gg_orbit_search_set_log(gg_orbit_search *s, int capacity);
int gg_orbit_search_get_log(gg_orbit_search *s, int64_t *buf, int capacity, int *count);

/* The same search over 2-D (`piece`, cppcode/piece.hpp:6-53) and 3-D (`box`,
 * cppcode/box.hpp:6-81) symmetric graphs, replacing the move generator and
 * loop of pieceannealing.cpp / pieceHillClimb.cpp (createNeighbour
 * pieceannealing.cpp:113-204, acceptance :252-260) and boxHillClimb.cpp
 * (:113-220, :252-273) with the evaluators of gg_eval_piece / gg_eval_box.
 * table = [Mx*My(*Mz)][degree][2 or 3] offsets in the evaluator's layout.
 * Edge-length changes are drawn only when every axis has more than one copy
 * (pieceannealing.cpp:115); lifts are drawn per axis, x first; a move whose
 * touched rows are not simple and symmetric is redrawn.  The handle is a
 * gg_orbit_search: run / table / log / destroy are shared. */
int gg_piece_search_create(int Nx, int Ny, int Mx, int My, int degree, const int32_t *table,
                           const gg_search_cfg *cfg, int device, gg_orbit_search **out);
int gg_box_search_create(int Nx, int Ny, int Nz, int Mx, int My, int Mz, int degree,
                         const int32_t *table, const gg_search_cfg *cfg, int device,
                         gg_orbit_search **out);
/* dims = 2 or 3; Ns / Ms per axis; the caller's exact table evaluator. */
int gg_lattice_search_create_cb(int dims, const int32_t *Ns, const int32_t *Ms, int degree,
                                const int32_t *table, const gg_search_cfg *cfg,
                                const gg_table_callbacks *cb, gg_orbit_search **out);

/* ---- seeded instances ----------------------------------------------------
 * gg_gen_random_regular: simple d-regular graph on n vertices as an n x d
 * adjacency with sorted rows; configuration-model pairing, resampled when
 * d <= 6, repaired by double-edge swaps otherwise (the reference's pairing
 * of initTemp.cpp:399-408 and RNDpiecegen.cpp:115-127 never rejects loops or
 * multi-edges).  gg_gen_orbit_part: cyclically symmetric d-regular graph on
 * g*m vertices as an m x d `part` table of orbit offsets (j - i) + m*s,
 * orbits drawn by stub pairing and lift s uniform in [0, g), resampled until
 * there are no self-loops, half orbits or repeated orbits.  Both draw from
 * mt19937_64(seed) with explicit bounded-integer mappings. */
int gg_gen_random_regular(int n, int d, uint64_t seed, int32_t *adj_out);
int gg_gen_orbit_part(int g, int m, int d, uint64_t seed, int32_t *table_out);
/* Cyclically symmetric d-regular table of `dims` axes (1 = part, 2 = piece,
 * 3 = box): orders Ns and base extents Ms per axis, table [prod Ms][d][dims]
 * in the evaluators' layout.  The stub pairing with a lift per axis of
 * RNDpiecegen.cpp:89-127 / RNDboxgen.cpp:97-140, with pairs that give a
 * self-loop or a repeated offset repaired (the reference keeps them). */
int gg_gen_lattice(int dims, const int32_t *Ns, const int32_t *Ms, int d, uint64_t seed,
                   int32_t *table_out);

/* ---- initial temperature ------------------------------------------------
 * cppcode/initTemp.cpp:399-430 (pieceinitTemp.cpp:228-260): over `graphs`
 * random graphs, the mean increase of the distance total over the
 * `neighbours` moves of each that make it worse (unconnected results
 * skipped), and T = -mean / ln 0.4.  gg_init_temp: random regular graphs and
 * 2-opt swaps, one batched evaluation per graph; gg_init_temp_part: random
 * orbit graphs (N = g*M, g >= 2) and orbit moves.  Graph i comes from seed
 * seed + 1 + i, the moves from one generator seeded with `seed`.
 * temp_total is the reference's inittemp (annealing.cpp -t, total-distance
 * units); temp_aspl = temp_total / (N(N-1)) is gg_search_cfg.T0's scale. */
typedef struct gg_init_temp_result {
    int64_t count;               /* worsening moves seen                        */
    double mean_increase_total;  /* their mean increase of the distance total   */
    double temp_total;
    double temp_aspl;
} gg_init_temp_result;
int gg_init_temp(int N, int degree, int graphs, int neighbours, uint64_t seed, int device,
                 gg_init_temp_result *out);
int gg_init_temp_part(int N, int M, int degree, int graphs, int neighbours, uint64_t seed,
                      int device, gg_init_temp_result *out);

#ifdef __cplusplus
}
#endif
#endif
