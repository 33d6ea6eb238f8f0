/*
 * graphgolf_mi355x.h -- C ABI of the MI355X exact ASPL / diameter evaluator.
 *
 * Plain pointers and sizes only.  Every call is synchronous: results are valid
 * on return and caller memory is read only during the call.  One handle is
 * used by one host thread at a time (handles own their device buffers and
 * stream; distinct handles may be used concurrently).
 *
 * Each entry point names the reference interface it replaces
 * (confused-uec/graphgolf-cuda, file:line).
 */
#ifndef GRAPHGOLF_MI355X_H
#define GRAPHGOLF_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------
 * The reference checks no device status and reports graph conditions through
 * sentinels (cudacode/cudaASPLconv.cu:182,187 calc; :229,234 diameterASPL;
 * :293,298 WVCdiameterASPL; cudaASPLgraph.cu:182,190; cppcode/
 * cpuASPLqueue.cpp:92).  Here graph conditions are positive codes carried in
 * gg_result.status; errors are negative return values. */
#define GG_OK               0
#define GG_UNCONNECTED      1   /* some source does not reach every vertex          */
#define GG_DIAMETER_LIMIT   2   /* level 99 still found new vertices (cudaASPLconv.cu:208) */
#define GG_BAD_ARG        (-1)
#define GG_HIP_ERROR      (-2)
#define GG_OOM            (-3)
#define GG_NO_DEVICE      (-4)
#define GG_UNSUPPORTED    (-5)
#define GG_SCRATCH_LIMIT  (-6)   /* input larger than gg_eval_part_noalloc's fixed scratch  */

/* Highest BFS level evaluated by default: the reference loops step = 1..99
 * (cudacode/cudaASPLconv.cu:208) and reports a sentinel when step 99 still
 * finds new vertices.  WVCdiameterASPL loops to 100 (cudaASPLconv.cu:264,296)
 * and cudaASPLgraph::calc to 199 (cudaASPLgraph.cu:163,187); GG_OPT_LEVEL_LIMIT sets
 * the limit per evaluator, up to GG_LEVEL_LIMIT_MAX. */
#define GG_LEVEL_LIMIT 99
#define GG_LEVEL_LIMIT_MAX 199

typedef struct gg_eval gg_eval;
typedef struct gg_batch gg_batch;

/* Result of one exact evaluation. */
typedef struct gg_result {
    int64_t total_reps;   /* sum over the M sources s of sum_v d(s,v) (pairs reached) */
    int64_t total_scaled; /* total_reps * (N / M), integer N/M as cudaASPLconv.cu:240 */
    int32_t diameter;     /* max eccentricity over the sources (reached part)         */
    int32_t wvc;          /* WVC as cudaASPLconv.cu:256-285 computes it (GG_OPT_WVC)  */
    int32_t levels;       /* BFS levels executed (diameter + 1 when connected)        */
    int32_t status;       /* GG_OK, GG_UNCONNECTED or GG_DIAMETER_LIMIT                */
} gg_result;

/* One 2-opt edge swap as four adjacency-slot replacements on a shared base:
 * row v[j], slot slot[j] becomes nbr[j].  (a,b),(c,d) -> (a,d),(b,c) is
 *   {a, idx(a->b), d}, {b, idx(b->a), c}, {c, idx(c->d), b}, {d, idx(d->c), a}.
 * Rows may repeat (entries are applied in order j = 0..3).  Unused entries have
 * v = -1.  No reference counterpart: the reference re-uploads the whole table
 * per move (cudacode/cudaASPLconv.cu:198-202). */
typedef struct gg_swap {
    int32_t v[4];
    int32_t slot[4];
    int32_t nbr[4];
} gg_swap;

/* ---- errors / device ----------------------------------------------------- */
const char *gg_strerror(int status);
int gg_device_count(int *count);
/* ASPL exactly as the reference derives it: double(total)/(int64(N)*(N-1))
 * (cudacode/cudaASPLconv.cu:242; cppcode/cpuASPLqueue.cpp:94). */
double gg_aspl(const gg_result *r, int N);

/* ---- single-graph evaluators ---------------------------------------------
 * Replace the constructors cudaASPLconv(N,M,degree,device)
 * (cudacode/cudaASPLconv.hpp:18, .cu:122-137), cudaASPLpiece(Nx,Ny,Mx,My,
 * degree,device) (cudaASPLpiece.hpp:18, .cu:128-145), cudaASPLbox(Nx,Ny,Nz,Mx,
 * My,Mz,degree,device) (cudaASPLbox.hpp:18, .cu:135-152) and
 * cudaASPLgraph(N,degree_max,device) (cudaASPLgraph.hpp:16, .cu:114-127).
 * N need not be a multiple of M: the scaled total uses integer N/M as the
 * reference does (cudaASPLconv.cu:240). */
int gg_create_part(int N, int M, int degree, int device, gg_eval **out);
int gg_create_piece(int Nx, int Ny, int Mx, int My, int degree, int device, gg_eval **out);
int gg_create_box(int Nx, int Ny, int Nz, int Mx, int My, int Mz, int degree,
                  int device, gg_eval **out);
int gg_create_adj(int N, int degree_max, int device, gg_eval **out);
/* Replaces the destructors (cudacode/cudaASPLconv.cu:138-150, cudaASPLpiece.cu:146-158,
 * cudaASPLbox.cu:153-165, cudaASPLgraph.cu:128-140). */
void gg_destroy(gg_eval *ev);

/* Replaces cudaASPLconv::calc / diameterASPL / WVCdiameterASPL
 * (cudacode/cudaASPLconv.cu:151-196, 197-243, 244-308).  table is the
 * reference's packed h_edges: M rows of `degree` int32 offsets, short rows
 * padded with 0 (cudaASPLconv.cu:198-202).  Vertex v's neighbours are (v + off + N) % N for
 * the offsets of row v % M. */
int gg_eval_part(gg_eval *ev, const int32_t *table, gg_result *out);
/* Replaces cudaASPLpiece::diameterASPL (cudacode/cudaASPLpiece.cu:159-216);
 * table = h_edges[((x*My+y)*degree+e)*2 + {0,1}] (cudaASPLpiece.cu:163-174). */
int gg_eval_piece(gg_eval *ev, const int32_t *table, gg_result *out);
/* Replaces cudaASPLbox::diameterASPL (cudacode/cudaASPLbox.cu:166-226);
 * table = h_edges[((x*My*Mz+y*Mz+z)*degree+e)*3 + {0,1,2}] (cudaASPLbox.cu:170-183). */
int gg_eval_box(gg_eval *ev, const int32_t *table, gg_result *out);
/* Replaces cudaASPLgraph::calc (cudacode/cudaASPLgraph.cu:141-199): explicit
 * adjacency N x degree_max, entries < 0 are padding (the reference pads with
 * self-loops, cudaASPLgraph.cu:144-153).  All N vertices are sources. */
int gg_eval_adj(gg_eval *ev, const int32_t *adj, gg_result *out);

/* gg_create_part + gg_eval_part + gg_destroy as one call that allocates
 * nothing (no hipMalloc, hipFree, pinned or library allocation): no
 * counterpart in the reference; for callers that cannot allocate (serving
 * loops, code that forbids device allocation).  Same table and result as
 * gg_eval_part with the default options (level limit 99, WVC computed).  Its
 * buffers live in one fixed 4 MiB device array per GPU, kept until the
 * process ends: N <= 16384 and M * degree <= 524288, else GG_SCRATCH_LIMIT.
 * Runs on the device's null stream and returns when the result is on the
 * host.  Calls that overlap on one device share that array and race. */
int gg_eval_part_noalloc(int N, int M, int degree, int device, const int32_t *table,
                         gg_result *out);

/* Plain single-move path on a resident adjacency (the table of the last
 * gg_eval_adj call stays on the device): evaluate base + one swap in one
 * launch without changing the base, and commit an accepted swap to the base.
 * Replaces the per-move cudaASPLgraph/cudaASPLconv call pattern of the drivers
 * (cppcode/annealing.cpp:399-403), which re-uploads the whole table. */
int gg_eval_adj_swap(gg_eval *ev, const gg_swap *sw, gg_result *out);
int gg_eval_adj_commit(gg_eval *ev, const gg_swap *sw);

/* Per-source eccentricities (M entries, source order: part/adj = vertex id;
 * piece = x*My+y; box = (x*My+y)*Mz+z) and per-level reach counts (pairs at
 * distance exactly l, l < reach_cap) of the most recent evaluation.  The BFS
 * pulls from each vertex's listed neighbours (as cudaASPLconv.cu:35-48 does),
 * so on a table that describes a directed graph these are in-eccentricities;
 * on undirected graphs (all the reference's drivers produce) in = out.  Either
 * pointer may be NULL.  No reference counterpart (the reference never exposes
 * them); used by the proposal features. */
int gg_eval_levels(gg_eval *ev, int32_t *ecc_out, int64_t *reach_out, int reach_cap);

/* Device time (ms, HIP events on the evaluator's stream) of the BFS kernel of
 * the most recent evaluation. */
int gg_eval_last_kernel_ms(gg_eval *ev, float *ms);
/* Mean time per launch of the last evaluation's BFS kernel, launched `reps`
 * times back to back between one event pair (no reference counterpart; for
 * the roofline measurement). */
int gg_eval_time_bfs(gg_eval *ev, int reps, float *ms);

/* Tuning knobs (0 = automatic choice). */
#define GG_OPT_BLOCK      1   /* threads per workgroup: 256 or 1024                  */
#define GG_OPT_LDS_LIMIT  2   /* max LDS bytes for the column bitset; a smaller value
                                 forces the global-memory kernel (used by tests)      */
#define GG_OPT_GENERIC    3   /* 1 = runtime-degree kernel even when the degree has a
                                 compile-time specialisation (used by tests)          */
#define GG_OPT_LEVEL_LIMIT 4  /* highest level evaluated (1..GG_LEVEL_LIMIT_MAX)      */
#define GG_OPT_PUSH       5   /* 1 (default) = levels 1-2 push from the frontier when
                                 the graph is undirected (adjacency evaluators verify
                                 it on upload and per swap); 0 = pull only           */
#define GG_OPT_WIDE       6   /* 1 (default) = a single evaluation whose column bitset
                                 does not fit in LDS runs level-synchronously over the
                                 whole grid (one launch per level, passes of 256
                                 columns = 16384 sources, as the reference's
                                 cudaASPLgraph runs passes of 1024 sources,
                                 cudaASPLgraph.cu:155-160; GG_INFO_VARIANT 4); 0 = one
                                 workgroup per column (GG_INFO_VARIANT 3)            */
#define GG_OPT_SWAP_GRAPH 7   /* adjacency evaluators: how gg_eval_adj_swap runs its
                                 BFS and finalize kernels.  0 (default) = swap and
                                 result copied, kernel-time events recorded; 1 = a
                                 captured HIP graph of the two kernels, which read
                                 the swap from and write the result to pinned host
                                 memory; 2 = the same two kernels launched directly
                                 (no graph).  1 and 2 record no events
                                 (gg_eval_last_kernel_ms is then unavailable),
                                 queue gg_eval_adj_commit without waiting, and
                                 apply where the column bitset fits in LDS          */
#define GG_OPT_LEVELS     8   /* 1 (default) = keep per-source eccentricities and
                                 per-level counts for gg_eval_levels; 0 = skip them
                                 (gg_eval_levels then returns GG_UNSUPPORTED)        */
#define GG_OPT_WVC        9   /* 1 = compute gg_result.wvc (default 1 for part, 0 for
                                 adjacency evaluators; piece and box have no WVC)    */
/* Read-only facts about the most recent launch (gg_eval_get_option). */
#define GG_INFO_VARIANT   100 /* 1 = LDS bitset / 256 threads, 2 = LDS / 1024, 3 = HBM,
                                 4 = HBM level-synchronous                          */
#define GG_INFO_VPT       101 /* vertices per thread of the LDS kernel             */
#define GG_INFO_FIXED     102 /* compile-time degree of the kernel (0 = runtime)     */
#define GG_INFO_PUSH_CAP  103 /* frontier list capacity (0 = pull-only launch)      */
int gg_eval_set_option(gg_eval *ev, int option, int value);
int gg_eval_get_option(gg_eval *ev, int option, int *value);

/* Autotuner: times the kernel variants (workgroup size 256/1024 x
 * fixed-degree / runtime-degree gather) on `table` (the layout of the
 * evaluator's eval call), `reps` launches each, applies the fastest and
 * records it in the cache file named by GG_TUNE_CACHE (default
 * $HOME/.cache/gg_mi355x_tune.txt), keyed by device architecture, evaluator
 * kind, N, M and degree.  Evaluators created later with the same key start
 * from the cached choice.  No reference counterpart. */
typedef struct gg_tune_report {
    int32_t count;            /* variants timed                         */
    int32_t block[4], generic[4], variant[4];
    float ms[4];              /* median kernel time of each variant     */
    int32_t best_block, best_generic;
    float best_ms;
} gg_tune_report;
int gg_eval_autotune(gg_eval *ev, const int32_t *table, int reps, gg_tune_report *report);

/* ---- batched evaluation ----------------------------------------------------
 * K swap candidates against one resident base adjacency (N x degree, explicit
 * neighbours, all N vertices are sources), exact per-candidate totals and
 * diameters from one launch.  No reference counterpart (the reference
 * evaluates one move per host round trip, cppcode/annealing.cpp:397-403). */
int gg_batch_create(int N, int degree, int max_k, int device, gg_batch **out);
void gg_batch_destroy(gg_batch *b);
/* Upload and evaluate the base graph. */
int gg_batch_set_base(gg_batch *b, const int32_t *adj, gg_result *base_out);
/* Evaluate base + swaps[k] for k < K (K <= max_k); the base is unchanged. */
int gg_batch_eval(gg_batch *b, const gg_swap *swaps, int K, gg_result *out);
/* Apply one swap to the base and re-evaluate it.  When the swap is one of
 * the candidates of the last gg_batch_eval and that launch was incremental,
 * its result is already known and the stored base levels are rewritten in
 * place from what the launch kept (no new BFS; see GG_BATCH_OPT_COMMIT). */
int gg_batch_commit(gg_batch *b, const gg_swap *sw, gg_result *base_out);
/* Device time (ms) of the most recent gg_batch_eval launch. */
int gg_batch_last_kernel_ms(gg_batch *b, float *ms);
/* Per-vertex eccentricity of the current base (N entries). */
int gg_batch_base_ecc(gg_batch *b, int32_t *ecc_out);
/* Order-independent 64-bit checksum of the state the incremental path starts
 * from: the stored words of every column below its eccentricity (levels at
 * or above it count as full), the column parts and per-level counts.  Equal
 * for equal bases however they were reached (set_base or commits).  No
 * reference counterpart; for tests and diagnostics. */
int gg_batch_levels_digest(gg_batch *b, uint64_t *digest);

/* Evaluation path: 0 = automatic (incremental when the base allows it),
 * 1 = dense re-evaluation of every candidate, 2 = incremental when possible. */
#define GG_BATCH_OPT_PATH          1
/* 1 = replay the incremental step as one captured HIP graph (default). */
#define GG_BATCH_OPT_GRAPH         2
/* read-only: 1 if the last gg_batch_eval used the incremental path */
#define GG_BATCH_INFO_LAST_SPARSE  100
/* read-only: 1 if the current base allows the incremental path */
#define GG_BATCH_INFO_SPARSE_OK    101
/* Commit of an evaluated candidate: 0 = rewrite the stored levels in place
 * when the last launch kept them (default), 1 = always a new base BFS. */
#define GG_BATCH_OPT_COMMIT        3
/* read-only: 1 if the last commit rewrote the stored levels in place */
#define GG_BATCH_INFO_LAST_COMMIT_INPLACE 102
/* Delta entries a (candidate, column) task may hold before it overflows to the
 * dense fallback queue (0 = the allocated capacity, min(2048, N)); a small
 * value sends most tasks through the queue (used by tests). */
#define GG_BATCH_OPT_DELTA_LIMIT   4
/* Launch shape of the incremental step; 0 restores the automatic value (all
 * candidates the buffers hold in one launch, min(64, that) per workgroup,
 * 1024 threads, as many staged keys as the LDS left over holds).  A shape
 * gg_batch_autotune cached for the device, N, degree and max_k applies at
 * creation.  The reference has no such knob: its launch shape is fixed
 * (cudacode/cudaASPLconv.cu:123-124). */
#define GG_BATCH_OPT_LAUNCH_K      5   /* candidates per launch (<= GG_BATCH_INFO_LAUNCH_MAX) */
#define GG_BATCH_OPT_GROUP_K       6   /* candidates per column workgroup (1..64)            */
#define GG_BATCH_OPT_SPARSE_BLOCK  7   /* threads per workgroup: 512 or 1024 (512: degrees
                                          4, 8, 12, 16, 24 only)                             */
#define GG_BATCH_OPT_STAGED_KEYS   8   /* Delta keys staged in LDS per column and level     */
#define GG_BATCH_OPT_BASE_BLOCK    9   /* dense base BFS threads per workgroup: 256 / 1024  */
#define GG_BATCH_OPT_BASE_PUSH    10   /* dense base BFS push levels: 1 allowed, 0 pull only */
/* read-only: candidates per launch the buffers hold */
#define GG_BATCH_INFO_LAUNCH_MAX   103
/* read-only: staged keys the LDS left over allows */
#define GG_BATCH_INFO_STAGED_MAX   104
int gg_batch_set_option(gg_batch *b, int option, int value);
int gg_batch_get_option(gg_batch *b, int option, int *value);

/* Times the incremental step on the caller's K candidates (the resident base
 * must allow it) over candidates per launch {K, K/2, K/4}, candidates per
 * workgroup {8, 16, 32, 64}, 512 / 1024 threads (fixed degrees) and the
 * staged-key count (all the LDS left over, or what leaves room for two
 * workgroups per CU), and the dense base BFS over 256 / 1024 threads (the
 * vertices per thread follow) with and without push levels.  Every
 * configuration's results must equal the first one's, bit for bit
 * (`mismatches` counts those that did not; they are never chosen).  Applies
 * the fastest step and base configurations and caches them (GG_TUNE_CACHE,
 * default $HOME/.cache/gg_mi355x_tune.txt) under the device architecture, N,
 * degree and max_k.  The base is unchanged.  No reference counterpart. */
#define GG_BATCH_TUNE_MAX 64
typedef struct gg_batch_tune_report {
    int32_t count;                       /* step configurations timed        */
    int32_t launch_k[GG_BATCH_TUNE_MAX], group_k[GG_BATCH_TUNE_MAX];
    int32_t block[GG_BATCH_TUNE_MAX], staged_keys[GG_BATCH_TUNE_MAX];
    float ms[GG_BATCH_TUNE_MAX];         /* median step time (ms)            */
    int32_t mismatches;
    int32_t best_launch_k, best_group_k, best_block, best_staged_keys;
    float best_ms, default_ms;           /* fastest and automatic shapes     */
    int32_t base_count;
    int32_t base_block[4], base_push[4];
    float base_ms[4];                    /* median base BFS kernel time (ms) */
    int32_t best_base_block, best_base_push;
    float best_base_ms;
} gg_batch_tune_report;
int gg_batch_autotune(gg_batch *b, const gg_swap *swaps, int K, int reps,
                      gg_batch_tune_report *report);

/* ---- measurement helper (no reference counterpart) ------------------------
 * LDS read bandwidth of the whole device: 16 workgroups per CU of 1024
 * threads read 64 KiB of LDS each, wave-consecutive (conflict-free) reads of
 * bytes_per_read (4, 8 or 16) bytes.  The ceiling the LDS-resident BFS
 * kernel's roofline is quoted against. */
typedef struct gg_lds_probe {
    double gbs;              /* bytes read per second / 1e9                  */
    double kernel_ms;
    double clock_mhz;        /* shader clock during the run (cycle counter vs wall clock) */
    double bytes_per_clk_cu; /* gbs / (clock * CUs)                          */
    int32_t cus;
} gg_lds_probe;
int gg_probe_lds_read(int device, int bytes_per_read, gg_lds_probe *out);

#ifdef __cplusplus
}
#endif
#endif
