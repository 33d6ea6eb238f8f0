// gg_batch.hpp -- the batch evaluator object (host side).
#pragma once
#include <vector>
#include "gg_internal.hpp"

struct SparseArgs;

struct gg_batch {
    int N = 0, degree = 0, max_k = 0, device = 0, W = 0;
    bool narrow = true;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    gg::DenseCtx dc;
    void *d_tab = nullptr, *h_tab = nullptr;  // transposed offset table (dense kernels)
    size_t tab_bytes = 0;
    int *d_adjR = nullptr;                    // row-major neighbours (sparse kernel)
    int *d_src = nullptr;
    ColPart *d_parts = nullptr;               // [(max_k + 1) * W]; slot max_k = base
    ResultRec *d_res = nullptr, *h_res = nullptr;  // [max_k + 1]
    gg_swap *d_swaps = nullptr, *h_swaps = nullptr;  // [max_k + 1]
    int *d_ecc = nullptr;                     // [N] base per-source eccentricity
    long long *d_base_reach = nullptr;        // [W][LEVEL_LIMIT+1]
    // incremental (sparse) evaluation state
    int snap_cap = 0;                         // base levels whose words are stored
    unsigned long long *d_snap = nullptr;     // [snap_cap][W][N]
    // the same words row-major for the first rows_levels levels, full levels
    // stored as all ones: rows[l-1][x][w] (features read a vertex's ball row)
    unsigned long long *d_rows = nullptr;     // [rows_levels][N][W]
    int rows_levels = 0;
    int cap = 0;                              // Delta list capacity per task
    int skcap = 0;                            // Delta keys staged in LDS per column and level
    int nqctl = 0;                            // d_qctl words: queue count/next + endpoint counters
    int chunk_max = 0;                        // candidates per sparse launch the buffers hold
    int chunk = 0;                            // candidates per sparse launch (<= chunk_max)
    int kg = 0;                               // candidates per column workgroup
    int sparse_block = 1024;                  // threads per sparse workgroup (512 / 1024)
    int skcap_max = 0;                        // staged keys the LDS left over allows
    int *d_dl_key = nullptr;
    unsigned *d_dl_bits = nullptr;            // [chunk_max*W][2][ceil(N/32)]
    unsigned long long *d_dl_ovr = nullptr;   // [chunk_max*W][2][N]
    int *d_task_state = nullptr, *d_task_end = nullptr,
        *d_task_flags = nullptr, *d_task_ecc = nullptr;
    long long *d_dcnt = nullptr;
    int2 *d_queue = nullptr;
    int *d_qctl = nullptr;                    // queue count, queue next, endpoint counters
    ColPart *d_dense_parts = nullptr;         // [max_k * W]
    // kept by the sparse launch for an in-place commit (null: not kept)
    int *d_lg_key = nullptr;                  // [chunk_max*W][snap_cap][cap]
    unsigned long long *d_lg_word = nullptr;  // [chunk_max*W][snap_cap][cap]
    int *d_lg_n = nullptr;                    // [chunk_max*W][snap_cap]
    unsigned long long *d_ep_snap = nullptr;  // [4*chunk_max][snap_cap][N]
    long long *d_ep_reach = nullptr;          // [4*chunk_max][LEVEL_LIMIT+1]
    // single-source BFSes of the swap endpoints (sparse kernels with a tail)
    int *d_ss_cnt = nullptr;                  // [4*max_k][LEVEL_LIMIT+1]
    int *d_ss_ecc = nullptr;                  // [4*max_k]
    int *d_mb_cnt = nullptr;                  // [max_k][W][snap_cap+1]
    int *d_ss_flags = nullptr;                // [4*max_k]
    int *d_oldecc = nullptr;                  // [W] base column eccentricities read by k_sparse
    unsigned *d_ss_vis = nullptr;             // [4*chunk_max][snap_cap][ceil(N/32)] (commit)
    int log_valid = 0;                        // the last eval was sparse with logs kept
    int log_k0 = 0, log_K = 0;                // candidates the logs cover: [log_k0, log_K)
    int ecc_stale = 0;                        // d_ecc predates an in-place commit
    int commit_mode = 0;                      // 0 in place when possible, 1 always a new BFS
    int dlimit_opt = 0;                       // Delta list limit below cap (0 = cap)
    int last_commit_inplace = 0;
    int mode_opt = 0;                         // 0 auto, 1 dense only, 2 sparse when possible
    int sparse_ok = 0;                        // base allows the sparse path
    int last_sparse = 0;                      // last gg_batch_eval used the sparse path
    int use_graph = 1;
    ResultRec base_res{};
    std::vector<int32_t> h_adj;               // host copy of the base adjacency
    bool base_sym = false;                    // the base is undirected
    int have_base = 0, have_timing = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_k = 0;
    SparseArgs *fin_args = nullptr;           // set: enqueue_sparse leaves the finalize to the
    int fin_nvalid = 0;                       // caller, handing over its arguments here
};
