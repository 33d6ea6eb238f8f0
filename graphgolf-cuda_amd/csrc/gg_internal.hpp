// gg_internal.hpp -- host-side types shared by the evaluator translation units.
#pragma once
#include "gg_common.hpp"
#include <string>
#include "gg_dense.hpp"

namespace gg {

enum EvalKind { EVAL_PART = 0, EVAL_PIECE = 1, EVAL_BOX = 2, EVAL_ADJ = 3 };
enum TopoKind { TOPO_RINGQ = 0, TOPO_RING32 = 1, TOPO_PIECE = 2, TOPO_BOX = 3 };

// Kernel-selection state of one evaluator (or batch evaluator).
struct DenseCtx {
    TopoKind topo_kind = TOPO_RINGQ;
    TopoRingQ ringq{};
    TopoRing<uint32_t> ring32{};
    TopoPiece piece{};
    TopoBox box{};
    int lds_device = 65536;   // max LDS bytes per workgroup on the device
    int cu_count = 256;
    int block_opt = 0;        // 0 = auto, else 256 / 1024
    int lds_limit_opt = 0;    // 0 = device limit
    int generic_opt = 0;      // 1 = runtime-degree kernels even when a specialisation exists
    int push_off = 0;         // 1 = pull from level 1 (no push levels)
    int push_ok = 0;          // the graph of this launch is known to be undirected
    int wide_ok = 0;          // level-synchronous path allowed (never under graph capture:
                              // the host reads the level counters between launches)
    int wide_off = 0;         // 1 = persistent global kernel only (GG_OPT_WIDE 0)
    int wide_push_ok = 0;     // level-synchronous path may push (graph verified undirected)
    hipEvent_t bfs_done = nullptr;  // if set, recorded between the BFS and finalize kernels
    int bfs_only = 0;         // 1 = launch the BFS kernel alone (no finalize): timing
    int last_variant = 0;     // 1 = LDS/256, 2 = LDS/1024, 3 = global, 4 = level-synchronous
    int last_vpt = 0;         // vertices per thread of the last LDS launch
    int last_fixed = 0;       // fixed-degree instance of the last launch (0 = generic)
    int last_push_cap = 0;    // frontier list capacity of the last launch (0 = pull only)
    unsigned long long *gscratch = nullptr;
    size_t gscratch_bytes = 0;
    FuseAcc *fuse = nullptr;  // accumulators (zero between launches): single-candidate LDS
                              // launches without ecc / level outputs finalize in the BFS kernel
    int seq = 0;              // sequence number the next fused launch writes with its result
    int last_fused = 0;       // the last launch finalized in its BFS kernel
};

int run_dense(DenseCtx &c, const DenseArgs &a, hipStream_t st);
int run_dense_queue(DenseCtx &c, const DenseArgs &a, int grid, hipStream_t st);
int dense_fits_lds(const DenseCtx &c, int N);
int pick_vpt(int bs, int need);
int device_lds_limit(int device);
int device_cu_count(int device);

// Frontier list capacity of the dense kernels' push levels (0 = pull only)
// for a column bitset of N words within lds_cap bytes of LDS.
int push_capacity(const DenseCtx &c, int N, int D, int lds_cap);

// Undirected-graph checks that gate the push levels (host side).
bool adj_symmetric(const int32_t *adj, int N, int D);
bool swap_keeps_symmetry(const int32_t *adj, int D, const gg_swap &sw);

// Tuning cache (GG_TUNE_CACHE, default $HOME/.cache/gg_mi355x_tune.txt): one
// line per key, "key values".
std::string device_arch(int device);
std::string tune_cache_get(const std::string &key);
void tune_cache_put(const std::string &key, const std::string &values);

// Quad-packed table size / index helpers (TopoRingQ).
inline int quads(int D) { return (D + 3) / 4; }
__host__ __device__ inline size_t quad_index(int slot, int R, int row) {  // uint16 index into the table
    return ((size_t)(slot >> 2) * R + row) * 4 + (slot & 3);
}

}  // namespace gg
