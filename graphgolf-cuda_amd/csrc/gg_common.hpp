// gg_common.hpp -- shared host/device definitions of the MI355X evaluator.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include "graphgolf_mi355x.h"

// Check a HIP call; on failure remember the message and return GG_HIP_ERROR
// (the reference checks no CUDA status at all, cudacode/cudaASPLconv.cu).
void gg_set_last_hip_error(hipError_t e, const char *what, const char *file, int line);
#define GG_HIP(call)                                                        \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess) {                                             \
            gg_set_last_hip_error(e_, #call, __FILE__, __LINE__);           \
            return GG_HIP_ERROR;                                            \
        }                                                                   \
    } while (0)

// Result of one candidate, written on the device by the last workgroup that
// finishes a column of that candidate.
struct ResultRec {
    long long total;    // sum over sources of sum over reached v of d(s,v)
    long long reached;  // (source, vertex) pairs reached, distance 0 included
    int diameter;       // max column eccentricity
    int levels;         // max levels executed by a column
    int wvc;
    int flags;          // bit 0: level limit hit
    int seq;            // fused single evaluations: the launch's sequence number, stored
                        // last (after the other fields are visible system-wide)
    int pad;
};

// Ints of zero padding after the last row of a device adjacency array
// (gg_batch's row-major neighbours): a row read as 32 ints in 16-byte loads
// stays in bounds for every degree up to 32.
#define GG_ADJ_ROW_PAD 32

#define GG_FLAG_LIMIT 1
#define GG_FLAG_QUEUED 2  // a column ran in the dense fallback queue (no stored levels)

// The candidate an in-place commit writes into the stored base, chosen on the
// device: k < 0 commits nothing.  k indexes the last incremental launch.
struct CommitSel {
    int k;
    int pad[3];
};

// Per (candidate, column) partial written by the workgroup that ran it.
struct ColPart {
    long long total;
    long long reached;
    int ecc;
    int levels;
    int flags;
    int pad;
};

// ---- wave64 reductions -------------------------------------------------------
__device__ __forceinline__ unsigned long long wave_or64(unsigned long long x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o, 64);
    return x;
}
__device__ __forceinline__ unsigned wave_sum32(unsigned x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}
__device__ __forceinline__ long long wave_sum64(long long x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}
__device__ __forceinline__ int wave_max32(int x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { int y = __shfl_xor(x, o, 64); x = y > x ? y : x; }
    return x;
}

// Loads that bypass the non-coherent per-CU cache, for data written by other
// workgroups (read after an agent-scope acquire).
template <class T>
__device__ __forceinline__ T ld_agent(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Stores another workgroup reads in the same launch: written through to the
// device-coherent level (no L2 writeback fence), complete after
// __builtin_amdgcn_s_waitcnt(0); the reader drops stale lines of its own
// caches with an agent-scope acquire fence before plain loads.
template <class T>
__device__ __forceinline__ void st_agent(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
