// gg_eval.hip -- single-graph exact evaluators behind the C ABI.
//
// Replaces cudaASPLconv / cudaASPLpiece / cudaASPLbox / cudaASPLgraph
// (cudacode/*.cu).  One call = pack the table on the host, one async H2D, one
// kernel launch (all levels, all columns, final reduction), one 40-byte D2H.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cstring>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>
#include "gg_common.hpp"
#include "gg_dense.hpp"
#include "gg_internal.hpp"
#include "gg_fixed_impl.hpp"

// ---- error bookkeeping -----------------------------------------------------------
static thread_local std::string g_last_hip_error;

void gg_set_last_hip_error(hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e),
             what, file, line);
    g_last_hip_error = buf;
}

extern "C" const char *gg_strerror(int status) {
    switch (status) {
        case GG_OK: return "ok";
        case GG_UNCONNECTED: return "graph is unconnected";
        case GG_DIAMETER_LIMIT: return "diameter reaches the level limit (99)";
        case GG_BAD_ARG: return "bad argument";
        case GG_HIP_ERROR: return g_last_hip_error.empty() ? "HIP error" : g_last_hip_error.c_str();
        case GG_OOM: return "out of device memory";
        case GG_NO_DEVICE: return "no HIP device";
        case GG_UNSUPPORTED: return "unsupported shape";
        case GG_SCRATCH_LIMIT: return "input exceeds the fixed scratch of the no-allocation evaluator";
        default: return "unknown status";
    }
}

extern "C" int gg_device_count(int *count) {
    if (!count) return GG_BAD_ARG;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) { *count = 0; return GG_NO_DEVICE; }
    *count = n;
    return GG_OK;
}

extern "C" double gg_aspl(const gg_result *r, int N) {
    if (!r || N < 2) return 0.0;
    return double(r->total_scaled) / (int64_t(N) * (N - 1));
}

// ---- kernel dispatch ------------------------------------------------------------------
namespace gg {

hipError_t ensure_dynamic_lds(const void *fn, size_t lds, LdsAttr &attr) {
    static std::mutex mu;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)lds <= attr.bytes[dev]) return hipSuccess;
    e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) attr.bytes[dev] = (int)lds;
    return e;
}

template <class Topo, int BS, int VPT, int DT>
static hipError_t launch_lds_inst(const Topo &topo, const DenseArgs &a, int grid, size_t lds,
                                  hipStream_t st) {
    static LdsAttr attr;
    auto fn = &k_dense_lds<Topo, BS, VPT, DT>;
    hipError_t e = ensure_dynamic_lds((const void *)fn, lds, attr);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(BS), lds, st, topo, a);
    return hipGetLastError();
}

template <class Topo>
static hipError_t launch_lds(const Topo &topo, const DenseArgs &a, int grid, size_t lds, int bs,
                             int vpt, hipStream_t st) {
#define GG_CASE(B, V) \
    if (bs == B && vpt == V) return launch_lds_inst<Topo, B, V, 0>(topo, a, grid, lds, st);
    GG_CASE(256, 1) GG_CASE(256, 2) GG_CASE(256, 4) GG_CASE(256, 8) GG_CASE(256, 16)
    GG_CASE(1024, 1) GG_CASE(1024, 2) GG_CASE(1024, 4) GG_CASE(1024, 8) GG_CASE(1024, 16)
    GG_CASE(1024, 20)
#undef GG_CASE
    return hipErrorInvalidValue;
}

// Fixed-degree kernels exist for the 16-bit ring topology, 1024-thread
// workgroups and these degrees (gg_fixed_d*.hip).
static hipError_t launch_fixed(const TopoRingQ &topo, const DenseArgs &a, int grid,
                               size_t lds, int vpt, bool queue, hipStream_t st, bool *done) {
    *done = true;
    switch (a.D) {
        case 4: return launch_fixed_d4(topo, a, grid, lds, vpt, queue, st);
        case 8: return launch_fixed_d8(topo, a, grid, lds, vpt, queue, st);
        case 12: return launch_fixed_d12(topo, a, grid, lds, vpt, queue, st);
        case 16: return launch_fixed_d16(topo, a, grid, lds, vpt, queue, st);
        case 24: return launch_fixed_d24(topo, a, grid, lds, vpt, queue, st);
        default: *done = false; return hipSuccess;
    }
}

template <class Topo, int BS, int VPT, int DT>
static hipError_t launch_lds_queue_inst(const Topo &topo, const DenseArgs &a, int grid, size_t lds,
                                        hipStream_t st) {
    static LdsAttr attr;
    auto fn = &k_dense_lds_queue<Topo, BS, VPT, DT>;
    hipError_t e = ensure_dynamic_lds((const void *)fn, lds, attr);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(BS), lds, st, topo, a);
    return hipGetLastError();
}

template <class Topo>
static hipError_t launch_lds_queue(const Topo &topo, const DenseArgs &a, int grid, size_t lds,
                                   int vpt, hipStream_t st) {
#define GG_CASE(V) \
    if (vpt == V) return launch_lds_queue_inst<Topo, 1024, V, 0>(topo, a, grid, lds, st);
    GG_CASE(1) GG_CASE(2) GG_CASE(4) GG_CASE(8) GG_CASE(16) GG_CASE(20)
#undef GG_CASE
    return hipErrorInvalidValue;
}

template <class Topo>
static hipError_t launch_global(const Topo &topo, const DenseArgs &a, int grid, int tasks,
                                hipStream_t st) {
    hipLaunchKernelGGL(k_dense_global<Topo>, dim3(grid), dim3(1024), 0, st, topo, a, tasks);
    return hipGetLastError();
}

// ---- level-synchronous BFS over the whole grid (N too large for LDS, few columns) ----
// The persistent kernel above runs each column on one workgroup, so with a
// handful of columns (a symmetric graph of 10^6 vertices has M <= 128
// representatives: two columns) all but a few CUs idle.  Here one launch is
// one level of every column: thread (v, w) owns word w of vertex v, the words
// of a vertex are contiguous ([N][Wp], Wp = the power of two >= W), so the
// lanes reading one neighbour's words share its cache lines.  Per-level
// new-pair counts and new-bit masks per column go to cnt/lor [lmax+1][Wp].
struct WideArgs {
    int N, M, D, W, wsh, lmax;          // Wp = 1 << wsh
    const int *src_vertex;              // [M]
    const gg_swap *swaps;               // [1] or nullptr
    unsigned long long *cnt, *lor;      // [lmax+1][Wp]
    // push levels (undirected graphs only; fcap = 0: pull from level 1).
    // Frontier lists [2][fcap] of (word index v*Wp + w, new bits); list
    // (l & 1) holds what level l added, list 0 starts with the sources.
    // ctl[2]: entries appended to each list (may exceed fcap: overflow).
    // mode[l] = 1: level l pushes from list (l-1) & 1; a level that pulls, or
    // whose push overflows its list, sets mode[l+1] = 0, so pulling persists.
    int fcap;
    unsigned *fv;
    unsigned long long *fb;
    int *ctl, *mode;
    // vertex order of the words: v itself (pcopies = 0), or for a table of
    // 2^pshift rows repeated pcopies = N >> pshift times, vertex v = c*R + r
    // (copy c of row r) at position r*pcopies + c.  A row's offsets are the
    // same in every copy, so the lanes holding consecutive copies of one row
    // gather consecutive words of a neighbour row.
    int pshift;
    unsigned pcopies;
};

__device__ __forceinline__ size_t wide_pos(const WideArgs &a, unsigned v) {
    if (!a.pcopies) return v;
    return (size_t)(v & ((1u << a.pshift) - 1u)) * a.pcopies + (v >> a.pshift);
}

__device__ __forceinline__ unsigned wide_vertex(const WideArgs &a, size_t p) {
    if (!a.pcopies) return (unsigned)p;
    const unsigned r = (unsigned)(p / a.pcopies);
    const unsigned c = (unsigned)(p - (size_t)r * a.pcopies);
    return (c << a.pshift) + r;
}

// cur := padding bits (columns >= W: all ones), counters of levels 1..lmax
// cleared, level 0 holds each column's valid source bits and their count.
__global__ void __launch_bounds__(256) k_wide_init(WideArgs a, unsigned long long *cur) {
    const size_t Wp = (size_t)1 << a.wsh;
    const size_t total = (size_t)a.N << a.wsh;
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < total; t += stride) {
        const int w = (int)(t & (Wp - 1));
        cur[t] = w < a.W ? pad_mask(w, a.M) : ~0ull;
    }
    const size_t nc = (size_t)(a.lmax + 1) * Wp;
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < nc; t += stride) {
        const int w = (int)(t & (Wp - 1));
        const unsigned long long src = w < a.W ? ~pad_mask(w, a.M) : 0ull;
        const bool l0 = t < Wp;
        a.cnt[t] = l0 ? (unsigned long long)__popcll(src) : 0ull;
        a.lor[t] = l0 ? src : 0ull;
    }
    if (a.fcap) {
        const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
        if (t < (size_t)a.lmax + 2) a.mode[t] = t >= 1 ? 1 : 0;
        if (t < 2) a.ctl[t] = t == 0 ? a.M : 0;
    }
}

// Source s sets bit s & 63 of word s >> 6 of its vertex (after k_wide_init).
__global__ void __launch_bounds__(256) k_wide_sources(WideArgs a, unsigned long long *cur) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= a.M) return;
    const size_t i = (wide_pos(a, (unsigned)a.src_vertex[s]) << a.wsh) + (size_t)(s >> 6);
    cur[i] |= 1ull << (s & 63);
    if (a.fcap) {
        a.fv[s] = (unsigned)i;
        a.fb[s] = 1ull << (s & 63);
    }
}

// Push level: nxt := cur (the push ORs into it), and the list this level
// fills starts empty.
__global__ void __launch_bounds__(256) k_wide_copy(WideArgs a, int level,
                                                   const unsigned long long *cur,
                                                   unsigned long long *nxt) {
    if (!a.fcap || !a.mode[level]) return;
    const size_t total = (size_t)a.N << a.wsh;
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < total; t += stride) nxt[t] = cur[t];
    if (blockIdx.x == 0 && threadIdx.x == 0) a.ctl[level & 1] = 0;
}

// Push level `level`: every (entry, neighbour) item ORs the entry's new bits
// into the neighbour's word of the same column.  The old word returned by
// the atomic tells which bits this item set first, so counts are exact and
// each bit new at this level is appended (wave-aggregated) exactly once.
// Equal to the pull level only on undirected graphs (pull reads listed
// neighbours, push follows listed edges), which the host checks.
template <class Topo>
__global__ void __launch_bounds__(256) k_wide_push(Topo topo, WideArgs a, int level,
                                                   unsigned long long *nxt) {
    if (!a.fcap || !a.mode[level]) return;  // uniform
    __shared__ unsigned long long s_cnt[256], s_lor[256];
    const int tid = threadIdx.x, lane = tid & 63;
    const int Wp = 1 << a.wsh;
    if (tid < Wp) { s_cnt[tid] = 0; s_lor[tid] = 0; }
    __syncthreads();
    SwapRegs sw;
    bool has_sw;
    load_swap(a.swaps, 0, sw, has_sw);
    const int pin = (level - 1) & 1, pout = level & 1;
    const unsigned *fv_in = a.fv + (size_t)pin * a.fcap;
    const unsigned long long *fb_in = a.fb + (size_t)pin * a.fcap;
    unsigned *fv_out = a.fv + (size_t)pout * a.fcap;
    unsigned long long *fb_out = a.fb + (size_t)pout * a.fcap;
    const size_t items = (size_t)a.ctl[pin] * a.D;
    const size_t nthr = (size_t)gridDim.x * 256;
    bool over = false;
    // wave-uniform trip count: the ballot below sees all 64 lanes
    for (size_t base = (size_t)blockIdx.x * 256 + (size_t)(tid & ~63); base < items; base += nthr) {
        const size_t it = base + (size_t)lane;
        unsigned long long nb = 0;
        unsigned ui = 0;
        if (it < items) {
            const size_t e = it / (size_t)a.D;
            const int j = (int)(it - e * (size_t)a.D);
            const unsigned idx = fv_in[e];
            const unsigned long long bits = fb_in[e];
            const unsigned v = wide_vertex(a, idx >> a.wsh);
            const int w = (int)(idx & (unsigned)(Wp - 1));
            const unsigned u = nbr_patched(topo, v, j, sw, has_sw);
            ui = (unsigned)(wide_pos(a, u) << a.wsh) + (unsigned)w;
            const unsigned long long old = atomicOr(&nxt[ui], bits);
            nb = bits & ~old;
            if (nb) {
                atomicAdd(&s_cnt[w], (unsigned long long)__popcll(nb));
                atomicOr(&s_lor[w], nb);
            }
        }
        const unsigned long long m = __ballot(nb != 0);
        if (m) {
            const int first = __builtin_ctzll(m);
            int pos0 = 0;
            if (lane == first) pos0 = atomicAdd(&a.ctl[pout], __popcll(m));
            pos0 = __shfl(pos0, first, 64);
            if (nb) {
                const int pos = pos0 + __popcll(m & ((1ull << lane) - 1ull));
                if (pos < a.fcap) {
                    fv_out[pos] = ui;
                    fb_out[pos] = nb;
                } else {
                    over = true;
                }
            }
        }
    }
    if (over) a.mode[level + 1] = 0;  // the next level pulls (benign race: all store 0)
    __syncthreads();
    if (tid < Wp && tid < a.W && s_cnt[tid]) {
        atomicAdd(&a.cnt[(size_t)level * Wp + tid], s_cnt[tid]);
        atomicOr(&a.lor[(size_t)level * Wp + tid], s_lor[tid]);
    }
}

// Level `level` of every column: nxt = cur | (OR of the listed neighbours'
// words & ~cur).  A column with no new bits at the previous level is closed
// and skipped (its counters stay zero); a launch after every column closed
// returns at once.  Each thread keeps one column w and strides over vertices,
// so counts are summed in registers, then over the lanes and waves holding
// the same w, then added to the level's counters (one atomic per column and
// workgroup).
template <class Topo>
__global__ void __launch_bounds__(256) k_wide_level(Topo topo, WideArgs a, int level,
                                                    const unsigned long long *cur,
                                                    unsigned long long *nxt) {
    if (a.fcap && a.mode[level]) return;  // a push level (uniform)
    __shared__ unsigned long long s_cnt[256], s_lor[256];
    __shared__ int s_live;
    const int tid = threadIdx.x, lane = tid & 63;
    if (a.fcap && blockIdx.x == 0 && tid == 0) a.mode[level + 1] = 0;  // pulling persists
    const int Wp = 1 << a.wsh;
    const size_t t0 = (size_t)blockIdx.x * 256 + tid;
    const int w = (int)(t0 & (size_t)(Wp - 1));
    const unsigned long long *prev_lor = a.lor + (size_t)(level - 1) * Wp;
    if (tid == 0) s_live = 0;
    if (tid < Wp) { s_cnt[tid] = 0; s_lor[tid] = 0; }
    __syncthreads();
    for (int i = tid; i < a.W; i += 256)
        if (prev_lor[i]) s_live = 1;  // benign race: every writer stores 1
    __syncthreads();
    if (!s_live) return;  // uniform: every column closed
    SwapRegs sw;
    bool has_sw;
    load_swap(a.swaps, 0, sw, has_sw);
    const bool live = w < a.W && prev_lor[w] != 0;
    unsigned long long cnt = 0, lor = 0;
    if (live) {
        const size_t vstride = ((size_t)gridDim.x * 256) >> a.wsh;
        for (size_t p = t0 >> a.wsh; p < (size_t)a.N; p += vstride) {
            const unsigned v = wide_vertex(a, p);
            const size_t i = (p << a.wsh) + (size_t)w;
            const unsigned long long own = cur[i];
            unsigned long long n = 0;
            if (~own) {
                // neighbours in groups of eight loads issued together (slots
                // past the degree repeat the last one); stop once every
                // missing bit has been found (the OR cannot add more)
                const unsigned long long need = ~own;
                unsigned long long acc = 0;
                for (int j0 = 0; j0 < a.D; j0 += 8) {
                    unsigned long long x[8];
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        const int j = j0 + k < a.D ? j0 + k : a.D - 1;
                        const unsigned u = nbr_patched(topo, v, j, sw, has_sw);
                        x[k] = cur[(wide_pos(a, u) << a.wsh) + (size_t)w];
                    }
#pragma unroll
                    for (int k = 0; k < 8; k++) acc |= x[k];
                    if ((acc & need) == need) break;
                }
                n = acc & need;
            }
            nxt[i] = own | n;
            cnt += (unsigned long long)__popcll(n);
            lor |= n;
        }
    }
    // lanes l and l ^ o (o >= Wp) hold the same column
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        if (o < Wp) break;
        cnt += __shfl_xor(cnt, o, 64);
        lor |= __shfl_xor(lor, o, 64);
    }
    if (live && lane < Wp && (cnt || lor)) {
        atomicAdd(&s_cnt[w], cnt);
        atomicOr(&s_lor[w], lor);
    }
    __syncthreads();
    if (tid < Wp && tid < a.W && s_cnt[tid]) {
        atomicAdd(&a.cnt[(size_t)level * Wp + tid], s_cnt[tid]);
        atomicOr(&a.lor[(size_t)level * Wp + tid], s_lor[tid]);
    }
}

// Column parts (and per-source eccentricities / per-level counts) from the
// counters of levels 1..lexec (the levels launched), in the form the
// persistent kernels write them.
__global__ void __launch_bounds__(256) k_wide_parts(WideArgs a, int lexec, ColPart *parts,
                                                    int *ecc_out, long long *reach_col) {
    const int Wp = 1 << a.wsh;
    for (int w = threadIdx.x; w < a.W; w += 256) {
        long long total = 0, reached = (long long)a.cnt[w];
        int ecc = 0;
        for (int l = 1; l <= a.lmax; l++) {
            const long long c = l <= lexec ? (long long)a.cnt[(size_t)l * Wp + w] : 0;
            total += (long long)l * c;
            reached += c;
            if (c) ecc = l;
            if (reach_col) reach_col[(size_t)w * (a.lmax + 1) + l] = c;
        }
        if (reach_col) reach_col[(size_t)w * (a.lmax + 1)] = (long long)a.cnt[w];
        ColPart p;
        p.total = total;
        p.reached = reached;
        p.ecc = ecc;
        p.levels = ecc < a.lmax ? ecc + 1 : a.lmax;
        p.flags = ecc >= a.lmax ? GG_FLAG_LIMIT : 0;
        p.pad = 0;
        parts[w] = p;
    }
    if (ecc_out) {
        for (int s = threadIdx.x; s < a.M; s += 256) {
            const int w = s >> 6;
            const unsigned long long bit = 1ull << (s & 63);
            int e = 0;
            for (int l = lexec; l >= 1; l--)
                if (a.lor[(size_t)l * Wp + w] & bit) { e = l; break; }
            ecc_out[s] = e;
        }
    }
}

// Level-synchronous path: at most 256 words per vertex in one pass.
static const int kWideMaxCols = 256;
static const int kWideChunk = 4;  // levels launched between reads of the counters

// Grows the evaluator's global scratch to `need` bytes.
static int ensure_gscratch(DenseCtx &c, size_t need) {
    if (need <= c.gscratch_bytes) return GG_OK;
    if (c.gscratch) { (void)hipFree(c.gscratch); c.gscratch = nullptr; c.gscratch_bytes = 0; }
    if (hipMalloc((void **)&c.gscratch, need) != hipSuccess) return GG_OOM;
    c.gscratch_bytes = need;
    return GG_OK;
}

// Rows of a ring table repeated over the vertices (0: other topologies).
template <class Topo> static unsigned topo_rows(const Topo &) { return 0; }
template <> unsigned topo_rows(const TopoRingQ &t) { return t.R; }
template <> unsigned topo_rows(const TopoRing<uint32_t> &t) { return t.R; }

// One launch per level over all (vertex, column) words; between chunks of
// kWideChunk levels the host reads the last level's new-bit masks and stops
// once every column is closed.  Words [N][Wp], then the level counters.
template <class Topo>
static int run_wide(DenseCtx &c, const DenseArgs &a, const Topo &topo, hipStream_t st) {
    int wsh = 0;
    while ((1 << wsh) < a.W) wsh++;
    const size_t Wp = (size_t)1 << wsh;
    const size_t words = (size_t)a.N * Wp, nctr = (size_t)(a.lmax + 1) * Wp;
    // push levels: a level pushes while the level before it fit its list
    // (fcap entries); the first level whose list overflows makes every later
    // level pull.  fcap = M*D (at most N*Wp/8, at least M): level 1 adds at
    // most M*D entries, so on expander graphs level 2 (M*D*D device-scope
    // atomics: the workgroups span every XCD) is the one that overflows and
    // pulling starts at level 3, where the atomics would cost more than the
    // pull gather (N = 2^18, d = 16: 0.38 ms pulling only, 0.58 ms pushing
    // with lists of N*Wp/8 entries); graphs with slowly growing frontiers
    // (long cycles with few chords) keep pushing
    const bool push = c.wide_push_ok && !c.push_off && words < (1ull << 31);
    const size_t fcap = push ? std::max((size_t)a.M,
                                        std::min(words / 8, (size_t)a.M * (size_t)a.D)) : 0;
    const size_t ints = 2 * fcap + 2 + (size_t)a.lmax + 2;
    int rc = ensure_gscratch(c, (2 * words + 2 * nctr + 2 * fcap) * sizeof(unsigned long long) +
                                    ints * sizeof(int));
    if (rc != GG_OK) return rc;
    unsigned long long *cur = c.gscratch, *nxt = cur + words;
    WideArgs wa;
    wa.N = a.N; wa.M = a.M; wa.D = a.D; wa.W = a.W; wa.wsh = wsh; wa.lmax = a.lmax;
    wa.src_vertex = a.src_vertex;
    wa.swaps = a.swaps;
    wa.cnt = nxt + words;
    wa.lor = wa.cnt + nctr;
    wa.fcap = (int)fcap;
    wa.fb = wa.lor + nctr;
    wa.fv = reinterpret_cast<unsigned *>(wa.fb + 2 * fcap);
    wa.ctl = reinterpret_cast<int *>(wa.fv + 2 * fcap);
    wa.mode = wa.ctl + 2;
    wa.pshift = 0;
    wa.pcopies = 0;
    const unsigned R = topo_rows(topo);
    if (R >= 1 && R < (unsigned)a.N && (R & (R - 1)) == 0 && a.N % R == 0) {
        while ((1u << wa.pshift) < R) wa.pshift++;
        wa.pcopies = (unsigned)a.N / R;
    }
    c.last_push_cap = (int)fcap;
    const size_t want = (words + 255) / 256;
    const int grid = (int)(want < (size_t)8 * c.cu_count ? want : (size_t)8 * c.cu_count);
    hipLaunchKernelGGL(k_wide_init, dim3(grid), dim3(256), 0, st, wa, cur);
    hipLaunchKernelGGL(k_wide_sources, dim3((a.M + 255) / 256), dim3(256), 0, st, wa, cur);
    std::vector<unsigned long long> h(a.W);
    int level = 1;
    while (level <= a.lmax) {
        // the first read of the counters after 2 * kWideChunk levels (most
        // graphs this path serves close by then), then every kWideChunk
        const int chunk = level == 1 ? 2 * kWideChunk : kWideChunk;
        const int end = level + chunk - 1 < a.lmax ? level + chunk - 1 : a.lmax;
        for (; level <= end; level++) {
            if (fcap) {
                hipLaunchKernelGGL(k_wide_copy, dim3(grid), dim3(256), 0, st, wa, level,
                                   (const unsigned long long *)cur, nxt);
                hipLaunchKernelGGL(k_wide_push<Topo>, dim3(grid), dim3(256), 0, st, topo, wa, level,
                                   nxt);
            }
            hipLaunchKernelGGL(k_wide_level<Topo>, dim3(grid), dim3(256), 0, st, topo, wa, level,
                               (const unsigned long long *)cur, nxt);
            unsigned long long *t = cur; cur = nxt; nxt = t;
        }
        if (level > a.lmax) break;
        if (hipMemcpyAsync(h.data(), wa.lor + (size_t)(level - 1) * Wp, sizeof(h[0]) * a.W,
                           hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            gg_set_last_hip_error(hipGetLastError(), "level counters", __FILE__, __LINE__);
            return GG_HIP_ERROR;
        }
        bool any = false;
        for (int w = 0; w < a.W; w++) any = any || h[w] != 0;
        if (!any) break;
    }
    const int lexec = level - 1 < a.lmax ? level - 1 : a.lmax;  // last level launched
    hipLaunchKernelGGL(k_wide_parts, dim3(1), dim3(256), 0, st, wa, lexec, a.parts, a.ecc_out,
                       a.reach_col);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        gg_set_last_hip_error(e, "level-synchronous BFS launch", __FILE__, __LINE__);
        return GG_HIP_ERROR;
    }
    return GG_OK;
}

// Smallest supported VPT >= need for block size bs (0 if none).
int pick_vpt(int bs, int need) {
    static const int v256[] = {1, 2, 4, 8, 16};
    static const int v1024[] = {1, 2, 4, 8, 16, 20};
    const int *v = bs == 256 ? v256 : v1024;
    int n = bs == 256 ? 5 : 6;
    for (int i = 0; i < n; i++)
        if (v[i] >= need) return v[i];
    return 0;
}

int device_lds_limit(int device) {
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
        lds = 65536;
    return lds;
}

int device_cu_count(int device) {
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        cu = 256;
    return cu;
}

// Static LDS used by the dense kernels besides the bitset (reduction scratch).
static const int kStaticLdsReserve = 2048;

// Level-1 frontier list capacity for the push levels (0 = pull only): the
// list follows the column bitset in LDS and packs (vertex, bit) in 17+6 bits.
int push_capacity(const DenseCtx &c, int N, int D, int lds_cap) {
    if (c.push_off || !c.push_ok || N > (1 << 17)) return 0;
    long long room = ((long long)lds_cap - (long long)N * 8) / 4;
    long long cap = 64LL * D;
    if (room < cap) cap = room;
    return cap >= 64 ? (int)cap : 0;
}

// Chooses the kernel variant and launches one dense evaluation of K candidates.
int run_dense(DenseCtx &c, const DenseArgs &a_in, hipStream_t st) {
    DenseArgs a = a_in;
    const int tasks = a.K * a.W;
    int lds_cap = c.lds_limit_opt > 0 ? c.lds_limit_opt : c.lds_device - kStaticLdsReserve;
    a.push_cap = push_capacity(c, a.N, a.D, lds_cap);
    const size_t lds = (size_t)a.N * 8 + (size_t)a.push_cap * 4;
    int bs = c.block_opt;
    if (bs != 256 && bs != 1024) bs = (a.N <= 4096 && tasks >= 4 * c.cu_count) ? 256 : 1024;
    int vpt = pick_vpt(bs, (a.N + bs - 1) / bs);
    if (vpt == 0 && bs == 256) { bs = 1024; vpt = pick_vpt(bs, (a.N + bs - 1) / bs); }
    hipError_t e;
    if ((size_t)a.N * 8 <= (size_t)lds_cap && vpt > 0) {
        // one candidate: its last column workgroup reduces the parts (no
        // separate finalize launch); timing launches keep the BFS alone
        if (c.fuse && a.K == 1 && !c.bfs_only && !a.ecc_out && !a.reach_col && !a.reach_out &&
            !a.snap) {
            a.fuse = c.fuse;
            a.seq = c.seq;
        }
        c.last_variant = bs == 256 ? 1 : 2;
        c.last_vpt = vpt;
        c.last_fixed = 0;
        c.last_push_cap = a.push_cap;
        switch (c.topo_kind) {
            case TOPO_RINGQ: {
                bool done = false;
                e = hipSuccess;
                if (bs == 1024 && !c.generic_opt && c.ringq.R == c.ringq.N)
                    e = launch_fixed(c.ringq, a, tasks, lds, vpt, false, st, &done);
                if (done) c.last_fixed = a.D;
                else e = launch_lds(c.ringq, a, tasks, lds, bs, vpt, st);
                break;
            }
            case TOPO_RING32: e = launch_lds(c.ring32, a, tasks, lds, bs, vpt, st); break;
            case TOPO_PIECE: e = launch_lds(c.piece, a, tasks, lds, bs, vpt, st); break;
            default: e = launch_lds(c.box, a, tasks, lds, bs, vpt, st); break;
        }
    } else if (c.wide_ok && !c.wide_off && a.K == 1 && !a.queue && !a.snap) {
        // a large graph: every level spread over the whole grid, the words
        // of up to kWideMaxCols columns of a vertex side by side (so the
        // lanes gathering one neighbour read its consecutive words); more
        // columns run as passes of kWideMaxCols, each writing its own parts
        c.last_variant = 4;
        c.last_vpt = 0;
        c.last_fixed = 0;
        c.last_push_cap = 0;  // run_wide sets its frontier capacity
        for (int w0 = 0; w0 < a.W; w0 += kWideMaxCols) {
            DenseArgs p = a;
            p.W = a.W - w0 < kWideMaxCols ? a.W - w0 : kWideMaxCols;
            p.M = a.M - 64 * w0 < 64 * p.W ? a.M - 64 * w0 : 64 * p.W;
            p.src_vertex = a.src_vertex + (size_t)64 * w0;
            p.parts = a.parts + w0;
            p.ecc_out = a.ecc_out ? a.ecc_out + (size_t)64 * w0 : nullptr;
            p.reach_col = a.reach_col ? a.reach_col + (size_t)w0 * (a.lmax + 1) : nullptr;
            int rc;
            switch (c.topo_kind) {
                case TOPO_RINGQ: rc = run_wide(c, p, c.ringq, st); break;
                case TOPO_RING32: rc = run_wide(c, p, c.ring32, st); break;
                case TOPO_PIECE: rc = run_wide(c, p, c.piece, st); break;
                default: rc = run_wide(c, p, c.box, st); break;
            }
            if (rc != GG_OK) return rc;
        }
        e = hipSuccess;
    } else {
        // global-memory variant: persistent grid, 2*N words of scratch per workgroup
        int grid = tasks < 2 * c.cu_count ? tasks : 2 * c.cu_count;
        size_t need = (size_t)grid * 2 * a.N * sizeof(unsigned long long);
        int rc = ensure_gscratch(c, need);
        if (rc != GG_OK) return rc;
        DenseArgs g = a;
        g.gscratch = c.gscratch;
        g.push_cap = 0;
        c.last_variant = 3;
        c.last_vpt = 0;
        c.last_fixed = 0;
        c.last_push_cap = 0;
        switch (c.topo_kind) {
            case TOPO_RINGQ: e = launch_global(c.ringq, g, grid, tasks, st); break;
            case TOPO_RING32: e = launch_global(c.ring32, g, grid, tasks, st); break;
            case TOPO_PIECE: e = launch_global(c.piece, g, grid, tasks, st); break;
            default: e = launch_global(c.box, g, grid, tasks, st); break;
        }
    }
    if (e != hipSuccess) {
        gg_set_last_hip_error(e, "dense BFS launch", __FILE__, __LINE__);
        return GG_HIP_ERROR;
    }
    if (c.bfs_done) {
        e = hipEventRecord(c.bfs_done, st);
        if (e != hipSuccess) {
            gg_set_last_hip_error(e, "event record", __FILE__, __LINE__);
            return GG_HIP_ERROR;
        }
    }
    c.last_fused = a.fuse ? 1 : 0;
    if (c.bfs_only || a.fuse) return GG_OK;
    hipLaunchKernelGGL(k_dense_finalize<1024>, dim3(a.K), dim3(1024), 0, st, a);
    e = hipGetLastError();
    if (e != hipSuccess) {
        gg_set_last_hip_error(e, "dense finalize launch", __FILE__, __LINE__);
        return GG_HIP_ERROR;
    }
    return GG_OK;
}

// Dense column BFS of the tasks in a device-side queue (ring topologies only;
// 1024-thread workgroups, bitset in LDS).  Returns GG_UNSUPPORTED when the
// column does not fit in LDS.
int run_dense_queue(DenseCtx &c, const DenseArgs &a_in, int grid, hipStream_t st) {
    DenseArgs a = a_in;
    int lds_cap = c.lds_limit_opt > 0 ? c.lds_limit_opt : c.lds_device - kStaticLdsReserve;
    a.push_cap = push_capacity(c, a.N, a.D, lds_cap);
    const size_t lds = (size_t)a.N * 8 + (size_t)a.push_cap * 4;
    int vpt = pick_vpt(1024, (a.N + 1023) / 1024);
    if ((size_t)a.N * 8 > (size_t)lds_cap || vpt == 0) return GG_UNSUPPORTED;
    hipError_t e;
    switch (c.topo_kind) {
        case TOPO_RINGQ: {
            bool done = false;
            e = hipSuccess;
            if (!c.generic_opt && c.ringq.R == c.ringq.N)
                e = launch_fixed(c.ringq, a, grid, lds, vpt, true, st, &done);
            if (!done) e = launch_lds_queue(c.ringq, a, grid, lds, vpt, st);
            break;
        }
        case TOPO_RING32: e = launch_lds_queue(c.ring32, a, grid, lds, vpt, st); break;
        default: return GG_UNSUPPORTED;
    }
    if (e != hipSuccess) {
        gg_set_last_hip_error(e, "dense queue launch", __FILE__, __LINE__);
        return GG_HIP_ERROR;
    }
    return GG_OK;
}

int dense_fits_lds(const DenseCtx &c, int N) {
    int lds_cap = c.lds_limit_opt > 0 ? c.lds_limit_opt : c.lds_device - kStaticLdsReserve;
    return (size_t)N * 8 <= (size_t)lds_cap && pick_vpt(1024, (N + 1023) / 1024) > 0;
}

// Adjacency (entries < 0 = self-loop padding) is undirected: the directed
// edge multiset equals its reverse.
bool adj_symmetric(const int32_t *adj, int N, int D) {
    std::vector<unsigned long long> fwd((size_t)N * D), rev((size_t)N * D);
    for (int v = 0; v < N; v++)
        for (int i = 0; i < D; i++) {
            long long u = adj[(size_t)v * D + i];
            if (u < 0) u = v;
            if (u >= N) return false;
            fwd[(size_t)v * D + i] = ((unsigned long long)v << 32) | (unsigned long long)u;
            rev[(size_t)v * D + i] = ((unsigned long long)u << 32) | (unsigned long long)v;
        }
    std::sort(fwd.begin(), fwd.end());
    std::sort(rev.begin(), rev.end());
    return fwd == rev;
}

// Sufficient condition for a swap to keep an undirected graph undirected:
// distinct (row, slot) entries whose removed and added directed edges are
// each closed under reversal.
bool swap_keeps_symmetry(const int32_t *adj, int D, const gg_swap &sw) {
    std::vector<std::pair<int, int>> rem, add;
    for (int j = 0; j < 4; j++) {
        if (sw.v[j] < 0) continue;
        for (int i = 0; i < j; i++)
            if (sw.v[i] == sw.v[j] && sw.slot[i] == sw.slot[j]) return false;
        int old = adj[(size_t)sw.v[j] * D + sw.slot[j]];
        if (old < 0) old = sw.v[j];
        rem.emplace_back(sw.v[j], old);
        add.emplace_back(sw.v[j], sw.nbr[j]);
    }
    for (auto *e : {&rem, &add}) {
        std::vector<std::pair<int, int>> r;
        for (auto &p : *e) r.emplace_back(p.second, p.first);
        std::vector<std::pair<int, int>> f = *e;
        std::sort(f.begin(), f.end());
        std::sort(r.begin(), r.end());
        if (f != r) return false;
    }
    return true;
}

}  // namespace gg

using namespace gg;

// ---- evaluator object ------------------------------------------------------------------
struct gg_eval {
    EvalKind kind;
    int N, M, degree;
    int Nx, Ny, Nz, Mx, My, Mz;
    int device;
    int W;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DenseArgs last_a{};  // launch arguments of the last evaluation's BFS
    DenseCtx dc;
    void *d_tab = nullptr;
    size_t tab_bytes = 0;
    void *h_tab = nullptr;  // pinned staging
    int *d_src = nullptr;
    ColPart *d_parts = nullptr;
    ResultRec *d_res = nullptr;
    ResultRec *h_res = nullptr;  // pinned
    int *d_ecc = nullptr;
    long long *d_reach_col = nullptr;
    long long *d_reach = nullptr;
    gg_swap *d_swap = nullptr;
    gg_swap *h_swap = nullptr;  // pinned
    gg_swap *h_commit = nullptr;  // pinned: the swap a queued commit kernel reads
    int commit_pending = 0;       // a commit kernel may not have run yet
    int seq = 0;                  // sequence number of the last polled evaluation
    std::vector<int32_t> h_adj;  // adjacency evaluators: host copy of the resident base
    bool base_sym = false;       // the resident base is undirected
    int have_base = 0;  // adjacency evaluators: a table is resident
    int lmax = GG_LEVEL_LIMIT;
    int have_result = 0;
    int last_levels = 0;
    int levels_on = 1;        // GG_OPT_LEVELS
    int wvc_on = 0;           // GG_OPT_WVC
    int swap_graph = 0;       // GG_OPT_SWAP_GRAPH: 0 copies + events, 1 graph, 2 launches
    int levels_valid = 0;     // the last evaluation kept eccentricities and level counts
    int events_valid = 0;     // the last evaluation recorded its kernel-time events
    // captured single-swap evaluations, one per push setting of the swap
    hipGraph_t sg_graph[2] = {nullptr, nullptr};
    hipGraphExec_t sg_exec[2] = {nullptr, nullptr};
};

// Captured swap graphs hold the launch arguments chosen at capture time:
// every option change drops them.
static void drop_swap_graphs(gg_eval *ev) {
    for (int k = 0; k < 2; k++) {
        if (ev->sg_exec[k]) (void)hipGraphExecDestroy(ev->sg_exec[k]);
        if (ev->sg_graph[k]) (void)hipGraphDestroy(ev->sg_graph[k]);
        ev->sg_exec[k] = nullptr;
        ev->sg_graph[k] = nullptr;
    }
}

static int eval_alloc(gg_eval *ev, size_t tab_elem_bytes) {
    GG_HIP(hipSetDevice(ev->device));
    GG_HIP(hipStreamCreateWithFlags(&ev->stream, hipStreamNonBlocking));
    GG_HIP(hipEventCreate(&ev->ev0));
    GG_HIP(hipEventCreate(&ev->ev1));
    ev->dc.bfs_done = ev->ev1;  // the timed span ends with the BFS kernel
    int R = ev->kind == EVAL_ADJ ? ev->N : ev->M;
    ev->tab_bytes = ev->dc.topo_kind == TOPO_RINGQ ? (size_t)quads(ev->degree) * R * 8
                                                   : (size_t)R * ev->degree * tab_elem_bytes;
    size_t tb = ev->tab_bytes ? ev->tab_bytes : 8;
    if (hipMalloc(&ev->d_tab, tb) != hipSuccess) return GG_OOM;
    if (hipHostMalloc(&ev->h_tab, tb, hipHostMallocDefault) != hipSuccess) return GG_OOM;
    ev->W = (ev->M + 63) / 64;
    const int L = GG_LEVEL_LIMIT_MAX + 1;
    if (hipMalloc((void **)&ev->d_src, sizeof(int) * ev->M) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&ev->d_parts, sizeof(ColPart) * ev->W) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&ev->d_res, sizeof(ResultRec)) != hipSuccess) return GG_OOM;
    if (hipHostMalloc((void **)&ev->h_res, sizeof(ResultRec), hipHostMallocDefault) != hipSuccess)
        return GG_OOM;
    memset(ev->h_res, 0, sizeof(ResultRec));  // seq 0: no evaluation polled yet
    if (hipMalloc((void **)&ev->d_ecc, sizeof(int) * ev->M) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&ev->d_reach_col, sizeof(long long) * (size_t)ev->W * L) != hipSuccess)
        return GG_OOM;
    if (hipMalloc((void **)&ev->d_reach, sizeof(long long) * L) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&ev->d_swap, sizeof(gg_swap)) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&ev->dc.fuse, sizeof(FuseAcc)) != hipSuccess) return GG_OOM;
    GG_HIP(hipMemsetAsync(ev->dc.fuse, 0, sizeof(FuseAcc), ev->stream));
    if (hipHostMalloc((void **)&ev->h_swap, sizeof(gg_swap), hipHostMallocDefault) != hipSuccess)
        return GG_OOM;
    if (hipHostMalloc((void **)&ev->h_commit, sizeof(gg_swap), hipHostMallocDefault) != hipSuccess)
        return GG_OOM;
    // source -> vertex map
    std::vector<int> src(ev->M);
    for (int s = 0; s < ev->M; s++) {
        switch (ev->kind) {
            case EVAL_PIECE: {
                int x = s / ev->My, y = s % ev->My;
                src[s] = x * ev->Ny + y;
                break;
            }
            case EVAL_BOX: {
                int x = s / (ev->My * ev->Mz), r = s % (ev->My * ev->Mz);
                int y = r / ev->Mz, z = r % ev->Mz;
                src[s] = (x * ev->Ny + y) * ev->Nz + z;
                break;
            }
            default: src[s] = s;
        }
    }
    GG_HIP(hipMemcpyAsync(ev->d_src, src.data(), sizeof(int) * ev->M, hipMemcpyHostToDevice,
                          ev->stream));
    GG_HIP(hipStreamSynchronize(ev->stream));
    ev->dc.lds_device = device_lds_limit(ev->device);
    ev->dc.cu_count = device_cu_count(ev->device);
    ev->dc.wide_ok = 1;  // single evaluations, never captured
    return GG_OK;
}

static int check_device(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return GG_NO_DEVICE;
    if (device < 0 || device >= n) return GG_BAD_ARG;
    return GG_OK;
}

static void tune_lookup(gg_eval *ev);

static int finish_create(gg_eval *ev, size_t elem, gg_eval **out) {
    int st = eval_alloc(ev, elem);
    if (st != GG_OK) { gg_destroy(ev); return st; }
    tune_lookup(ev);
    *out = ev;
    return GG_OK;
}

extern "C" int gg_create_part(int N, int M, int degree, int device, gg_eval **out) {
    if (!out) return GG_BAD_ARG;
    *out = nullptr;
    if (N < 1 || M < 1 || M > N || degree < 0) return GG_BAD_ARG;
    int st = check_device(device);
    if (st != GG_OK) return st;
    gg_eval *ev = new gg_eval();
    ev->kind = EVAL_PART;
    ev->N = N; ev->M = M; ev->degree = degree; ev->device = device;
    ev->wvc_on = 1;  // cudaASPLconv::WVCdiameterASPL
    bool narrow = N <= 65535;
    ev->dc.topo_kind = narrow ? TOPO_RINGQ : TOPO_RING32;
    return finish_create(ev, narrow ? 2 : 4, out);
}

extern "C" int gg_create_adj(int N, int degree_max, int device, gg_eval **out) {
    if (!out) return GG_BAD_ARG;
    *out = nullptr;
    if (N < 1 || degree_max < 0) return GG_BAD_ARG;
    int st = check_device(device);
    if (st != GG_OK) return st;
    gg_eval *ev = new gg_eval();
    ev->kind = EVAL_ADJ;
    ev->N = N; ev->M = N; ev->degree = degree_max; ev->device = device;
    bool narrow = N <= 65535;
    ev->dc.topo_kind = narrow ? TOPO_RINGQ : TOPO_RING32;
    return finish_create(ev, narrow ? 2 : 4, out);
}

extern "C" int gg_create_piece(int Nx, int Ny, int Mx, int My, int degree, int device,
                               gg_eval **out) {
    if (!out) return GG_BAD_ARG;
    *out = nullptr;
    if (Nx < 1 || Ny < 1 || Mx < 1 || My < 1 || Mx > Nx || My > Ny || degree < 0) return GG_BAD_ARG;
    if (Nx > 65535 || Ny > 65535 || (long long)Nx * Ny > 0x7FFFFFFF) return GG_UNSUPPORTED;
    int st = check_device(device);
    if (st != GG_OK) return st;
    gg_eval *ev = new gg_eval();
    ev->kind = EVAL_PIECE;
    ev->Nx = Nx; ev->Ny = Ny; ev->Mx = Mx; ev->My = My;
    ev->N = Nx * Ny; ev->M = Mx * My; ev->degree = degree; ev->device = device;
    ev->dc.topo_kind = TOPO_PIECE;
    return finish_create(ev, 4, out);
}

extern "C" int gg_create_box(int Nx, int Ny, int Nz, int Mx, int My, int Mz, int degree,
                             int device, gg_eval **out) {
    if (!out) return GG_BAD_ARG;
    *out = nullptr;
    if (Nx < 1 || Ny < 1 || Nz < 1 || Mx < 1 || My < 1 || Mz < 1 || Mx > Nx || My > Ny ||
        Mz > Nz || degree < 0)
        return GG_BAD_ARG;
    if (Nx >= (1 << 21) || Ny >= (1 << 21) || Nz >= (1 << 21) ||
        (long long)Nx * Ny * Nz > 0x7FFFFFFF)
        return GG_UNSUPPORTED;
    int st = check_device(device);
    if (st != GG_OK) return st;
    gg_eval *ev = new gg_eval();
    ev->kind = EVAL_BOX;
    ev->Nx = Nx; ev->Ny = Ny; ev->Nz = Nz; ev->Mx = Mx; ev->My = My; ev->Mz = Mz;
    ev->N = Nx * Ny * Nz; ev->M = Mx * My * Mz; ev->degree = degree; ev->device = device;
    ev->dc.topo_kind = TOPO_BOX;
    return finish_create(ev, 8, out);
}

extern "C" void gg_destroy(gg_eval *ev) {
    if (!ev) return;
    (void)hipSetDevice(ev->device);
    if (ev->stream) (void)hipStreamSynchronize(ev->stream);
    drop_swap_graphs(ev);
    (void)hipFree(ev->d_tab);
    (void)hipHostFree(ev->h_tab);
    (void)hipFree(ev->d_src);
    (void)hipFree(ev->d_parts);
    (void)hipFree(ev->d_res);
    (void)hipHostFree(ev->h_res);
    (void)hipFree(ev->d_ecc);
    (void)hipFree(ev->d_reach_col);
    (void)hipFree(ev->d_reach);
    (void)hipFree(ev->d_swap);
    (void)hipHostFree(ev->h_swap);
    (void)hipHostFree(ev->h_commit);
    if (ev->dc.gscratch) (void)hipFree(ev->dc.gscratch);
    if (ev->dc.fuse) (void)hipFree(ev->dc.fuse);
    if (ev->ev0) (void)hipEventDestroy(ev->ev0);
    if (ev->ev1) (void)hipEventDestroy(ev->ev1);
    if (ev->stream) (void)hipStreamDestroy(ev->stream);
    delete ev;
}

// ---- packing ---------------------------------------------------------------------------
static inline long long norm_mod(long long x, long long n) {
    long long r = x % n;
    return r < 0 ? r + n : r;
}

// The reference adds N before % (cudaASPLconv.cu:44); for every offset where
// that is well defined the value equals the mathematical residue used here.
static int pack_ring(gg_eval *ev, const int32_t *table, int R, bool adj) {
    const int N = ev->N, D = ev->degree;
    if (ev->dc.topo_kind == TOPO_RINGQ) {
        uint16_t *h = (uint16_t *)ev->h_tab;
        memset(h, 0, ev->tab_bytes);  // padding slots: offset 0 (self-loop)
        for (int r = 0; r < R; r++)
            for (int i = 0; i < D; i++) {
                long long x = table[(size_t)r * D + i];
                long long off;
                if (adj) {
                    if (x < 0) off = 0;  // padding = self-loop (cudaASPLgraph.cu:148-151)
                    else if (x >= N) return GG_BAD_ARG;
                    else off = norm_mod(x - r, N);
                } else {
                    off = norm_mod(x, N);
                }
                // every vertex its own row: store the neighbour id itself
                if (R == N) off = norm_mod(r + off, N);
                h[quad_index(i, R, r)] = (uint16_t)off;
            }
    } else {
        uint32_t *h = (uint32_t *)ev->h_tab;
        for (int r = 0; r < R; r++)
            for (int i = 0; i < D; i++) {
                long long x = table[(size_t)r * D + i];
                long long off;
                if (adj) {
                    if (x < 0) off = 0;
                    else if (x >= N) return GG_BAD_ARG;
                    else off = norm_mod(x - r, N);
                } else {
                    off = norm_mod(x, N);
                }
                h[(size_t)i * R + r] = (uint32_t)off;
            }
    }
    return GG_OK;
}

static int pack_piece(gg_eval *ev, const int32_t *table) {
    const int R = ev->M, D = ev->degree;
    uint32_t *h = (uint32_t *)ev->h_tab;
    for (int r = 0; r < R; r++)
        for (int i = 0; i < D; i++) {
            long long dx = norm_mod(table[((size_t)r * D + i) * 2], ev->Nx);
            long long dy = norm_mod(table[((size_t)r * D + i) * 2 + 1], ev->Ny);
            h[(size_t)i * R + r] = (uint32_t)dx | ((uint32_t)dy << 16);
        }
    return GG_OK;
}

static int pack_box(gg_eval *ev, const int32_t *table) {
    const int R = ev->M, D = ev->degree;
    uint64_t *h = (uint64_t *)ev->h_tab;
    for (int r = 0; r < R; r++)
        for (int i = 0; i < D; i++) {
            const int32_t *o = table + ((size_t)r * D + i) * 3;
            uint64_t dx = (uint64_t)norm_mod(o[0], ev->Nx);
            uint64_t dy = (uint64_t)norm_mod(o[1], ev->Ny);
            uint64_t dz = (uint64_t)norm_mod(o[2], ev->Nz);
            h[(size_t)i * R + r] = dx | (dy << 21) | (dz << 42);
        }
    return GG_OK;
}

static void bind_topology(gg_eval *ev) {
    DenseCtx &c = ev->dc;
    unsigned R = ev->kind == EVAL_ADJ ? ev->N : ev->M;
    c.ringq = TopoRingQ{(const uint64_t *)ev->d_tab, (unsigned)ev->N, R};
    c.ring32 = TopoRing<uint32_t>{(const uint32_t *)ev->d_tab, (unsigned)ev->N, R};
    c.piece = TopoPiece{(const uint32_t *)ev->d_tab, (unsigned)ev->Nx, (unsigned)ev->Ny,
                        (unsigned)ev->Mx, (unsigned)ev->My, (unsigned)ev->M};
    c.box = TopoBox{(const uint64_t *)ev->d_tab, (unsigned)ev->Nx, (unsigned)ev->Ny,
                    (unsigned)ev->Nz, (unsigned)ev->Mx, (unsigned)ev->My, (unsigned)ev->Mz,
                    (unsigned)ev->M};
}

static void fill_result(const gg_eval *ev, const ResultRec &r, gg_result *out) {
    out->total_reps = r.total;
    out->total_scaled = r.total * (long long)(ev->N / ev->M);
    out->diameter = r.diameter;
    out->wvc = r.wvc;
    out->levels = r.levels;
    const long long pairs = (long long)ev->N * ev->M;
    if (r.flags & GG_FLAG_LIMIT) out->status = GG_DIAMETER_LIMIT;
    else if (r.reached != pairs) out->status = GG_UNCONNECTED;
    else out->status = GG_OK;
}

static int eval_common(gg_eval *ev, gg_result *out, bool upload = true,
                       const gg_swap *d_swap = nullptr) {
    if (upload)
        GG_HIP(hipMemcpyAsync(ev->d_tab, ev->h_tab, ev->tab_bytes, hipMemcpyHostToDevice,
                              ev->stream));
    bind_topology(ev);
    DenseArgs a{};
    a.N = ev->N; a.M = ev->M; a.D = ev->degree; a.W = ev->W; a.K = 1;
    a.lmax = ev->lmax;
    a.src_vertex = ev->d_src;
    a.swaps = d_swap;
    a.parts = ev->d_parts;
    a.res = ev->d_res;
    a.ecc_out = ev->levels_on || ev->wvc_on ? ev->d_ecc : nullptr;
    a.reach_col = ev->levels_on ? ev->d_reach_col : nullptr;
    a.reach_out = ev->levels_on ? ev->d_reach : nullptr;
    a.want_wvc = ev->wvc_on;
    ev->last_a = a;
    GG_HIP(hipEventRecord(ev->ev0, ev->stream));
    int st = run_dense(ev->dc, a, ev->stream);  // records ev1 after the BFS kernel
    if (st != GG_OK) return st;
    GG_HIP(hipMemcpyAsync(ev->h_res, ev->d_res, sizeof(ResultRec), hipMemcpyDeviceToHost,
                          ev->stream));
    GG_HIP(hipStreamSynchronize(ev->stream));
    fill_result(ev, *ev->h_res, out);
    ev->have_result = 1;
    ev->levels_valid = ev->levels_on;
    ev->events_valid = 1;
    ev->last_levels = ev->h_res->levels;
    return GG_OK;
}

// A part table is undirected when every offset o of row r is matched, with
// the same multiplicity, by -o in row (r + o) mod M (the row of vertex r + o).
// Checked only where the level-synchronous path could push (O(M D^2)).  The
// row of vertex v is v mod M, so checking rows 0..M-1 covers every vertex
// only when M divides N: past the last full copy the rows wrap differently,
// and such tables are treated as directed (they always pull).
static bool part_table_undirected(const int32_t *tab, int N, int M, int D) {
    if (M < 1 || N % M != 0) return false;
    std::vector<int32_t> t((size_t)M * D);  // offsets reduced to [0, N)
    for (size_t i = 0; i < t.size(); i++) t[i] = (int32_t)((((long long)tab[i] % N) + N) % N);
    for (int r = 0; r < M; r++) {
        const int32_t *row = t.data() + (size_t)r * D;
        for (int i = 0; i < D; i++) {
            const int32_t o = row[i];
            const int ru = (int)(((long long)r + o) % N % M);
            const int32_t neg = o ? N - o : 0;
            const int32_t *back = t.data() + (size_t)ru * D;
            int c1 = 0, c2 = 0;
            for (int j = 0; j < D; j++) {
                c1 += row[j] == o;
                c2 += back[j] == neg;
            }
            if (c1 != c2) return false;
        }
    }
    return true;
}

extern "C" int gg_eval_part(gg_eval *ev, const int32_t *table, gg_result *out) {
    if (!ev || !out || (!table && ev->degree > 0) || ev->kind != EVAL_PART) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(ev->device));
    int st = pack_ring(ev, table, ev->M, false);
    if (st != GG_OK) return st;
    ev->dc.wide_push_ok = (size_t)ev->N * 8 > (size_t)(ev->dc.lds_limit_opt > 0 ? ev->dc.lds_limit_opt
                                                                   : ev->dc.lds_device) &&
                          (long long)ev->M * ev->degree * ev->degree <= (1ll << 24) &&
                          part_table_undirected(table, ev->N, ev->M, ev->degree);
    return eval_common(ev, out);
}

extern "C" int gg_eval_adj(gg_eval *ev, const int32_t *adj, gg_result *out) {
    if (!ev || !out || (!adj && ev->degree > 0) || ev->kind != EVAL_ADJ) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(ev->device));
    int st = pack_ring(ev, adj, ev->N, true);
    if (st != GG_OK) return st;
    ev->h_adj.assign(adj, adj + (size_t)ev->N * ev->degree);
    ev->base_sym = adj_symmetric(adj, ev->N, ev->degree);
    ev->dc.push_ok = ev->base_sym;
    ev->dc.wide_push_ok = ev->dc.push_ok;
    st = eval_common(ev, out);
    if (st == GG_OK) ev->have_base = 1;
    return st;
}

// Validates a swap against the evaluator's shape (rows, slots, neighbours).
static int check_swap(const gg_eval *ev, const gg_swap *sw) {
    for (int j = 0; j < 4; j++) {
        if (sw->v[j] < 0) continue;
        if (sw->v[j] >= ev->N || sw->slot[j] < 0 || sw->slot[j] >= ev->degree ||
            sw->nbr[j] < 0 || sw->nbr[j] >= ev->N)
            return GG_BAD_ARG;
    }
    return GG_OK;
}

// One swap evaluation as a replayed HIP graph: the BFS kernel reads the swap
// from pinned host memory and finalize writes the result there, so a call is
// a host store, one graph launch and one wait.  The graph is captured on
// first use for each push setting (the launch arguments depend on it) and
// dropped by any option change.  Only where the column bitset fits in LDS
// (the persistent global and level-synchronous kernels allocate scratch or
// read counters back between launches).
// Waits until the pinned result carries sequence number `seq` (the launch
// that writes it, and everything queued before it on `st`, has then
// finished its work); the stream's status is checked every few thousand
// polls so that a failed launch is reported instead of waited on forever.
static hipError_t poll_result(const ResultRec *h_res, int seq, hipStream_t st) {
    const volatile int *flag = &h_res->seq;
    for (unsigned it = 1;; it++) {
        if (*flag == seq) break;
        if ((it & 4095) == 0) {
            const hipError_t q = hipStreamQuery(st);
            if (q == hipSuccess) {
                if (*flag == seq) break;
                return hipErrorUnknown;  // the stream drained without the mark
            }
            if (q != hipErrorNotReady) return q;
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return hipSuccess;
}

static int eval_swap_graph(gg_eval *ev, gg_result *out) {
    const int key = ev->dc.push_ok ? 1 : 0;
    DenseArgs a{};
    a.N = ev->N; a.M = ev->M; a.D = ev->degree; a.W = ev->W; a.K = 1;
    a.lmax = ev->lmax;
    a.src_vertex = ev->d_src;
    a.swaps = ev->h_swap;
    a.parts = ev->d_parts;
    a.res = ev->h_res;
    a.ecc_out = ev->levels_on || ev->wvc_on ? ev->d_ecc : nullptr;
    a.reach_col = ev->levels_on ? ev->d_reach_col : nullptr;
    a.reach_out = ev->levels_on ? ev->d_reach : nullptr;
    a.want_wvc = ev->wvc_on;
    bool polled = false;
    if (ev->swap_graph == 2) {
        // the same kernels launched one by one (no graph); a BFS launch that
        // finalizes in its last workgroup marks the pinned result with its
        // sequence number, which the host polls instead of waiting for the
        // stream
        bind_topology(ev);
        DenseCtx &c = ev->dc;
        hipEvent_t done = c.bfs_done;
        c.bfs_done = nullptr;
        const int wide = c.wide_ok;
        c.wide_ok = 0;
        ev->seq = ev->seq == INT_MAX ? 1 : ev->seq + 1;
        c.seq = ev->seq;
        const int st = run_dense(c, a, ev->stream);
        c.bfs_done = done;
        c.wide_ok = wide;
        if (st != GG_OK) return st;
        if (c.last_fused) {
            GG_HIP(poll_result(ev->h_res, ev->seq, ev->stream));
            polled = true;
        }
    } else if (!ev->sg_exec[key]) {
        bind_topology(ev);
        DenseCtx &c = ev->dc;
        hipEvent_t done = c.bfs_done;
        c.bfs_done = nullptr;  // no event nodes in the graph
        const int wide = c.wide_ok;
        c.wide_ok = 0;
        GG_HIP(hipStreamBeginCapture(ev->stream, hipStreamCaptureModeThreadLocal));
        int st = run_dense(c, a, ev->stream);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ev->stream, &g);
        c.bfs_done = done;
        c.wide_ok = wide;
        if (st != GG_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        GG_HIP(e);
        ev->sg_graph[key] = g;
        GG_HIP(hipGraphInstantiate(&ev->sg_exec[key], g, nullptr, nullptr, 0));
    }
    if (ev->swap_graph == 1) GG_HIP(hipGraphLaunch(ev->sg_exec[key], ev->stream));
    if (!polled) GG_HIP(hipStreamSynchronize(ev->stream));
    ev->commit_pending = 0;
    fill_result(ev, *ev->h_res, out);
    ev->last_a = a;
    ev->have_result = 1;
    ev->levels_valid = ev->levels_on;
    ev->events_valid = 0;
    ev->last_levels = ev->h_res->levels;
    return GG_OK;
}

extern "C" int gg_eval_adj_swap(gg_eval *ev, const gg_swap *sw, gg_result *out) {
    if (!ev || !sw || !out || ev->kind != EVAL_ADJ || !ev->have_base) return GG_BAD_ARG;
    int st = check_swap(ev, sw);
    if (st != GG_OK) return st;
    GG_HIP(hipSetDevice(ev->device));
    ev->dc.push_ok = ev->base_sym && swap_keeps_symmetry(ev->h_adj.data(), ev->degree, *sw);
    ev->dc.wide_push_ok = ev->dc.push_ok;
    // the previous call's kernels have finished (every call waits), so the
    // pinned swap can be rewritten
    *ev->h_swap = *sw;
    if (ev->swap_graph && dense_fits_lds(ev->dc, ev->N)) return eval_swap_graph(ev, out);
    GG_HIP(hipMemcpyAsync(ev->d_swap, ev->h_swap, sizeof(gg_swap), hipMemcpyHostToDevice,
                          ev->stream));
    return eval_common(ev, out, false, ev->d_swap);
}

// Applies a swap to the resident table (ring topologies, R = N: the 16-bit
// table holds neighbour ids, the 32-bit one offsets).
__global__ void k_commit_swap(uint16_t *tabq, uint32_t *tab32, int N, const gg_swap *sw) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int j = 0; j < 4; j++) {
        int v = sw->v[j];
        if (v < 0) continue;
        int off = sw->nbr[j] - v;
        if (off < 0) off += N;
        if (tabq) tabq[quad_index(sw->slot[j], N, v)] = (uint16_t)sw->nbr[j];
        else tab32[(size_t)sw->slot[j] * N + v] = (uint32_t)off;
    }
}

extern "C" int gg_eval_adj_commit(gg_eval *ev, const gg_swap *sw) {
    if (!ev || !sw || ev->kind != EVAL_ADJ || !ev->have_base) return GG_BAD_ARG;
    int st = check_swap(ev, sw);
    if (st != GG_OK) return st;
    ev->base_sym = ev->base_sym && swap_keeps_symmetry(ev->h_adj.data(), ev->degree, *sw);
    for (int j = 0; j < 4; j++)
        if (sw->v[j] >= 0) ev->h_adj[(size_t)sw->v[j] * ev->degree + sw->slot[j]] = sw->nbr[j];
    GG_HIP(hipSetDevice(ev->device));
    const bool q = ev->dc.topo_kind == TOPO_RINGQ;
    if (ev->swap_graph) {
        // queued, not waited for: the kernel reads the swap from pinned
        // memory, which is rewritten only after it has run (the next
        // evaluation waits on the stream, a second commit waits here)
        if (ev->commit_pending) GG_HIP(hipStreamSynchronize(ev->stream));
        *ev->h_commit = *sw;
        hipLaunchKernelGGL(k_commit_swap, dim3(1), dim3(64), 0, ev->stream,
                           q ? (uint16_t *)ev->d_tab : nullptr,
                           q ? nullptr : (uint32_t *)ev->d_tab, ev->N,
                           (const gg_swap *)ev->h_commit);
        GG_HIP(hipGetLastError());
        ev->commit_pending = 1;
        return GG_OK;
    }
    *ev->h_swap = *sw;
    GG_HIP(hipMemcpyAsync(ev->d_swap, ev->h_swap, sizeof(gg_swap), hipMemcpyHostToDevice,
                          ev->stream));
    hipLaunchKernelGGL(k_commit_swap, dim3(1), dim3(64), 0, ev->stream,
                       q ? (uint16_t *)ev->d_tab : nullptr, q ? nullptr : (uint32_t *)ev->d_tab,
                       ev->N, (const gg_swap *)ev->d_swap);
    GG_HIP(hipGetLastError());
    GG_HIP(hipStreamSynchronize(ev->stream));
    return GG_OK;
}

extern "C" int gg_eval_piece(gg_eval *ev, const int32_t *table, gg_result *out) {
    if (!ev || !out || (!table && ev->degree > 0) || ev->kind != EVAL_PIECE) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(ev->device));
    int st = pack_piece(ev, table);
    if (st != GG_OK) return st;
    return eval_common(ev, out);
}

extern "C" int gg_eval_box(gg_eval *ev, const int32_t *table, gg_result *out) {
    if (!ev || !out || (!table && ev->degree > 0) || ev->kind != EVAL_BOX) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(ev->device));
    int st = pack_box(ev, table);
    if (st != GG_OK) return st;
    return eval_common(ev, out);
}

extern "C" int gg_eval_levels(gg_eval *ev, int32_t *ecc_out, int64_t *reach_out, int reach_cap) {
    if (!ev || reach_cap < 0) return GG_BAD_ARG;
    if (!ev->have_result) return GG_BAD_ARG;
    if (!ev->levels_valid) return GG_UNSUPPORTED;  // evaluated with GG_OPT_LEVELS 0
    GG_HIP(hipSetDevice(ev->device));
    if (ecc_out)
        GG_HIP(hipMemcpyAsync(ecc_out, ev->d_ecc, sizeof(int) * ev->M, hipMemcpyDeviceToHost,
                              ev->stream));
    if (reach_out && reach_cap > 0) {
        int L = ev->lmax + 1;
        int n = reach_cap < L ? reach_cap : L;
        std::vector<long long> tmp(L, 0);
        GG_HIP(hipMemcpyAsync(tmp.data(), ev->d_reach, sizeof(long long) * L,
                              hipMemcpyDeviceToHost, ev->stream));
        GG_HIP(hipStreamSynchronize(ev->stream));
        for (int l = 0; l < reach_cap; l++)
            reach_out[l] = (l < n && l <= ev->last_levels) ? tmp[l] : 0;
    }
    GG_HIP(hipStreamSynchronize(ev->stream));
    return GG_OK;
}

// The last evaluation's BFS kernel launched `reps` times back to back
// between one event pair (the kernel is idempotent: it rewrites the same
// column parts); the mean per launch leaves out the event overhead that a
// single bracketed launch carries.
extern "C" int gg_eval_time_bfs(gg_eval *ev, int reps, float *ms) {
    if (!ev || !ms || reps < 1 || reps > 100000 || !ev->have_result) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(ev->device));
    DenseCtx &c = ev->dc;
    hipEvent_t done = c.bfs_done;
    c.bfs_done = nullptr;
    c.bfs_only = 1;
    int st = hipEventRecord(ev->ev0, ev->stream) == hipSuccess ? GG_OK : GG_HIP_ERROR;
    for (int r = 0; r < reps && st == GG_OK; r++) st = run_dense(c, ev->last_a, ev->stream);
    c.bfs_only = 0;
    c.bfs_done = done;
    if (st != GG_OK) return st;
    GG_HIP(hipEventRecord(ev->ev1, ev->stream));
    GG_HIP(hipEventSynchronize(ev->ev1));
    float t = 0;
    GG_HIP(hipEventElapsedTime(&t, ev->ev0, ev->ev1));
    *ms = t / (float)reps;
    return GG_OK;
}

extern "C" int gg_eval_last_kernel_ms(gg_eval *ev, float *ms) {
    if (!ev || !ms || !ev->have_result) return GG_BAD_ARG;
    if (!ev->events_valid) return GG_UNSUPPORTED;  // a captured swap evaluation
    GG_HIP(hipEventElapsedTime(ms, ev->ev0, ev->ev1));
    return GG_OK;
}

extern "C" int gg_eval_set_option(gg_eval *ev, int option, int value) {
    if (!ev) return GG_BAD_ARG;
    drop_swap_graphs(ev);
    switch (option) {
        case GG_OPT_BLOCK:
            if (value != 0 && value != 256 && value != 1024) return GG_BAD_ARG;
            ev->dc.block_opt = value;
            return GG_OK;
        case GG_OPT_LDS_LIMIT:
            if (value < 0) return GG_BAD_ARG;
            ev->dc.lds_limit_opt = value;
            return GG_OK;
        case GG_OPT_GENERIC:
            ev->dc.generic_opt = value ? 1 : 0;
            return GG_OK;
        case GG_OPT_LEVEL_LIMIT:
            if (value < 1 || value > GG_LEVEL_LIMIT_MAX) return GG_BAD_ARG;
            ev->lmax = value;
            return GG_OK;
        case GG_OPT_PUSH:
            ev->dc.push_off = value ? 0 : 1;
            return GG_OK;
        case GG_OPT_WIDE:
            ev->dc.wide_off = value ? 0 : 1;
            return GG_OK;
        case GG_OPT_SWAP_GRAPH:
            if (value < 0 || value > 2) return GG_BAD_ARG;
            if (value && ev->kind != EVAL_ADJ) return GG_UNSUPPORTED;
            ev->swap_graph = value;
            return GG_OK;
        case GG_OPT_LEVELS:
            ev->levels_on = value ? 1 : 0;
            return GG_OK;
        case GG_OPT_WVC:
            if (value && ev->kind != EVAL_PART && ev->kind != EVAL_ADJ) return GG_UNSUPPORTED;
            ev->wvc_on = value ? 1 : 0;
            return GG_OK;
        default: return GG_BAD_ARG;
    }
}

extern "C" int gg_eval_get_option(gg_eval *ev, int option, int *value) {
    if (!ev || !value) return GG_BAD_ARG;
    switch (option) {
        case GG_OPT_BLOCK: *value = ev->dc.block_opt; return GG_OK;
        case GG_OPT_LDS_LIMIT: *value = ev->dc.lds_limit_opt; return GG_OK;
        case GG_OPT_GENERIC: *value = ev->dc.generic_opt; return GG_OK;
        case GG_OPT_LEVEL_LIMIT: *value = ev->lmax; return GG_OK;
        case GG_OPT_PUSH: *value = ev->dc.push_off ? 0 : 1; return GG_OK;
        case GG_OPT_WIDE: *value = ev->dc.wide_off ? 0 : 1; return GG_OK;
        case GG_OPT_SWAP_GRAPH: *value = ev->swap_graph; return GG_OK;
        case GG_OPT_LEVELS: *value = ev->levels_on; return GG_OK;
        case GG_OPT_WVC: *value = ev->wvc_on; return GG_OK;
        case 100: *value = ev->dc.last_variant; return GG_OK;  // kernel variant of the last call
        case 101: *value = ev->dc.last_vpt; return GG_OK;      // its vertices per thread
        case 102: *value = ev->dc.last_fixed; return GG_OK;    // its fixed degree (0 = generic)
        case 103: *value = ev->dc.last_push_cap; return GG_OK; // its frontier list capacity
        default: return GG_BAD_ARG;
    }
}

// ---- autotuner -----------------------------------------------------------------------------
// Times the dense kernel variants (workgroup size x fixed/generic gather) on
// the caller's graph, applies the fastest and records it in a cache file
// (GG_TUNE_CACHE, default $HOME/.cache/gg_mi355x_tune.txt), keyed by device
// architecture, evaluator kind, N, M and degree; evaluators created later with
// the same key start from the cached choice.
static std::mutex g_tune_mu;

static std::string tune_cache_path() {
    const char *e = getenv("GG_TUNE_CACHE");
    if (e && *e) return e;
    const char *h = getenv("HOME");
    return std::string(h ? h : "/tmp") + "/.cache/gg_mi355x_tune.txt";
}

namespace gg {
std::string device_arch(int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) return prop.gcnArchName;
    return "unknown";
}

// The values stored under `key` in the tuning cache ("" if none).
std::string tune_cache_get(const std::string &key) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::ifstream f(tune_cache_path());
    std::string line, found;
    while (f && std::getline(f, line))
        if (line.size() > key.size() && line.compare(0, key.size() + 1, key + " ") == 0)
            found = line.substr(key.size() + 1);
    return found;
}

// Replaces the line of `key` in the tuning cache.
void tune_cache_put(const std::string &key, const std::string &values) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const std::string path = tune_cache_path();
    std::vector<std::string> keep;
    {
        std::ifstream f(path);
        std::string line;
        while (f && std::getline(f, line))
            if (line.compare(0, key.size() + 1, key + " ") != 0) keep.push_back(line);
    }
    const size_t slash = path.rfind('/');
    if (slash != std::string::npos) {
        std::string dir = path.substr(0, slash);
        std::string cmd = "mkdir -p '" + dir + "'";
        if (system(cmd.c_str()) != 0) return;
    }
    std::ofstream f(path, std::ios::trunc);
    for (auto &l : keep) f << l << '\n';
    f << key << ' ' << values << '\n';
}
}  // namespace gg

static std::string tune_key(const gg_eval *ev) {
    std::ostringstream k;
    k << device_arch(ev->device) << ' ' << (int)ev->kind << ' ' << ev->N << ' ' << ev->M << ' '
      << ev->degree;
    return k.str();
}

static void tune_lookup(gg_eval *ev) {
    std::istringstream in(tune_cache_get(tune_key(ev)));
    int block = 0, generic = 0;
    if (in >> block >> generic) {
        ev->dc.block_opt = block;
        ev->dc.generic_opt = generic;
    }
}

static void tune_store(gg_eval *ev, int block, int generic, float ms) {
    std::ostringstream v;
    v << block << ' ' << generic << ' ' << ms;
    tune_cache_put(tune_key(ev), v.str());
}

extern "C" int gg_eval_autotune(gg_eval *ev, const int32_t *table, int reps,
                                gg_tune_report *report) {
    if (!ev || !table || reps < 1) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(ev->device));
    drop_swap_graphs(ev);  // the kernel choice changes
    int st;
    switch (ev->kind) {
        case EVAL_PART: st = pack_ring(ev, table, ev->M, false); break;
        case EVAL_ADJ: st = pack_ring(ev, table, ev->N, true); break;
        case EVAL_PIECE: st = pack_piece(ev, table); break;
        default: st = pack_box(ev, table); break;
    }
    if (st != GG_OK) return st;
    const int saved_block = ev->dc.block_opt, saved_generic = ev->dc.generic_opt;
    const int blocks[2] = {256, 1024};
    float best_ms = 1e30f;
    int best_block = 1024, best_generic = 0, n = 0;
    gg_tune_report rep{};
    for (int bi = 0; bi < 2; bi++)
        for (int generic = 0; generic < 2; generic++) {
            ev->dc.block_opt = blocks[bi];
            ev->dc.generic_opt = generic;
            std::vector<float> t;
            gg_result r;
            for (int i = 0; i < reps + 1; i++) {
                st = eval_common(ev, &r, i == 0);
                if (st != GG_OK) { ev->dc.block_opt = saved_block; ev->dc.generic_opt = saved_generic; return st; }
                float ms = 0;
                GG_HIP(hipEventElapsedTime(&ms, ev->ev0, ev->ev1));
                if (i > 0) t.push_back(ms);
            }
            std::sort(t.begin(), t.end());
            const float med = t[t.size() / 2];
            if (n < 4) {
                rep.block[n] = blocks[bi];
                rep.generic[n] = generic;
                rep.variant[n] = ev->dc.last_variant;
                rep.ms[n] = med;
                n++;
            }
            if (med < best_ms) { best_ms = med; best_block = blocks[bi]; best_generic = generic; }
        }
    ev->dc.block_opt = best_block;
    ev->dc.generic_opt = best_generic;
    rep.count = n;
    rep.best_block = best_block;
    rep.best_generic = best_generic;
    rep.best_ms = best_ms;
    tune_store(ev, best_block, best_generic, best_ms);
    if (report) *report = rep;
    return GG_OK;
}
