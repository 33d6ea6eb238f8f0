// gg_batch.hip -- batched exact evaluation of K swap candidates against one
// resident base adjacency.  No reference counterpart: the reference evaluates
// one move per host round trip (cppcode/annealing.cpp:397-403).
//
// Two paths, both exact:
//   dense   one launch of K*W column BFS workgroups (gg_dense.hpp)
//   sparse  difference propagation against the stored per-level base state
//           (gg_sparse.hpp) with a dense fallback queue, captured as one HIP
//           graph: H2D swaps -> k_sparse -> k_dense_lds_queue ->
//           k_sparse_finalize -> queue reset -> D2H results.
#include <cstring>
#include <vector>
#include "gg_common.hpp"
#include "gg_dense.hpp"
#include "gg_sparse.hpp"
#include "gg_internal.hpp"
#include "gg_batch.hpp"

using namespace gg;

static const int kSparseBlock = 1024;
static const int kSnapCapDefault = 8;
static const size_t kSnapBytesMax = (size_t)8 << 30;
static const size_t kOverrideBytesMax = (size_t)4 << 30;
static const int kDeltaCap = 2048;

// The sparse kernel keeps one column (N words) in LDS plus small bookkeeping.
// Static LDS of k_sparse (bookkeeping and the launch's swaps), rounded up.
static const int kSparseStaticLds = 8192;
static const int kStagedKeysMax = 8192;

static bool sparse_fits(const gg_batch *b) {
    return sparse_lds_bytes(b->N, 0) + kSparseStaticLds <= (size_t)b->dc.lds_device;
}

// Delta keys the sparse kernel can stage in the LDS left over.
static int sparse_staged_keys(const gg_batch *b) {
    long long room = ((long long)b->dc.lds_device - kSparseStaticLds -
                      (long long)sparse_lds_bytes(b->N, 0)) / 4;
    if (room < 0) room = 0;
    return room < kStagedKeysMax ? (int)room : kStagedKeysMax;
}

static int batch_alloc(gg_batch *b) {
    GG_HIP(hipSetDevice(b->device));
    GG_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    GG_HIP(hipEventCreate(&b->ev0));
    GG_HIP(hipEventCreate(&b->ev1));
    const int N = b->N, D = b->degree, K = b->max_k;
    const int L = GG_LEVEL_LIMIT + 1;
    b->W = (N + 63) / 64;
    const int W = b->W;
    b->narrow = N <= 65535;
    size_t tb = b->narrow ? (size_t)quads(D) * N * 8 : (size_t)N * D * 4;
    if (hipMalloc(&b->d_tab, tb) != hipSuccess) return GG_OOM;
    if (hipHostMalloc(&b->h_tab, tb, hipHostMallocDefault) != hipSuccess) return GG_OOM;
    b->tab_bytes = tb;
    if (hipMalloc((void **)&b->d_adjR, sizeof(int) * (size_t)N * D) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&b->d_src, sizeof(int) * N) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&b->d_parts, sizeof(ColPart) * (size_t)W * (K + 1)) != hipSuccess)
        return GG_OOM;
    if (hipMalloc((void **)&b->d_res, sizeof(ResultRec) * (K + 1)) != hipSuccess) return GG_OOM;
    if (hipHostMalloc((void **)&b->h_res, sizeof(ResultRec) * (K + 1), hipHostMallocDefault) !=
        hipSuccess)
        return GG_OOM;
    if (hipMalloc((void **)&b->d_swaps, sizeof(gg_swap) * (K + 1)) != hipSuccess) return GG_OOM;
    if (hipHostMalloc((void **)&b->h_swaps, sizeof(gg_swap) * (K + 1), hipHostMallocDefault) !=
        hipSuccess)
        return GG_OOM;
    if (hipMalloc((void **)&b->d_ecc, sizeof(int) * N) != hipSuccess) return GG_OOM;
    if (hipMalloc((void **)&b->d_base_reach, sizeof(long long) * (size_t)W * L) != hipSuccess)
        return GG_OOM;
    std::vector<int> src(N);
    for (int s = 0; s < N; s++) src[s] = s;
    GG_HIP(hipMemcpyAsync(b->d_src, src.data(), sizeof(int) * N, hipMemcpyHostToDevice, b->stream));
    b->dc.lds_device = device_lds_limit(b->device);
    b->dc.cu_count = device_cu_count(b->device);
    b->dc.topo_kind = b->narrow ? TOPO_RINGQ : TOPO_RING32;
    b->dc.ringq = TopoRingQ{(const uint64_t *)b->d_tab, (unsigned)N, (unsigned)N};
    b->dc.ring32 = TopoRing<uint32_t>{(const uint32_t *)b->d_tab, (unsigned)N, (unsigned)N};

    // sparse-path buffers (only when the column bitset fits in LDS)
    if (sparse_fits(b) && dense_fits_lds(b->dc, N)) {
        size_t per_level = (size_t)W * N * 8;
        int cap = kSnapCapDefault;
        while (cap > 1 && per_level * cap > kSnapBytesMax) cap--;
        b->snap_cap = cap;
        b->cap = kDeltaCap < N ? kDeltaCap : N;
        b->skcap = sparse_staged_keys(b);
        const int NB = (N + 31) / 32;
        const size_t per_task = 2 * ((size_t)N * 8 + (size_t)NB * 4 + (size_t)b->cap * 4);
        int chunk = K < GG_SPARSE_MAX_K ? K : GG_SPARSE_MAX_K;
        while (chunk > 1 && per_task * W * chunk > kOverrideBytesMax) chunk--;
        b->chunk = chunk;
        const size_t T = (size_t)K * W, TC = (size_t)chunk * W;
        if (hipMalloc((void **)&b->d_snap, per_level * cap) != hipSuccess) return GG_OOM;
        if (hipMalloc((void **)&b->d_dl_key, sizeof(int) * TC * 2 * b->cap) != hipSuccess)
            return GG_OOM;
        if (hipMalloc((void **)&b->d_dl_bits, 4 * TC * 2 * NB) != hipSuccess) return GG_OOM;
        if (hipMalloc((void **)&b->d_dl_ovr, 8 * TC * 2 * (size_t)N) != hipSuccess) return GG_OOM;
        if (hipMalloc((void **)&b->d_task_state, sizeof(int) * T) != hipSuccess) return GG_OOM;
        if (hipMalloc((void **)&b->d_task_end, sizeof(int) * T) != hipSuccess) return GG_OOM;
        if (hipMalloc((void **)&b->d_task_flags, sizeof(int) * T) != hipSuccess) return GG_OOM;
        if (hipMalloc((void **)&b->d_dcnt, sizeof(long long) * T * L) != hipSuccess) return GG_OOM;
        if (hipMalloc((void **)&b->d_queue, sizeof(int2) * T) != hipSuccess) return GG_OOM;
        b->nqctl = 2 + (K + chunk - 1) / chunk;  // queue count, queue next, endpoint counters
        if (hipMalloc((void **)&b->d_qctl, sizeof(int) * b->nqctl) != hipSuccess) return GG_OOM;
        if (hipMalloc((void **)&b->d_dense_parts, sizeof(ColPart) * T) != hipSuccess)
            return GG_OOM;
        GG_HIP(hipMemsetAsync(b->d_qctl, 0, sizeof(int) * b->nqctl, b->stream));
        GG_HIP(hipMemsetAsync(b->d_dl_bits, 0, 4 * TC * 2 * NB, b->stream));
    }
    GG_HIP(hipStreamSynchronize(b->stream));
    return GG_OK;
}

extern "C" int gg_batch_create(int N, int degree, int max_k, int device, gg_batch **out) {
    if (!out) return GG_BAD_ARG;
    *out = nullptr;
    if (N < 2 || degree < 1 || max_k < 1 || max_k > (1 << 16)) return GG_BAD_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return GG_NO_DEVICE;
    if (device < 0 || device >= n) return GG_BAD_ARG;
    gg_batch *b = new gg_batch();
    b->N = N; b->degree = degree; b->max_k = max_k; b->device = device;
    int st = batch_alloc(b);
    if (st != GG_OK) { gg_batch_destroy(b); return st; }
    *out = b;
    return GG_OK;
}

static void drop_graph(gg_batch *b) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    b->graph_k = 0;
}

extern "C" void gg_batch_destroy(gg_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    void *bufs[] = {b->d_tab, b->d_adjR, b->d_src, b->d_parts, b->d_res, b->d_swaps,
                    b->d_ecc, b->d_base_reach, b->d_snap, b->d_dl_key, b->d_dl_bits, b->d_dl_ovr,
                    b->d_task_state, b->d_task_end, b->d_task_flags, b->d_dcnt, b->d_queue,
                    b->d_qctl, b->d_dense_parts};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (b->h_tab) (void)hipHostFree(b->h_tab);
    if (b->h_res) (void)hipHostFree(b->h_res);
    if (b->h_swaps) (void)hipHostFree(b->h_swaps);
    if (b->dc.gscratch) (void)hipFree(b->dc.gscratch);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

static void fill(const gg_batch *b, const ResultRec &r, gg_result *out) {
    out->total_reps = r.total;
    out->total_scaled = r.total;
    out->diameter = r.diameter;
    out->wvc = r.wvc;
    out->levels = r.levels;
    const long long pairs = (long long)b->N * b->N;
    if (r.flags & GG_FLAG_LIMIT) out->status = GG_DIAMETER_LIMIT;
    else if (r.reached != pairs) out->status = GG_UNCONNECTED;
    else out->status = GG_OK;
}

static DenseArgs dense_args(gg_batch *b, int K, const gg_swap *d_swaps, ColPart *parts,
                            ResultRec *res, int *ecc) {
    DenseArgs a{};
    a.N = b->N; a.M = b->N; a.D = b->degree; a.W = b->W; a.K = K;
    a.lmax = GG_LEVEL_LIMIT;
    a.src_vertex = b->d_src;
    a.swaps = d_swaps;
    a.parts = parts;
    a.res = res;
    a.ecc_out = ecc;
    return a;
}

// Evaluates the resident base (slot max_k), storing the per-level snapshots
// and per-column level counts the sparse path starts from.
static int eval_base(gg_batch *b, gg_result *out) {
    const int K = b->max_k;
    DenseArgs a = dense_args(b, 1, nullptr, b->d_parts + (size_t)K * b->W, b->d_res + K,
                             b->d_ecc);
    a.reach_col = b->d_base_reach;
    a.snap = b->d_snap;
    a.snap_levels = b->d_snap ? b->snap_cap : 0;
    int st = run_dense(b->dc, a, b->stream);
    if (st != GG_OK) return st;
    GG_HIP(hipMemcpyAsync(b->h_res + K, b->d_res + K, sizeof(ResultRec), hipMemcpyDeviceToHost,
                          b->stream));
    GG_HIP(hipStreamSynchronize(b->stream));
    gg_result r;
    fill(b, b->h_res[K], &r);
    if (out) *out = r;
    b->base_res = b->h_res[K];
    // the sparse path needs a connected base whose levels below the diameter
    // are all stored, and the LDS-resident kernels
    b->sparse_ok = b->d_snap && r.status == GG_OK && r.diameter - 1 <= b->snap_cap &&
                   b->dc.last_variant != 3;
    return GG_OK;
}

extern "C" int gg_batch_set_base(gg_batch *b, const int32_t *adj, gg_result *out) {
    if (!b || !adj) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(b->device));
    const int N = b->N, D = b->degree;
    for (size_t i = 0; i < (size_t)N * D; i++)
        if (adj[i] < 0 || adj[i] >= N) return GG_BAD_ARG;
    if (b->narrow) memset(b->h_tab, 0, b->tab_bytes);
    for (int v = 0; v < N; v++)
        for (int i = 0; i < D; i++) {
            long long u = adj[(size_t)v * D + i];
            long long off = ((u - v) % N + N) % N;
            if (b->narrow) ((uint16_t *)b->h_tab)[quad_index(i, N, v)] = (uint16_t)u;
            else ((uint32_t *)b->h_tab)[(size_t)i * N + v] = (uint32_t)off;
        }
    GG_HIP(hipMemcpyAsync(b->d_tab, b->h_tab, b->tab_bytes, hipMemcpyHostToDevice, b->stream));
    GG_HIP(hipMemcpyAsync(b->d_adjR, adj, sizeof(int) * (size_t)N * D, hipMemcpyHostToDevice,
                          b->stream));
    b->h_adj.assign(adj, adj + (size_t)N * D);
    b->base_sym = adj_symmetric(adj, N, D);
    b->have_base = 1;
    b->dc.push_ok = b->base_sym;
    return eval_base(b, out);
}

static int check_swaps(const gg_batch *b, const gg_swap *sw, int K) {
    for (int k = 0; k < K; k++)
        for (int j = 0; j < 4; j++) {
            if (sw[k].v[j] < 0) continue;
            if (sw[k].v[j] >= b->N || sw[k].slot[j] < 0 || sw[k].slot[j] >= b->degree ||
                sw[k].nbr[j] < 0 || sw[k].nbr[j] >= b->N)
                return GG_BAD_ARG;
        }
    return GG_OK;
}

// Sparse kernel for the degree (fixed-degree instances with a dense tail
// when vpt > 0, the runtime-degree one without).
static hipError_t launch_sparse(const SparseArgs &s, const TopoRingQ &topo, const DenseArgs &qa,
                                size_t lds, int vpt, hipStream_t st) {
    switch (s.D) {
        case 4: return launch_sparse_d4(s, topo, qa, lds, vpt, st);
        case 8: return launch_sparse_d8(s, topo, qa, lds, vpt, st);
        case 12: return launch_sparse_d12(s, topo, qa, lds, vpt, st);
        case 16: return launch_sparse_d16(s, topo, qa, lds, vpt, st);
        case 24: return launch_sparse_d24(s, topo, qa, lds, vpt, st);
        default: return launch_sparse_deg<0>(s, topo, qa, lds, 0, st);
    }
}

// Enqueues the sparse pipeline for K candidates whose swaps are in d_swaps.
static int enqueue_sparse(gg_batch *b, int K) {
    const int W = b->W, N = b->N;
    SparseArgs s{};
    s.N = N; s.D = b->degree; s.W = W; s.K = K; s.lmax = GG_LEVEL_LIMIT;
    s.adjR = b->d_adjR;
    s.swaps = b->d_swaps;
    s.snap = b->d_snap;
    s.S = b->snap_cap;
    s.base_parts = b->d_parts + (size_t)b->max_k * W;
    s.base_reach = b->d_base_reach;
    s.cap = b->cap;
    s.skcap = b->skcap;
    s.dl_key = b->d_dl_key;
    s.dl_bits = b->d_dl_bits;
    s.dl_ovr = b->d_dl_ovr;
    s.dcnt = b->d_dcnt;
    s.task_state = b->d_task_state;
    s.task_end = b->d_task_end;
    s.task_flags = b->d_task_flags;
    s.queue = b->d_queue;
    s.queue_count = b->d_qctl;
    s.queue_next = b->d_qctl + 1;
    s.dense_parts = b->d_dense_parts;
    s.res = b->d_res;
    const size_t lds = sparse_lds_bytes(N, b->skcap);
    // dense tail of the sparse kernel (endpoint columns): fixed-degree ring
    // table, its vertices-per-thread instance, a push list within the LDS
    // the sparse layout leaves after B_{l-1}
    int vpt = 0;
    if (b->narrow && b->dc.ringq.R == b->dc.ringq.N && !b->dc.generic_opt) {
        const int v = pick_vpt(1024, (N + 1023) / 1024);
        if (v == 4 || v == 8 || v == 16) vpt = v;
    }
    DenseArgs qa = dense_args(b, K, b->d_swaps, b->d_dense_parts, nullptr, nullptr);
    qa.push_cap = push_capacity(b->dc, N, b->degree, (int)lds);
    for (int k0 = 0, c_i = 0; k0 < K; k0 += b->chunk, c_i++) {
        SparseArgs c = s;
        c.k0 = k0;
        c.K = K - k0 < b->chunk ? K - k0 : b->chunk;
        c.ep_next = b->d_qctl + 2 + c_i;
        hipError_t e = launch_sparse(c, b->dc.ringq, qa, lds, vpt, b->stream);
        if (e != hipSuccess) {
            gg_set_last_hip_error(e, "k_sparse", __FILE__, __LINE__);
            return GG_HIP_ERROR;
        }
    }
    // dense fallback for queued tasks
    DenseArgs a = dense_args(b, K, b->d_swaps, b->d_dense_parts, nullptr, nullptr);
    a.queue = b->d_queue;
    a.queue_count = b->d_qctl;
    a.queue_next = b->d_qctl + 1;
    int per_cu = (int)((size_t)b->dc.lds_device / (lds + 4096));
    if (per_cu < 1) per_cu = 1;
    if (per_cu > 2) per_cu = 2;
    int grid = b->dc.cu_count * per_cu;
    if (grid > K * W) grid = K * W;
    int st = run_dense_queue(b->dc, a, grid, b->stream);
    if (st != GG_OK) return st;
    hipLaunchKernelGGL(k_sparse_finalize<256>, dim3(K), dim3(256), 0, b->stream, s,
                       N - 64 * (W - 1));
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, b->stream, b->d_qctl, b->nqctl);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

static int enqueue_step(gg_batch *b, int K, bool sparse) {
    GG_HIP(hipMemcpyAsync(b->d_swaps, b->h_swaps, sizeof(gg_swap) * K, hipMemcpyHostToDevice,
                          b->stream));
    int st;
    if (sparse) {
        st = enqueue_sparse(b, K);
    } else {
        DenseArgs a = dense_args(b, K, b->d_swaps, b->d_parts, b->d_res, nullptr);
        st = run_dense(b->dc, a, b->stream);
    }
    if (st != GG_OK) return st;
    GG_HIP(hipMemcpyAsync(b->h_res, b->d_res, sizeof(ResultRec) * K, hipMemcpyDeviceToHost,
                          b->stream));
    return GG_OK;
}

extern "C" int gg_batch_eval(gg_batch *b, const gg_swap *swaps, int K, gg_result *out) {
    if (!b || !swaps || !out || K < 1 || K > b->max_k || !b->have_base) return GG_BAD_ARG;
    int st = check_swaps(b, swaps, K);
    if (st != GG_OK) return st;
    GG_HIP(hipSetDevice(b->device));
    memcpy(b->h_swaps, swaps, sizeof(gg_swap) * K);
    bool sym = b->base_sym;
    for (int k = 0; k < K && sym; k++) sym = swap_keeps_symmetry(b->h_adj.data(), b->degree, swaps[k]);
    if ((b->dc.push_ok != 0) != sym) drop_graph(b);  // the captured launches bake the push choice
    b->dc.push_ok = sym;
    const bool sparse = b->sparse_ok && b->mode_opt != 1;
    GG_HIP(hipEventRecord(b->ev0, b->stream));
    if (sparse && b->use_graph) {
        if (b->graph_k != K) {
            drop_graph(b);
            GG_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            st = enqueue_step(b, K, true);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (st != GG_OK) { if (g) (void)hipGraphDestroy(g); return st; }
            GG_HIP(e);
            b->graph = g;
            GG_HIP(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
            b->graph_k = K;
        }
        GG_HIP(hipGraphLaunch(b->graph_exec, b->stream));
    } else {
        st = enqueue_step(b, K, sparse);
        if (st != GG_OK) return st;
    }
    GG_HIP(hipEventRecord(b->ev1, b->stream));
    GG_HIP(hipStreamSynchronize(b->stream));
    for (int k = 0; k < K; k++) fill(b, b->h_res[k], &out[k]);
    b->last_sparse = sparse ? 1 : 0;
    b->have_timing = 1;
    return GG_OK;
}

// The swap travels as a kernel argument: no host buffer outlives the launch.
__global__ void k_batch_commit(uint16_t *tab16, uint32_t *tab32, int *adjR, int N, int D,
                               gg_swap sw) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int j = 0; j < 4; j++) {
        int v = sw.v[j];
        if (v < 0) continue;
        int off = sw.nbr[j] - v;
        if (off < 0) off += N;
        if (tab16) tab16[quad_index(sw.slot[j], N, v)] = (uint16_t)sw.nbr[j];
        else tab32[(size_t)sw.slot[j] * N + v] = (uint32_t)off;
        adjR[(size_t)v * D + sw.slot[j]] = sw.nbr[j];
    }
}

// Applies the swap to the resident tables (host mirror and device).
static int apply_commit(gg_batch *b, const gg_swap *sw) {
    b->base_sym = b->base_sym && swap_keeps_symmetry(b->h_adj.data(), b->degree, *sw);
    for (int j = 0; j < 4; j++)
        if (sw->v[j] >= 0) b->h_adj[(size_t)sw->v[j] * b->degree + sw->slot[j]] = sw->nbr[j];
    b->dc.push_ok = b->base_sym;
    hipLaunchKernelGGL(k_batch_commit, dim3(1), dim3(64), 0, b->stream,
                       b->narrow ? (uint16_t *)b->d_tab : nullptr,
                       b->narrow ? nullptr : (uint32_t *)b->d_tab, b->d_adjR, b->N, b->degree,
                       *sw);
    GG_HIP(hipGetLastError());
    return GG_OK;
}

extern "C" int gg_batch_commit(gg_batch *b, const gg_swap *sw, gg_result *out) {
    if (!b || !sw || !b->have_base) return GG_BAD_ARG;
    int st = check_swaps(b, sw, 1);
    if (st != GG_OK) return st;
    GG_HIP(hipSetDevice(b->device));
    st = apply_commit(b, sw);
    if (st != GG_OK) return st;
    return eval_base(b, out);
}

// Commit of a candidate whose exact result the caller already holds (one of
// the last batch's): the base re-evaluation that refreshes the stored levels
// is queued on the batch stream, which orders every later launch after it,
// and not waited for.  Errors of that launch surface at the next call.
extern "C" int gg_batch_commit_known(gg_batch *b, const gg_swap *sw, const gg_result *known) {
    if (!b || !sw || !known || !b->have_base) return GG_BAD_ARG;
    int st = check_swaps(b, sw, 1);
    if (st != GG_OK) return st;
    GG_HIP(hipSetDevice(b->device));
    st = apply_commit(b, sw);
    if (st != GG_OK) return st;
    const int K = b->max_k;
    DenseArgs a = dense_args(b, 1, nullptr, b->d_parts + (size_t)K * b->W, b->d_res + K,
                             b->d_ecc);
    a.reach_col = b->d_base_reach;
    a.snap = b->d_snap;
    a.snap_levels = b->d_snap ? b->snap_cap : 0;
    st = run_dense(b->dc, a, b->stream);
    if (st != GG_OK) return st;
    ResultRec r{};
    r.total = known->total_reps;
    r.reached = known->status == GG_OK ? (long long)b->N * b->N : 0;
    r.diameter = known->diameter;
    r.levels = known->levels;
    r.flags = known->status == GG_DIAMETER_LIMIT ? GG_FLAG_LIMIT : 0;
    r.wvc = known->wvc;
    b->base_res = r;
    b->sparse_ok = b->d_snap && known->status == GG_OK && known->diameter - 1 <= b->snap_cap &&
                   b->dc.last_variant != 3;
    return GG_OK;
}

extern "C" int gg_batch_last_kernel_ms(gg_batch *b, float *ms) {
    if (!b || !ms || !b->have_timing) return GG_BAD_ARG;
    GG_HIP(hipEventElapsedTime(ms, b->ev0, b->ev1));
    return GG_OK;
}

extern "C" int gg_batch_base_ecc(gg_batch *b, int32_t *ecc_out) {
    if (!b || !ecc_out || !b->have_base) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(b->device));
    GG_HIP(hipMemcpyAsync(ecc_out, b->d_ecc, sizeof(int) * b->N, hipMemcpyDeviceToHost, b->stream));
    GG_HIP(hipStreamSynchronize(b->stream));
    return GG_OK;
}

extern "C" int gg_batch_set_option(gg_batch *b, int option, int value) {
    if (!b) return GG_BAD_ARG;
    switch (option) {
        case GG_BATCH_OPT_PATH:
            if (value < 0 || value > 2) return GG_BAD_ARG;
            b->mode_opt = value;
            return GG_OK;
        case GG_BATCH_OPT_GRAPH:
            b->use_graph = value ? 1 : 0;
            if (!value) drop_graph(b);
            return GG_OK;
        default: return GG_BAD_ARG;
    }
}

extern "C" int gg_batch_get_option(gg_batch *b, int option, int *value) {
    if (!b || !value) return GG_BAD_ARG;
    switch (option) {
        case GG_BATCH_OPT_PATH: *value = b->mode_opt; return GG_OK;
        case GG_BATCH_OPT_GRAPH: *value = b->use_graph; return GG_OK;
        case GG_BATCH_INFO_LAST_SPARSE: *value = b->last_sparse; return GG_OK;
        case GG_BATCH_INFO_SPARSE_OK: *value = b->sparse_ok; return GG_OK;
        default: return GG_BAD_ARG;
    }
}
