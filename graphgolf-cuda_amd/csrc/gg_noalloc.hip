// gg_noalloc.hip -- the part evaluator (gg_create_part + gg_eval_part) as one
// call that allocates nothing, for callers that cannot call hipMalloc.
//
// Every buffer the evaluation needs (the raw offset table, its packed form,
// the source map, the per-column partial results, the per-source
// eccentricities and the result record) is carved out of `scratch`, a
// file-scope device array of kScratchBytes.  Each GPU holds its own copy of
// the array for the life of the process; its address on a device is taken
// once with hipGetSymbolAddress while that device is current and only that
// address is handed to kernels and copies.  Every call on a device reuses the
// same array, so calls that overlap on one device race; no region is read
// before the same call writes it, so nothing is cleared between calls.
#include <mutex>
#include "gg_common.hpp"
#include "gg_internal.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(4) << 20;
constexpr int kMaxN = 16384;                      // column bitset of N words fits LDS
constexpr long long kMaxEntries = 1LL << 19;      // M * degree
constexpr size_t kAlign = 256;

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

struct Layout {
    size_t raw, packed, src, parts, ecc, res, end;
};

Layout layout(int M, int D) {
    Layout l{};
    const int W = (M + 63) / 64;
    l.raw = 0;
    l.packed = align_up(l.raw + (size_t)M * D * sizeof(int32_t));
    l.src = align_up(l.packed + (size_t)gg::quads(D) * M * 8);
    l.parts = align_up(l.src + (size_t)M * sizeof(int));
    l.ecc = align_up(l.parts + (size_t)W * sizeof(ColPart));
    l.res = align_up(l.ecc + (size_t)M * sizeof(int));
    l.end = align_up(l.res + sizeof(ResultRec));
    return l;
}

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` per device (null until the first call there).
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_dev[gg::kMaxDevices];

// Must run with `device` current.
static int scratch_address(int device, unsigned char **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            gg_set_last_hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol,
                                  "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
            return GG_HIP_ERROR;
        }
        g_scratch_dev[device] = (unsigned char *)p;
    }
    *out = g_scratch_dev[device];
    return GG_OK;
}

// Packs the reference's offset table (M rows of D int32 offsets) into the
// 16-bit quad layout of TopoRingQ and writes the identity source map; the
// host packing of the handle-based evaluator does the same on the CPU.
__global__ void __launch_bounds__(256) k_pack_part_noalloc(const int32_t *raw, int N, int M, int D,
                                                           uint16_t *packed, int *src) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t slots = (size_t)((D + 3) / 4) * 4 * M;
    if (i < (size_t)M) src[i] = (int)i;
    if (i >= slots) return;
    const int slot = (int)(i / (4 * (size_t)M)) * 4 + (int)(i & 3);
    const int row = (int)((i >> 2) % (size_t)M);
    uint16_t v = 0;  // padding slots: offset 0 (a self-loop), never read
    if (slot < D) {
        long long off = (long long)raw[(size_t)row * D + slot] % N;
        if (off < 0) off += N;
        if (M == N) off = (row + off) % N;  // every vertex its own row: neighbour ids
        v = (uint16_t)off;
    }
    packed[gg::quad_index(slot, M, row)] = v;
}

extern "C" int gg_eval_part_noalloc(int N, int M, int degree, int device, const int32_t *table,
                                    gg_result *out) {
    using namespace gg;
    if (!out || N < 1 || M < 1 || M > N || degree < 0 || (!table && degree > 0))
        return GG_BAD_ARG;
    if (N > kMaxN || (long long)M * degree > kMaxEntries) return GG_SCRATCH_LIMIT;
    const Layout l = layout(M, degree);
    if (l.end > kScratchBytes) return GG_SCRATCH_LIMIT;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return GG_NO_DEVICE;
    if (device < 0 || device >= n || device >= kMaxDevices) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(device));
    unsigned char *base = nullptr;
    int st = scratch_address(device, &base);
    if (st != GG_OK) return st;

    DenseCtx c;
    c.topo_kind = TOPO_RINGQ;
    c.lds_device = device_lds_limit(device);
    c.cu_count = device_cu_count(device);
    c.wide_ok = 0;  // only the LDS kernels: the other paths allocate their scratch
    c.wide_off = 1;
    if (!dense_fits_lds(c, N)) return GG_SCRATCH_LIMIT;

    int32_t *raw = (int32_t *)(base + l.raw);
    uint16_t *packed = (uint16_t *)(base + l.packed);
    int *src = (int *)(base + l.src);
    const hipStream_t stream = nullptr;  // the device's null stream
    if (degree > 0)
        GG_HIP(hipMemcpy(raw, table, (size_t)M * degree * sizeof(int32_t), hipMemcpyHostToDevice));
    const size_t slots = (size_t)quads(degree) * 4 * M;
    const size_t work = slots > (size_t)M ? slots : (size_t)M;
    hipLaunchKernelGGL(k_pack_part_noalloc, dim3((unsigned)((work + 255) / 256)), dim3(256), 0,
                       stream, raw, N, M, degree, packed, src);
    GG_HIP(hipGetLastError());

    c.ringq = TopoRingQ{(const uint64_t *)packed, (unsigned)N, (unsigned)M};
    DenseArgs a{};
    a.N = N; a.M = M; a.D = degree; a.W = (M + 63) / 64; a.K = 1;
    a.lmax = GG_LEVEL_LIMIT;
    a.src_vertex = src;
    a.parts = (ColPart *)(base + l.parts);
    a.res = (ResultRec *)(base + l.res);
    a.ecc_out = (int *)(base + l.ecc);
    a.want_wvc = 1;  // as gg_create_part: WVCdiameterASPL's figure
    st = run_dense(c, a, stream);
    if (st != GG_OK) return st;
    ResultRec r;
    GG_HIP(hipMemcpy(&r, a.res, sizeof r, hipMemcpyDeviceToHost));
    out->total_reps = r.total;
    out->total_scaled = r.total * (long long)(N / M);
    out->diameter = r.diameter;
    out->wvc = r.wvc;
    out->levels = r.levels;
    if (r.flags & GG_FLAG_LIMIT) out->status = GG_DIAMETER_LIMIT;
    else if (r.reached != (long long)N * M) out->status = GG_UNCONNECTED;
    else out->status = GG_OK;
    return GG_OK;
}
