// gg_probe.hip -- LDS read-bandwidth probe: the measured ceiling against which
// the LDS-resident BFS kernel's roofline is quoted.
//
// Every workgroup (1024 threads) fills 64 KiB of LDS, then each wave reads
// 64 consecutive elements per instruction (conflict-free), eight distinct
// windows per iteration, moving every iteration so no load can be hoisted.
// Bandwidth = bytes read by all workgroups / kernel time (HIP events); the
// shader clock during the run comes from the cycle counter against the
// constant-rate wall clock of workgroup 0, which gives bytes per clock per CU.
#include "gg_common.hpp"

namespace {

constexpr int kProbeBytes = 64 * 1024;
constexpr int kProbeThreads = 1024;

__device__ __forceinline__ unsigned int fold(uint4 v) { return v.x ^ v.y ^ v.z ^ v.w; }
__device__ __forceinline__ unsigned int fold(uint2 v) { return v.x ^ v.y; }
__device__ __forceinline__ unsigned int fold(unsigned int v) { return v; }

template <typename T>
__global__ void __launch_bounds__(kProbeThreads) k_lds_read(int iters, unsigned int *out,
                                                             long long *clk) {
    __shared__ uint4 buf[kProbeBytes / 16];
    const int tid = threadIdx.x;
    for (int i = tid; i < kProbeBytes / 16; i += kProbeThreads)
        buf[i] = make_uint4(i, 3 * i + 1, 5 * i + 2, 7 * i + 3);
    __syncthreads();
    const long long c0 = clock64(), w0 = wall_clock64();
    const T *b = reinterpret_cast<const T *>(buf);
    constexpr int NW = kProbeBytes / (int)sizeof(T);  // power of two
    unsigned int acc = 0;
    for (int it = 0; it < iters; it++) {
        const int base = tid + it * 64;
#pragma unroll
        for (int u = 0; u < 8; u++) acc ^= fold(b[(base + u * (NW / 8)) & (NW - 1)]);
    }
    __syncthreads();
    out[(size_t)blockIdx.x * kProbeThreads + tid] = acc;
    if (blockIdx.x == 0 && tid == 0) {
        clk[0] = clock64() - c0;
        clk[1] = wall_clock64() - w0;
    }
}

template <typename T>
int timed_launches(int blocks, int iters, hipStream_t st, unsigned int *d_out, long long *d_clk,
                   hipEvent_t e0, hipEvent_t e1, float *ms) {
    hipLaunchKernelGGL(k_lds_read<T>, dim3(blocks), dim3(kProbeThreads), 0, st, 8, d_out, d_clk);
    GG_HIP(hipGetLastError());
    GG_HIP(hipEventRecord(e0, st));
    hipLaunchKernelGGL(k_lds_read<T>, dim3(blocks), dim3(kProbeThreads), 0, st, iters, d_out,
                       d_clk);
    GG_HIP(hipGetLastError());
    GG_HIP(hipEventRecord(e1, st));
    GG_HIP(hipEventSynchronize(e1));
    GG_HIP(hipEventElapsedTime(ms, e0, e1));
    return GG_OK;
}

// A warm-up launch, then `iters` iterations timed between two events.
template <typename T>
int run_probe(int blocks, int iters, hipStream_t st, unsigned int *d_out, long long *d_clk,
              float *ms) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = GG_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = GG_HIP_ERROR;
    if (rc == GG_OK) rc = timed_launches<T>(blocks, iters, st, d_out, d_clk, e0, e1, ms);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

}  // namespace

extern "C" int gg_probe_lds_read(int device, int bytes_per_read, gg_lds_probe *out) {
    if (!out || (bytes_per_read != 4 && bytes_per_read != 8 && bytes_per_read != 16))
        return GG_BAD_ARG;
    int n = 0;
    GG_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return GG_BAD_ARG;
    GG_HIP(hipSetDevice(device));
    int cus = 0, wall_khz = 0;
    GG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    GG_HIP(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device));
    const int blocks = cus * 16, iters = 2048;
    unsigned int *d_out = nullptr;
    long long *d_clk = nullptr, h_clk[2] = {0, 0};
    hipStream_t st = nullptr;
    int rc = GG_OK;
    float ms = 0;
    if (hipMalloc((void **)&d_out, sizeof(unsigned int) * (size_t)blocks * kProbeThreads) !=
            hipSuccess ||
        hipMalloc((void **)&d_clk, sizeof(h_clk)) != hipSuccess)
        rc = GG_OOM;
    if (rc == GG_OK && hipStreamCreate(&st) != hipSuccess) rc = GG_HIP_ERROR;
    if (rc == GG_OK) {
        if (bytes_per_read == 16) rc = run_probe<uint4>(blocks, iters, st, d_out, d_clk, &ms);
        else if (bytes_per_read == 8) rc = run_probe<uint2>(blocks, iters, st, d_out, d_clk, &ms);
        else rc = run_probe<unsigned int>(blocks, iters, st, d_out, d_clk, &ms);
    }
    if (rc == GG_OK &&
        hipMemcpy(h_clk, d_clk, sizeof(h_clk), hipMemcpyDeviceToHost) != hipSuccess)
        rc = GG_HIP_ERROR;
    if (st) (void)hipStreamDestroy(st);
    if (d_out) (void)hipFree(d_out);
    if (d_clk) (void)hipFree(d_clk);
    if (rc != GG_OK) return rc;
    const double bytes = (double)blocks * kProbeThreads * iters * 8.0 * bytes_per_read;
    out->gbs = bytes / (ms * 1e-3) / 1e9;
    out->kernel_ms = ms;
    out->cus = cus;
    out->clock_mhz = (h_clk[1] > 0 && wall_khz > 0)
                         ? (double)h_clk[0] / ((double)h_clk[1] / (wall_khz * 1e3)) / 1e6
                         : 0.0;
    out->bytes_per_clk_cu =
        out->clock_mhz > 0 ? out->gbs * 1e9 / (out->clock_mhz * 1e6 * cus) : 0.0;
    return GG_OK;
}
