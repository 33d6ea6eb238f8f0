// gg_fixed_impl.hpp -- launchers of the fixed-degree dense kernels for the
// quad-packed 16-bit ring topology (part / adjacency graphs with N <= 65535).  Each degree
// is instantiated in its own translation unit (gg_fixed_d*.hip) so the
// unrolled kernels compile in parallel.
#pragma once
#include "gg_dense.hpp"

namespace gg {


template <int VPT, int DT>
static hipError_t launch_fixed_inst(const TopoRingQ &topo, const DenseArgs &a, int grid,
                                    size_t lds, bool queue, hipStream_t st) {
    static LdsAttr cfg_lds, cfg_queue;
    hipError_t e;
    if (queue) {
        auto fn = &k_dense_lds_queue<TopoRingQ, 1024, VPT, DT>;
        if ((e = ensure_dynamic_lds((const void *)fn, lds, cfg_queue)) != hipSuccess) return e;
        hipLaunchKernelGGL(fn, dim3(grid), dim3(1024), lds, st, topo, a);
    } else {
        auto fn = &k_dense_lds<TopoRingQ, 1024, VPT, DT>;
        if ((e = ensure_dynamic_lds((const void *)fn, lds, cfg_lds)) != hipSuccess) return e;
        hipLaunchKernelGGL(fn, dim3(grid), dim3(1024), lds, st, topo, a);
    }
    return hipGetLastError();
}

template <int DT>
static hipError_t launch_fixed_deg(const TopoRingQ &topo, const DenseArgs &a, int grid,
                                   size_t lds, int vpt, bool queue, hipStream_t st) {
    switch (vpt) {
        case 1: return launch_fixed_inst<1, DT>(topo, a, grid, lds, queue, st);
        case 2: return launch_fixed_inst<2, DT>(topo, a, grid, lds, queue, st);
        case 4: return launch_fixed_inst<4, DT>(topo, a, grid, lds, queue, st);
        case 8: return launch_fixed_inst<8, DT>(topo, a, grid, lds, queue, st);
        case 16: return launch_fixed_inst<16, DT>(topo, a, grid, lds, queue, st);
        case 20: return launch_fixed_inst<20, DT>(topo, a, grid, lds, queue, st);
        default: return hipErrorInvalidValue;
    }
}

// Defined in gg_fixed_d*.hip.
hipError_t launch_fixed_d4(const TopoRingQ &, const DenseArgs &, int, size_t, int, bool,
                           hipStream_t);
hipError_t launch_fixed_d8(const TopoRingQ &, const DenseArgs &, int, size_t, int, bool,
                           hipStream_t);
hipError_t launch_fixed_d12(const TopoRingQ &, const DenseArgs &, int, size_t, int, bool,
                            hipStream_t);
hipError_t launch_fixed_d16(const TopoRingQ &, const DenseArgs &, int, size_t, int, bool,
                            hipStream_t);
hipError_t launch_fixed_d24(const TopoRingQ &, const DenseArgs &, int, size_t, int, bool,
                            hipStream_t);

}  // namespace gg
