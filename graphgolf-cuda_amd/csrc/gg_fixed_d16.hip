// Fixed-degree (16) dense kernels for the quad-packed 16-bit ring topology.
#include "gg_fixed_impl.hpp"
namespace gg {
hipError_t launch_fixed_d16(const TopoRingQ &topo, const DenseArgs &a, int grid,
                            size_t lds, int vpt, bool queue, hipStream_t st) {
    return launch_fixed_deg<16>(topo, a, grid, lds, vpt, queue, st);
}
}  // namespace gg
