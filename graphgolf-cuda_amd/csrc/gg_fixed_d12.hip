// Fixed-degree (12) dense kernels for the quad-packed 16-bit ring topology.
#include "gg_fixed_impl.hpp"
namespace gg {
hipError_t launch_fixed_d12(const TopoRingQ &topo, const DenseArgs &a, int grid,
                            size_t lds, int vpt, bool queue, hipStream_t st) {
    return launch_fixed_deg<12>(topo, a, grid, lds, vpt, queue, st);
}
}  // namespace gg
