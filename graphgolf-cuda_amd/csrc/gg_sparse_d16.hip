// Sparse batch kernels of degree 16 (with and without the dense tail).
#include "gg_sparse.hpp"
namespace gg {
hipError_t launch_sparse_d16(const SparseArgs &s, const TopoRingQ &topo, const DenseArgs &qa,
                            size_t lds, int vpt, hipStream_t st) {
    return launch_sparse_deg<16>(s, topo, qa, lds, vpt, st);
}
}  // namespace gg
