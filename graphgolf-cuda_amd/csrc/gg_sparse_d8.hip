// Sparse batch kernels of degree 8 (with and without the dense tail).
#include "gg_sparse.hpp"
namespace gg {
hipError_t launch_sparse_d8(const SparseArgs &s, const TopoRingQ &topo, const DenseArgs &qa,
                            size_t lds, int vpt, hipStream_t st) {
    return launch_sparse_deg<8>(s, topo, qa, lds, vpt, st);
}
}  // namespace gg
