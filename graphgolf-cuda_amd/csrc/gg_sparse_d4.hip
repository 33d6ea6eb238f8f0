// Sparse batch kernels of degree 4 (with and without the dense tail).
#include "gg_sparse.hpp"
namespace gg {
hipError_t launch_sparse_d4(const SparseArgs &s, const TopoRingQ &topo, const DenseArgs &qa,
                            size_t lds, int vpt, hipStream_t st) {
    return launch_sparse_deg<4>(s, topo, qa, lds, vpt, st);
}
}  // namespace gg
