// Sparse batch kernels of degree 12 (with and without the dense tail).
#include "gg_sparse.hpp"
namespace gg {
hipError_t launch_sparse_d12(const SparseArgs &s, const TopoRingQ &topo, const DenseArgs &qa,
                            size_t lds, int vpt, hipStream_t st) {
    return launch_sparse_deg<12>(s, topo, qa, lds, vpt, st);
}
}  // namespace gg
