// Sparse batch kernels of degree 12 (with and without the single-source tail).
#include "gg_sparse.hpp"
namespace gg {
hipError_t launch_sparse_d12(const SparseArgs &s, size_t lds, int tail, hipStream_t st) {
    return launch_sparse_deg<12>(s, lds, tail, st);
}
}  // namespace gg
