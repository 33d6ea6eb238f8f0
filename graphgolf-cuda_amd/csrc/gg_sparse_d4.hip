// Sparse batch kernels of degree 4 (with and without the single-source tail).
#include "gg_sparse.hpp"
namespace gg {
hipError_t launch_sparse_d4(const SparseArgs &s, size_t lds, int tail, hipStream_t st) {
    return launch_sparse_deg<4>(s, lds, tail, st);
}
}  // namespace gg
