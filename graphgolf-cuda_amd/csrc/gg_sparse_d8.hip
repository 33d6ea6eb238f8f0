// Sparse batch kernels of degree 8 (with and without the single-source tail).
#include "gg_sparse.hpp"
namespace gg {
hipError_t launch_sparse_d8(const SparseArgs &s, size_t lds, int tail, hipStream_t st) {
    return launch_sparse_deg<8>(s, lds, tail, st);
}
}  // namespace gg
