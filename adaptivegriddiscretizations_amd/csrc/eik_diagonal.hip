// eik_diagonal.hip -- kernel instantiations of the Diagonal model (FP32 and FP64).
#include "eikonal_kernels.hpp"

namespace agd {

template <class S> bool lookup_diagonal(int d, const int32_t* si, ModelKernels<S>& out) {
  return lookup_model_kernels<S, AGD_DIAGONAL>(d, si, out);
}

template bool lookup_diagonal<float>(int, const int32_t*, ModelKernels<float>&);
template bool lookup_diagonal<double>(int, const int32_t*, ModelKernels<double>&);

}  // namespace agd
