// eik_rander.hip -- kernel instantiations of the Rander model (FP32 and FP64).
#include "eikonal_kernels.hpp"

namespace agd {

template <class S> bool lookup_rander(int d, const int32_t* si, ModelKernels<S>& out) {
  return lookup_model_kernels<S, AGD_RANDER>(d, si, out);
}

template bool lookup_rander<float>(int, const int32_t*, ModelKernels<float>&);
template bool lookup_rander<double>(int, const int32_t*, ModelKernels<double>&);

}  // namespace agd
