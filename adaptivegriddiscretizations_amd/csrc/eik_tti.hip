// eik_tti.hip -- kernel instantiations of the TTI model (FP32 and FP64).
#include "eikonal_kernels.hpp"

namespace agd {

template <class S> bool lookup_tti(int d, const int32_t* si, ModelKernels<S>& out) {
  return lookup_model_kernels<S, AGD_TTI>(d, si, out);
}

template bool lookup_tti<float>(int, const int32_t*, ModelKernels<float>&);
template bool lookup_tti<double>(int, const int32_t*, ModelKernels<double>&);

}  // namespace agd
