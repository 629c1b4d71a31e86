// eik_riemann.hip -- kernel instantiations of the Riemann model (FP32 and FP64).
#include "eikonal_kernels.hpp"

namespace agd {

template <class S> bool lookup_riemann(int d, const int32_t* si, ModelKernels<S>& out) {
  return lookup_model_kernels<S, AGD_RIEMANN>(d, si, out);
}

template bool lookup_riemann<float>(int, const int32_t*, ModelKernels<float>&);
template bool lookup_riemann<double>(int, const int32_t*, ModelKernels<double>&);

}  // namespace agd
