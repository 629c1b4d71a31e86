// eik_isotropic.hip -- kernel instantiations of the Isotropic model (FP32 and FP64).
#include "eikonal_kernels.hpp"

namespace agd {

template <class S> bool lookup_isotropic(int d, const int32_t* si, ModelKernels<S>& out) {
  return lookup_model_kernels<S, AGD_ISOTROPIC>(d, si, out);
}

template bool lookup_isotropic<float>(int, const int32_t*, ModelKernels<float>&);
template bool lookup_isotropic<double>(int, const int32_t*, ModelKernels<double>&);

}  // namespace agd
