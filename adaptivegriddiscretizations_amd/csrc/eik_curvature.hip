// eik_curvature.hip -- kernel instantiations of the curvature-penalised
// models on (x, y, theta) grids (FP32 and FP64; 3D tiles only).
#include "eikonal_kernels.hpp"

namespace agd {

template <class S, int MODEL> static bool lookup_curv3(int d, const int32_t* si, ModelKernels<S>& out) {
  if (d != 3) return false;
  switch (tile_key(3, si)) {
    case 40404: out = make_model_kernels<S, MODEL, 3, 4, 4, 4>(); return true;
    case 80804: out = make_model_kernels<S, MODEL, 3, 8, 8, 4>(); return true;
    case 40408: out = make_model_kernels<S, MODEL, 3, 4, 4, 8>(); return true;
  }
  return false;
}

template <class S> bool lookup_curvature(int model, int d, const int32_t* si, ModelKernels<S>& out) {
  switch (model) {
    case AGD_REEDSSHEPP2: return lookup_curv3<S, AGD_REEDSSHEPP2>(d, si, out);
    case AGD_REEDSSHEPPFWD2: return lookup_curv3<S, AGD_REEDSSHEPPFWD2>(d, si, out);
    case AGD_ELASTICA2: return lookup_curv3<S, AGD_ELASTICA2>(d, si, out);
    case AGD_DUBINS2: return lookup_curv3<S, AGD_DUBINS2>(d, si, out);
  }
  return false;
}

template bool lookup_curvature<float>(int, int, const int32_t*, ModelKernels<float>&);
template bool lookup_curvature<double>(int, int, const int32_t*, ModelKernels<double>&);

}  // namespace agd
