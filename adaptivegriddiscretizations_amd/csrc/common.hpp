// common.hpp -- host-side error handling shared by the library's translation
// units.  Every C-ABI entry point returns an agd_status and records a
// human-readable message retrievable with agd_last_error().
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "../../include/agd_hip.h"

namespace agd {

void set_error(const std::string& msg);

#define AGD_HIP_TRY(expr)                                                          \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::agd::set_error(std::string(#expr) + " failed: " + hipGetErrorString(e_) + \
                       " (" __FILE__ ":" + std::to_string(__LINE__) + ")");        \
      return AGD_EHIP;                                                             \
    }                                                                              \
  } while (0)

#define AGD_REQUIRE(cond, msg)                      \
  do {                                              \
    if (!(cond)) {                                  \
      ::agd::set_error(std::string("invalid argument: ") + (msg)); \
      return AGD_EINVAL;                            \
    }                                               \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Geometry components per model (agd_model) and dimension.
inline int ngeom_of(int model, int d) {
  const int sd = d * (d + 1) / 2;
  switch (model) {
    case AGD_ISOTROPIC: return 0;
    case AGD_DIAGONAL: return d;
    case AGD_RIEMANN: return sd;
    case AGD_RANDER: return sd + d;
    case AGD_TTI: return 5 + d * d;
    case AGD_REEDSSHEPP2:
    case AGD_REEDSSHEPPFWD2:
    case AGD_ELASTICA2:
    case AGD_DUBINS2: return 4;
  }
  return -1;
}
inline bool curvature_model(int model) { return model >= AGD_REEDSSHEPP2 && model <= AGD_DUBINS2; }
inline int nmix_of(int model) { return model == AGD_TTI ? 7 : (model == AGD_DUBINS2 ? 2 : 1); }
// Offsets of one sub-scheme, each stored as a pair of neighbour codes (the
// forward offsets of the curvature models included).
inline int nsym_of(int model, int d) {
  switch (model) {
    case AGD_ISOTROPIC:
    case AGD_DIAGONAL: return d;
    case AGD_REEDSSHEPP2: return 6;
    case AGD_REEDSSHEPPFWD2: return 7;
    case AGD_ELASTICA2: return 30;
    case AGD_DUBINS2: return 6;
  }
  return d * (d + 1) / 2;
}
// Models whose scheme is always precomputed (agd_eikonal_scheme).
inline bool stored_scheme_model(int model) { return model == AGD_TTI || curvature_model(model); }

}  // namespace agd
