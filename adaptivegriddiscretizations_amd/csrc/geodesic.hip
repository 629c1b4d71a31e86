// geodesic.hip -- geodesic backtracking from tips along the upwind geodesic
// flow read from the arrays of the flow pass (Kernel_GeodesicODE.h:116-413
// with the imported flow of GeodesicODE_Opt.h:222-238, driven by
// _GetGeodesics.py:9-147; the ODE is in geodesic_kernels.hpp, its variant
// with the flow recomputed is instantiated with the model kernels), and the
// capped Euclidean distance to the seeds that its PastSeed criterion reads
// (_GetGeodesics.py:50-67: inf_convolution.py:24-93, Kernel_InfConvolution.h).
//
// A geodesic is a sequential ODE: one lane integrates one geodesic with the
// midpoint rule, reading values, flow vectors and flow weight sums at the
// corners of the cell it is in (tiled layout, the solver's own buffers).
// The per-lane caches and short histories live in registers and private
// memory; the work is latency bound and tiny next to the solve.
#include <cmath>
#include <utility>

#include "common.hpp"
#include "device.hpp"
#include "geodesic_kernels.hpp"

namespace agd {

// ------------------------------------------------------------------ distance to the seeds
// eucl = 0 on the seed cells, 255 elsewhere (and in the tile padding).
template <int D>
__global__ void eucl_init_kernel(GeoGrid g, const uint32_t* seeds, uint8_t* out) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= g.size_tot) return;
  int y[3];
  const bool in = geo_pos<D>(n, g, y);
  const bool seed = (seeds[n >> 5] >> (n & 31)) & 1u;
  out[n] = (in && seed) ? uint8_t(0) : uint8_t(255);
}

// One min-plus pass over the 3^D neighbourhood with the integer kernel
// kern[m] (m = number of non-zero offset components), saturating at 255
// (Kernel_InfConvolution.h: Times with upper saturation, Plus = min).
template <int D>
__global__ void eucl_step_kernel(GeoGrid g, const uint8_t* in, uint8_t* out, int k1, int k2, int k3) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= g.size_tot) return;
  int y[3] = {0, 0, 0};
  if (!geo_pos<D>(n, g, y)) {
    out[n] = 255;
    return;
  }
  constexpr int NC = D == 2 ? 9 : 27;
  int r = 255;
  for (int c = 0; c < NC; ++c) {
    int z[3], m = 0, cc = c;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const int o = cc % 3 - 1;
      cc /= 3;
      z[k] = y[k] + o;
      m += o != 0;
    }
    const int64_t nz = geo_index<D>(z, g);
    if (nz < 0) continue;
    const int a = in[nz];
    const int b = m == 0 ? 0 : (m == 1 ? k1 : (m == 2 ? k2 : k3));
    const int prod = (b >= 255 - a) ? 255 : a + b;
    r = prod < r ? prod : r;
  }
  out[n] = uint8_t(r);
}

// Values beyond the target tolerance are dropped (_GetGeodesics.py:64).
__global__ void eucl_cut_kernel(int64_t size_tot, uint8_t* e, int bound) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n < size_tot && e[n] > bound) e[n] = 255;
}

}  // namespace agd

using namespace agd;

extern "C" {

int agd_geodesic_eucl(const agd_eikonal_desc* desc, const uint32_t* seeds, int32_t niter,
                      int32_t mult, int32_t bound, uint8_t* eucl, uint8_t* tmp, void* stream) {
  if (int st = check_geo_desc(desc)) return st;
  AGD_REQUIRE(seeds && eucl && (tmp || niter == 0), "missing buffers");
  AGD_REQUIRE(niter >= 0 && mult >= 0 && bound >= 0, "niter, mult and bound must be >= 0");
  const GeoGrid g = geo_grid(desc);
  hipStream_t s = as_stream(stream);
  const dim3 grid(unsigned(ceil_div(g.size_tot, 256))), block(256);
  // kernel values round(mult * |offset|), |offset|^2 = 1, 2, 3 (inf_convolution.py:24-30)
  const int k1 = int(std::nearbyint(mult * 1.)), k2 = int(std::nearbyint(mult * std::sqrt(2.))),
            k3 = int(std::nearbyint(mult * std::sqrt(3.)));
  // ping-pong so that the last pass writes `eucl`
  uint8_t* a = (niter % 2 == 0) ? eucl : tmp;
  uint8_t* b = (niter % 2 == 0) ? tmp : eucl;
  if (desc->ndim == 2) hipLaunchKernelGGL(eucl_init_kernel<2>, grid, block, 0, s, g, seeds, a);
  else hipLaunchKernelGGL(eucl_init_kernel<3>, grid, block, 0, s, g, seeds, a);
  for (int i = 0; i < niter; ++i) {
    if (desc->ndim == 2) hipLaunchKernelGGL(eucl_step_kernel<2>, grid, block, 0, s, g, a, b, k1, k2, k3);
    else hipLaunchKernelGGL(eucl_step_kernel<3>, grid, block, 0, s, g, a, b, k1, k2, k3);
    std::swap(a, b);
  }
  hipLaunchKernelGGL(eucl_cut_kernel, grid, block, 0, s, g.size_tot, eucl, bound);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_geodesic_ode(const agd_eikonal_desc* desc, const agd_geodesic_desc* gd, const void* values,
                     const void* flow_vector, const void* flow_weightsum, const uint8_t* eucl,
                     void* x, int32_t* len, int8_t* stop, void* stream) {
  if (int st = check_geo_desc(desc)) return st;
  int hlen = 0;
  if (int st = check_geodesic_desc(gd, hlen)) return st;
  AGD_REQUIRE(gd->chart_ndim_s <= desc->ndim, "chart_ndim_s must be in [0, ndim]");
  AGD_REQUIRE(values && flow_vector && flow_weightsum && eucl && x && len && stop, "missing buffers");
  if (gd->ngeo == 0) return AGD_OK;
  const GeoGrid g = geo_grid(desc);
  hipStream_t s = as_stream(stream);
  const dim3 grid(unsigned(ceil_div(gd->ngeo, 64))), block(64);
  if (desc->scalar == AGD_F32) {
    GeoArgs<float> a = geodesic_args<float>(g, gd, hlen, values, eucl, x, len, stop);
    a.fvec = static_cast<const float*>(flow_vector);
    a.wsum = static_cast<const float*>(flow_weightsum);
    if (desc->ndim == 2) hipLaunchKernelGGL((geodesic_kernel<float, 2>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((geodesic_kernel<float, 3>), grid, block, 0, s, a);
  } else {
    GeoArgs<double> a = geodesic_args<double>(g, gd, hlen, values, eucl, x, len, stop);
    a.fvec = static_cast<const double*>(flow_vector);
    a.wsum = static_cast<const double*>(flow_weightsum);
    if (desc->ndim == 2) hipLaunchKernelGGL((geodesic_kernel<double, 2>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((geodesic_kernel<double, 3>), grid, block, 0, s, a);
  }
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

}  // extern "C"
