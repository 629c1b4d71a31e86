// chart.hip -- value pasting between the local charts of a manifold
// (ChartPaste.h:71-167, run by the stopping criterion of the iterative solver,
// _StopCriterion.py:60-80, after each converged solve).
//
// Two passes over the tiled grid, one lane per cell: the first interpolates,
// for every cell, the values at the point its chart identifies it with (into
// a scratch array); the second lowers each cell to that value where smaller,
// sets its trigger bit (so that a following solve starts from its tile) and
// counts the lowered cells.  Splitting the passes makes the result independent
// of the order in which lanes run (the reference pastes in place).  Both
// passes are HBM streams over the values; the interpolation's corner reads
// are gathers within a few cells of the mapped point.
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "device.hpp"
#include "geodesic_kernels.hpp"

namespace agd {

template <class S> struct ChartArgs {
  GeoGrid g;         // periodicity cleared: the paste does not wrap axes
  const S* u;
  const S* mapping;  // [ndim_s][size_s], continuous grid indices
  int32_t ndim_s;
  int64_t size_s;
  S* out;            // interpolated value per tiled cell, +inf where none
  int32_t sh[3];     // log2 of the tile shape per axis, when every axis is a power of two
};

// Position of tiled index n (< 2^31) in 32-bit arithmetic; false for the
// padding of partial tiles.  P2: tile sides are powers of two (a.sh), so the
// in-tile splits are shifts and masks.
template <int D, bool P2> __device__ inline bool chart_pos(uint32_t n, const GeoGrid& g, const int32_t* sh, int* x) {
  uint32_t n_o, n_i;
  if constexpr (P2) {
    const int shi = D == 2 ? sh[0] + sh[1] : sh[0] + sh[1] + sh[2];
    n_o = n >> shi;
    n_i = n & ((1u << shi) - 1u);
  } else {
    n_o = n / uint32_t(g.size_i);
    n_i = n % uint32_t(g.size_i);
  }
  bool in = true;
#pragma unroll
  for (int k = D - 1; k >= 0; --k) {
    const uint32_t so = uint32_t(g.shape_o[k]);
    if constexpr (P2) {
      x[k] = int(((n_o % so) << sh[k]) | (n_i & ((1u << sh[k]) - 1u)));
      n_i >>= sh[k];
    } else {
      const uint32_t si = uint32_t(g.shape_i[k]);
      x[k] = int((n_o % so) * si + n_i % si);
      n_i /= si;
    }
    n_o /= so;
    in = in && x[k] < g.shape[k];
  }
  return in;
}

// Tiled index of position y (no wrap: the paste clears periodicity), -1
// outside the grid; 32-bit, since the host rejects grids of 2^31 cells.
template <int D, bool P2> __device__ inline int chart_index(const int* y, const GeoGrid& g, const int32_t* sh) {
  uint32_t n_o = 0, n_i = 0;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    if (y[k] < 0 || y[k] >= g.shape[k]) return -1;
    const uint32_t v = uint32_t(y[k]);
    if constexpr (P2) {
      n_o = n_o * uint32_t(g.shape_o[k]) + (v >> sh[k]);
      n_i = (n_i << sh[k]) | (v & ((1u << sh[k]) - 1u));
    } else {
      n_o = n_o * uint32_t(g.shape_o[k]) + v / uint32_t(g.shape_i[k]);
      n_i = n_i * uint32_t(g.shape_i[k]) + v % uint32_t(g.shape_i[k]);
    }
  }
  return int(n_o * uint32_t(g.size_i) + n_i);
}

// ChartPaste.h:99-149 for the cell at tiled index n.
template <class S, int D, bool P2> __global__ void chart_eval_kernel(ChartArgs<S> a) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= a.g.size_tot) return;
  const S inf = inf_v<S>();
  int x[3] = {0, 0, 0};
  if (!chart_pos<D, P2>(uint32_t(n), a.g, a.sh, x)) {
    a.out[n] = inf;
    return;
  }
  const int nb = D - a.ndim_s;
  // index of x among the cells of the last ndim_s axes
  int64_t n_s = 0;
  for (int i = nb; i < D; ++i) n_s = n_s * a.g.shape[i] + x[i];
  const S tol = S(1e-4);  // boundary_tol
  int q[3];
  S r[3];
  for (int i = 0; i < a.ndim_s; ++i) {
    const S m = a.mapping[int64_t(i) * a.size_s + n_s];
    const int qmax = a.g.shape[nb + i] - 1;
    // NaN, or far outside the grid (kept out of the integer conversion)
    if (!(m >= S(-2) && m <= S(qmax + 2))) {
      a.out[n] = inf;
      return;
    }
    const S f = floor(m);
    q[i] = int(f);
    r[i] = m - f;
    // points just outside the last layer are moved onto it
    if (q[i] == qmax && r[i] <= tol) {
      q[i] -= 1;
      r[i] = S(1);
    }
    if (q[i] == -1 && r[i] >= S(1) - tol) {
      q[i] += 1;
      r[i] = S(0);
    }
    if (q[i] >= qmax || q[i] < 0) {
      a.out[n] = inf;
      return;
    }
  }
  // multilinear interpolation over the 2^ndim_s corners (a corner of zero
  // weight and infinite value gives NaN, hence no paste, as in the reference)
  S val = S(0), wsum = S(0);
  int y[3] = {x[0], x[1], x[2]};
  const int ncorner = 1 << a.ndim_s;
  for (int c = 0; c < ncorner; ++c) {
    S w = S(1);
    for (int i = 0; i < a.ndim_s; ++i) {
      const int e = (c >> i) & 1;
      y[nb + i] = q[i] + e;
      w *= e ? r[i] : S(1) - r[i];
    }
    const int ny = chart_index<D, P2>(y, a.g, a.sh);
    if (ny < 0) continue;
    val += w * a.u[ny];
    wsum += w;
  }
  a.out[n] = wsum > S(0) ? val / wsum : inf;
}

// ChartPaste.h:151-165: lower the cells whose pasted value is smaller.
// Grid-stride over 256-cell slices, so that the count of lowered cells takes
// one global atomic per block (a paste that lowers whole regions otherwise
// serialises one atomic per wave on that one word).
template <class S>
__global__ void chart_apply_kernel(int64_t size_tot, S* u, const S* cand, S atol, S rtol, uint32_t* trigger,
                                   int32_t* count) {
  __shared__ int32_t block_count;
  if (threadIdx.x == 0) block_count = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  int32_t lowered = 0;  // by this wave
  // the bound is uniform over the block: every lane runs every iteration
  for (int64_t base = int64_t(blockIdx.x) * blockDim.x; base < size_tot; base += int64_t(gridDim.x) * blockDim.x) {
    const int64_t n = base + threadIdx.x;
    bool upd = false;
    if (n < size_tot) {
      const S c = cand[n];
      if (c < u[n] - (atol + rtol * abs_(c))) {  // false for NaN and +inf
        u[n] = c;
        upd = true;
      }
    }
    // One atomic per trigger word: a wave's first lane sits at a multiple of
    // 64, so its lanes 0-31 and 32-63 own one 32-bit word each, and bit L of
    // the ballot is the bit of lane L.
    const unsigned long long bal = __ballot(upd);
    if (bal) {
      if (lane == 0 && uint32_t(bal)) atomicOr(trigger + (n >> 5), uint32_t(bal));
      if (lane == 32 && uint32_t(bal >> 32)) atomicOr(trigger + (n >> 5), uint32_t(bal >> 32));
      lowered += int32_t(__popcll(bal));
    }
  }
  if (lane == 0 && lowered) atomicAdd(&block_count, lowered);
  __syncthreads();
  if (threadIdx.x == 0 && block_count) atomicAdd(count, block_count);
}

template <class S>
int chart_paste_impl(const agd_eikonal_desc* d, void* values, const void* mapping, int32_t ndim_s, double atol,
                     double rtol, void* scratch, uint32_t* trigger, int32_t* count, hipStream_t s) {
  ChartArgs<S> a{};
  a.g = geo_grid(d);
  a.g.periodic = 0;
  a.u = static_cast<const S*>(values);
  a.mapping = static_cast<const S*>(mapping);
  a.ndim_s = ndim_s;
  a.size_s = 1;
  for (int i = d->ndim - ndim_s; i < d->ndim; ++i) a.size_s *= d->shape[i];
  a.out = static_cast<S*>(scratch);
  bool p2 = true;
  for (int k = 0; k < 3; ++k) {
    const int t = k < d->ndim ? d->shape_i[k] : 1;
    int l = 0;
    while ((int64_t(1) << l) < t) ++l;
    a.sh[k] = l;
    p2 = p2 && (int64_t(1) << l) == t;
  }
  const dim3 grid(unsigned(ceil_div(a.g.size_tot, 256))), block(256);
  if (d->ndim == 2) {
    if (p2) hipLaunchKernelGGL((chart_eval_kernel<S, 2, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((chart_eval_kernel<S, 2, false>), grid, block, 0, s, a);
  } else {
    if (p2) hipLaunchKernelGGL((chart_eval_kernel<S, 3, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((chart_eval_kernel<S, 3, false>), grid, block, 0, s, a);
  }
  const dim3 grid_apply(unsigned(std::min<int64_t>(ceil_div(a.g.size_tot, 256), 2048)));
  hipLaunchKernelGGL(chart_apply_kernel<S>, grid_apply, block, 0, s, a.g.size_tot, static_cast<S*>(values),
                     static_cast<const S*>(scratch), S(atol), S(rtol), trigger, count);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

}  // namespace agd

using namespace agd;

extern "C" {

int agd_chart_paste(const agd_eikonal_desc* desc, void* values, const void* mapping, int32_t ndim_s,
                    double atol, double rtol, void* scratch, uint32_t* trigger, int32_t* count,
                    void* stream) {
  if (int st = check_geo_desc(desc)) return st;
  AGD_REQUIRE(geo_grid(desc).size_tot < (int64_t(1) << 31), "grids of 2^31 tiled cells or more are not supported");
  AGD_REQUIRE(ndim_s >= 1 && ndim_s <= desc->ndim, "ndim_s must be in [1, ndim]");
  AGD_REQUIRE(atol >= 0. && rtol >= 0., "atol and rtol must be >= 0");
  AGD_REQUIRE(values && mapping && scratch && trigger && count, "missing buffers");
  if (desc->scalar == AGD_F32)
    return chart_paste_impl<float>(desc, values, mapping, ndim_s, atol, rtol, scratch, trigger, count,
                                   as_stream(stream));
  return chart_paste_impl<double>(desc, values, mapping, ndim_s, atol, rtol, scratch, trigger, count,
                                  as_stream(stream));
}

}  // extern "C"
