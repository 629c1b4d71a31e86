// geodesic_kernels.hpp -- the geodesic ODE of the backtracking from tips
// (Kernel_GeodesicODE.h:116-413): one lane integrates one geodesic with the
// midpoint rule along the causally filtered, normalised interpolation of the
// flow at the corners of the cell it is in.  The flow at a cell comes from a
// source: the arrays exported by the flow pass (geodesic.hip), or the scheme
// itself, recomputed at each cell the ODE enters (geodesic_online_flow,
// GeodesicODE_Opt.h:101-238; instantiated with the model kernels).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"
#include "device.hpp"

namespace agd {

// Tiled grid with a run-time tile shape.
struct GeoGrid {
  int32_t shape[3];
  int32_t shape_i[3];
  int32_t shape_o[3];
  int32_t periodic;  // bit k: axis k periodic
  int32_t size_i;
  int64_t size_tot;
};

// Grid.h:15-35 (InRange_per, Index_tot): tiled index of position y (wrapped
// along periodic axes), -1 outside the grid.
template <int D> __device__ inline int64_t geo_index(const int* y_in, const GeoGrid& g) {
  int64_t n_o = 0, n_i = 0;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    int y = y_in[k];
    if ((g.periodic >> k) & 1) y = mod_pos(y, g.shape[k]);
    else if (y < 0 || y >= g.shape[k]) return -1;
    n_o = n_o * g.shape_o[k] + y / g.shape_i[k];
    n_i = n_i * g.shape_i[k] + y % g.shape_i[k];
  }
  return n_o * g.size_i + n_i;
}

// Position of tiled index n; false for the padding of partial tiles.
template <int D> __device__ inline bool geo_pos(int64_t n, const GeoGrid& g, int* y) {
  int64_t n_o = n / g.size_i, n_i = n % g.size_i;
  bool in = true;
#pragma unroll
  for (int k = D - 1; k >= 0; --k) {
    const int qo = int(n_o % g.shape_o[k]), qi = int(n_i % g.shape_i[k]);
    n_o /= g.shape_o[k];
    n_i /= g.shape_i[k];
    y[k] = qo * g.shape_i[k] + qi;
    in = in && y[k] < g.shape[k];
  }
  return in;
}

// ------------------------------------------------------------------ ODE
constexpr int kGeoHist = 256;  // largest history length 1 + max(delays)

enum GeoStop : int8_t {
  kGeoContinue = 0, kGeoAtSeed = 1, kGeoInWall = 2, kGeoStationnary = 3, kGeoPastSeed = 4,
  kGeoVanishingFlow = 5,
};

template <class S> struct GeoArgs {
  GeoGrid g;
  const S* u;
  const S* fvec;  // D components (flow read from arrays; null when recomputed)
  const S* wsum;
  const uint8_t* eucl;
  S* x;           // ngeo x max_len x D; row start holds the starting point
  int32_t* len;
  int8_t* stop;
  int32_t ngeo, max_len, hlen, eucl_delay, nymin_delay, pastseed_delay;
  S step, causality;
  // charts: mapping [ndim_s][n_s] of the last ndim_s axes (null: no charts)
  const S* chart;
  int32_t chart_ndim_s;
  S chart_var;
};

template <class S, int D> struct GeoCell {
  int xq[D];
  S flow[1 << D][D];
  S dist[1 << D];
  S thr;
  int64_t nymin;
};

// NormalizedFlow (Kernel_GeodesicODE.h:160-260): multilinear interpolation
// of the flow at the corners of the cell of x, leaving out corners whose
// value reaches the causality threshold set when the cell was entered
// (smallest value among the corners of weight >= 1/(2 * 2^D), plus
// causality / its flow weight sum), then Euclidean normalisation.
// Flow read from the arrays of agd_eikonal_flow (the default).
template <class S, int D> struct GeoFlowArrays {
  const GeoArgs<S>* a;
  __device__ void at(int64_t n, S (&vec)[D], S& ws) const {
    ws = a->wsum[n];
#pragma unroll
    for (int k = 0; k < D; ++k) vec[k] = a->fvec[k * a->g.size_tot + n];
  }
};

template <class S, int D, class Src>
__device__ int geo_flow(const GeoArgs<S>& a, const Src& src, const S (&x)[D], S (&flow)[D],
                        GeoCell<S, D>& c) {
  constexpr int NC = 1 << D;
  const S inf = inf_v<S>();
  bool fresh = false;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    constexpr S big = S(1 << 30);  // keeps the conversion defined far outside the grid
    S f = floor(x[k]);
    f = f < -big ? -big : (f > big ? big : f);
    const int q = int(f);
    fresh = fresh || q != c.xq[k];
    c.xq[k] = q;
  }
  S dx[D];
#pragma unroll
  for (int k = 0; k < D; ++k) dx[k] = x[k] - S(c.xq[k]);
  S w[NC];
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    w[i] = S(1);
#pragma unroll
    for (int k = 0; k < D; ++k) w[i] *= ((i >> k) & 1) ? dx[k] : S(1) - dx[k];
  }
  if (fresh) {
    S ws[NC];
    int64_t nc[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      int yq[3];
#pragma unroll
      for (int k = 0; k < D; ++k) yq[k] = c.xq[k] + ((i >> k) & 1);
      nc[i] = geo_index<D>(yq, a.g);
      ws[i] = S(0);
      if (nc[i] < 0) {
        c.dist[i] = inf;
        continue;
      }
      c.dist[i] = a.u[nc[i]];
      src.at(nc[i], c.flow[i], ws[i]);
    }
    const S wthr = S(0.5) / S(NC);
    S dmin = inf;
    int imin = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      if (w[i] < wthr) continue;
      if (c.dist[i] < dmin) {
        imin = i;
        dmin = c.dist[i];
      }
    }
    if (dmin == inf) return kGeoInWall;
    S wsmin = ws[0];
    int64_t nmin = nc[0];
#pragma unroll
    for (int i = 1; i < NC; ++i) {
      if (i == imin) {
        wsmin = ws[i];
        nmin = nc[i];
      }
    }
    c.nymin = nmin;
    if (wsmin == S(0)) return kGeoAtSeed;
    c.thr = dmin + a.causality / wsmin;
  }
#pragma unroll
  for (int k = 0; k < D; ++k) flow[k] = S(0);
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    if (c.dist[i] >= c.thr) continue;
#pragma unroll
    for (int k = 0; k < D; ++k) flow[k] += w[i] * c.flow[i][k];
  }
  S n2 = S(0);
#pragma unroll
  for (int k = 0; k < D; ++k) n2 += flow[k] * flow[k];
  const S nrm = sqrt(n2);
  if (!(nrm > S(0))) return kGeoVanishingFlow;
#pragma unroll
  for (int k = 0; k < D; ++k) flow[k] /= nrm;
  return kGeoContinue;
}

// ChartJump (GeodesicODE_Opt.h:43-99): the last ndim_s coordinates of x
// replaced by the multilinear interpolation of the chart mapping at the
// corners of their cell (clamped to the grid), or, where the corners' mapped
// points are spread (summed variance >= chart_var), by the mapping of the
// nearest corner.  A NaN mapping leaves x unchanged.
template <class S, int D> __device__ inline void chart_jump(const GeoArgs<S>& a, S (&x)[D]) {
  const int ns = a.chart_ndim_s, nb = D - ns;
  int64_t size_s = 1;
  for (int i = nb; i < D; ++i) size_s *= a.g.shape[i];
  int q[D];
  S r[D], xs[D], sum[D], sq[D];
  for (int i = 0; i < ns; ++i) {
    const S f = floor(x[nb + i]);
    const S lim = S(a.g.shape[nb + i]);
    const S fc = f < S(-1) ? S(-1) : (f > lim ? lim : f);  // NaN-free, in int range
    q[i] = fc == fc ? int(fc) : 0;
    r[i] = x[nb + i] - f;
    xs[i] = S(0);
    sum[i] = S(0);
    sq[i] = S(0);
  }
  const int ncorner = 1 << ns;
  auto corner_index = [&](int icorner, bool nearest) {
    int64_t n = 0;
    for (int i = 0; i < ns; ++i) {
      int y = q[i] + (nearest ? (r[i] > S(0.5) ? 1 : 0) : ((icorner >> i) & 1));
      const int hi = a.g.shape[nb + i] - 1;
      y = y < 0 ? 0 : (y > hi ? hi : y);
      n = n * a.g.shape[nb + i] + y;
    }
    return n;
  };
  for (int c = 0; c < ncorner; ++c) {
    S w = S(1);
    for (int i = 0; i < ns; ++i) w *= ((c >> i) & 1) ? r[i] : S(1) - r[i];
    const int64_t n = corner_index(c, false);
    for (int i = 0; i < ns; ++i) {
      const S m = a.chart[int64_t(i) * size_s + n];
      xs[i] += w * m;
      sum[i] += m;
      sq[i] += m * m;
    }
  }
  S var = S(0);
  for (int i = 0; i < ns; ++i) var += sq[i] / S(ncorner) - sum[i] * sum[i] / S(ncorner * ncorner);
  if (var >= a.chart_var) {
    const int64_t n = corner_index(0, true);
    for (int i = 0; i < ns; ++i) xs[i] = a.chart[int64_t(i) * size_s + n];
  }
  bool ok = true;
  for (int i = 0; i < ns; ++i) ok = ok && xs[i] == xs[i];
  if (ok)
    for (int i = 0; i < ns; ++i) x[nb + i] = xs[i];
}

// GeodesicODE (Kernel_GeodesicODE.h:265-413): from the row's first point,
// midpoint steps x <- x + step * flow(x + step/2 * flow(x)) until a
// stopping criterion or max_len points.  Histories of the corner of least
// value (Stationnary: the same corner nymin_delay steps ago) and of its
// distance to the seeds (PastSeed) are kept over hlen steps.
template <class S, int D, class Src>
__device__ inline void geodesic_run(const GeoArgs<S>& a, const Src& src) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= a.ngeo) return;
  uint8_t eucl_p[kGeoHist];
  int64_t nymin_p[kGeoHist];
  for (int l = 0; l < a.hlen; ++l) {
    eucl_p[l] = 255;
    nymin_p[l] = INT64_MAX;
  }
  S* row = a.x + int64_t(tid) * a.max_len * D;
  S x[D];
#pragma unroll
  for (int k = 0; k < D; ++k) x[k] = row[k];
  GeoCell<S, D> c;
#pragma unroll
  for (int k = 0; k < D; ++k) c.xq[k] = INT32_MAX;
  c.thr = S(0);
  c.nymin = INT64_MAX;
  int len;
  int stop = kGeoContinue;
  for (len = 1; len < a.max_len; ++len) {
    const int l = len % a.hlen;
    S xp[D], flow[D];
#pragma unroll
    for (int k = 0; k < D; ++k) xp[k] = x[k];
    stop = geo_flow<S, D>(a, src, x, flow, c);
    if (stop != kGeoContinue) break;
    nymin_p[l] = c.nymin;
    eucl_p[l] = a.eucl[c.nymin];
    // a chart cell keeps the previous distance for the tests and jumps after
    // the step (Kernel_GeodesicODE.h:317-320, 406)
    const bool chart = a.chart != nullptr && eucl_p[l] == AGD_EUCL_CHART;
    if (chart) eucl_p[l] = eucl_p[(l - 1 + a.hlen) % a.hlen];
    if (c.nymin == nymin_p[(l - a.nymin_delay + a.hlen) % a.hlen]) {
      stop = kGeoStationnary;
      break;
    }
    // Kernel_GeodesicODE.h:338 indexes this history with the boolean
    // ((l - eucl_delay + hlen) % hlen && len >= pastseed_delay), that is
    // slot 1 or slot 0; restated as the reference computes it.
    const int ip = (((l - a.eucl_delay + a.hlen) % a.hlen) != 0 && len >= a.pastseed_delay) ? 1 : 0;
    if (eucl_p[l] > eucl_p[ip]) {
      stop = kGeoPastSeed;
      break;
    }
    S xm[D];
#pragma unroll
    for (int k = 0; k < D; ++k) xm[k] = S(0.5) * a.step * flow[k] + xp[k];
    stop = geo_flow<S, D>(a, src, xm, flow, c);
    if (stop != kGeoContinue) break;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      x[k] = a.step * flow[k] + xp[k];
      row[int64_t(len) * D + k] = x[k];
    }
    if (chart) chart_jump<S, D>(a, x);
  }
  a.len[tid] = len;
  a.stop[tid] = int8_t(stop);
}

template <class S, int D>
__global__ void __launch_bounds__(64) geodesic_kernel(GeoArgs<S> a) {
  GeoFlowArrays<S, D> src{&a};
  geodesic_run<S, D>(a, src);
}

// Checks of a geodesic descriptor; hlen receives the history length.
inline int check_geodesic_desc(const agd_geodesic_desc* gd, int& hlen) {
  AGD_REQUIRE(gd, "missing geodesic descriptor");
  AGD_REQUIRE(gd->ngeo >= 0 && gd->max_len >= 1, "ngeo >= 0 and max_len >= 1 required");
  AGD_REQUIRE(gd->eucl_delay >= 0 && gd->nymin_delay >= 0, "delays must be >= 0");
  hlen = 1 + (gd->eucl_delay < gd->nymin_delay ? gd->nymin_delay : gd->eucl_delay);
  AGD_REQUIRE(hlen <= kGeoHist, "PastSeed / Stationnary delays must be below 256");
  AGD_REQUIRE(gd->step > 0. && gd->step < __builtin_huge_val(), "step must be positive");
  AGD_REQUIRE(gd->chart_ndim_s >= 0 && gd->chart_ndim_s <= 3, "chart_ndim_s must be in [0, ndim]");
  AGD_REQUIRE(gd->chart_ndim_s == 0 || gd->chart_mapping, "chart_ndim_s > 0 needs chart_mapping");
  AGD_REQUIRE(!(gd->chart_jump_variance < 0.), "chart_jump_variance must be >= 0");
  return AGD_OK;
}

// ODE arguments of a descriptor (the flow source aside).
template <class S>
inline GeoArgs<S> geodesic_args(const GeoGrid& g, const agd_geodesic_desc* gd, int hlen, const void* values,
                                const uint8_t* eucl, void* x, int32_t* len, int8_t* stop) {
  GeoArgs<S> a{};
  a.g = g;
  a.u = static_cast<const S*>(values);
  a.eucl = eucl;
  a.x = static_cast<S*>(x);
  a.len = len;
  a.stop = stop;
  a.ngeo = gd->ngeo;
  a.max_len = gd->max_len;
  a.hlen = hlen;
  a.eucl_delay = gd->eucl_delay;
  a.nymin_delay = gd->nymin_delay;
  a.pastseed_delay = gd->pastseed_delay;
  a.step = S(gd->step);
  a.causality = S(gd->causality_tolerance);
  a.chart = gd->chart_ndim_s > 0 ? static_cast<const S*>(gd->chart_mapping) : nullptr;
  a.chart_ndim_s = gd->chart_ndim_s;
  a.chart_var = S(gd->chart_jump_variance);
  return a;
}

// Tiled grid of a solve descriptor.
inline GeoGrid geo_grid(const agd_eikonal_desc* d) {
  GeoGrid g{};
  g.size_i = 1;
  int64_t ntiles = 1;
  for (int k = 0; k < 3; ++k) {
    g.shape[k] = k < d->ndim ? int32_t(d->shape[k]) : 1;
    g.shape_i[k] = k < d->ndim ? d->shape_i[k] : 1;
    g.shape_o[k] = int32_t(ceil_div(g.shape[k], g.shape_i[k]));
    g.size_i *= g.shape_i[k];
    ntiles *= g.shape_o[k];
    if (k < d->ndim && d->periodic[k]) g.periodic |= 1 << k;
  }
  g.size_tot = ntiles * g.size_i;
  return g;
}

inline int check_geo_desc(const agd_eikonal_desc* d) {
  AGD_REQUIRE(d && (d->ndim == 2 || d->ndim == 3), "ndim must be 2 or 3");
  AGD_REQUIRE(d->scalar == AGD_F32 || d->scalar == AGD_F64, "scalar must be AGD_F32 or AGD_F64");
  for (int k = 0; k < d->ndim; ++k)
    AGD_REQUIRE(d->shape[k] >= 1 && d->shape_i[k] >= 1, "shape and shape_i must be positive");
  return AGD_OK;
}

}  // namespace agd
