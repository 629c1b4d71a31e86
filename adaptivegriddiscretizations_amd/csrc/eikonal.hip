// eikonal.hip -- host side of the eikonal tile solver for gfx950: argument
// checks, the per-device workspace, the grid-level worklist drivers (single,
// batched), the flow and precomputed-scheme entry points, and the
// model-independent kernels (initial lists, seeds, walls, stencil export).
// The model kernels are in eikonal_kernels.hpp.
//
// Worklist (replaces the host round trip of _solvers.py:122-193): the tile
// list of pass k+1 is built on the device while pass k runs (worklist.hpp).
// A tile is appended at most once per pass thanks to a per-tile stamp
// (atomicMax(stamp, k+1) < k+1 <=> first to tag it), so neither the list
// nor the flags need clearing.  Every launch uses the same grid; workgroups
// stride over the list.  The host reads the next list length every
// `check_interval` passes (one small copy + sync), or after each replay of
// a captured chunk of passes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "eikonal_kernels.hpp"
#include "worklist.hpp"

static_assert(sizeof(agd_eikonal_desc) == 320, "agd_eikonal_desc layout changed");
static_assert(sizeof(agd_eikonal_io) == 80, "agd_eikonal_io layout changed");

namespace agd {

// Initial list: tiles holding at least one trigger bit (or every tile).
// 64-lane workgroups, one tile per lane, wave-aggregated appends into
// sub-list blockIdx % kShards of pass 0.
template <int D, int NI>
__global__ void __launch_bounds__(64) eik_init_list(const uint32_t* trigger, GridDesc g, int32_t all,
                                                   int32_t* stamp, int32_t* lists, int32_t* counts,
                                                   int32_t shard_cap) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  bool on = false;
  int xo[3] = {0, 0, 0};
  if (t < g.ntiles) {
    on = all != 0;
    const int64_t b0 = int64_t(t) * NI, b1 = b0 + NI;
    for (int64_t b = b0; b < b1 && !on; ) {
      const int64_t wi = b >> 5;
      const int lo = int(b & 31);
      const int64_t lim = std::min<int64_t>(b1, (wi + 1) << 5);
      const int hi = int(lim - (wi << 5));
      uint32_t word = trigger[wi];
      uint32_t mask = (hi >= 32 ? 0xffffffffu : ((1u << hi) - 1u)) & ~((1u << lo) - 1u);
      on = (word & mask) != 0;
      b = lim;
    }
    tile_pos<D>(t, g, xo);
  }
  const int shard = blockIdx.x % kShards;
  enqueue_wave(on, on ? t : 0, pack_tile<D>(xo), 0, stamp, lists + int64_t(shard) * shard_cap,
               counts + shard * kCountStride);
}

// Batched solves: the same with tile ids as entries.
template <int NI>
__global__ void __launch_bounds__(64) init_list_ids(const uint32_t* trigger, int32_t ntiles, int32_t all,
                                                   int32_t* stamp, int32_t* lists, int32_t* counts,
                                                   int32_t shard_cap) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  bool on = false;
  if (t < ntiles) {
    on = all != 0;
    const int64_t b0 = int64_t(t) * NI, b1 = b0 + NI;
    if constexpr (NI >= 32) {
      for (int64_t b = b0; b < b1 && !on; b += 32) {
        // NI is a multiple of 32: whole words
        on = trigger[b >> 5] != 0u;
      }
    } else {
      // NI (16) divides 32: the tile's bits within one word
      on = on || ((trigger[b0 >> 5] >> (b0 & 31)) & ((1u << NI) - 1u)) != 0u;
    }
  }
  const int shard = blockIdx.x % kShards;
  enqueue_wave(on, on ? t : 0, t, 0, stamp, lists + int64_t(shard) * shard_cap,
               counts + shard * kCountStride);
}

// Every tile id (batched global-iteration passes).
__global__ void all_ids_kernel(int32_t ntiles, int32_t* list, int32_t* count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntiles) list[t] = t;
  if (t == 0) *count = ntiles;
}

// tile -> instance map: the last instance whose tile_base is <= t.
__global__ void tile_inst_kernel(int32_t ninst, const InstDesc* inst, int32_t ntiles,
                                 int32_t* tile_inst) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  int lo = 0, hi = ninst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (inst[mid].tile_base <= t) lo = mid;
    else hi = mid - 1;
  }
  tile_inst[t] = lo;
}

// bound_active_blocks state at the start of a solve: no pending change,
// thresholds +inf, counters cleared (update requests `needs` are zeroed by
// the caller: every tile of the initial list is a real update).
__global__ void bab_init_kernel(int32_t ntiles, double* pending, int32_t* ctl) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntiles) pending[t] = __builtin_huge_val();
  if (blockIdx.x == 0 && threadIdx.x < kShards) {
    const int sh = threadIdx.x * kCountStride;
    ctl[kBabProp + sh] = 0;
    ctl[kBabFrozen + sh] = 0;
    ctl[kBabMin + sh] = ord_int(__builtin_huge_valf());
    ctl[kBabMax + sh] = ord_int(-__builtin_huge_valf());
    if (threadIdx.x < 2) reinterpret_cast<double*>(ctl + kBabThr)[threadIdx.x] = __builtin_huge_val();
  }
}

// Every tile, as packed coordinates (global-iteration passes).
template <int D>
__global__ void all_tiles_kernel(GridDesc g, int32_t* list, int32_t* count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < g.ntiles) {
    int xo[3];
    tile_pos<D>(t, g, xo);
    list[t] = pack_tile<D>(xo);
  }
  if (t == 0) *count = g.ntiles;
}

// ------------------------------------------------------------------ stencil export
template <class S, int MODEL, int D>
__global__ void stencil_export_kernel(int64_t n, const S* geom, S iw0, S iw1, S iw2, S* wout,
                                      int32_t* eout) {
  using MT = ModelTraits<MODEL, D>;
  constexpr int NS = MT::NSYM;
  constexpr int NG = MT::NGEOM;
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  S geo[NG > 0 ? NG : 1];
#pragma unroll
  for (int k = 0; k < NG; ++k) geo[k] = geom[k * n + c];
  S iw[3] = {iw0, iw1, iw2};
  Stencil<S, MODEL, D> st;
  make_stencil<S, MODEL, D>(geo, iw, st);
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    wout[c * NS + k] = st.w[k];
#pragma unroll
    for (int l = 0; l < D; ++l) eout[(c * NS + k) * D + l] = st.e[k][l];
  }
}

// ------------------------------------------------------------------ seeds / walls
template <class S>
__global__ void scatter_seeds_kernel(int64_t n, const int64_t* idx, const double* val, S* rhs,
                                     uint32_t* seeds, uint32_t* trigger) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t m = idx[i];
  rhs[m] = S(val[i]);
  atomicOr(seeds + (m >> 5), 1u << (m & 31));
  if (trigger) atomicOr(trigger + (m >> 5), 1u << (m & 31));
}

template <class S>
__global__ void walls_mark_kernel(int64_t N, const uint8_t* walls, S* rhs, uint32_t* seeds,
                                  uint8_t* dist) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const bool w = walls[i] != 0;
  dist[i] = w ? 0 : 255;
  if (w) {
    rhs[i] = inf_v<S>();
    atomicOr(seeds + (i >> 5), 1u << (i & 31));
  }
}

// One min-plus pass with the L1 unit kernel (Kernel_InfConvolution.h:66-91
// with distance_kernel(1, ndim, ord=1)), saturated at 255.  Operates on the
// tiled layout; padded cells stay at 255.
template <int D, int T0, int T1, int T2>
__global__ void wall_dist_pass(GridDesc g, const uint8_t* in, uint8_t* out) {
  using TL = Tile<D, T0, T1, T2>;
  const int n_o = blockIdx.x;
  const int tid = threadIdx.x;
  int xi[3], xo[3], xt[3] = {0, 0, 0};
  TL::pos(tid, xi);
  tile_pos<D>(n_o, g, xo);
  bool dom = true;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    xt[k] = xo[k] * TL::T[k] + xi[k];
    dom = dom && xt[k] < g.shape_tot[k];
  }
  const int64_t n_t = int64_t(n_o) * TL::NI + tid;
  if (!dom) { out[n_t] = 255; return; }
  int best = in[n_t];
  // 3^D neighbourhood, L1 kernel weights.
  const int r2 = (D == 3) ? 1 : 0;
  for (int d0 = -1; d0 <= 1; ++d0)
    for (int d1 = -1; d1 <= 1; ++d1)
      for (int d2 = -r2; d2 <= r2; ++d2) {
        int y[3] = {xt[0] + d0, xt[1] + d1, xt[2] + d2};
        const int64_t m = index_tot<D, TL>(y, g);
        if (m < 0) continue;
        const int kw = (d0 < 0 ? -d0 : d0) + (d1 < 0 ? -d1 : d1) + (d2 < 0 ? -d2 : d2);
        const int vv = in[m];
        const int p = (vv >= 255 - kw) ? 255 : vv + kw;
        best = p < best ? p : best;
      }
  out[n_t] = uint8_t(best);
}

__global__ void wall_dist_cap(int64_t N, uint8_t* d, int bound) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < N && d[i] > bound) d[i] = 255;
}

// ------------------------------------------------------------------ dispatch
// Model-independent kernels of one tile shape.
struct TileKernels {
  void (*init)(const uint32_t*, GridDesc, int32_t, int32_t*, int32_t*, int32_t*, int32_t);
  void (*init_ids)(const uint32_t*, int32_t, int32_t, int32_t*, int32_t*, int32_t*, int32_t);
  void (*all_tiles)(GridDesc, int32_t*, int32_t*);
  void (*wallpass)(GridDesc, const uint8_t*, uint8_t*);
};

template <int D, int T0, int T1, int T2> TileKernels make_tile_kernels() {
  TileKernels k;
  k.init_ids = init_list_ids<T0 * T1 * T2>;
  if constexpr (T0 * T1 * T2 < 64) {
    // tiles below a wave: batched solves only
    k.init = nullptr;
    k.all_tiles = nullptr;
    k.wallpass = nullptr;
  } else {
    k.init = eik_init_list<D, T0 * T1 * T2>;
    k.all_tiles = all_tiles_kernel<D>;
    k.wallpass = wall_dist_pass<D, T0, T1, T2>;
  }
  return k;
}

static bool lookup_tile_kernels(int d, const int32_t* si, TileKernels& out) {
  const int key = tile_key(d, si);
  if (d == 2) {
    switch (key) {
      case 808: out = make_tile_kernels<2, 8, 8, 1>(); return true;
      case 1616: out = make_tile_kernels<2, 16, 16, 1>(); return true;
      case 816: out = make_tile_kernels<2, 8, 16, 1>(); return true;
      case 404: out = make_tile_kernels<2, 4, 4, 1>(); return true;
      case 408: out = make_tile_kernels<2, 4, 8, 1>(); return true;
      case 804: out = make_tile_kernels<2, 8, 4, 1>(); return true;
    }
  } else if (d == 3) {
    switch (key) {
      case 40404: out = make_tile_kernels<3, 4, 4, 4>(); return true;
      case 80804: out = make_tile_kernels<3, 8, 8, 4>(); return true;
      case 40408: out = make_tile_kernels<3, 4, 4, 8>(); return true;
    }
  }
  return false;
}

template <class S> struct KernelSet : ModelKernels<S>, TileKernels {};

template <class S>
static bool lookup(int model, int d, const int32_t* si, KernelSet<S>& out) {
  ModelKernels<S>& mk = out;
  bool ok = false;
  switch (model) {
    case AGD_ISOTROPIC: ok = lookup_isotropic<S>(d, si, mk); break;
    case AGD_DIAGONAL: ok = lookup_diagonal<S>(d, si, mk); break;
    case AGD_RIEMANN: ok = lookup_riemann<S>(d, si, mk); break;
    case AGD_RANDER: ok = lookup_rander<S>(d, si, mk); break;
    case AGD_TTI: ok = lookup_tti<S>(d, si, mk); break;
    case AGD_REEDSSHEPP2:
    case AGD_REEDSSHEPPFWD2:
    case AGD_ELASTICA2:
    case AGD_DUBINS2: ok = lookup_curvature<S>(model, d, si, mk); break;
  }
  return ok && lookup_tile_kernels(d, si, out);
}

static int check_desc(const agd_eikonal_desc* d) {
  AGD_REQUIRE(d != nullptr, "null descriptor");
  AGD_REQUIRE(d->ndim == 2 || d->ndim == 3, "ndim must be 2 or 3");
  AGD_REQUIRE(d->model >= AGD_ISOTROPIC && d->model <= AGD_DUBINS2, "unknown model");
  if (curvature_model(d->model)) {
    AGD_REQUIRE(d->ndim == 3, "the curvature models solve on (x, y, theta) grids (ndim 3)");
    AGD_REQUIRE(d->curvature_relax > 0., "curvature_relax must be > 0 (eps^2, 0.01 by default)");
    AGD_REQUIRE(d->factor_radius2 == 0., "source factorization is not provided for the curvature models");
    AGD_REQUIRE(d->order <= 1, "the curvature models use the first-order scheme");
    AGD_REQUIRE(d->convex == 0 || d->model != AGD_REEDSSHEPP2,
                "ReedsShepp2 has no Convex variant (ReedsSheppForward2, Elastica2, Dubins2 do)");
  } else {
    AGD_REQUIRE(d->convex == 0, "convex applies to the curvature models only");
  }
  AGD_REQUIRE(d->scalar == AGD_F32 || d->scalar == AGD_F64, "unknown scalar type");
  for (int k = 0; k < d->ndim; ++k) {
    AGD_REQUIRE(d->shape[k] >= 1 && d->shape[k] < (1 << 30), "shape out of range");
    AGD_REQUIRE(d->shape_i[k] >= 1, "tile shape must be positive");
  }
  AGD_REQUIRE(d->niter_i >= 1, "niter_i must be >= 1");
  AGD_REQUIRE(d->nitermax_o >= 1, "nitermax_o must be >= 1");
  int64_t size_tot = 1;
  for (int k = 0; k < d->ndim; ++k) size_tot *= ceil_div(d->shape[k], d->shape_i[k]) * d->shape_i[k];
  AGD_REQUIRE(size_tot < (int64_t(1) << 31), "grids of 2^31 padded cells or more are not supported");
  for (int k = 0; k < d->ndim; ++k)
    AGD_REQUIRE(ceil_div(d->shape[k], d->shape_i[k]) <= (d->ndim == 2 ? 32768 : 1024),
                "too many tiles along an axis (2D: 32768, 3D: 1024)");
  AGD_REQUIRE(d->factor_radius2 >= 0., "factor_radius2 must be >= 0 (0 disables factorization)");
  return AGD_OK;
}

// Source factorization parameters of a solve (disabled when factor_radius2
// is 0).
template <class S> static FactorDesc<S> make_factor(const agd_eikonal_desc* d) {
  FactorDesc<S> f{};
  f.r2 = S(d->factor_radius2);
  for (int k = 0; k < 3; ++k) f.origin[k] = S(d->factor_origin[k]);
  for (int k = 0; k < 14; ++k) f.m[k] = S(d->factor_metric[k]);
  if (d->factor_radius2 > 0.) {
    // The box of the ball, one cell wider on each side than the exact one
    // (the device tests membership in the scalar type), clipped to the grid.
    const double r = std::sqrt(d->factor_radius2);
    f.nbox = 1;
    for (int k = 0; k < 3; ++k) {
      f.lo[k] = 0;
      f.n[k] = 1;
    }
    for (int k = 0; k < d->ndim; ++k) {
      const double lo = std::max(0., std::floor(d->factor_origin[k] - r) - 1.);
      const double hi = std::min(double(d->shape[k] - 1), std::ceil(d->factor_origin[k] + r) + 1.);
      f.lo[k] = int32_t(lo);
      f.n[k] = hi >= lo ? int32_t(hi - lo + 1.) : 0;
      f.nbox *= f.n[k];
    }
  }
  return f;
}

// Scalars of the TTI shift table: NMIX * 2 * nsym shifts (twice that with
// order 2) per cell of the factor box.
template <class S> static int64_t tti_factor_table_len(const agd_eikonal_desc* d, const FactorDesc<S>& f) {
  const int nsym = d->ndim * (d->ndim + 1) / 2;
  return int64_t(kTTIMix) * 2 * nsym * (d->order == 2 ? 2 : 1) * f.nbox;
}

// TTI scheme codes: NMIX * 2 * nsym neighbour codes, the mix flag, then
// (second order) NMIX * 2 * nsym further-neighbour codes.
static int64_t tti_codes2_offset(const agd_eikonal_desc* d) {
  const int nsym = d->ndim * (d->ndim + 1) / 2;
  return int64_t(kTTIMix) * 2 * nsym + 1;
}

// Components of the precomputed scheme's weight and code buffers (the io
// struct's layout comment in agd_hip.h).
static void scheme_components(int model, int ndim, int order, int64_t& wc, int64_t& cc) {
  const int64_t nsym = nsym_of(model, ndim);
  const int64_t nmix = nmix_of(model);
  wc = nmix * nsym;
  cc = nmix * 2 * nsym;
  if (stored_scheme_model(model)) cc += 1 + (model == AGD_TTI && order == 2 ? nmix * 2 * nsym : 0);
}

// The precomputed scheme buffers hold what the descriptor's layout needs.
static int check_scheme_lengths(const agd_eikonal_desc* d, const agd_eikonal_io* io, int64_t size_tot) {
  int64_t wc = 0, cc = 0;
  scheme_components(d->model, d->ndim, d->order, wc, cc);
  AGD_REQUIRE(io->scheme_weights_len >= wc * size_tot,
              "scheme_weights_len is shorter than the scheme layout of this model and order");
  AGD_REQUIRE(io->scheme_codes_len >= cc * size_tot,
              "scheme_codes_len is shorter than the scheme layout of this model and order "
              "(TTI with order 2 also stores the further-neighbour codes)");
  return AGD_OK;
}

static GridDesc make_grid(const agd_eikonal_desc* d) {
  GridDesc g{};
  g.ntiles = 1;
  int64_t ni = 1;
  for (int k = 0; k < 3; ++k) {
    g.shape_tot[k] = k < d->ndim ? int32_t(d->shape[k]) : 1;
    g.shape_o[k] = k < d->ndim ? int32_t(ceil_div(d->shape[k], d->shape_i[k])) : 1;
    if (k < d->ndim) {
      g.ntiles *= g.shape_o[k];
      ni *= d->shape_i[k];
    }
  }
  g.periodic = 0;
  for (int k = 0; k < d->ndim; ++k)
    if (d->periodic[k]) g.periodic |= (1 << k);
  g.size_tot = int64_t(g.ntiles) * ni;
  return g;
}

// ------------------------------------------------------------------ workspace
template <class T> static void free_dev(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void Workspace::release() {
  if (!fixed) {
    free_dev(ftab);
    free_dev(stamp);
    free_dev(needs);
    free_dev(pending);
    free_dev(bab);
    free_dev(lists);
    free_dev(counts);
    free_dev(hist);
    free_dev(score);
    free_dev(check_count);
    free_dev(inst);
    free_dev(tile_inst);
    if (host_counts) (void)hipHostFree(host_counts);
    if (exec) (void)hipGraphExecDestroy(exec);
    for (hipEvent_t e : {ev0, ev1, chunk_ev[0], chunk_ev[1]})
      if (e) (void)hipEventDestroy(e);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
  }
  stamp = needs = bab = lists = counts = hist = score = check_count = tile_inst = nullptr;
  pending = nullptr;
  inst = nullptr;
  host_counts = nullptr;
  exec = nullptr;
  exec_key.clear();
  ev0 = ev1 = chunk_ev[0] = chunk_ev[1] = nullptr;
  capture_stream = nullptr;
  tile_cap = shard_cap = hist_cap = inst_cap = tinst_cap = 0;
  ftab = nullptr;
  ftab_bytes = 0;
  fixed = false;
}

int Workspace::ensure_ftab(size_t bytes) {
  AGD_REQUIRE(!fixed, "source factorization of the TTI model needs an allocating workspace");
  if (ftab_bytes < bytes) {
    free_dev(ftab);
    AGD_HIP_TRY(hipMalloc(&ftab, bytes));
    ftab_bytes = bytes;
  }
  return AGD_OK;
}

int Workspace::ensure(int64_t ntiles, int64_t scap, int64_t nhist) {
  if (fixed) {
    AGD_REQUIRE(ntiles <= tile_cap && scap <= shard_cap && nhist <= hist_cap,
                "the solve does not fit the library's fixed device region");
    return AGD_OK;
  }
  if (!ev0) {
    AGD_HIP_TRY(hipEventCreate(&ev0));
    AGD_HIP_TRY(hipEventCreate(&ev1));
    AGD_HIP_TRY(hipHostMalloc((void**)&host_counts, 2 * kRow * sizeof(int32_t), hipHostMallocDefault));
    AGD_HIP_TRY(hipEventCreateWithFlags(&chunk_ev[0], hipEventDisableTiming));
    AGD_HIP_TRY(hipEventCreateWithFlags(&chunk_ev[1], hipEventDisableTiming));
    AGD_HIP_TRY(hipStreamCreateWithFlags(&capture_stream, hipStreamNonBlocking));
    AGD_HIP_TRY(hipMalloc((void**)&counts, kCountsAlloc * sizeof(int32_t)));
    AGD_HIP_TRY(hipMalloc((void**)&check_count, 64));
    AGD_HIP_TRY(hipMalloc((void**)&bab, kBabLen * sizeof(int32_t)));
  }
  if (tile_cap < ntiles) {
    free_dev(stamp);
    free_dev(score);
    free_dev(needs);
    free_dev(pending);
    const int64_t cap = std::max<int64_t>(ntiles, 1024);
    AGD_HIP_TRY(hipMalloc((void**)&stamp, cap * sizeof(int32_t)));
    AGD_HIP_TRY(hipMalloc((void**)&score, 2 * cap * sizeof(int32_t)));
    AGD_HIP_TRY(hipMalloc((void**)&needs, cap * sizeof(int32_t)));
    AGD_HIP_TRY(hipMalloc((void**)&pending, cap * sizeof(double)));
    tile_cap = cap;
  }
  if (shard_cap < scap) {
    free_dev(lists);
    const int64_t cap = std::max<int64_t>(scap, 256);
    AGD_HIP_TRY(hipMalloc((void**)&lists, 2 * kShards * cap * sizeof(int32_t)));
    shard_cap = cap;
  }
  if (hist_cap < nhist) {
    free_dev(hist);
    const int64_t cap = std::max<int64_t>(nhist, 4096);
    AGD_HIP_TRY(hipMalloc((void**)&hist, cap * sizeof(int32_t)));
    hist_cap = cap;
  }
  return AGD_OK;
}

int Workspace::ensure_batch(int64_t ninst, int64_t ntiles, size_t inst_bytes) {
  AGD_REQUIRE(!fixed, "batched solves need an allocating workspace");
  if (inst_cap < ninst) {
    free_dev(inst);
    AGD_HIP_TRY(hipMalloc(&inst, size_t(ninst) * inst_bytes));
    inst_cap = ninst;
  }
  if (tinst_cap < ntiles) {
    free_dev(tile_inst);
    AGD_HIP_TRY(hipMalloc((void**)&tile_inst, size_t(ntiles) * sizeof(int32_t)));
    tinst_cap = ntiles;
  }
  return AGD_OK;
}

// Fixed-region layout: every buffer starts on a 256-byte boundary.
static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

size_t Workspace::fixed_bytes(int64_t ntiles, int64_t scap, int64_t nhist) {
  const size_t t = size_t(ntiles);
  return align256(t * 4) + align256(t * 4) + align256(t * 8) + align256(kBabLen * 4) +
         align256(size_t(2) * kShards * size_t(scap) * 4) + align256(kCountsAlloc * 4) +
         align256(size_t(nhist) * 4) + align256(2 * t * 4) + align256(64);
}

int Workspace::bind_fixed(void* base, size_t bytes, int64_t ntiles, int64_t scap, int64_t nhist) {
  AGD_REQUIRE(fixed_bytes(ntiles, scap, nhist) <= bytes,
              "the solve does not fit the library's fixed device region");
  release();
  fixed = true;
  char* p = static_cast<char*>(base);
  auto take = [&](size_t b) {
    char* q = p;
    p += align256(b);
    return q;
  };
  stamp = reinterpret_cast<int32_t*>(take(size_t(ntiles) * 4));
  needs = reinterpret_cast<int32_t*>(take(size_t(ntiles) * 4));
  pending = reinterpret_cast<double*>(take(size_t(ntiles) * 8));
  bab = reinterpret_cast<int32_t*>(take(kBabLen * 4));
  lists = reinterpret_cast<int32_t*>(take(size_t(2) * kShards * size_t(scap) * 4));
  counts = reinterpret_cast<int32_t*>(take(kCountsAlloc * 4));
  hist = reinterpret_cast<int32_t*>(take(size_t(nhist) * 4));
  score = reinterpret_cast<int32_t*>(take(2 * size_t(ntiles) * 4));
  check_count = reinterpret_cast<int32_t*>(take(64));
  tile_cap = ntiles;
  shard_cap = scap;
  hist_cap = nhist;
  host_counts = host_local;
  return AGD_OK;
}

int blocks_for(int ntiles, int lanes_per_block) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  // Enough resident waves per CU to hide the neighbour-gather latency; the
  // workgroups stride over the tile list.
  const int waves_per_block = (lanes_per_block + 63) / 64;
  const int target = cus * std::max(1, 32 / waves_per_block);
  return std::max(1, std::min(ntiles, target));
}

template <class S>
static int solve_impl(Workspace& ws, const agd_eikonal_desc* d, const agd_eikonal_io* io,
                      agd_eikonal_stats* stats, hipStream_t stream) {
  KernelSet<S> ks;
  AGD_REQUIRE(lookup<S>(d->model, d->ndim, d->shape_i, ks), "unsupported tile shape for this model");
  AGD_REQUIRE(ks.tiles_per_block == 1, "tiles of fewer than 64 cells are for batched solves only");
  AGD_REQUIRE(io && io->values && io->rhs && io->seeds, "values, rhs and seeds are required");
  AGD_REQUIRE(d->model == AGD_ISOTROPIC || io->geom, "geometry required for this model");
  AGD_REQUIRE(d->solver != AGD_FIM || d->fim_front_width >= 2, "fim_front_width must be >= 2");
  const GridDesc g = make_grid(d);
  const int ni = ks.ni;
  const int blocks = blocks_for(g.ntiles, ni);
  const bool global = d->solver == AGD_GLOBAL;

  const bool pre = io->scheme_weights != nullptr || io->scheme_codes != nullptr;
  AGD_REQUIRE(!pre || (io->scheme_weights && io->scheme_codes),
              "scheme_weights and scheme_codes go together");
  if (pre) {
    const int stl = check_scheme_lengths(d, io, g.size_tot);
    if (stl) return stl;
  }
  AGD_REQUIRE(!pre || ks.update_pre, "no precomputed-scheme variant for this model");
  AGD_REQUIRE(pre || ks.update, "this model runs from a precomputed scheme (agd_eikonal_scheme)");
  const bool order2 = d->order == 2;
  AGD_REQUIRE(d->order >= 0 && d->order <= 2, "order must be 1 or 2 (0 means 1)");
  const bool tti = d->model == AGD_TTI;
  AGD_REQUIRE(!order2 || (!pre && ks.update_order2) || (pre && tti),
              "the second-order scheme has no precomputed-scheme variant (except TTI's)");
  const bool factor = d->factor_radius2 > 0.;
  AGD_REQUIRE(!factor || !pre || tti,
              "source factorization has no precomputed-scheme variant (except TTI's)");
  const bool bab = d->bound_active_blocks > 0;
  AGD_REQUIRE(!bab || d->solver == AGD_AGSI, "bound_active_blocks runs with the AGSI solver");
  AGD_REQUIRE(!bab || d->minchg_delta_min > 0., "bound_active_blocks needs minchg_delta_min > 0");
  AGD_REQUIRE(!bab || !pre, "bound_active_blocks runs without a precomputed scheme");
  // Fields of 4 GiB or more go to the general variant (64-bit addressing).
  const bool big = uint64_t(g.size_tot) * sizeof(S) > uint64_t(0xffffffffu);
  AGD_REQUIRE(!pre || !big, "a precomputed scheme needs fields below 4 GiB (2^30 FP32 cells)");
  // Zeroed bytes, padding included: the argument bytes key the graph cache.
  EikArgs<S> a;
  std::memset(&a, 0, sizeof a);
  a.u = static_cast<S*>(io->values);
  a.geom = static_cast<const S*>(io->geom);
  a.sw = static_cast<const S*>(io->scheme_weights);
  a.scode = io->scheme_codes;
  // TTI, second order: the further-neighbour codes follow the mix flag
  a.scode2 = (pre && tti && order2) ? io->scheme_codes + tti_codes2_offset(d) * g.size_tot : nullptr;
  a.rhs = static_cast<const S*>(io->rhs);
  a.seeds = io->seeds;
  a.wall = io->wall_dist;
  a.g = g;
  a.atol = S(d->atol);
  a.rtol = S(d->rtol);
  for (int k = 0; k < 3; ++k) a.iso_w[k] = S(d->iso_weights[k]);
  a.niter_i = d->niter_i;
  a.solver = global ? AGD_AGSI : d->solver;
  a.fim_front = d->fim_front_width;
  a.order2_thr = S(d->order2_threshold);
  a.fac = make_factor<S>(d);
  if (tti && factor && a.fac.nbox > 0) {
    // TTI: the factorization shifts of the ball's cells, once per solve.
    const int64_t len = tti_factor_table_len<S>(d, a.fac);
    const int st = ws.ensure_ftab(size_t(len) * sizeof(S));
    if (st) return st;
    a.fac.tab = static_cast<const S*>(ws.ftab);
    TTIFactorArgs<S> fa;
    std::memset(&fa, 0, sizeof fa);
    fa.geom = a.geom;
    fa.g = g;
    fa.fac = a.fac;
    fa.tab = static_cast<S*>(ws.ftab);
    fa.order2 = order2 ? 1 : 0;
    hipLaunchKernelGGL(ks.tti_factor, dim3(unsigned(g.ntiles)), dim3(ni), 0, stream, fa);
    AGD_HIP_TRY(hipGetLastError());
  }

  auto init = [&](Workspace* ws) -> int {
    a.w.lists = ws->lists;
    a.w.counts = ws->counts;
    a.w.stamp = ws->stamp;
    a.w.hist = ws->hist;
    a.w.score = ws->score;
    a.w.tile_cap = int32_t(ws->tile_cap);
    a.w.shard_cap = int32_t(ws->shard_cap);
    a.w.fixed = nullptr;
    fill_i32(ws->score, 2 * ws->tile_cap, 0x7f7f7f7f, stream);
    if (global) {
      // Every pass visits the full list (kept in the score buffer, unused
      // by this solver); the AGSI tags only decide when to stop.
      hipLaunchKernelGGL(ks.all_tiles, dim3(unsigned(ceil_div(g.ntiles, 256))), dim3(256), 0,
                         stream, g, ws->score, ws->counts + kFixedIdx);
      a.w.fixed = ws->score;
    } else {
      const uint32_t* trig = io->trigger ? io->trigger : io->seeds;
      hipLaunchKernelGGL(ks.init, dim3(unsigned(ceil_div(g.ntiles, 64))), dim3(64), 0, stream, trig,
                         g, d->activate_all, ws->stamp, ws->lists, ws->counts, a.w.shard_cap);
    }
    if (bab) {
      a.bab.ctl = ws->bab;
      a.bab.needs = ws->needs;
      a.bab.pending = ws->pending;
      fill_i32(ws->needs, g.ntiles, 0, stream);
      hipLaunchKernelGGL(bab_init_kernel, dim3(unsigned(ceil_div(g.ntiles, 256))), dim3(256), 0,
                         stream, g.ntiles, ws->pending, ws->bab);
    }
    return AGD_OK;
  };
  // walls, periodic axes, source factorization and bound_active_blocks run
  // in the general variant
  auto kern = pre ? ks.update_pre
                  : order2 ? ks.update_order2
                           : ((io->wall_dist || g.periodic || factor || big || bab)
                                  ? ks.update_walls
                                  : ks.update);
  auto launch = [&](int32_t koff, hipStream_t s, bool phase) {
    a.w.koff = koff;
    a.w.phase_from_koff = phase ? 1 : 0;
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(ni), 0, s, a);
    if (bab)
      hipLaunchKernelGGL(bab_threshold_kernel, dim3(1), dim3(64), 0, s, a.bab.ctl,
                         int64_t(d->bound_active_blocks), d->minchg_delta_min);
  };
  auto key = [&](std::vector<unsigned char>& k) {
    key_append(k, kern);
    key_append(k, a);
    key_append(k, d->bound_active_blocks);
    key_append(k, d->minchg_delta_min);
  };
  WorklistResult r;
  int st = run_worklist(ws, g.ntiles, blocks, d->nitermax_o, d->check_interval, stream, init, launch,
                        key, r, d->time_launches != 0, d->use_graph != 0);
  if (st) return st;
  if (stats) {
    stats->kernel_ms = r.kernel_ms;
    stats->niter_o = r.niter;
    stats->tile_updates = r.visits;
    stats->cell_updates = r.visits * ni;
    stats->launches = r.passes;
    stats->gpu_ms = r.ms;
    stats->converged = r.converged ? 1 : 0;
    stats->grid_blocks = blocks;
  }
  if (!r.converged) {
    set_error("eikonal solver did not converge within nitermax_o passes");
    return AGD_NOCONV;
  }
  return AGD_OK;
}

// Flow-pass arguments of a solve: values, geometry and constants, with no
// output buffers.
template <class S> static FlowArgs<S> flow_args(const agd_eikonal_desc* d, const agd_eikonal_io* io) {
  FlowArgs<S> a{};
  a.u = static_cast<S*>(io->values);
  a.geom = static_cast<const S*>(io->geom);
  a.rhs = static_cast<const S*>(io->rhs);
  a.seeds = io->seeds;
  a.wall = io->wall_dist;
  a.g = make_grid(d);
  for (int k = 0; k < 3; ++k) a.iso_w[k] = S(d->iso_weights[k]);
  a.order2 = d->order == 2 ? 1 : 0;
  a.order2_thr = S(d->order2_threshold);
  a.fac = make_factor<S>(d);
  a.relax = S(d->curvature_relax);
  a.convex = d->convex;
  return a;
}

template <class S>
static int flow_impl(const agd_eikonal_desc* d, const agd_eikonal_io* io, const agd_flow_io* f,
                     hipStream_t stream, int32_t* violations = nullptr, double chk_atol = 0.,
                     double chk_rtol = 0., bool repair = false, uint32_t* stale_bits = nullptr,
                     bool stale = false) {
  KernelSet<S> ks;
  AGD_REQUIRE(lookup<S>(d->model, d->ndim, d->shape_i, ks), "unsupported tile shape for this model");
  AGD_REQUIRE(ks.flow, "tiles of fewer than 64 cells are for batched solves only");
  AGD_REQUIRE(io && io->values && io->rhs && io->seeds && (f || violations), "missing buffers");
  AGD_REQUIRE(d->order >= 0 && d->order <= 2, "order must be 1 or 2 (0 means 1)");
  const GridDesc g = make_grid(d);
  FlowArgs<S> a = flow_args<S>(d, io);
  a.fw = f ? static_cast<S*>(f->flow_weights) : nullptr;
  a.wsum = f ? static_cast<S*>(f->flow_weightsum) : nullptr;
  a.fidx = f ? f->flow_indices : nullptr;
  a.fvec = f ? static_cast<S*>(f->flow_vector) : nullptr;
  a.violations = violations;
  a.chk_atol = S(chk_atol);
  a.chk_rtol = S(chk_rtol);
  a.repair = repair ? 1 : 0;
  a.stale = stale ? 1 : 0;
  a.stale_bits = stale_bits;
  hipLaunchKernelGGL(ks.flow, dim3(g.ntiles), dim3(ks.ni), 0, stream, a);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

template <class S>
static int geodesic_online_impl(const agd_eikonal_desc* d, const agd_eikonal_io* io,
                                const agd_geodesic_desc* gd, const uint8_t* eucl, void* x,
                                int32_t* len, int8_t* stop, hipStream_t stream) {
  KernelSet<S> ks;
  AGD_REQUIRE(lookup<S>(d->model, d->ndim, d->shape_i, ks), "unsupported tile shape for this model");
  AGD_REQUIRE(ks.geodesic_online, "tiles of fewer than 64 cells are for batched solves only");
  AGD_REQUIRE(io && io->values && io->rhs && io->seeds, "values, rhs and seeds are required");
  AGD_REQUIRE(d->model == AGD_ISOTROPIC || io->geom, "geometry required for this model");
  AGD_REQUIRE(d->order >= 0 && d->order <= 2, "order must be 1 or 2 (0 means 1)");
  AGD_REQUIRE(eucl && x && len && stop, "missing buffers");
  int hlen = 0;
  if (int st = check_geodesic_desc(gd, hlen)) return st;
  AGD_REQUIRE(gd->chart_ndim_s <= d->ndim, "chart_ndim_s must be in [0, ndim]");
  if (gd->ngeo == 0) return AGD_OK;
  GeoArgs<S> a = geodesic_args<S>(geo_grid(d), gd, hlen, io->values, eucl, x, len, stop);
  const FlowArgs<S> fa = flow_args<S>(d, io);
  hipLaunchKernelGGL(ks.geodesic_online, dim3(unsigned(ceil_div(gd->ngeo, 64))), dim3(64), 0, stream,
                     a, fa);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

template <class S>
static int solve_batched_impl(Workspace& ws, const agd_eikonal_desc* d, int32_t ninst,
                              const agd_batch_instance* bi, const agd_eikonal_io* io,
                              agd_eikonal_stats* stats, hipStream_t stream) {
  KernelSet<S> ks;
  AGD_REQUIRE(lookup<S>(d->model, d->ndim, d->shape_i, ks), "unsupported tile shape for this model");
  AGD_REQUIRE(ks.update_batch, "batched solves are FP32 only");
  AGD_REQUIRE(ninst >= 1 && bi, "at least one instance is required");
  AGD_REQUIRE(io && io->values && io->rhs && io->seeds, "values, rhs and seeds are required");
  AGD_REQUIRE(d->model == AGD_ISOTROPIC || io->geom, "geometry required for this model");
  AGD_REQUIRE(!io->wall_dist && !io->scheme_weights && !io->scheme_codes,
              "batched solves take no walls and no precomputed scheme");
  AGD_REQUIRE(d->solver != AGD_FIM || d->fim_front_width >= 2, "fim_front_width must be >= 2");
  for (int k = 0; k < d->ndim; ++k) AGD_REQUIRE(!d->periodic[k], "batched solves are not periodic");
  AGD_REQUIRE(d->order <= 1, "batched solves use the first-order scheme");
  AGD_REQUIRE(d->factor_radius2 == 0., "batched solves take no source factorization");
  AGD_REQUIRE(d->bound_active_blocks == 0, "batched solves take no bound_active_blocks");
  const int ni = ks.ni;
  std::vector<InstDesc> h(static_cast<size_t>(ninst));
  int64_t tiles = 0;
  for (int i = 0; i < ninst; ++i) {
    InstDesc& I = h[size_t(i)];
    I = InstDesc{};
    I.g.ntiles = 1;
    for (int k = 0; k < 3; ++k) {
      const int64_t n = k < d->ndim ? bi[i].shape[k] : 1;
      AGD_REQUIRE(n >= 1 && n < (1 << 30), "instance shape out of range");
      const int64_t no = ceil_div(n, k < d->ndim ? d->shape_i[k] : 1);
      AGD_REQUIRE(no <= (d->ndim == 2 ? 32768 : 1024), "too many tiles along an axis");
      I.g.shape_tot[k] = int32_t(n);
      I.g.shape_o[k] = int32_t(no);
      I.g.ntiles *= int32_t(no);
      I.iso_w[k] = bi[i].iso_weights[k];
    }
    I.g.periodic = 0;
    I.g.size_tot = int64_t(I.g.ntiles) * ni;
    I.tile_base = int32_t(tiles);
    I.atol = bi[i].atol;
    I.rtol = bi[i].rtol;
    tiles += I.g.ntiles;
    AGD_REQUIRE(uint64_t(tiles * ni) * sizeof(S) <= uint64_t(0xffffffffu),
                "batches hold less than 4 GiB per field (2^30 FP32 cells)");
  }
  GridDesc gb{};
  gb.ntiles = int32_t(tiles);
  gb.size_tot = tiles * ni;
  for (int k = 0; k < 3; ++k) gb.shape_tot[k] = gb.shape_o[k] = 1;

  int st = ws.ensure_batch(ninst, tiles, sizeof(InstDesc));
  if (st) return st;
  InstDesc* inst_d = static_cast<InstDesc*>(ws.inst);
  AGD_HIP_TRY(hipMemcpyAsync(inst_d, h.data(), size_t(ninst) * sizeof(InstDesc),
                             hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(tile_inst_kernel, dim3(unsigned(ceil_div(tiles, 256))), dim3(256), 0, stream,
                     ninst, inst_d, int32_t(tiles), ws.tile_inst);
  AGD_HIP_TRY(hipGetLastError());
  AGD_HIP_TRY(hipStreamSynchronize(stream));  // h is released on return

  // workgroups of ks.block lanes, each visiting ks.tiles_per_block tiles at once
  const int kb = ks.tiles_per_block;
  const int blocks = blocks_for(int(ceil_div(tiles, kb)), ks.block);
  const bool global = d->solver == AGD_GLOBAL;
  EikArgs<S> a;
  std::memset(&a, 0, sizeof a);
  a.u = static_cast<S*>(io->values);
  a.geom = static_cast<const S*>(io->geom);
  a.rhs = static_cast<const S*>(io->rhs);
  a.seeds = io->seeds;
  a.g = gb;
  a.niter_i = d->niter_i;
  a.solver = global ? AGD_AGSI : d->solver;
  a.fim_front = d->fim_front_width;
  a.inst = inst_d;
  a.tile_inst = ws.tile_inst;
  auto init = [&](Workspace* ws) -> int {
    a.w.lists = ws->lists;
    a.w.counts = ws->counts;
    a.w.stamp = ws->stamp;
    a.w.hist = ws->hist;
    a.w.score = ws->score;
    a.w.tile_cap = int32_t(ws->tile_cap);
    a.w.shard_cap = int32_t(ws->shard_cap);
    a.w.fixed = nullptr;
    fill_i32(ws->score, 2 * ws->tile_cap, 0x7f7f7f7f, stream);
    if (global) {
      hipLaunchKernelGGL(all_ids_kernel, dim3(unsigned(ceil_div(tiles, 256))), dim3(256), 0, stream,
                         int32_t(tiles), ws->score, ws->counts + kFixedIdx);
      a.w.fixed = ws->score;
    } else {
      const uint32_t* trig = io->trigger ? io->trigger : io->seeds;
      hipLaunchKernelGGL(ks.init_ids, dim3(unsigned(ceil_div(tiles, 64))), dim3(64), 0, stream, trig,
                         int32_t(tiles), d->activate_all, ws->stamp, ws->lists, ws->counts,
                         a.w.shard_cap);
    }
    return AGD_OK;
  };
  auto launch = [&](int32_t koff, hipStream_t s, bool phase) {
    a.w.koff = koff;
    a.w.phase_from_koff = phase ? 1 : 0;
    hipLaunchKernelGGL(ks.update_batch, dim3(blocks), dim3(ks.block), 0, s, a);
  };
  auto key = [&](std::vector<unsigned char>& k) {
    key_append(k, ks.update_batch);
    key_append(k, a);
  };
  WorklistResult r;
  st = run_worklist(ws, int32_t(tiles), blocks, d->nitermax_o, d->check_interval, stream, init,
                    launch, key, r, d->time_launches != 0, d->use_graph != 0, kb);
  if (st) return st;
  if (stats) {
    stats->kernel_ms = r.kernel_ms;
    stats->niter_o = r.niter;
    stats->tile_updates = r.visits;
    stats->cell_updates = r.visits * ni;
    stats->launches = r.passes;
    stats->gpu_ms = r.ms;
    stats->converged = r.converged ? 1 : 0;
    stats->grid_blocks = blocks;
  }
  if (!r.converged) {
    set_error("batched eikonal solve did not converge within nitermax_o passes");
    return AGD_NOCONV;
  }
  return AGD_OK;
}

template <class S>
static int scheme_impl(const agd_eikonal_desc* d, const agd_eikonal_io* io, hipStream_t stream) {
  KernelSet<S> ks;
  AGD_REQUIRE(lookup<S>(d->model, d->ndim, d->shape_i, ks), "unsupported tile shape for this model");
  AGD_REQUIRE(ks.scheme, "no precomputed-scheme variant for this model");
  AGD_REQUIRE(io && io->geom && io->scheme_weights && io->scheme_codes,
              "geom, scheme_weights and scheme_codes are required");
  const GridDesc g = make_grid(d);
  const int stl = check_scheme_lengths(d, io, g.size_tot);
  if (stl) return stl;
  SchemeArgs<S> a{};
  a.geom = static_cast<const S*>(io->geom);
  a.wall = io->wall_dist;
  a.g = g;
  for (int k = 0; k < 3; ++k) a.iso_w[k] = S(d->iso_weights[k]);
  a.sw = static_cast<S*>(io->scheme_weights);
  a.scode = io->scheme_codes;
  a.scode2 = (d->model == AGD_TTI && d->order == 2) ? io->scheme_codes + tti_codes2_offset(d) * g.size_tot
                                                     : nullptr;
  a.relax = S(d->curvature_relax);
  a.convex = d->convex;
  hipLaunchKernelGGL((io->wall_dist || g.periodic) ? ks.scheme_general : ks.scheme, dim3(g.ntiles),
                     dim3(ks.ni), 0, stream, a);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

}  // namespace agd

using namespace agd;

// A solver handle: the workspace of the solves run through it, on the device
// current at creation and on one stream.
struct agd_solver {
  int device = 0;
  hipStream_t stream = nullptr;
  Workspace ws;
};

namespace agd {

static int check_handle(const agd_solver* s) {
  AGD_REQUIRE(s != nullptr, "null solver handle");
  int dev = 0;
  AGD_HIP_TRY(hipGetDevice(&dev));
  AGD_REQUIRE(dev == s->device, "the solver handle belongs to another device than the current one");
  return AGD_OK;
}

static int solve_entry(Workspace& ws, const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                       agd_eikonal_stats* stats, hipStream_t stream) {
  int st = check_desc(desc);
  if (st) return st;
  if (desc->scalar == AGD_F32) return solve_impl<float>(ws, desc, io, stats, stream);
  return solve_impl<double>(ws, desc, io, stats, stream);
}

static int batched_entry(Workspace& ws, const agd_eikonal_desc* desc, int32_t ninst,
                         const agd_batch_instance* instances, const agd_eikonal_io* io,
                         agd_eikonal_stats* stats, hipStream_t stream) {
  int st = check_desc(desc);
  if (st) return st;
  if (desc->scalar == AGD_F32)
    return solve_batched_impl<float>(ws, desc, ninst, instances, io, stats, stream);
  return solve_batched_impl<double>(ws, desc, ninst, instances, io, stats, stream);
}

// Residual check with the counter `cnt` (one int32 of device memory); with
// stale, the converse test (cells above their update), marked in stale_bits.
static int check_entry(int32_t* cnt, const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                       double atol, double rtol, int64_t* violations, hipStream_t s,
                       bool stale = false, uint32_t* stale_bits = nullptr) {
  int st = check_desc(desc);
  if (st) return st;
  AGD_REQUIRE(violations, "null output");
  fill_i32(cnt, 1, 0, s);
  st = desc->scalar == AGD_F32
           ? flow_impl<float>(desc, io, nullptr, s, cnt, atol, rtol, false, stale_bits, stale)
           : flow_impl<double>(desc, io, nullptr, s, cnt, atol, rtol, false, stale_bits, stale);
  if (st) return st;
  int32_t n = 0;
  AGD_HIP_TRY(hipMemcpyAsync(&n, cnt, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  AGD_HIP_TRY(hipStreamSynchronize(s));
  *violations = n;
  return AGD_OK;
}

// ------------------------------------------------------------------ fixed region
// Worklist of agd_eikonal_solve_noalloc: one array per device, held until
// the process ends.  Host code reaches it only through the address that
// hipGetSymbolAddress returns on each device (scratch_ptr), never by name.
constexpr size_t kScratchBytes = size_t(12) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::once_flag g_scratch_once[kMaxDevices];
static void* g_scratch_addr[kMaxDevices];
static hipError_t g_scratch_err[kMaxDevices];

// The device address of `scratch` on the current device (looked up once per
// device, with that device current).
static int scratch_ptr(void*& out) {
  int dev = 0;
  AGD_HIP_TRY(hipGetDevice(&dev));
  AGD_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index beyond the fixed-region table");
  std::call_once(g_scratch_once[dev], [dev] {
    void* p = nullptr;
    g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_addr[dev] = p;
  });
  AGD_HIP_TRY(g_scratch_err[dev]);
  AGD_REQUIRE(g_scratch_addr[dev] != nullptr, "no device address for the fixed region");
  out = g_scratch_addr[dev];
  return AGD_OK;
}

// Repair rounds with the counter `cnt`: each counts the cells below their
// scheme update and raises them to +inf; stops after a round that counts
// nothing or after max_rounds rounds.
static int repair_entry(int32_t* cnt, const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                        double atol, double rtol, int32_t max_rounds, int64_t* left,
                        int32_t* rounds, hipStream_t s) {
  int st = check_desc(desc);
  if (st) return st;
  AGD_REQUIRE(left && rounds, "null output");
  AGD_REQUIRE(max_rounds >= 1, "max_rounds must be >= 1");
  int32_t n = 0, r = 0;
  do {
    fill_i32(cnt, 1, 0, s);
    st = desc->scalar == AGD_F32 ? flow_impl<float>(desc, io, nullptr, s, cnt, atol, rtol, true)
                                 : flow_impl<double>(desc, io, nullptr, s, cnt, atol, rtol, true);
    if (st) return st;
    AGD_HIP_TRY(hipMemcpyAsync(&n, cnt, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    AGD_HIP_TRY(hipStreamSynchronize(s));
    ++r;
  } while (n > 0 && r < max_rounds);
  *left = n;
  *rounds = r;
  return AGD_OK;
}

}  // namespace agd

extern "C" {

int agd_eikonal_solve_noalloc(const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                              agd_eikonal_stats* stats, void* stream) {
  int st = check_desc(desc);
  if (st) return st;
  // Sizes first (no device call): the worklist this solve needs.
  const GridDesc g = make_grid(desc);
  int ni = 1;
  for (int k = 0; k < desc->ndim; ++k) ni *= desc->shape_i[k];
  const int64_t scap = shard_capacity(g.ntiles, blocks_for(g.ntiles, ni));
  const int64_t nhist = hist_len(desc->nitermax_o, desc->check_interval);
  const size_t need = Workspace::fixed_bytes(g.ntiles, scap, nhist);
  if (need > kScratchBytes) {
    set_error("invalid argument: the solve needs " + std::to_string(need) +
              " bytes of worklist, beyond the fixed region of " + std::to_string(kScratchBytes) +
              " bytes (about 150,000 tiles and 2,000 passes); use agd_eikonal_solve");
    return AGD_EINVAL;
  }
  void* base = nullptr;
  st = scratch_ptr(base);
  if (st) return st;
  Workspace ws;  // views the fixed region: nothing is allocated or freed
  st = ws.bind_fixed(base, kScratchBytes, g.ntiles, scap, nhist);
  if (st) return st;
  return solve_entry(ws, desc, io, stats, as_stream(stream));
}

int agd_solver_create(const agd_eikonal_desc* desc, void* stream, agd_solver** out) {
  AGD_REQUIRE(out != nullptr, "null output");
  *out = nullptr;
  if (desc) {
    const int st = check_desc(desc);
    if (st) return st;
  }
  agd_solver* s = new (std::nothrow) agd_solver;
  if (!s) {
    set_error("out of host memory");
    return AGD_ENOMEM;
  }
  s->stream = as_stream(stream);
  if (hipGetDevice(&s->device) != hipSuccess) {
    delete s;
    set_error("hipGetDevice failed");
    return AGD_EHIP;
  }
  if (desc) {
    // Size the workspace for the descriptor's grid (it still grows on demand).
    const GridDesc g = make_grid(desc);
    int ni = 1;
    for (int k = 0; k < desc->ndim; ++k) ni *= desc->shape_i[k];
    const int st = s->ws.ensure(g.ntiles, shard_capacity(g.ntiles, blocks_for(g.ntiles, ni)),
                                hist_len(desc->nitermax_o, desc->check_interval));
    if (st) {
      delete s;
      return st;
    }
  }
  *out = s;
  return AGD_OK;
}

int agd_solver_destroy(agd_solver* s) {
  if (!s) return AGD_OK;
  int cur = 0;
  const bool swap = hipGetDevice(&cur) == hipSuccess && cur != s->device &&
                    hipSetDevice(s->device) == hipSuccess;
  delete s;
  if (swap) (void)hipSetDevice(cur);
  return AGD_OK;
}

int agd_solver_solve(agd_solver* s, const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                     agd_eikonal_stats* stats) {
  const int st = check_handle(s);
  if (st) return st;
  return solve_entry(s->ws, desc, io, stats, s->stream);
}

int agd_solver_solve_batched(agd_solver* s, const agd_eikonal_desc* desc, int32_t ninst,
                             const agd_batch_instance* instances, const agd_eikonal_io* io,
                             agd_eikonal_stats* stats) {
  const int st = check_handle(s);
  if (st) return st;
  return batched_entry(s->ws, desc, ninst, instances, io, stats, s->stream);
}

int agd_solver_check(agd_solver* s, const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                     double atol, double rtol, int64_t* violations) {
  int st = check_handle(s);
  if (st) return st;
  st = s->ws.ensure(1, 1, 1);
  if (st) return st;
  return check_entry(s->ws.check_count, desc, io, atol, rtol, violations, s->stream);
}

int agd_solver_check_stale(agd_solver* s, const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                           double atol, double rtol, uint32_t* stale_bits, int64_t* stale) {
  int st = check_handle(s);
  if (st) return st;
  st = s->ws.ensure(1, 1, 1);
  if (st) return st;
  return check_entry(s->ws.check_count, desc, io, atol, rtol, stale, s->stream, true, stale_bits);
}

int agd_solver_repair(agd_solver* s, const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                      double atol, double rtol, int32_t max_rounds, int64_t* violations_left,
                      int32_t* rounds) {
  int st = check_handle(s);
  if (st) return st;
  st = s->ws.ensure(1, 1, 1);
  if (st) return st;
  return repair_entry(s->ws.check_count, desc, io, atol, rtol, max_rounds, violations_left, rounds,
                      s->stream);
}

int agd_eikonal_solve_batched(const agd_eikonal_desc* desc, int32_t ninst,
                              const agd_batch_instance* instances, const agd_eikonal_io* io,
                              agd_eikonal_stats* stats, void* stream) {
  Workspace ws;  // this call's own, released on return
  return batched_entry(ws, desc, ninst, instances, io, stats, as_stream(stream));
}

int agd_scheme_supported(int32_t model, int32_t ndim, const int32_t* shape_i) {
  KernelSet<float> ks;
  return (shape_i && lookup<float>(model, ndim, shape_i, ks) && ks.scheme) ? 1 : 0;
}

int agd_scheme_lengths(int32_t model, int32_t ndim, int32_t order, int64_t* weight_components,
                       int64_t* code_components) {
  AGD_REQUIRE(ndim == 2 || ndim == 3, "ndim must be 2 or 3");
  AGD_REQUIRE(model >= AGD_ISOTROPIC && model <= AGD_DUBINS2, "unknown model");
  AGD_REQUIRE(weight_components && code_components, "null output");
  scheme_components(model, ndim, order, *weight_components, *code_components);
  return AGD_OK;
}

int agd_eikonal_scheme(const agd_eikonal_desc* desc, const agd_eikonal_io* io, void* stream) {
  int st = check_desc(desc);
  if (st) return st;
  if (desc->scalar == AGD_F32) return scheme_impl<float>(desc, io, as_stream(stream));
  return scheme_impl<double>(desc, io, as_stream(stream));
}

int agd_tile_supported(int32_t model, int32_t ndim, const int32_t* shape_i) {
  KernelSet<float> ks;
  return (shape_i && lookup<float>(model, ndim, shape_i, ks) && ks.tiles_per_block == 1) ? 1 : 0;
}

int agd_batch_tile_supported(int32_t model, int32_t ndim, const int32_t* shape_i) {
  KernelSet<float> ks;
  return (shape_i && lookup<float>(model, ndim, shape_i, ks) && ks.update_batch) ? 1 : 0;
}

int agd_eikonal_solve(const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                      agd_eikonal_stats* stats, void* stream) {
  Workspace ws;  // this call's own, released on return
  return solve_entry(ws, desc, io, stats, as_stream(stream));
}

int agd_geodesic_ode_online(const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                            const agd_geodesic_desc* gd, const uint8_t* eucl, void* x, int32_t* len,
                            int8_t* stop, void* stream) {
  int st = check_desc(desc);
  if (st) return st;
  if (desc->scalar == AGD_F32)
    return geodesic_online_impl<float>(desc, io, gd, eucl, x, len, stop, as_stream(stream));
  return geodesic_online_impl<double>(desc, io, gd, eucl, x, len, stop, as_stream(stream));
}

int agd_eikonal_flow(const agd_eikonal_desc* desc, const agd_eikonal_io* io,
                     const agd_flow_io* flow, void* stream) {
  int st = check_desc(desc);
  if (st) return st;
  if (desc->scalar == AGD_F32) return flow_impl<float>(desc, io, flow, as_stream(stream));
  return flow_impl<double>(desc, io, flow, as_stream(stream));
}

int agd_eikonal_check(const agd_eikonal_desc* desc, const agd_eikonal_io* io, double atol,
                      double rtol, int64_t* violations, void* stream) {
  hipStream_t s = as_stream(stream);
  int32_t* cnt = nullptr;
  AGD_HIP_TRY(hipMalloc((void**)&cnt, sizeof(int32_t)));
  const int st = check_entry(cnt, desc, io, atol, rtol, violations, s);
  (void)hipStreamSynchronize(s);
  (void)hipFree(cnt);
  return st;
}

int agd_stencil_export(int32_t model, int32_t ndim, int32_t scalar, int64_t n, const void* geom,
                       const double* iso_weights, void* weights, int32_t* offsets, void* stream) {
  AGD_REQUIRE(ndim == 2 || ndim == 3, "ndim must be 2 or 3");
  AGD_REQUIRE(n >= 0 && weights && offsets, "bad buffers");
  AGD_REQUIRE(model == AGD_ISOTROPIC || geom, "geometry required");
  if (n == 0) return AGD_OK;
  const double iw[3] = {iso_weights ? iso_weights[0] : 0., iso_weights ? iso_weights[1] : 0.,
                        iso_weights ? iso_weights[2] : 0.};
  hipStream_t s = as_stream(stream);
  const dim3 grid(unsigned(ceil_div(n, 256))), block(256);
#define AGD_EXPORT(S, M, D)                                                                  \
  hipLaunchKernelGGL((stencil_export_kernel<S, M, D>), grid, block, 0, s, n,                 \
                     static_cast<const S*>(geom), S(iw[0]), S(iw[1]), S(iw[2]),              \
                     static_cast<S*>(weights), offsets)
#define AGD_EXPORT_M(S)                                                         \
  switch (model * 10 + ndim) {                                                  \
    case AGD_ISOTROPIC * 10 + 2: AGD_EXPORT(S, AGD_ISOTROPIC, 2); break;        \
    case AGD_ISOTROPIC * 10 + 3: AGD_EXPORT(S, AGD_ISOTROPIC, 3); break;        \
    case AGD_DIAGONAL * 10 + 2: AGD_EXPORT(S, AGD_DIAGONAL, 2); break;          \
    case AGD_DIAGONAL * 10 + 3: AGD_EXPORT(S, AGD_DIAGONAL, 3); break;          \
    case AGD_RIEMANN * 10 + 2: AGD_EXPORT(S, AGD_RIEMANN, 2); break;            \
    case AGD_RIEMANN * 10 + 3: AGD_EXPORT(S, AGD_RIEMANN, 3); break;            \
    case AGD_RANDER * 10 + 2: AGD_EXPORT(S, AGD_RANDER, 2); break;              \
    case AGD_RANDER * 10 + 3: AGD_EXPORT(S, AGD_RANDER, 3); break;              \
    default: set_error("invalid argument: unknown model"); return AGD_EINVAL;   \
  }
  if (scalar == AGD_F32) { AGD_EXPORT_M(float) }
  else { AGD_EXPORT_M(double) }
#undef AGD_EXPORT_M
#undef AGD_EXPORT
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_scatter_seeds(int32_t scalar, int64_t nseeds, const int64_t* tiled_index,
                      const double* seed_values, void* rhs, uint32_t* seeds, uint32_t* trigger,
                      void* stream) {
  AGD_REQUIRE(nseeds >= 0, "negative seed count");
  if (nseeds == 0) return AGD_OK;
  AGD_REQUIRE(tiled_index && seed_values && rhs && seeds, "null buffer");
  hipStream_t s = as_stream(stream);
  const dim3 grid(unsigned(ceil_div(nseeds, 256))), block(256);
  if (scalar == AGD_F32)
    hipLaunchKernelGGL(scatter_seeds_kernel<float>, grid, block, 0, s, nseeds, tiled_index,
                       seed_values, static_cast<float*>(rhs), seeds, trigger);
  else
    hipLaunchKernelGGL(scatter_seeds_kernel<double>, grid, block, 0, s, nseeds, tiled_index,
                       seed_values, static_cast<double*>(rhs), seeds, trigger);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_prepare_walls(const agd_eikonal_desc* desc, const uint8_t* walls, int32_t bound,
                      void* rhs, uint32_t* seeds, uint8_t* wall_dist, void* stream) {
  int st = check_desc(desc);
  if (st) return st;
  AGD_REQUIRE(walls && rhs && seeds && wall_dist, "null buffer");
  AGD_REQUIRE(bound >= 0 && bound < 255, "wallDistBound must lie in [0,255)");
  KernelSet<float> ks;
  AGD_REQUIRE(lookup<float>(desc->model, desc->ndim, desc->shape_i, ks), "unsupported tile shape");
  const GridDesc g = make_grid(desc);
  hipStream_t s = as_stream(stream);
  const int64_t N = g.size_tot;
  const dim3 grid(unsigned(ceil_div(N, 256))), block(256);
  if (desc->scalar == AGD_F32)
    hipLaunchKernelGGL(walls_mark_kernel<float>, grid, block, 0, s, N, walls,
                       static_cast<float*>(rhs), seeds, wall_dist);
  else
    hipLaunchKernelGGL(walls_mark_kernel<double>, grid, block, 0, s, N, walls,
                       static_cast<double*>(rhs), seeds, wall_dist);
  AGD_HIP_TRY(hipGetLastError());
  if (bound > 0) {
    uint8_t* tmp = nullptr;
    AGD_HIP_TRY(hipMalloc((void**)&tmp, N));
    uint8_t* src = wall_dist;
    uint8_t* dst = tmp;
    for (int i = 0; i < bound; ++i) {
      hipLaunchKernelGGL(ks.wallpass, dim3(g.ntiles), dim3(ks.ni), 0, s, g, src, dst);
      std::swap(src, dst);
    }
    if (src != wall_dist)
      AGD_HIP_TRY(hipMemcpyAsync(wall_dist, src, N, hipMemcpyDeviceToDevice, s));
    AGD_HIP_TRY(hipStreamSynchronize(s));
    AGD_HIP_TRY(hipFree(tmp));
  }
  hipLaunchKernelGGL(wall_dist_cap, grid, block, 0, s, N, wall_dist, bound);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

}  // extern "C"
