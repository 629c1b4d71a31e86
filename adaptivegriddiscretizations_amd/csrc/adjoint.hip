// adjoint.hip -- reverse-mode sensitivities of the eikonal solve.
//
// The converged upwind scheme, linearised, reads (D - W) du = c dc with
// W(x, y) = flow weight of x towards its upwind neighbour y and D(x) = sum of
// x's flow weights (1 on seeds / walls, where it vanishes).  Reverse mode
// solves (D - W)^T lambda = s:
//     lambda(y) = ( s(y) + sum_{x : y upwind of x} W(x, y) lambda(x) ) / D(y).
// (reference: _PostProcess.py:155-219 SolveAD, graph_reverse.py:10-77,
// GraphReverse.h:35-82, Kernel_LinearUpdate.h:99-216.)
//
// MI355X formulation:
//  1. transpose the flow graph into CSR rows indexed by the upwind cell y:
//     in-degree count (atomics), exclusive scan (hipCUB), scatter (atomic
//     cursors), then each row sorted by source index so the summation order,
//     hence the FP32 result, is deterministic;
//  2. each row stably partitioned into the sources inside the row's tile,
//     then those outside (rowloc marks the split);
//  3. tile Gauss-Seidel on the transposed system with the same device
//     worklist as the eikonal solver (AGSI activation, as the reference's
//     SolveLinear switches FIM to AGSI, _PostProcess.py:117-118).  The tile's
//     lambda values sit in LDS; the sources outside the tile do not change
//     during a visit, so their part of each row is summed once per visit,
//     and the in-tile part (slots and weights) is held in registers.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <cstring>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "worklist.hpp"

namespace agd {

template <class S>
__global__ void adj_count(int64_t N, int nsym, const S* fw, const int32_t* fidx, int32_t* deg) {
  const int64_t x = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (x >= N) return;
  for (int k = 0; k < nsym; ++k) {
    const int32_t y = fidx[k * N + x];
    if (fw[k * N + x] != S(0) && y != INT32_MAX) atomicAdd(deg + y, 1);
  }
}

template <class S>
__global__ void adj_fill(int64_t N, int nsym, const S* fw, const int32_t* fidx, int32_t* cursor,
                         int32_t* src, S* wt) {
  const int64_t x = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (x >= N) return;
  for (int k = 0; k < nsym; ++k) {
    const int32_t y = fidx[k * N + x];
    const S w = fw[k * N + x];
    if (w != S(0) && y != INT32_MAX) {
      const int32_t p = atomicAdd(cursor + y, 1);
      src[p] = int32_t(x);
      wt[p] = w;
    }
  }
}

// Sort each row by source index (rows hold a handful of entries).
template <class S>
__global__ void adj_sort_rows(int64_t N, const int32_t* rowptr, int32_t* src, S* wt) {
  const int64_t y = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (y >= N) return;
  const int32_t r0 = rowptr[y], r1 = rowptr[y + 1];
  for (int32_t i = r0 + 1; i < r1; ++i) {
    const int32_t s = src[i];
    const S w = wt[i];
    int32_t j = i - 1;
    while (j >= r0 && src[j] > s) {
      src[j + 1] = src[j];
      wt[j + 1] = wt[j];
      --j;
    }
    src[j + 1] = s;
    wt[j + 1] = w;
  }
}

// Stable partition of every row: sources in the row's own tile (of NI
// cells) first, then the others; rowloc[y] = end of the in-tile part.
template <class S>
__global__ void lin_partition_rows(int64_t N, int ni, const int32_t* rowptr, int32_t* src, S* wt,
                                   int32_t* rowloc) {
  const int64_t y = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (y >= N) return;
  const int32_t r0 = rowptr[y], r1 = rowptr[y + 1];
  const int64_t tile = y / ni;
  auto out = [&](int32_t s) { return int64_t(s) / ni != tile; };
  int32_t nloc = r0;
  for (int32_t i = r0; i < r1; ++i) {
    const int32_t s = src[i];
    if (out(s)) continue;
    const S w = wt[i];
    for (int32_t j = i; j > nloc; --j) {  // shift the outside entries right
      src[j] = src[j - 1];
      wt[j] = wt[j - 1];
    }
    src[nloc] = s;
    wt[nloc] = w;
    ++nloc;
  }
  rowloc[y] = nloc;
}

// Forward system: row x holds x's own upwind neighbours in stencil order
// (already deterministic, no sort needed).
template <class S>
__global__ void fwd_count(int64_t N, int nsym, const S* fw, const int32_t* fidx, int32_t* deg) {
  const int64_t x = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (x >= N) return;
  int c = 0;
  for (int k = 0; k < nsym; ++k) c += (fw[k * N + x] != S(0) && fidx[k * N + x] != INT32_MAX);
  deg[x] = c;
}

template <class S>
__global__ void fwd_fill(int64_t N, int nsym, const S* fw, const int32_t* fidx,
                         const int32_t* rowptr, int32_t* src, S* wt) {
  const int64_t x = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (x >= N) return;
  int32_t p = rowptr[x];
  for (int k = 0; k < nsym; ++k) {
    const int32_t y = fidx[k * N + x];
    const S w = fw[k * N + x];
    if (w != S(0) && y != INT32_MAX) {
      src[p] = y;
      wt[p] = w;
      ++p;
    }
  }
}

// Trigger: cells of the domain (finite value) whose row is empty
// (_PostProcess.py:143-144): in reverse mode no cell uses them as upwind
// neighbour, in forward mode they have no upwind neighbour (seeds).
template <class S>
__global__ void adj_trigger(int64_t N, const S* u, const int32_t* rowptr, uint32_t* trig) {
  const int64_t y = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (y >= N) return;
  const S v = u[y];
  if (rowptr[y + 1] == rowptr[y] && v == v && v != inf_v<S>())
    atomicOr(trig + (y >> 5), 1u << (y & 31));
}

template <class S> struct LinArgs {
  S* lam;
  const S* rhs;
  const S* diag;
  const int32_t* rowptr;
  const int32_t* rowloc;  // end of each row's in-tile part
  const int32_t* src;
  const S* wt;
  GridDesc g;
  S atol, rtol;
  int32_t niter_i;
  WorkCtx w;
};

template <class S> __global__ void fill_inf_kernel(int64_t n, S* x) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[i] = inf_v<S>();
}

template <class S> __device__ inline S init_val(S x) { return x == inf_v<S>() ? S(0) : x; }

template <class S, int D, int T0, int T1, int T2>
__global__ void __launch_bounds__(T0* T1* T2) adj_update(LinArgs<S> a) {
  static_assert((T0 * T1 * T2) % 64 == 0, "tiles are whole waves (pass_begin scans per wave)");
  using TL = Tile<D, T0, T1, T2>;
  constexpr int NI = TL::NI;
  static_assert((NI & (NI - 1)) == 0, "in-tile slots are taken as src & (NI - 1)");
  __shared__ S l_s[NI];
  const int tid = threadIdx.x;
  const PassView pv = pass_begin(a.w);
  const int count = pv.count;
  const int32_t tag = pv.iter + 1;
  for (int li = blockIdx.x; li < count; li += gridDim.x) {
    const int n_o = pass_entry(pv, li);
    int xo[3];
    tile_pos<D>(n_o, a.g, xo);
    const int64_t n_t = int64_t(n_o) * NI + tid;
    const S lold = a.lam[n_t];
    const S rhs = a.rhs[n_t];
    const S dg = a.diag[n_t];
    const S diag = dg == S(0) ? S(1) : dg;
    const int32_t r0 = a.rowptr[n_t], rl = a.rowloc[n_t], r1 = a.rowptr[n_t + 1];
    // sources outside the tile: fixed during the visit, summed once
    S sum_out = rhs;
    for (int32_t j = rl; j < r1; ++j) sum_out += a.wt[j] * init_val(a.lam[a.src[j]]);
    // in-tile sources: the first K held in registers (a zero weight pads
    // shorter rows), any further ones re-read at every sweep
    constexpr int K = 8;
    int slot[K];
    S wk[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const bool in = r0 + q < rl;
      slot[q] = in ? (a.src[r0 + q] & (NI - 1)) : tid;
      wk[q] = in ? a.wt[r0 + q] : S(0);
    }
    l_s[tid] = init_val(lold);
    __syncthreads();
    for (int it = 0; it < a.niter_i; ++it) {
      S sum = sum_out;
#pragma unroll
      for (int q = 0; q < K; ++q) sum += wk[q] * l_s[slot[q]];
      for (int32_t j = r0 + K; j < rl; ++j) sum += a.wt[j] * l_s[a.src[j] & (NI - 1)];
      const S nv = sum / diag;
      __syncthreads();
      const bool ch = !(nv == l_s[tid]);
      l_s[tid] = nv;
      if (!__syncthreads_or(ch)) break;
    }
    const S val = l_s[tid];
    a.lam[n_t] = val;
    const S tol = abs_(val) * a.rtol + a.atol;
    const bool changed = abs_(val - lold) > tol;
    const bool any = __syncthreads_or(changed);
    if (any && tid < 2 * D + 1) {
      const int nb = (tid == 2 * D) ? n_o : tile_neighbor<D>(xo, tid >> 1, 2 * (tid & 1) - 1, a.g);
      if (nb >= 0) enqueue(nb, tag, a.w.stamp, pv.list_out, pv.count_out);
    }
    __syncthreads();
  }
}

template <int NI>
__global__ void __launch_bounds__(64) adj_init_list(const uint32_t* trigger, int32_t ntiles,
                                                   int32_t* stamp, int32_t* lists,
                                                   int32_t* counts, int32_t shard_cap) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  bool on = false;
  if (t < ntiles) {
    const int64_t b0 = int64_t(t) * NI, b1 = b0 + NI;
    for (int64_t b = b0; b < b1 && !on; ++b) on = (trigger[b >> 5] >> (b & 31)) & 1u;
  }
  const int shard = blockIdx.x % kShards;
  enqueue_wave(on, on ? t : 0, t, 0, stamp, lists + int64_t(shard) * shard_cap,
               counts + shard * kCountStride);
}

template <class S> struct LinSet {
  void (*update)(LinArgs<S>);
  void (*init)(const uint32_t*, int32_t, int32_t*, int32_t*, int32_t*, int32_t);
  int ni;
};

template <class S, int D, int T0, int T1, int T2> LinSet<S> make_lin() {
  LinSet<S> s;
  s.update = adj_update<S, D, T0, T1, T2>;
  s.init = adj_init_list<T0 * T1 * T2>;
  s.ni = T0 * T1 * T2;
  return s;
}

template <class S> static bool lin_lookup(int d, const int32_t* si, LinSet<S>& out) {
  if (d == 2) {
    const int key = si[0] * 100 + si[1];
    switch (key) {
      case 808: out = make_lin<S, 2, 8, 8, 1>(); return true;
      case 1616: out = make_lin<S, 2, 16, 16, 1>(); return true;
      case 816: out = make_lin<S, 2, 8, 16, 1>(); return true;
    }
  } else if (d == 3) {
    const int key = si[0] * 10000 + si[1] * 100 + si[2];
    switch (key) {
      case 40404: out = make_lin<S, 3, 4, 4, 4>(); return true;
      case 80804: out = make_lin<S, 3, 8, 8, 4>(); return true;
      case 40408: out = make_lin<S, 3, 4, 4, 8>(); return true;
    }
  }
  return false;
}

template <class S>
static int linear_impl(bool forward, const agd_eikonal_desc* d, const S* values, const S* fw,
                       const int32_t* fidx, const S* wsum, int32_t nrhs, const S* sens, S* lam,
                       double atol, double rtol, agd_eikonal_stats* stats, hipStream_t stream) {
  LinSet<S> ls;
  AGD_REQUIRE(lin_lookup<S>(d->ndim, d->shape_i, ls), "unsupported tile shape");
  AGD_REQUIRE(values && fw && fidx && wsum && sens && lam && nrhs >= 1, "null buffer");
  GridDesc g{};
  g.ntiles = 1;
  for (int k = 0; k < 3; ++k) {
    g.shape_tot[k] = k < d->ndim ? int32_t(d->shape[k]) : 1;
    g.shape_o[k] = k < d->ndim ? int32_t(ceil_div(d->shape[k], d->shape_i[k])) : 1;
    g.ntiles *= g.shape_o[k];
  }
  g.periodic = 0;
  for (int k = 0; k < d->ndim; ++k)
    if (d->periodic[k]) g.periodic |= (1 << k);
  g.size_tot = int64_t(g.ntiles) * ls.ni;
  const int64_t N = g.size_tot;
  const int nsym = nsym_of(d->model, d->ndim);
  const dim3 grid(unsigned(ceil_div(N, 256))), block(256);

  // ---- 1. the system's rows (reverse: transposed flow graph)
  int32_t *deg = nullptr, *rowptr = nullptr, *rowloc = nullptr, *src = nullptr;
  uint32_t* trig = nullptr;
  S* wt = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  AGD_HIP_TRY(hipMalloc((void**)&deg, (N + 1) * sizeof(int32_t)));
  AGD_HIP_TRY(hipMalloc((void**)&rowptr, (N + 1) * sizeof(int32_t)));
  AGD_HIP_TRY(hipMalloc((void**)&rowloc, N * sizeof(int32_t)));
  AGD_HIP_TRY(hipMalloc((void**)&trig, ceil_div(N, 32) * sizeof(uint32_t)));
  AGD_HIP_TRY(hipMemsetAsync(deg, 0, (N + 1) * sizeof(int32_t), stream));
  AGD_HIP_TRY(hipMemsetAsync(trig, 0, ceil_div(N, 32) * sizeof(uint32_t), stream));
  if (forward)
    hipLaunchKernelGGL(fwd_count<S>, grid, block, 0, stream, N, nsym, fw, fidx, deg);
  else
    hipLaunchKernelGGL(adj_count<S>, grid, block, 0, stream, N, nsym, fw, fidx, deg);
  AGD_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, deg, rowptr, int(N + 1), stream));
  AGD_HIP_TRY(hipMalloc(&tmp, tmp_bytes));
  AGD_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, deg, rowptr, int(N + 1), stream));
  int32_t nnz = 0;
  AGD_HIP_TRY(hipMemcpyAsync(&nnz, rowptr + N, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
  AGD_HIP_TRY(hipStreamSynchronize(stream));
  AGD_HIP_TRY(hipMalloc((void**)&src, std::max<int64_t>(nnz, 1) * sizeof(int32_t)));
  AGD_HIP_TRY(hipMalloc((void**)&wt, std::max<int64_t>(nnz, 1) * sizeof(S)));
  if (forward) {
    hipLaunchKernelGGL(fwd_fill<S>, grid, block, 0, stream, N, nsym, fw, fidx, rowptr, src, wt);
  } else {
    AGD_HIP_TRY(hipMemcpyAsync(deg, rowptr, (N + 1) * sizeof(int32_t), hipMemcpyDeviceToDevice,
                               stream));
    hipLaunchKernelGGL(adj_fill<S>, grid, block, 0, stream, N, nsym, fw, fidx, deg, src, wt);
    hipLaunchKernelGGL(adj_sort_rows<S>, grid, block, 0, stream, N, rowptr, src, wt);
  }
  hipLaunchKernelGGL(lin_partition_rows<S>, grid, block, 0, stream, N, ls.ni, rowptr, src, wt, rowloc);
  hipLaunchKernelGGL(adj_trigger<S>, grid, block, 0, stream, N, values, rowptr, trig);
  AGD_HIP_TRY(hipGetLastError());

  // ---- 2. tile Gauss-Seidel, one right-hand side at a time
  const int blocks = blocks_for(g.ntiles, ls.ni);
  Workspace ws;  // this call's own worklist, released on return
  int64_t niter = 0, visits = 0, passes = 0;
  float ms = 0.f;
  double kernel_ms = 0.;
  bool converged = true;
  int status = AGD_OK;
  for (int r = 0; r < nrhs && status == AGD_OK; ++r) {
    LinArgs<S> a;
    std::memset(&a, 0, sizeof a);  // zeroed padding: the bytes key the graph cache
    a.lam = lam + int64_t(r) * N;
    a.rhs = sens + int64_t(r) * N;
    a.diag = wsum;
    a.rowptr = rowptr;
    a.rowloc = rowloc;
    a.src = src;
    a.wt = wt;
    a.g = g;
    a.atol = S(atol);
    a.rtol = S(rtol);
    a.niter_i = d->niter_i;
    // Dummy initialisation to +inf marks unvisited cells (read as 0).
    hipLaunchKernelGGL(fill_inf_kernel<S>, grid, block, 0, stream, N, a.lam);
    auto init = [&](Workspace* ws) -> int {
      a.w.lists = ws->lists;
      a.w.counts = ws->counts;
      a.w.stamp = ws->stamp;
      a.w.hist = ws->hist;
      a.w.score = ws->score;
      a.w.tile_cap = int32_t(ws->tile_cap);
      a.w.shard_cap = int32_t(ws->shard_cap);
      a.w.fixed = nullptr;
      hipLaunchKernelGGL(ls.init, dim3(unsigned(ceil_div(g.ntiles, 64))), dim3(64), 0, stream,
                         trig, g.ntiles, ws->stamp, ws->lists, ws->counts, a.w.shard_cap);
      return AGD_OK;
    };
    auto launch = [&](int32_t koff, hipStream_t s, bool phase) {
      a.w.koff = koff;
      a.w.phase_from_koff = phase ? 1 : 0;
      hipLaunchKernelGGL(ls.update, dim3(blocks), dim3(ls.ni), 0, s, a);
    };
    auto key = [&](std::vector<unsigned char>& k) {
      key_append(k, ls.update);
      key_append(k, a);
    };
    WorklistResult res;
    status = run_worklist(ws, g.ntiles, blocks, d->nitermax_o, d->check_interval, stream, init,
                          launch, key, res, d->time_launches != 0, d->use_graph != 0);
    kernel_ms += res.kernel_ms;
    niter += res.niter;
    visits += res.visits;
    passes += res.passes;
    ms += res.ms;
    converged = converged && res.converged;
  }
  (void)hipStreamSynchronize(stream);
  (void)hipFree(deg);
  (void)hipFree(rowptr);
  (void)hipFree(rowloc);
  (void)hipFree(trig);
  (void)hipFree(src);
  (void)hipFree(wt);
  (void)hipFree(tmp);
  if (status) return status;
  if (stats) {
    stats->niter_o = niter;
    stats->tile_updates = visits;
    stats->cell_updates = visits * ls.ni;
    stats->launches = passes;
    stats->gpu_ms = ms;
    stats->converged = converged ? 1 : 0;
    stats->grid_blocks = blocks;
    stats->kernel_ms = kernel_ms;
  }
  if (!converged) {
    set_error(forward ? "forward AD solver did not converge within nitermax_o passes"
                      : "adjoint solver did not converge within nitermax_o passes");
    return AGD_NOCONV;
  }
  return AGD_OK;
}

}  // namespace agd

using namespace agd;

namespace {
int linear_entry(bool forward, const agd_eikonal_desc* desc, const void* values,
                 const void* flow_weights, const int32_t* flow_indices, const void* flow_weightsum,
                 int32_t nrhs, const void* rhs, void* sol, double atol, double rtol,
                 agd_eikonal_stats* stats, void* stream) {
  AGD_REQUIRE(desc && (desc->ndim == 2 || desc->ndim == 3), "bad descriptor");
  AGD_REQUIRE(desc->niter_i >= 1 && desc->nitermax_o >= 1, "bad iteration counts");
  if (desc->scalar == AGD_F32)
    return linear_impl<float>(forward, desc, static_cast<const float*>(values),
                              static_cast<const float*>(flow_weights), flow_indices,
                              static_cast<const float*>(flow_weightsum), nrhs,
                              static_cast<const float*>(rhs), static_cast<float*>(sol), atol, rtol,
                              stats, as_stream(stream));
  return linear_impl<double>(forward, desc, static_cast<const double*>(values),
                             static_cast<const double*>(flow_weights), flow_indices,
                             static_cast<const double*>(flow_weightsum), nrhs,
                             static_cast<const double*>(rhs), static_cast<double*>(sol), atol, rtol,
                             stats, as_stream(stream));
}
}  // namespace

extern "C" int agd_eikonal_adjoint(const agd_eikonal_desc* desc, const void* values,
                                   const void* flow_weights, const int32_t* flow_indices,
                                   const void* flow_weightsum, int32_t nrhs, const void* sens,
                                   void* lambda, double atol, double rtol,
                                   agd_eikonal_stats* stats, void* stream) {
  return linear_entry(false, desc, values, flow_weights, flow_indices, flow_weightsum, nrhs, sens,
                      lambda, atol, rtol, stats, stream);
}

extern "C" int agd_eikonal_forward_ad(const agd_eikonal_desc* desc, const void* values,
                                      const void* flow_weights, const int32_t* flow_indices,
                                      const void* flow_weightsum, int32_t nrhs, const void* rhs,
                                      void* dvalues, double atol, double rtol,
                                      agd_eikonal_stats* stats, void* stream) {
  return linear_entry(true, desc, values, flow_weights, flow_indices, flow_weightsum, nrhs, rhs,
                      dvalues, atol, rtol, stats, stream);
}
