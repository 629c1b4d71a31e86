// worklist.hpp -- device-resident tile worklist shared by the eikonal and the
// adjoint tile solvers.
//
// Pass k reads list[k%2] with length count[k%3], appends the tiles it
// activates to list[(k+1)%2] / count[(k+1)%3] (deduplicated by a per-tile
// stamp: atomicMax(stamp, k+1) returns a smaller value exactly once per pass)
// and zeroes count[(k+2)%3], the counter of pass k+1's successor.  No pass
// ever clears a list or a flag array, so a solve costs one launch per pass
// plus one 4-byte read-back every `check_interval` passes.  This replaces
// the reference's per-pass host round trip (np.any + flatnonzero + fill,
// _solvers.py:184-190, _StopCriterion.py:62-82).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"

namespace agd {

struct Workspace {
  int32_t* stamp = nullptr;
  int32_t* lists = nullptr;   // 2 * cap
  int32_t* counts = nullptr;  // 8
  int32_t* hist = nullptr;    // hist_cap: list length of every pass
  int32_t* score = nullptr;   // 2 * cap (FIM scores)
  int32_t* host_count = nullptr;  // pinned
  int64_t cap = 0, hist_cap = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

// Per-device cached workspace (grown on demand).  Defined in eikonal.hip.
int get_workspace(int64_t ntiles, int64_t nhist, Workspace*& out);
int blocks_for(int ntiles, int lanes_per_block);

__device__ inline void enqueue(int32_t tile, int32_t tag, int32_t* stamp, int32_t* list,
                               int32_t* count) {
  const int32_t old = atomicMax(stamp + tile, tag);
  if (old < tag) {
    const int32_t pos = atomicAdd(count, 1);
    list[pos] = tile;
  }
}

struct PassPointers {
  const int32_t* list_in;
  const int32_t* count_in;
  int32_t* list_out;
  int32_t* count_out;
  int32_t* count_clear;
  int32_t iter;
};

struct WorklistResult {
  int64_t passes = 0;      // launches issued
  int64_t niter = 0;       // passes with a non-empty list
  int64_t visits = 0;      // tile visits
  float ms = 0.f;
  double kernel_ms = 0.;   // sum of per-launch durations (time_launches only)
  bool converged = false;
};

// Runs passes until the list empties or nitermax passes were issued.
// `init(ws)` enqueues the initial list (stamp 0) into list0/counts[0];
// `launch(PassPointers)` enqueues one pass.  Stamps, counters and the pass
// histogram are reset here.
template <class Init, class Launch>
int run_worklist(int32_t ntiles, int64_t nitermax, int check_interval, hipStream_t stream,
                 Init init, Launch launch, WorklistResult& res, bool time_launches = false) {
  Workspace* ws = nullptr;
  int st = get_workspace(ntiles, nitermax + 8, ws);
  if (st) return st;
  int32_t* list0 = ws->lists;
  int32_t* list1 = ws->lists + ws->cap;
  int32_t* counts = ws->counts;
  AGD_HIP_TRY(hipMemsetAsync(ws->stamp, 0xff, size_t(ntiles) * sizeof(int32_t), stream));
  AGD_HIP_TRY(hipMemsetAsync(counts, 0, 8 * sizeof(int32_t), stream));
  AGD_HIP_TRY(hipMemsetAsync(ws->hist, 0, size_t(nitermax + 8) * sizeof(int32_t), stream));
  st = init(ws, list0, counts);
  if (st) return st;
  AGD_HIP_TRY(hipGetLastError());
  AGD_HIP_TRY(hipEventRecord(ws->ev0, stream));
  const int ci = std::max(1, check_interval);
  // Optional per-launch timing: one event pair per pass, read back at every
  // convergence check (the stream is synchronised there anyway).
  std::vector<hipEvent_t> evs;
  if (time_launches) {
    evs.resize(2 * size_t(ci));
    for (auto& e : evs) AGD_HIP_TRY(hipEventCreate(&e));
  }
  res.kernel_ms = 0.;
  int pending = 0;
  auto drain = [&]() -> int {
    for (int i = 0; i < pending; ++i) {
      float t = 0.f;
      AGD_HIP_TRY(hipEventElapsedTime(&t, evs[2 * i], evs[2 * i + 1]));
      res.kernel_ms += t;
    }
    pending = 0;
    return AGD_OK;
  };
  int64_t k = 0;
  bool converged = false;
  while (k < nitermax) {
    PassPointers p;
    p.list_in = (k % 2 == 0) ? list0 : list1;
    p.list_out = (k % 2 == 0) ? list1 : list0;
    p.count_in = counts + (k % 3);
    p.count_out = counts + ((k + 1) % 3);
    p.count_clear = counts + ((k + 2) % 3);
    p.iter = int32_t(k);
    if (time_launches) AGD_HIP_TRY(hipEventRecord(evs[2 * pending], stream));
    launch(p);
    if (time_launches) AGD_HIP_TRY(hipEventRecord(evs[2 * pending + 1], stream));
    ++pending;
    ++k;
    if (k % ci == 0 || k == nitermax) {
      AGD_HIP_TRY(hipMemcpyAsync(ws->host_count, counts + (k % 3), sizeof(int32_t),
                                 hipMemcpyDeviceToHost, stream));
      AGD_HIP_TRY(hipStreamSynchronize(stream));
      if (time_launches) {
        const int st2 = drain();
        if (st2) return st2;
      }
      pending = 0;
      if (*ws->host_count == 0) {
        converged = true;
        break;
      }
    }
  }
  AGD_HIP_TRY(hipGetLastError());
  AGD_HIP_TRY(hipEventRecord(ws->ev1, stream));
  AGD_HIP_TRY(hipEventSynchronize(ws->ev1));
  AGD_HIP_TRY(hipEventElapsedTime(&res.ms, ws->ev0, ws->ev1));
  if (time_launches) {
    const int st2 = drain();
    if (st2) return st2;
    for (auto& e : evs) AGD_HIP_TRY(hipEventDestroy(e));
  }
  std::vector<int32_t> hist(k);
  if (k > 0)
    AGD_HIP_TRY(hipMemcpy(hist.data(), ws->hist, size_t(k) * sizeof(int32_t), hipMemcpyDeviceToHost));
  res.passes = k;
  res.niter = 0;
  res.visits = 0;
  for (int64_t i = 0; i < k; ++i) {
    if (hist[i] > 0) ++res.niter;
    res.visits += hist[i];
  }
  res.converged = converged;
  return AGD_OK;
}

}  // namespace agd
