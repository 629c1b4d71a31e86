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
  // Private stream used only to record pass chunks into a graph (the legacy
  // null stream, torch's default, cannot be captured).  Nothing executes on
  // it: the instantiated graph is launched on the caller's stream.
  hipStream_t capture_stream = nullptr;
};

// Per-device cached workspace (grown on demand).  Defined in eikonal.hip.
int get_workspace(int64_t ntiles, int64_t nhist, Workspace*& out);
int blocks_for(int ntiles, int lanes_per_block);

// Per-lane append of `entry` (default: the tile index) if `tile` was not yet
// stamped with `tag`.
__device__ inline void enqueue(int32_t tile, int32_t tag, int32_t* stamp, int32_t* list,
                               int32_t* count, int32_t entry = -1) {
  const int32_t old = atomicMax(stamp + tile, tag);
  if (old < tag) {
    const int32_t pos = atomicAdd(count, 1);
    list[pos] = entry >= 0 ? entry : tile;
  }
}

// Wave-aggregated append: every lane of the calling wave must execute this
// (lanes with want == false contribute nothing).  Lanes whose tile was not yet
// stamped for this pass append `entry`; one atomicAdd on the counter per wave.
__device__ inline void enqueue_wave(bool want, int32_t tile, int32_t entry, int32_t tag,
                                    int32_t* stamp, int32_t* list, int32_t* count) {
  const bool first = want && atomicMax(stamp + tile, tag) < tag;
  const uint64_t mask = __ballot(first);
  if (mask == 0) return;
  const int lane = __lane_id();
  const int leader = __ffsll((unsigned long long)mask) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(count, __popcll(mask));
  base = __shfl(base, leader);
  if (first) list[base + __popcll(mask & ((uint64_t(1) << lane) - 1))] = entry;
}

// Worklist state seen by a pass kernel.  counts[0..2] are the rotating list
// lengths, counts[3] the length of a fixed list (global iteration),
// counts[4] the pass base.  A kernel runs pass iter = counts[4] + koff, so
// a chunk of passes with koff = 0..G-1 followed by an advance of counts[4]
// by G can be captured once as a graph and replayed.
struct WorkCtx {
  int32_t* lists;          // 2 * cap entries
  int32_t* counts;         // 8
  int32_t* stamp;          // ntiles
  int32_t* hist;           // pass histogram
  int32_t* score;          // 2 * cap (FIM scores)
  const int32_t* fixed;    // non-null: every pass reads this list (length counts[3])
  int32_t cap;
  int32_t koff;
};

struct PassView {
  int32_t iter, count;
  const int32_t* list_in;
  int32_t* list_out;
  int32_t* count_out;
  const int32_t* score_prev;
  int32_t* score_next;
};

// One scalar load brings the counters and the pass base together.  Block 0
// zeroes the counter of pass iter+2 and records the list length.
__device__ inline PassView pass_begin(const WorkCtx& c) {
  int4 q0 = *reinterpret_cast<const int4*>(c.counts);
  int4 q1 = *reinterpret_cast<const int4*>(c.counts + 4);
  const int cnt[5] = {q0.x, q0.y, q0.z, q0.w, q1.x};
  PassView v;
  v.iter = cnt[4] + c.koff;
  const int r = v.iter % 3;
  v.count = c.fixed ? cnt[3] : (r == 0 ? cnt[0] : (r == 1 ? cnt[1] : cnt[2]));
  v.list_in = c.fixed ? c.fixed : c.lists + (v.iter & 1) * c.cap;
  v.list_out = c.lists + ((v.iter + 1) & 1) * c.cap;
  v.count_out = c.counts + (v.iter + 1) % 3;
  v.score_prev = c.score + ((v.iter + 1) & 1) * c.cap;
  v.score_next = c.score + (v.iter & 1) * c.cap;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c.counts[(v.iter + 2) % 3] = 0;
    c.hist[v.iter] = v.count;
  }
  return v;
}

__global__ inline void advance_kernel(int32_t* counts, int32_t g) { counts[4] += g; }

struct WorklistResult {
  int64_t passes = 0;      // launches issued
  int64_t niter = 0;       // passes with a non-empty list
  int64_t visits = 0;      // tile visits
  float ms = 0.f;
  double kernel_ms = 0.;   // sum of per-launch durations (time_launches only)
  bool converged = false;
};

// Runs passes until the list empties or nitermax passes were issued.
// `init(ws, list0, counts)` enqueues the initial list into list0 /
// counts[0]; `launch(koff, stream)` enqueues one pass of iteration
// counts[4] + koff on `stream`.
// With `graph`, chunks of `check_interval` passes are captured once into a
// hipGraph and replayed; the host reads the next list length after each
// chunk.  Optional per-pass timing records an event pair around every pass
// (and then launches eagerly).
template <class Init, class Launch>
int run_worklist(int32_t ntiles, int64_t nitermax, int check_interval, hipStream_t stream,
                 Init init, Launch launch, WorklistResult& res, bool time_launches = false,
                 bool graph = false) {
  Workspace* ws = nullptr;
  const int ci = std::max(1, check_interval);
  int st = get_workspace(ntiles, nitermax + ci + 8, ws);
  if (st) return st;
  int32_t* list0 = ws->lists;
  int32_t* counts = ws->counts;
  AGD_HIP_TRY(hipMemsetAsync(ws->stamp, 0xff, size_t(ntiles) * sizeof(int32_t), stream));
  AGD_HIP_TRY(hipMemsetAsync(counts, 0, 8 * sizeof(int32_t), stream));
  AGD_HIP_TRY(hipMemsetAsync(ws->hist, 0, size_t(nitermax + ci + 8) * sizeof(int32_t), stream));
  st = init(ws, list0, counts);
  if (st) return st;
  AGD_HIP_TRY(hipGetLastError());

  // Events and the graph are released on every return path.
  struct Owned {
    std::vector<hipEvent_t> evs;
    hipGraphExec_t exec = nullptr;
    ~Owned() {
      for (auto e : evs)
        if (e) (void)hipEventDestroy(e);
      if (exec) (void)hipGraphExecDestroy(exec);
    }
  } own;
  std::vector<hipEvent_t>& evs = own.evs;
  if (time_launches) {
    evs.assign(2 * size_t(ci), nullptr);
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

  // Per-launch timing needs event pairs around every launch, which a graph
  // does not record: timed runs launch eagerly.
  graph = graph && !time_launches;
  hipGraphExec_t& exec = own.exec;
  if (graph) {
    hipGraph_t g = nullptr;
    hipStream_t cs = ws->capture_stream;
    AGD_HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < ci; ++i) launch(i, cs);
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(1), 0, cs, counts, ci);
    // End the capture whatever happened inside it, so the stream is usable again.
    const hipError_t cap_err = hipGetLastError();
    const hipError_t end_err = hipStreamEndCapture(cs, &g);
    if (cap_err != hipSuccess || end_err != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      AGD_HIP_TRY(cap_err);
      AGD_HIP_TRY(end_err);
    }
    const hipError_t inst_err = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    AGD_HIP_TRY(inst_err);
  }

  AGD_HIP_TRY(hipEventRecord(ws->ev0, stream));
  int64_t k = 0;
  bool converged = false;
  while (k < nitermax) {
    if (graph) {
      AGD_HIP_TRY(hipGraphLaunch(exec, stream));
      pending = ci;
      k += ci;
    } else {
      if (time_launches) AGD_HIP_TRY(hipEventRecord(evs[2 * pending], stream));
      launch(int32_t(k), stream);
      if (time_launches) AGD_HIP_TRY(hipEventRecord(evs[2 * pending + 1], stream));
      ++pending;
      ++k;
    }
    if (graph || k % ci == 0 || k >= nitermax) {
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
