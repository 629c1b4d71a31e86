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

// The list of a pass is split into kShards sub-lists: a workgroup appends to
// sub-list blockIdx % kShards.  One counter shared by the whole grid
// serialises its atomics (about 11 ns each on MI355X, i.e. ~100 us for 8k
// appending waves); 64 counters on separate cache lines do not.
constexpr int kShards = 64;
constexpr int kCountStride = 16;                         // one 64-byte line per counter
constexpr int kRow = kShards * kCountStride;             // one pass's counters
constexpr int kFixedIdx = 3 * kRow;                      // fixed-list length (global iteration)
constexpr int kBaseIdx = 3 * kRow + kCountStride;        // pass base
constexpr int kCountsLen = 3 * kRow + 2 * kCountStride;  // cleared at the start of a solve
constexpr int kSumIdx = kCountsLen;                      // int64[2]: non-empty passes, visits
constexpr int kCountsAlloc = kCountsLen + 4;

// bound_active_blocks control block (one per device workspace): per-shard
// counters of the pass just run (tiles that propagated, tiles frozen, and
// the order-preserving integer images of the smallest and largest pending
// change value), then the two thresholds as doubles.
constexpr int kBabProp = 0;
constexpr int kBabFrozen = kRow;
constexpr int kBabMin = 2 * kRow;
constexpr int kBabMax = 3 * kRow;
constexpr int kBabThr = 4 * kRow;                         // double[2]: prev, next
constexpr int kBabLen = 4 * kRow + 2 * kCountStride;

// Worklist workspace of the solves run through one owner: a solver handle
// (agd_solver), a single call of an entry point without a handle, or the
// fixed device region of the allocation-free entry point.  Nothing is shared
// between workspaces, so solves with different workspaces may run at once.
struct Workspace {
  int32_t* stamp = nullptr;
  int32_t* needs = nullptr;   // bound_active_blocks: last pass a real update was requested for
  double* pending = nullptr;  // bound_active_blocks: unpropagated smallest change (+inf: none)
  int32_t* bab = nullptr;     // bound_active_blocks control block (kBabLen)
  int32_t* lists = nullptr;   // 2 * kShards * shard_cap
  int32_t* counts = nullptr;  // kCountsLen
  int32_t* hist = nullptr;    // hist_cap: list length of every pass
  int32_t* score = nullptr;   // 2 * tile_cap (FIM scores; fixed list of global iteration)
  int32_t* check_count = nullptr;  // residual-check counter
  void* inst = nullptr;            // batched solves: instance table (inst_cap InstDesc)
  int32_t* tile_inst = nullptr;    // batched solves: tile -> instance (tinst_cap)
  void* ftab = nullptr;            // TTI source factorization: shift table (ftab_bytes)
  size_t ftab_bytes = 0;
  int64_t tile_cap = 0, shard_cap = 0, hist_cap = 0, inst_cap = 0, tinst_cap = 0;
  // Host side.  host_counts: pinned (owned) or host_local (fixed region),
  // 2 * kRow entries, two chunks in flight.
  int32_t* host_counts = nullptr;
  int32_t host_local[2 * kRow];
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t chunk_ev[2] = {nullptr, nullptr};  // list length of a chunk copied
  // Private stream used only to record pass chunks into a graph (the legacy
  // null stream, torch's default, cannot be captured).  Nothing executes on
  // it: the instantiated graph is launched on the caller's stream.
  hipStream_t capture_stream = nullptr;
  // Last instantiated pass chunk and the bytes (kernel, arguments, grid,
  // chunk length) it was captured with: a solve with identical bytes
  // replays it instead of capturing again.
  hipGraphExec_t exec = nullptr;
  std::vector<unsigned char> exec_key;
  // Fixed region (allocation-free entry point): nothing is ever allocated,
  // grown or freed, no events are recorded and passes launch eagerly.
  bool fixed = false;

  Workspace() = default;
  Workspace(const Workspace&) = delete;
  Workspace& operator=(const Workspace&) = delete;
  ~Workspace() { release(); }

  // Grows the device buffers to ntiles tiles, sub-lists of shard_cap entries
  // and nhist passes (allocating the host side on first use).
  int ensure(int64_t ntiles, int64_t shard_cap, int64_t nhist);
  // Batched solves: instance table and tile map of ninst instances, ntiles tiles.
  int ensure_batch(int64_t ninst, int64_t ntiles, size_t inst_bytes);
  // TTI source factorization: a shift table of at least `bytes`.
  int ensure_ftab(size_t bytes);
  // Frees everything owned (no-op for a fixed region).
  void release();
  // Lays the buffers out in [base, base + bytes) (fixed region); AGD_EINVAL
  // if they do not fit.
  int bind_fixed(void* base, size_t bytes, int64_t ntiles, int64_t shard_cap, int64_t nhist);
  // Bytes bind_fixed needs.
  static size_t fixed_bytes(int64_t ntiles, int64_t shard_cap, int64_t nhist);
};

// Workgroups of a pass launch: enough resident waves per CU of the current
// device to hide the gather latency (the workgroups stride over the list).
int blocks_for(int ntiles, int lanes_per_block);

// Fills n int32 words with v (a kernel rather than hipMemsetAsync, so that
// the solve stays capturable by a caller).
__global__ inline void fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
inline void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_i32_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

// Appends the bytes of `x` to a graph cache key.
template <class T> inline void key_append(std::vector<unsigned char>& k, const T& x) {
  const unsigned char* p = reinterpret_cast<const unsigned char*>(&x);
  k.insert(k.end(), p, p + sizeof(T));
}

// Capacity of one sub-list.  A pass appends at most 2d+1 <= 7 tiles per
// tile visit, and the workgroups b = s mod kShards of a grid of G, each
// visiting K tiles at once, visit at most
// ceil(G/kShards) * K * ceil(n/(G K)) <= n/kShards + G K/kShards + n/G + K
// tiles; the initial list (64-lane workgroups over the tiles) puts at most
// ntiles/kShards + 64 in a sub-list; no sub-list exceeds ntiles (one append
// per tile and pass).
inline int64_t shard_capacity(int64_t ntiles, int64_t grid, int64_t tiles_per_block = 1) {
  auto cd = [](int64_t a, int64_t b) { return (a + b - 1) / b; };
  grid = std::max<int64_t>(grid, 1);
  const int64_t k = std::max<int64_t>(tiles_per_block, 1);
  const int64_t visits = cd(ntiles, kShards) + cd(grid * k, kShards) + cd(ntiles, grid) + k;
  return std::min<int64_t>(ntiles, 7 * visits + 64);
}

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

// Worklist state seen by a pass kernel.  counts holds three rows of kShards
// rotating sub-list lengths, the fixed-list length and the pass base.  A
// kernel runs pass iter = counts[kBaseIdx] + koff, so a chunk of passes with
// koff = 0..G-1 followed by an advance of the base by G can be captured once
// as a graph and replayed.
struct WorkCtx {
  int32_t* lists;          // 2 * kShards * shard_cap entries
  int32_t* counts;         // kCountsLen
  int32_t* stamp;          // ntiles
  int32_t* hist;           // pass histogram
  int32_t* score;          // 2 * tile_cap (FIM scores)
  const int32_t* fixed;    // non-null: every pass reads this list (length counts[kFixedIdx])
  int32_t tile_cap;
  int32_t shard_cap;
  int32_t koff;
  // 1 when the pass number is koff modulo 6 (eager passes: base 0; graph
  // chunks of a multiple of 6 passes): the counter row and the list parity
  // then follow from koff, and a wave loads only its row.
  int32_t phase_from_koff;
};

struct PassView {
  int32_t iter, count;     // pass number, total list length
  int32_t incl;            // this lane's inclusive prefix of the sub-list lengths
  const int32_t* list_in;  // sub-lists of this pass (or the fixed list)
  bool fixed;
  int32_t shard_cap;
  int32_t* list_out;       // this workgroup's sub-list for the next pass
  int32_t* count_out;      // ... and its counter
  const int32_t* score_prev;
  int32_t* score_next;
};

// Every wave loads the three counter rows and the pass base at once (no
// dependent load), keeps its pass's row as a prefix sum across lanes (lane
// s = sub-list s).  Block 0 zeroes the row of pass iter+2 and records the
// list length.  Must be called by every thread of the block.
__device__ inline PassView pass_begin(const WorkCtx& c) {
  const int lane = threadIdx.x & 63;
  const int base = c.counts[kBaseIdx];
  const int fixed_n = c.counts[kFixedIdx];
  PassView v;
  v.iter = base + c.koff;
  int mine, par;
  if (c.phase_from_koff) {
    mine = c.counts[(c.koff % 3) * kRow + lane * kCountStride];
    par = c.koff & 1;
  } else {
    const int c0 = c.counts[0 * kRow + lane * kCountStride];
    const int c1 = c.counts[1 * kRow + lane * kCountStride];
    const int c2 = c.counts[2 * kRow + lane * kCountStride];
    const int r = v.iter % 3;
    mine = r == 0 ? c0 : (r == 1 ? c1 : c2);
    par = v.iter & 1;
  }
  int incl = mine;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(incl, d);
    if (lane >= d) incl += o;
  }
  v.incl = incl;
  v.fixed = c.fixed != nullptr;
  v.count = v.fixed ? fixed_n : __shfl(incl, 63);
  v.shard_cap = c.shard_cap;
  v.list_in = v.fixed ? c.fixed : c.lists + int64_t(par) * kShards * c.shard_cap;
  const int shard = blockIdx.x % kShards;
  v.list_out = c.lists + int64_t((v.iter + 1) & 1) * kShards * c.shard_cap +
               int64_t(shard) * c.shard_cap;
  v.count_out = c.counts + ((v.iter + 1) % 3) * kRow + shard * kCountStride;
  v.score_prev = c.score + ((v.iter + 1) & 1) * c.tile_cap;
  v.score_next = c.score + (v.iter & 1) * c.tile_cap;
  if (blockIdx.x == 0 && threadIdx.x < 64) {
    c.counts[((v.iter + 2) % 3) * kRow + lane * kCountStride] = 0;
    if (lane == 0) c.hist[v.iter] = v.count;
  }
  return v;
}

// Entry li (0 <= li < count, the same in every lane of the wave) of the
// pass's list: sub-list s = number of lanes whose inclusive prefix is <= li.
__device__ inline int32_t pass_entry(const PassView& v, int li) {
  if (v.fixed) return v.list_in[li];
  const int s = __popcll(__ballot(v.incl <= li));
  const int excl = s == 0 ? 0 : __shfl(v.incl, s - 1);
  return v.list_in[int64_t(s) * v.shard_cap + (li - excl)];
}

__global__ inline void advance_kernel(int32_t* counts, int32_t g) { counts[kBaseIdx] += g; }

// out[0] = passes with a non-empty list, out[1] = tile visits, over hist[0..n).
// One workgroup of 256 lanes.
__global__ inline void __launch_bounds__(256) hist_sum_kernel(const int32_t* hist, int64_t n,
                                                              int64_t* out) {
  __shared__ int64_t part[2][4];
  int64_t ne = 0, vis = 0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const int32_t h = hist[i];
    ne += h > 0 ? 1 : 0;
    vis += h;
  }
  for (int off = 32; off > 0; off >>= 1) {
    ne += __shfl_xor(ne, off);
    vis += __shfl_xor(vis, off);
  }
  if ((threadIdx.x & 63) == 0) {
    part[0][threadIdx.x >> 6] = ne;
    part[1][threadIdx.x >> 6] = vis;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = part[0][0] + part[0][1] + part[0][2] + part[0][3];
    out[1] = part[1][0] + part[1][1] + part[1][2] + part[1][3];
  }
}

// ---- bound_active_blocks (Propagation.h:53-99,120-126, _solvers.py:195-224)
// Order-preserving map of a float onto int32 (atomicMin/atomicMax on floats).
__device__ inline int32_t ord_int(float f) {
  const int32_t i = __float_as_int(f);
  return i >= 0 ? i : int32_t(0x80000000u - uint32_t(i));
}
__device__ inline float ord_float(int32_t i) {
  return __int_as_float(i >= 0 ? i : int32_t(0x80000000u - uint32_t(i)));
}

struct BabCtx {
  int32_t* ctl;            // control block (kBabLen), null: disabled
  int32_t* needs;          // per tile
  double* pending;         // per tile
};

// Threshold update after pass iter (one wave; the kernel of pass iter + 1
// reads the new value).  nConsidered = tiles the pass propagated or froze;
// below `bound` the threshold is lifted (+inf).  Otherwise the threshold
// grows by max(delta * clamp(bound / nActive, 0.7, 2), delta_min), where
// delta is the previous increase or, at the start, the spread of the
// pending change values.  Counters are cleared for the next pass.
__global__ inline void __launch_bounds__(64) bab_threshold_kernel(int32_t* ctl, int64_t bound,
                                                                  double delta_min) {
  const int lane = threadIdx.x;
  int prop = ctl[kBabProp + lane * kCountStride];
  int frz = ctl[kBabFrozen + lane * kCountStride];
  int mn = ctl[kBabMin + lane * kCountStride];
  int mx = ctl[kBabMax + lane * kCountStride];
  for (int off = 32; off > 0; off >>= 1) {
    prop += __shfl_xor(prop, off);
    frz += __shfl_xor(frz, off);
    mn = min(mn, __shfl_xor(mn, off));
    mx = max(mx, __shfl_xor(mx, off));
  }
  ctl[kBabProp + lane * kCountStride] = 0;
  ctl[kBabFrozen + lane * kCountStride] = 0;
  ctl[kBabMin + lane * kCountStride] = ord_int(__builtin_huge_valf());
  ctl[kBabMax + lane * kCountStride] = ord_int(-__builtin_huge_valf());
  if (lane != 0) return;
  double* thr = reinterpret_cast<double*>(ctl + kBabThr);
  double prev = thr[0];
  double next = thr[1];
  const double prev_old = prev;
  prev = next;
  const int64_t considered = int64_t(prop) + frz;
  if (considered < bound) {
    next = __builtin_huge_val();
  } else {
    double lo = prev_old;
    double delta = next - lo;
    if (!__builtin_isfinite(delta)) {
      lo = double(ord_float(mn));
      next = double(ord_float(mx));
      delta = next - lo;
      if (!__builtin_isfinite(delta)) delta = 0.;
    }
    const double act = double(prop > 1 ? prop : 1);
    double mult = double(bound) / act;
    mult = mult < 2. ? mult : 2.;
    mult = mult > 0.7 ? mult : 0.7;
    const double inc = delta * mult;
    next += inc > delta_min ? inc : delta_min;
  }
  thr[0] = prev;
  thr[1] = next;
}

struct WorklistResult {
  int64_t passes = 0;      // launches issued
  int64_t niter = 0;       // passes with a non-empty list
  int64_t visits = 0;      // tile visits
  float ms = 0.f;
  double kernel_ms = 0.;   // sum of per-launch durations (time_launches only)
  bool converged = false;
};

// Passes a solve needs room for in the pass histogram.
inline int64_t hist_len(int64_t nitermax, int check_interval) {
  return nitermax + std::max(1, check_interval) + 8;
}

// Runs passes until the list empties or nitermax passes were issued.
// `init(ws)` enqueues the initial list into the sub-lists of pass 0 (list
// row 0, counter row 0); `launch(koff, stream)` enqueues one pass of
// iteration base + koff on `stream`, on a grid of `grid` workgroups.
// With `graph`, chunks of `check_interval` passes are captured into a
// hipGraph and replayed; `key(k)` appends to k the bytes the captured
// launches depend on (kernel, arguments, grid), and the workspace's last
// graph is replayed when they are unchanged.  The host reads the next list
// length after each chunk.  Optional per-pass timing records an event pair
// around every pass (and then launches eagerly).  A fixed-region workspace
// launches eagerly and records no event.
template <class Init, class Launch, class Key>
int run_worklist(Workspace& wsr, int32_t ntiles, int32_t grid, int64_t nitermax, int check_interval,
                 hipStream_t stream, Init init, Launch launch, Key key, WorklistResult& res,
                 bool time_launches = false, bool graph = false, int tiles_per_block = 1) {
  Workspace* ws = &wsr;
  const int ci = std::max(1, check_interval);
  const int64_t nhist = hist_len(nitermax, ci);
  int st = ws->ensure(ntiles, shard_capacity(ntiles, grid, tiles_per_block), nhist);
  if (st) return st;
  if (ws->fixed) {
    time_launches = false;
    graph = false;
  }
  int32_t* counts = ws->counts;
  fill_i32(ws->stamp, ntiles, -1, stream);
  fill_i32(counts, kCountsLen, 0, stream);
  fill_i32(ws->hist, nhist, 0, stream);
  st = init(ws);
  if (st) return st;
  AGD_HIP_TRY(hipGetLastError());

  // Events are released on every return path.
  struct Owned {
    std::vector<hipEvent_t> evs;
    ~Owned() {
      for (auto e : evs)
        if (e) (void)hipEventDestroy(e);
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
  hipGraphExec_t exec = nullptr;
  if (graph) {
    std::vector<unsigned char> k;
    key(k);
    key_append(k, grid);
    key_append(k, ci);
    if (ws->exec && k == ws->exec_key) {
      exec = ws->exec;
    } else {
      if (ws->exec) {
        (void)hipGraphExecDestroy(ws->exec);
        ws->exec = nullptr;
        ws->exec_key.clear();
      }
      hipGraph_t g = nullptr;
      hipStream_t cs = ws->capture_stream;
      AGD_HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      for (int i = 0; i < ci; ++i) launch(i, cs, ci % 6 == 0);
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
      ws->exec = exec;
      ws->exec_key = std::move(k);
    }
  }

  const bool timed = ws->ev0 != nullptr;
  if (timed) AGD_HIP_TRY(hipEventRecord(ws->ev0, stream));
  int64_t k = 0;
  bool converged = false;
  auto list_length = [&](const int32_t* row) {
    int64_t n = 0;
    for (int sh = 0; sh < kShards; ++sh) n += row[sh * kCountStride];
    return n;
  };
  if (graph) {
    // Pipelined chunks: chunk c+1 is queued before the host looks at the
    // list length left by chunk c (copied right behind chunk c, ahead of
    // chunk c+1 in stream order), so the device never idles on the host's
    // check.  The chunk queued after convergence runs empty passes.
    int slot = 0;
    AGD_HIP_TRY(hipGraphLaunch(exec, stream));
    k += ci;
    AGD_HIP_TRY(hipMemcpyAsync(ws->host_counts, counts + (k % 3) * kRow, kRow * sizeof(int32_t),
                               hipMemcpyDeviceToHost, stream));
    AGD_HIP_TRY(hipEventRecord(ws->chunk_ev[0], stream));
    while (true) {
      const bool more = k < nitermax;
      if (more) {
        AGD_HIP_TRY(hipGraphLaunch(exec, stream));
        k += ci;
        AGD_HIP_TRY(hipMemcpyAsync(ws->host_counts + (1 - slot) * kRow, counts + (k % 3) * kRow,
                                   kRow * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        AGD_HIP_TRY(hipEventRecord(ws->chunk_ev[1 - slot], stream));
      }
      AGD_HIP_TRY(hipEventSynchronize(ws->chunk_ev[slot]));
      if (list_length(ws->host_counts + slot * kRow) == 0) {
        converged = true;
        break;
      }
      if (!more) break;
      slot = 1 - slot;
    }
  } else {
    while (k < nitermax) {
      if (time_launches) AGD_HIP_TRY(hipEventRecord(evs[2 * pending], stream));
      launch(int32_t(k), stream, true);
      if (time_launches) AGD_HIP_TRY(hipEventRecord(evs[2 * pending + 1], stream));
      ++pending;
      ++k;
      if (k % ci == 0 || k >= nitermax) {
        AGD_HIP_TRY(hipMemcpyAsync(ws->host_counts, counts + (k % 3) * kRow,
                                   kRow * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        AGD_HIP_TRY(hipStreamSynchronize(stream));
        if (time_launches) {
          const int st2 = drain();
          if (st2) return st2;
        }
        pending = 0;
        if (list_length(ws->host_counts) == 0) {
          converged = true;
          break;
        }
      }
    }
  }
  AGD_HIP_TRY(hipGetLastError());
  if (timed) {
    AGD_HIP_TRY(hipEventRecord(ws->ev1, stream));
    AGD_HIP_TRY(hipEventSynchronize(ws->ev1));
    AGD_HIP_TRY(hipEventElapsedTime(&res.ms, ws->ev0, ws->ev1));
  } else {
    res.ms = 0.f;
  }
  if (time_launches) {
    const int st2 = drain();
    if (st2) return st2;
  }
  // Non-empty passes and visits, summed on the device from the histogram.
  int64_t* sums = reinterpret_cast<int64_t*>(counts + kSumIdx);
  hipLaunchKernelGGL(hist_sum_kernel, dim3(1), dim3(256), 0, stream, ws->hist, k, sums);
  AGD_HIP_TRY(hipGetLastError());
  int64_t hs[2] = {0, 0};
  AGD_HIP_TRY(hipMemcpyAsync(hs, sums, sizeof(hs), hipMemcpyDeviceToHost, stream));
  AGD_HIP_TRY(hipStreamSynchronize(stream));
  res.passes = k;
  res.niter = hs[0];
  res.visits = hs[1];
  res.converged = converged;
  return AGD_OK;
}

}  // namespace agd
