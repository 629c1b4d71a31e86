// Host-thread check of the tile-queue protocol of the persistent solver
// (adaptivegriddiscretizations_amd/csrc/tile_queue.hpp): the same ring, state
// and termination code the device kernel runs, driven by std::thread
// workers on a relaxation problem with a known answer.
//
// Problem: integer shortest-path distances on a 3D grid graph of `side`^3
// nodes ("tiles") with random edge weights from a few sources.  Visiting a
// tile relaxes its value from its six neighbours (Bellman-Ford at one node);
// a decrease requests the six neighbours.  The solve must end with every
// value equal to Dijkstra's, every ring drained (head == tail), the
// termination detected once, no worker left busy, and several solves in a
// row on the same rings (positions keep growing across solves) must all be
// exact.  Workers claim positions from any non-empty ring, so 8 workers over
// 64 rings is a case a ring-bound consumer would not survive.  Prints "ok"
// and exits 0 on success.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <queue>
#include <random>
#include <thread>
#include <vector>

#include "../adaptivegriddiscretizations_amd/csrc/tile_queue.hpp"

using namespace agd;

struct HostAtomics {
  static uint32_t add(uint32_t* p, uint32_t v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
  static uint32_t load(uint32_t* p) { return __atomic_load_n(p, __ATOMIC_SEQ_CST); }
  static void store(uint32_t* p, uint32_t v) { __atomic_store_n(p, v, __ATOMIC_SEQ_CST); }
  static uint32_t exch(uint32_t* p, uint32_t v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
  static uint32_t fetch_or(uint32_t* p, uint32_t v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
  static uint64_t load64(uint64_t* p) { return __atomic_load_n(p, __ATOMIC_SEQ_CST); }
  static void store64(uint64_t* p, uint64_t v) { __atomic_store_n(p, v, __ATOMIC_SEQ_CST); }
  static bool cas64(uint64_t* p, uint64_t e, uint64_t d) {
    return __atomic_compare_exchange_n(p, &e, d, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
  }
  static void pause() { std::this_thread::yield(); }
};
using A = HostAtomics;

struct Sim {
  int side, ntiles, nworkers;
  uint32_t cap, mask;
  std::vector<uint64_t> rings;           // kTqShards * cap
  std::vector<uint32_t> ctl;             // kTqCtlLen
  std::vector<uint32_t> state, busy;
  std::vector<int64_t> val;
  std::vector<int32_t> wt;               // 6 weights per node (to the node from each neighbour)
  std::vector<uint32_t> visits;
  uint64_t* ring(int s) { return rings.data() + size_t(s) * cap; }
  uint32_t* c(int i) { return ctl.data() + i; }

  int nb(int t, int k) const {
    int x = t / (side * side), y = (t / side) % side, z = t % side;
    int d[3] = {0, 0, 0};
    d[k >> 1] = (k & 1) ? 1 : -1;
    x += d[0]; y += d[1]; z += d[2];
    if (x < 0 || y < 0 || z < 0 || x >= side || y >= side || z >= side) return -1;
    return (x * side + y) * side + z;
  }

  bool push(uint32_t shard, const uint32_t* e, int n) {
    if (n == 0) return true;
    const uint32_t base = A::add(c(kTqTail + shard * kTqLine), uint32_t(n));
    for (int j = 0; j < n; ++j)
      if (!tq_write<A>(ring(shard), mask, base + j, e[j], 1u << 26)) return false;
    return true;
  }

  bool terminated() {
    uint32_t t1[kTqShards];
    for (int s = 0; s < kTqShards; ++s) {
      t1[s] = A::load(c(kTqTail + s * kTqLine));
      if (A::load(c(kTqHead + s * kTqLine)) != t1[s]) return false;
    }
    for (int w = 0; w < nworkers; ++w)
      if (A::load(&busy[w])) return false;
    for (int s = 0; s < kTqShards; ++s)
      if (A::load(c(kTqTail + s * kTqLine)) != t1[s]) return false;
    return true;
  }

  // Claims the oldest position of a non-empty ring, scanning from `start`.
  // Returns false if every ring looked empty.
  bool claim(uint32_t start, uint32_t& shard, uint32_t& pos) {
    for (int i = 0; i < kTqShards; ++i) {
      const uint32_t s = (start + uint32_t(i)) % kTqShards;
      uint32_t h = A::load(c(kTqHead + s * kTqLine));
      while (h < A::load(c(kTqTail + s * kTqLine))) {
        if (__atomic_compare_exchange_n(c(kTqHead + s * kTqLine), &h, h + 1, false,
                                        __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
          shard = s;
          pos = h;
          return true;
        }
      }
    }
    return false;
  }

  void worker(int w) {
    uint32_t psh = uint32_t(w) % kTqShards, start = psh;
    uint64_t idle = 0;
    for (;;) {
      if (A::load(c(kTqDone)) || A::load(c(kTqAbort))) break;
      A::store(&busy[w], 1);
      uint32_t s = 0, pos = 0;
      if (!claim(start, s, pos)) {
        A::store(&busy[w], 0);
        if ((++idle & 15) == 0 && terminated()) {
          A::store(c(kTqDone), 1);
          break;
        }
        if (idle > (1ull << 26)) {
          A::store(c(kTqAbort), 1);
          break;
        }
        A::pause();
        continue;
      }
      idle = 0;
      start = (start + 1) % kTqShards;
      uint32_t e = kTqEmpty;
      for (uint32_t k = 0; k < (1u << 26) && e == kTqEmpty; ++k) {
        e = tq_peek<A>(ring(s), mask, pos);
        if (e == kTqEmpty) A::pause();
      }
      if (e == kTqEmpty) {
        A::store(c(kTqAbort), 1);
        A::store(&busy[w], 0);
        break;
      }
      tq_release<A>(ring(s), mask, pos);
      A::exch(&state[e], 0);
      // visit: relax from the neighbours
      ++visits[w];
      const int t = int(e);
      int64_t best = __atomic_load_n(&val[t], __ATOMIC_SEQ_CST);
      const int64_t old = best;
      for (int k = 0; k < 6; ++k) {
        const int m = nb(t, k);
        if (m < 0) continue;
        const int64_t v = __atomic_load_n(&val[m], __ATOMIC_SEQ_CST);
        if (v != INT64_MAX && v + wt[size_t(t) * 6 + k] < best) best = v + wt[size_t(t) * 6 + k];
      }
      uint32_t out[6];
      int n = 0;
      if (best < old) {
        // monotone store: keep the smaller value if another visit raced ahead
        int64_t cur = old;
        while (best < cur && !__atomic_compare_exchange_n(&val[t], &cur, best, false,
                                                          __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
        }
        for (int k = 0; k < 6; ++k) {
          const int m = nb(t, k);
          if (m >= 0 && A::fetch_or(&state[m], 1) == 0) out[n++] = uint32_t(m);
        }
      }
      if (!push(psh, out, n)) {
        A::store(c(kTqAbort), 1);
        A::store(&busy[w], 0);
        break;
      }
      psh = (psh + 1) % kTqShards;
      A::store(&busy[w], 0);
    }
  }
};

int main(int argc, char** argv) {
  const int nthreads = argc > 1 ? std::atoi(argv[1]) : 8;
  int failures = 0;
  for (int trial = 0; trial < 6; ++trial) {
    std::mt19937 rng(1234 + trial);
    Sim S;
    S.side = 12 + 4 * (trial % 3);
    S.ntiles = S.side * S.side * S.side;
    S.nworkers = nthreads * (1 + trial % 2);
    S.cap = 1;
    while (S.cap < uint32_t(S.ntiles)) S.cap <<= 1;
    S.mask = S.cap - 1;
    S.rings.resize(size_t(kTqShards) * S.cap);
    for (int s = 0; s < kTqShards; ++s)
      for (uint32_t i = 0; i < S.cap; ++i) S.ring(s)[i] = tq_slot(i, kTqEmpty);
    S.ctl.assign(kTqCtlLen, 0);
    S.wt.resize(size_t(S.ntiles) * 6);
    for (auto& x : S.wt) x = 1 + int(rng() % 50);
    for (int solve = 0; solve < 3; ++solve) {
      const int nsrc = 1 + solve;
      std::vector<int> src;
      for (int i = 0; i < nsrc; ++i) src.push_back(int(rng() % uint32_t(S.ntiles)));
      // exact answer (Dijkstra)
      std::vector<int64_t> ref(S.ntiles, INT64_MAX);
      std::priority_queue<std::pair<int64_t, int>, std::vector<std::pair<int64_t, int>>,
                          std::greater<std::pair<int64_t, int>>> pq;
      for (int t : src) { ref[t] = 0; pq.push({0, t}); }
      while (!pq.empty()) {
        auto [d, t] = pq.top();
        pq.pop();
        if (d > ref[t]) continue;
        for (int k = 0; k < 6; ++k) {
          const int m = S.nb(t, k);
          if (m < 0) continue;
          const int back = k ^ 1;  // the edge into m from t is m's neighbour slot `back`
          const int64_t nd = d + S.wt[size_t(m) * 6 + back];
          if (nd < ref[m]) { ref[m] = nd; pq.push({nd, m}); }
        }
      }
      // solve: reset per-solve state; rings keep their positions
      S.val.assign(S.ntiles, INT64_MAX);
      S.state.assign(S.ntiles, 0);
      S.busy.assign(S.nworkers, 0);
      S.visits.assign(S.nworkers, 0);
      S.ctl[kTqDone] = S.ctl[kTqAbort] = 0;
      // initial list: the sources' neighbours (sources fixed at 0)
      for (int t : src) { S.val[t] = 0; S.state[t] = 1; }  // sources never relax
      std::vector<uint32_t> init;
      for (int t : src)
        for (int k = 0; k < 6; ++k) {
          const int m = S.nb(t, k);
          if (m >= 0 && S.state[m] == 0) { S.state[m] = 1; init.push_back(uint32_t(m)); }
        }
      for (size_t i = 0; i < init.size(); ++i) S.push(uint32_t(i % kTqShards), &init[i], 1);
      std::vector<std::thread> th;
      for (int w = 0; w < S.nworkers; ++w) th.emplace_back([&S, w] { S.worker(w); });
      for (auto& t : th) t.join();
      // sources keep state 1 so that no request ever pushes them: clear for the check
      for (int t : src) S.state[t] = 0;
      int bad = 0;
      for (int t = 0; t < S.ntiles; ++t) bad += S.val[t] != ref[t];
      bool drained = true;
      for (int s = 0; s < kTqShards; ++s)
        drained = drained && S.ctl[kTqHead + s * kTqLine] == S.ctl[kTqTail + s * kTqLine];
      int busy = 0, queued = 0;
      for (auto b : S.busy) busy += b;
      for (auto q : S.state) queued += q;
      uint64_t vis = 0;
      for (auto v : S.visits) vis += v;
      const bool ok = bad == 0 && drained && busy == 0 && queued == 0 && S.ctl[kTqDone] == 1 &&
                      S.ctl[kTqAbort] == 0;
      std::printf("trial %d solve %d: %d^3 tiles, %d workers, cap %u, visits %llu, wrong %d, "
                  "drained %d, busy %d, queued %d, done %u, abort %u -> %s\n",
                  trial, solve, S.side, S.nworkers, S.cap, (unsigned long long)vis, bad, int(drained),
                  busy, queued, S.ctl[kTqDone], S.ctl[kTqAbort], ok ? "ok" : "FAIL");
      failures += !ok;
    }
  }
  if (failures) {
    std::printf("%d failures\n", failures);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
