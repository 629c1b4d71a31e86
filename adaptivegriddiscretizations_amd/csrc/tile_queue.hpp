// tile_queue.hpp -- the tile queue of the persistent (barrier-free) solver:
// sharded bounded multi-producer multi-consumer rings with per-slot
// sequence numbers, a one-bit "queued" state per tile, and termination by
// counting.  No counterpart in the reference, whose host loop launches one
// kernel per pass over a fresh tile list (_solvers.py:122-193).
//
// The protocol is written once over an atomics policy, so that the same code
// runs on the device (AGENT-scope atomics, one lane per wave) and in host
// threads (tests/tile_queue_sim.cpp drives it with std::thread workers).
//
// Rings.  Shard s holds a ring of C = 2^k slots (C >= number of tiles).  A
// slot is a 64-bit word: (sequence << 32) | entry, entry 0xffffffff = empty.
// Position p of shard s lives in slot p & (C-1).  The slot of position p is
// ready for writing when its sequence is p and its entry empty; the writer
// stores (p, entry); the reader of p waits for (p, entry), then releases
// the slot for position p + C by storing (p + C, empty).  Initially slot i
// holds (i, empty).  Positions only grow, across solves too, so rings are
// initialised once per workspace.
//
// Pops.  A consumer claims the oldest unclaimed position of a non-empty
// ring: it reads head[s] and tail[s] of every ring at once (one lane per
// ring on the device), picks a ring with head < tail and claims position
// head with a compare-and-swap of head[s]; the producer that reserved that
// position writes its slot shortly (a bounded wait).  A producer reserves
// positions p = tail[s] += n.  Nothing ties a consumer to a ring, so an
// entry in any ring is taken by the next idle consumer.
//
// Tile states.  Bit 0 of state[t] = t is in a ring or about to be read from
// one.  request(t): atomic OR; the caller that turns 0 into 1 pushes t.  The
// reader of t's entry clears the bit (atomic exchange) before reading any
// value of t's neighbourhood, so a request arriving after that pushes t
// again: no change to a neighbour is ever missed.  Each tile thus has at
// most one unread entry at any time, which bounds every ring by the tile
// count (with C >= ntiles a producer never waits for a lap: the slot of
// position p - C was claimed before position p could be reserved).
//
// Termination.  A worker marks itself busy (busy[w] = 1) before it tries to
// claim a position, clears the mark after its pushes (or after a failed
// try).  The solve is over when, in this order, (1) head[s] == tail[s] on
// every ring, (2) no worker is busy, (3) no tail moved since (1).  Any
// unclaimed entry fails (1); an entry claimed by a worker still busy fails
// (2); a worker that finished between (1) and (2) pushed, if at all, after
// (1) -- failing (3) -- or before (1), where its entries are covered by the
// same argument.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define AGD_HD __host__ __device__
#else
#define AGD_HD
#endif

namespace agd {

constexpr int kTqShards = 64;      // rings
constexpr int kTqLine = 16;        // int32 words per 64-byte line
constexpr uint32_t kTqEmpty = 0xffffffffu;

// Control block (int32 words, one counter per 64-byte line).
constexpr int kTqHead = 0;                               // [kTqShards] claimed positions
constexpr int kTqTail = kTqShards * kTqLine;             // [kTqShards] reserved positions
constexpr int kTqDone = 2 * kTqShards * kTqLine;         // termination detected
constexpr int kTqAbort = kTqDone + kTqLine;              // 1: spin limit, 2: visit limit
constexpr int kTqCtlLen = kTqAbort + kTqLine;

AGD_HD inline uint64_t tq_slot(uint32_t seq, uint32_t entry) { return (uint64_t(seq) << 32) | entry; }

// One producer push of entry at position pos of a ring (mask = C - 1).
// Returns false if the slot did not become ready within `limit` polls.
template <class A>
AGD_HD inline bool tq_write(uint64_t* ring, uint32_t mask, uint32_t pos, uint32_t entry, uint32_t limit) {
  uint64_t* slot = ring + (pos & mask);
  const uint64_t ready = tq_slot(pos, kTqEmpty);
  for (uint32_t i = 0; i < limit; ++i) {
    if (A::cas64(slot, ready, tq_slot(pos, entry))) return true;
    A::pause();
  }
  return false;
}

// Consumer poll of a claimed position pos: the entry if it was written,
// else kTqEmpty.
template <class A>
AGD_HD inline uint32_t tq_peek(uint64_t* ring, uint32_t mask, uint32_t pos) {
  const uint64_t v = A::load64(ring + (pos & mask));
  const uint32_t seq = uint32_t(v >> 32);
  const uint32_t e = uint32_t(v);
  return (seq == pos && e != kTqEmpty) ? e : kTqEmpty;
}

// Consumer release after a successful peek (the slot then serves position
// pos + C).
template <class A>
AGD_HD inline void tq_release(uint64_t* ring, uint32_t mask, uint32_t pos) {
  A::store64(ring + (pos & mask), tq_slot(pos + mask + 1, kTqEmpty));
}

}  // namespace agd
