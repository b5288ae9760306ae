// Device work queue of the persistent solve form, written once for both the
// GPU kernel and a CPU model (tests/workqueue_cases.cpp): the protocol is a
// template over an operations policy (atomics, clock, pause).
//
// Items are tiles; an item is on a queue at most once at a time, which the
// per-item activation flag (0 or 1) guarantees: only a visit that turns a flag
// from 0 to 1 pushes the item, and the worker that receives it turns it back
// to 0 before the item is visited.  `pend` counts outstanding work, raised
// before a flag is set and lowered after the visit that consumed it has made
// its own activations:
//
//   pend = (activation flags set) + (visits in progress) + (raises not yet
//          undone by an activation that found its flag already set)
//
// so pend never counts less than the outstanding work, and pend == 0 means
// no item is queued and no visit can push one: a worker waiting for an item
// may stop.
//
// The queue is a ring of tickets.  A worker takes the next ticket from
// `head` (one fetch-and-add, never retried) and waits on that ticket's slot;
// a push takes the next ticket from `tail` and fills its slot.  Tickets t
// and t + cap share a slot; the slot holds (epoch << 32) | (item + 1) with
// epoch t / cap + 1, and item field 0 once the entry has been received, so a
// worker recognises exactly its own entry and a push never overwrites an
// entry not yet received.  A worker waits alone on its slot and reads pend
// only now and then, backing off: an idle worker costs the memory system
// almost nothing.  Every wait (a worker's, a push's for a slot not yet
// received) is bounded by a clock limit after which the caller reports an
// error and stops, so no worker loops forever whatever goes wrong.
//
// The policy O supplies load / add / exch on u32, load64 / store64 on u64,
// now() (a clock), pause(k) (about 2^k idle steps) and order(): order()
// returns once every memory operation the caller issued before it has been
// performed, which keeps a raise of pend ahead of the flag it accounts for
// and a visit's own activations ahead of its done().
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mrf {
namespace wq {

struct Queue {
  uint64_t* slots;  // cap slots: (epoch << 32) | (item + 1), item field 0 = received
  uint32_t cap;     // more than the items that can be queued at once
  uint32_t* head;   // tickets taken by workers
  uint32_t* tail;   // tickets taken by pushes
  uint32_t* pend;
};

// Result of a wait for an item.
enum : uint32_t { GOT = 0u, STOP = 1u, TIMEOUT = 2u };

__host__ __device__ inline uint64_t entry(uint32_t ticket, uint32_t cap, uint32_t item1) {
  return ((uint64_t)(ticket / cap + 1u) << 32) | item1;
}

// A worker's ticket.
template <class O>
__host__ __device__ inline uint32_t take_ticket(const Queue& q) {
  return O::add(q.head, 1u);
}

// Wait for the entry of `ticket`: GOT with its item (the entry marked
// received), STOP when nothing is pending (no entry can come), TIMEOUT after
// `limit` clock ticks.
template <class O>
__host__ __device__ inline uint32_t wait_item(const Queue& q, uint32_t ticket, uint64_t limit, uint32_t& item) {
  uint64_t* s = q.slots + (ticket % q.cap);
  const uint64_t want = entry(ticket, q.cap, 0u);
  const uint64_t t0 = O::now();
  for (uint32_t k = 0;; ++k) {
    const uint64_t v = O::load64(s);
    if ((v >> 32) == (want >> 32) && (uint32_t)v != 0u) {
      item = (uint32_t)v - 1u;
      O::store64(s, want);  // received
      return GOT;
    }
    if ((k & 7u) == 7u && O::load(q.pend) == 0u) return STOP;
    if (O::now() - t0 > limit) return TIMEOUT;
    O::pause(k < 6u ? k : 6u);
  }
}

// Push an item; false when its slot's previous entry was not received within
// `limit` ticks.
template <class O>
__host__ __device__ inline bool push(const Queue& q, uint32_t item, uint64_t limit) {
  const uint32_t t = O::add(q.tail, 1u);
  uint64_t* s = q.slots + (t % q.cap);
  // the previous entry of this slot (ticket t - cap), received, or none yet
  const uint64_t before = t < q.cap ? 0ull : entry(t - q.cap, q.cap, 0u);
  const uint64_t t0 = O::now();
  for (uint32_t k = 0; O::load64(s) != before; ++k) {
    if (O::now() - t0 > limit) return false;
    O::pause(k < 6u ? k : 6u);
  }
  O::store64(s, entry(t, q.cap, item + 1u));
  return true;
}

// Activate an item: raise pend, then set its flag; a flag that was clear puts
// the item on the queue, one that was set already undoes the raise.
template <class O>
__host__ __device__ inline bool activate(const Queue& q, uint32_t* act, uint32_t item, uint64_t limit) {
  O::add(q.pend, 1u);
  O::order();
  if (O::exch(act + item, 1u) == 0u) return push<O>(q, item, limit);
  O::add(q.pend, 0xFFFFFFFFu);
  return true;
}

// Take a received item's activation (its flag back to 0); false when the flag
// was not set, which the protocol never allows.
template <class O>
__host__ __device__ inline bool take(uint32_t* act, uint32_t item) {
  return O::exch(act + item, 0u) == 1u;
}

// The visit of a taken item has made its activations.
template <class O>
__host__ __device__ inline void done(const Queue& q) {
  O::order();
  O::add(q.pend, 0xFFFFFFFFu);
}

}  // namespace wq
}  // namespace mrf
