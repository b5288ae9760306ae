// Device work queue of the persistent solve form, written once for both the
// GPU kernel and a CPU model (tests/workqueue_cases.cpp): the protocol is a
// template over an operations policy (atomics, clock, pause).
//
// Items are tiles.  Each item's activation flag has two bits: bit 0 "wanted"
// (queued, or to be visited again once the running visit ends) and bit 1
// "running":
//
//   0 idle   1 queued   2 running   3 running and wanted again
//
// An item is pushed only when its flag goes 0 -> 1 (activate) or 3 -> 1 (the
// end of the running visit), the worker that receives it turns 1 -> 2 before
// the visit, and the visit ends with 2 -> 0 or 3 -> 1; so an item is on a
// queue at most once and visited by at most one worker at a time (its labels
// need no atomic minimum).  `pend` counts outstanding work, raised before a
// flag is set and lowered after the visit that consumed it has made its own
// activations and ended:
//
//   pend = (queued items) + (visits in progress) + (wanted-again marks) +
//          (raises not yet undone by an activation that found the item
//          already queued or marked)
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
// The policy O supplies load / add / exch / orr / cas on u32, load64 /
// store64 on u64, now() (a clock), pause(k) (about 2^k idle steps) and
// order(): order()
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

// Activate an item: raise pend, then set its wanted bit.  An idle item goes
// on the queue; a running one is marked, and its visit's end pushes it (the
// raise counts the mark); a queued or marked one undoes the raise.
template <class O>
__host__ __device__ inline bool activate(const Queue& q, uint32_t* act, uint32_t item, uint64_t limit) {
  O::add(q.pend, 1u);
  O::order();
  const uint32_t was = O::orr(act + item, 1u);
  if (was == 0u) return push<O>(q, item, limit);
  if (was == 2u) return true;
  O::add(q.pend, 0xFFFFFFFFu);
  return true;
}

// Take a received item for its visit (queued -> running); false when it was
// not queued, which the protocol never allows.
template <class O>
__host__ __device__ inline bool take(uint32_t* act, uint32_t item) {
  return O::exch(act + item, 2u) == 1u;
}

// End an item's visit, after its writes and its own activations: running ->
// idle, or, when it was marked meanwhile, queued again (the mark's raise
// counts the entry).  False when the push timed out.
template <class O>
__host__ __device__ inline bool finish(const Queue& q, uint32_t* act, uint32_t item, uint64_t limit) {
  O::order();
  if (O::cas(act + item, 2u, 0u)) return true;
  O::exch(act + item, 1u);  // 3 -> 1: nothing else changes a marked flag
  return push<O>(q, item, limit);
}

// The visit of a taken item has made its activations and ended.
template <class O>
__host__ __device__ inline void done(const Queue& q) {
  O::order();
  O::add(q.pend, 0xFFFFFFFFu);
}

}  // namespace wq
}  // namespace mrf
