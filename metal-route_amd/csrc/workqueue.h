// Device work queue of the persistent solve form, written once for both the
// GPU kernel and a CPU model (tests/workqueue_cases.cpp): the protocol is a
// template over an operations policy (atomics, clock, pause).
//
// Items are tiles; an item is on a queue at most once at a time, which the
// per-item activation flag (0 or 1) guarantees: only a visit that turns a flag
// from 0 to 1 pushes the item, and the pop turns it back to 0 before the item
// is visited.  `pend` counts outstanding work, raised before a flag is set and
// lowered after the visit that consumed it has made its own activations:
//
//   pend = (activation flags set) + (visits in progress) + (raises not yet
//          undone by an activation that found its flag already set)
//
// so pend never counts less than the outstanding work, and pend == 0 means
// no item is queued and no visit can push one: a worker that finds its queue
// empty and pend == 0 may stop.  Every wait (a pop's slot not yet written, a
// push's slot not yet freed, an idle worker) is bounded by a clock limit
// after which the caller reports an error and stops, so no worker loops
// forever whatever goes wrong.
#pragma once

#include <cstdint>

namespace mrf {
namespace wq {

// One queue: a ring of slots (item + 1, 0 = empty) with monotonic head (pops
// reserved) and tail (pushes reserved) counters.  cap exceeds the number of
// items that can be queued at once.
struct Queue {
  uint32_t* slots;
  uint32_t cap;
  uint32_t* head;
  uint32_t* tail;
  uint32_t* pend;
};

enum : uint32_t { POP_EMPTY = 0u, POP_TIMEOUT = 0xFFFFFFFFu };

// Pop one item: item + 1, POP_EMPTY when nothing is queued (or the head is
// contended; the caller checks pend and tries again), POP_TIMEOUT when a
// reserved slot was not written within `limit` clock ticks.
template <class O>
inline uint32_t try_pop(const Queue& q, uint64_t limit) {
  for (int tries = 0; tries < 32; ++tries) {
    const uint32_t h = O::load(q.head);
    const uint32_t t = O::load(q.tail);
    if (h == t) return POP_EMPTY;  // every reserved push has been reserved by a pop
    if (O::cas(q.head, h, h + 1u) != h) continue;
    uint32_t* s = q.slots + (h % q.cap);
    const uint64_t t0 = O::now();
    for (;;) {
      const uint32_t v = O::load(s);
      if (v) {
        O::store(s, 0u);
        return v;
      }
      if (O::now() - t0 > limit) return POP_TIMEOUT;
      O::pause();
    }
  }
  return POP_EMPTY;
}

// Push an item; false when its slot was not freed within `limit` ticks.
template <class O>
inline bool push(const Queue& q, uint32_t item, uint64_t limit) {
  const uint32_t t = O::add(q.tail, 1u);
  uint32_t* s = q.slots + (t % q.cap);
  const uint64_t t0 = O::now();
  // the slot's previous entry was reserved by a pop; wait until it took it
  while (O::load(s) != 0u) {
    if (O::now() - t0 > limit) return false;
    O::pause();
  }
  O::store(s, item + 1u);
  return true;
}

// Activate an item: raise pend, then set its flag; a flag that was clear puts
// the item on the queue, one that was set already undoes the raise.
template <class O>
inline bool activate(const Queue& q, uint32_t* act, uint32_t item, uint64_t limit) {
  O::add(q.pend, 1u);
  if (O::exch(act + item, 1u) == 0u) return push<O>(q, item, limit);
  O::add(q.pend, 0xFFFFFFFFu);
  return true;
}

// Take a popped item's activation (its flag back to 0); false when the flag
// was not set, which the protocol never allows.
template <class O>
inline bool take(uint32_t* act, uint32_t item) {
  return O::exch(act + item, 0u) == 1u;
}

// The visit of a taken item has made its activations.
template <class O>
inline void done(const Queue& q) {
  O::add(q.pend, 0xFFFFFFFFu);
}

}  // namespace wq
}  // namespace mrf
