// CPU model of the persistent solve form's work queue
// (metal-route_amd/csrc/workqueue.h), run with the same protocol code the
// kernel uses, over a host operations policy.
//
// Worker threads play the kernel's workgroups: each serves one queue and
// loops ticket -> wait for its item -> take -> visit -> activations -> done
// until a wait finds nothing pending, exactly the control flow of the
// kernel's loop.  A visit is a
// label-correcting relaxation of a node of a random weighted graph (labels
// lowered with an atomic minimum, so visits of one node may overlap, as tile
// visits may); every improved neighbour is activated.  Checked:
//   * every worker stops, no wait times out, and the labels equal Dijkstra's
//     on every graph (one disjoint graph per queue, several workers per queue);
//   * with pend raised by one that no item carries, every worker reports the
//     wait timeout and stops instead of waiting forever;
//   * every push was received (each worker ends holding one unused ticket),
//     every flag is clear and pend is 0 at the end;
//   * with a ring of one slot more than the items (every slot reused many
//     times over, pushes waiting for slots not yet received) the results
//     still equal Dijkstra's.
// Exit status 0 when every case holds.
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <queue>
#include <random>
#include <thread>
#include <vector>

#include "../metal-route_amd/csrc/workqueue.h"

namespace wq = mrf::wq;

struct HostOps {
  static uint32_t load(uint32_t* p) { return __atomic_load_n(p, __ATOMIC_SEQ_CST); }
  static uint64_t load64(uint64_t* p) { return __atomic_load_n(p, __ATOMIC_SEQ_CST); }
  static void store64(uint64_t* p, uint64_t v) { __atomic_store_n(p, v, __ATOMIC_SEQ_CST); }
  static uint32_t exch(uint32_t* p, uint32_t v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
  static uint32_t add(uint32_t* p, uint32_t v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
  static uint32_t orr(uint32_t* p, uint32_t v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
  static bool cas(uint32_t* p, uint32_t expect, uint32_t want) {
    return __atomic_compare_exchange_n(p, &expect, want, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
  }
  static uint64_t now() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::microseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
  }
  static void pause(uint32_t) { std::this_thread::yield(); }
  static void order() {}  // every operation above is sequentially consistent
};

static int fails = 0;
static void expect(bool ok, const char* name) {
  if (!ok) {
    printf("FAIL %s\n", name);
    fails++;
  }
}

struct Graph {
  uint32_t n0, n;  // nodes [n0, n0 + n) of the shared arrays
  std::vector<std::vector<std::pair<uint32_t, uint64_t>>> adj;  // local node -> (global node, weight)
  uint32_t source;                                               // global
};

static void atomic_min64(uint64_t* p, uint64_t v, uint64_t& old) {
  old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (v < old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
}

struct Run {
  std::vector<uint64_t> slots;
  std::vector<uint32_t> act, ctl;  // ctl: per queue head, tail, pend (one line each)
  std::vector<uint64_t> dist;
  std::vector<wq::Queue> queues;
  std::atomic<uint32_t> errors{0}, pops{0}, exits{0}, overlaps{0};
  std::vector<uint32_t> in_visit;  // per node: workers inside its visit (at most one)
};

// The kernel's loop for one worker of queue q.
static void worker(Run& R, const std::vector<Graph>& gs, uint32_t q, uint64_t limit_us) {
  const wq::Queue& Q = R.queues[q];
  const Graph& G = gs[q];
  for (;;) {
    const uint32_t ticket = wq::take_ticket<HostOps>(Q);
    uint32_t u = 0;
    const uint32_t r = wq::wait_item<HostOps>(Q, ticket, limit_us, u);
    if (r == wq::TIMEOUT) {
      R.errors++;
      break;
    }
    if (r == wq::GOT) {
      R.pops++;
      if (!wq::take<HostOps>(R.act.data(), u)) R.errors++;
      if (__atomic_fetch_add(&R.in_visit[u], 1u, __ATOMIC_SEQ_CST) != 0u) R.overlaps++;
      const uint64_t du = __atomic_load_n(&R.dist[u], __ATOMIC_SEQ_CST);
      for (auto& e : G.adj[u - G.n0]) {
        uint64_t old;
        atomic_min64(&R.dist[e.first], du + e.second, old);
        if (du + e.second < old && !wq::activate<HostOps>(Q, R.act.data(), e.first, limit_us)) R.errors++;
      }
      __atomic_fetch_sub(&R.in_visit[u], 1u, __ATOMIC_SEQ_CST);
      if (!wq::finish<HostOps>(Q, R.act.data(), u, limit_us)) R.errors++;
      wq::done<HostOps>(Q);
      continue;
    }
    break;  // STOP: nothing pending
  }
  R.exits++;
}

static std::vector<uint64_t> dijkstra(const Graph& g) {
  std::vector<uint64_t> d(g.n, UINT64_MAX);
  using E = std::pair<uint64_t, uint32_t>;
  std::priority_queue<E, std::vector<E>, std::greater<E>> pq;
  d[g.source - g.n0] = 0;
  pq.push({0, g.source});
  while (!pq.empty()) {
    auto [du, u] = pq.top();
    pq.pop();
    if (du > d[u - g.n0]) continue;
    for (auto& e : g.adj[u - g.n0])
      if (du + e.second < d[e.first - g.n0]) {
        d[e.first - g.n0] = du + e.second;
        pq.push({du + e.second, e.first});
      }
  }
  return d;
}

static void setup(Run& R, std::vector<Graph>& gs, uint32_t nq, uint32_t nodes, uint32_t degree, uint64_t seed,
                  uint32_t spare = 8) {
  std::mt19937_64 rng(seed);
  gs.clear();
  uint32_t total = 0;
  for (uint32_t q = 0; q < nq; ++q) {
    Graph g;
    g.n0 = total;
    g.n = nodes;
    g.adj.resize(nodes);
    for (uint32_t u = 0; u < nodes; ++u)
      for (uint32_t k = 0; k < degree; ++k)
        g.adj[u].push_back({g.n0 + (uint32_t)(rng() % nodes), 1 + rng() % 100});
    g.source = g.n0 + (uint32_t)(rng() % nodes);
    total += nodes;
    gs.push_back(std::move(g));
  }
  R.act.assign(total, 0);
  R.in_visit.assign(total, 0);
  R.dist.assign(total, UINT64_MAX);
  R.ctl.assign(nq * 3 * 16, 0);
  uint32_t slot0 = 0;
  std::vector<uint32_t> caps;
  for (uint32_t q = 0; q < nq; ++q) caps.push_back(nodes + spare);
  uint32_t ns = 0;
  for (uint32_t c : caps) ns += c;
  R.slots.assign(ns, 0);
  R.queues.clear();
  for (uint32_t q = 0; q < nq; ++q) {
    wq::Queue Q;
    Q.slots = R.slots.data() + slot0;
    Q.cap = caps[q];
    Q.head = &R.ctl[(q * 3 + 0) * 16];
    Q.tail = &R.ctl[(q * 3 + 1) * 16];
    Q.pend = &R.ctl[(q * 3 + 2) * 16];
    slot0 += caps[q];
    R.queues.push_back(Q);
    // seed: the source at label 0, activated by the host (ticket 0)
    R.dist[gs[q].source] = 0;
    R.act[gs[q].source] = 1;
    Q.slots[0] = wq::entry(0, Q.cap, gs[q].source + 1);
    *Q.tail = 1;
    *Q.pend = 1;
  }
}

static bool run_case(uint32_t nq, uint32_t per_q, uint32_t nodes, uint32_t degree, uint64_t seed,
                     uint32_t spare = 8) {
  Run R;
  std::vector<Graph> gs;
  setup(R, gs, nq, nodes, degree, seed, spare);
  std::vector<std::thread> th;
  for (uint32_t w = 0; w < nq * per_q; ++w) th.emplace_back(worker, std::ref(R), std::cref(gs), w % nq, 5000000ull);
  for (auto& t : th) t.join();
  // no node is ever visited by two workers at once
  bool ok = R.errors.load() == 0 && R.exits.load() == nq * per_q && R.overlaps.load() == 0;
  for (uint32_t q = 0; q < nq; ++q) {
    const std::vector<uint64_t> ref = dijkstra(gs[q]);
    for (uint32_t u = 0; u < nodes; ++u) ok = ok && R.dist[gs[q].n0 + u] == ref[u];
    // every push received: each worker holds one abandoned ticket at the end
    ok = ok && *R.queues[q].pend == 0 && *R.queues[q].head == *R.queues[q].tail + per_q;
  }
  for (uint32_t a : R.act) ok = ok && a == 0;
  ok = ok && R.pops.load() > 0;
  return ok;
}

int main() {
  expect(run_case(1, 1, 200, 3, 1), "one queue, one worker");
  expect(run_case(1, 8, 2000, 4, 2), "one queue, eight workers");
  expect(run_case(4, 3, 1500, 3, 3), "four queues, three workers each");
  expect(run_case(8, 2, 500, 6, 4), "eight queues, dense graphs");
  for (uint64_t s = 10; s < 30; ++s) expect(run_case(3, 4, 300, 3, s), "random small graphs");
  // more workers than host cores: preemption anywhere in the protocol
  expect(run_case(1, 32, 3000, 3, 5), "one queue, thirty-two workers");
  expect(run_case(8, 6, 800, 4, 6), "eight queues, six workers each");
  for (uint64_t s = 40; s < 46; ++s) expect(run_case(2, 8, 60, 5, s, 1), "ring of one spare slot");
  {
    // pend raised with no queued item behind it: every worker must give up
    // after the wait limit and report it, not wait forever
    Run R;
    std::vector<Graph> gs;
    setup(R, gs, 2, 50, 2, 7);
    for (auto& Q : R.queues) *Q.pend += 1;
    std::vector<std::thread> th;
    for (uint32_t w = 0; w < 6; ++w) th.emplace_back(worker, std::ref(R), std::cref(gs), w % 2, 20000ull);
    for (auto& t : th) t.join();
    expect(R.exits.load() == 6 && R.errors.load() == 6, "wait limit ends every worker with an error");
  }
  printf("%s\n", fails ? "FAILED" : "ok");
  return fails ? 1 : 0;
}
