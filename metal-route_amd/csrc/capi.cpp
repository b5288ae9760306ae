// C ABI of the field engine (include/mrfield.h): validation, the provider
// adapter's window -> full-board retry, chunking by the label budget, and the
// fail-closed host re-validation of every returned path.
#include <algorithm>
#include <climits>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mrfield.h"
#include "engine.h"
#include "hostcheck.h"
#include "hostpool.h"

using mrf::Error;
using mrf::kMax;
using namespace mrf::host;

struct mrf_ctx {
  mrf::Ctx c;
};

// Page-locked host memory for the board's additive plane: the negotiated
// batches rebuild and upload it every iteration, and a copy from pageable
// memory would be staged through the driver first.
template <typename T>
struct PinnedAlloc {
  using value_type = T;
  PinnedAlloc() = default;
  template <typename U>
  PinnedAlloc(const PinnedAlloc<U>&) {}
  T* allocate(size_t n) {
    void* p = nullptr;
    if (hipHostMalloc(&p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault) != hipSuccess) throw std::bad_alloc();
    return (T*)p;
  }
  void deallocate(T* p, size_t) { (void)hipHostFree(p); }
  template <typename U>
  bool operator==(const PinnedAlloc<U>&) const { return true; }
  template <typename U>
  bool operator!=(const PinnedAlloc<U>&) const { return false; }
};

struct mrf_board {
  mrf_ctx* ctx = nullptr;
  mrf::DevBoard b;
  mrf_dims dims{};
  // Host copies needed for request validation and path checks.
  std::vector<uint32_t> cost;
  std::vector<uint8_t> vf, board;
  std::vector<uint32_t> xe, ye, ve;
  std::vector<uint32_t, PinnedAlloc<uint32_t>> add;
  std::vector<uint8_t> va;
  std::vector<int32_t> halo;
  uint32_t halo_pen = 0;
  // Legal-search planes of the running mrf_route_legal_batch call (empty
  // otherwise).
  std::vector<int32_t> owner, lhalo;
};

namespace {

void set_err(mrf_error* err, int code, const char* msg, uint32_t net = 0) {
  if (!err) return;
  err->code = code;
  err->net = net;
  snprintf(err->message, sizeof(err->message), "%s", msg);
}

int fail(mrf_error* err, const Error& e) {
  set_err(err, e.code(), e.what(), e.net());
  return e.code();
}


// Solve one pass of nets (each inside its window) on a resident board, chunked
// by the label budget, re-validating every path on the host.
// `prices` (optional, indexed like nets) carries per-net additive overrides.
void solve_nets(mrf::Ctx& c, const mrf::DevBoard& b, const HostGrid& g, const mrf_net* nets,
                const std::vector<size_t>& idx, const std::vector<mrf_window>& wins, bool lee,
                const int32_t* net_ids, Batch& out, const std::vector<NetPlan>* prices = nullptr) {
  size_t n = idx.size();
  out.routed.assign(n, 0);
  out.cost.assign(n, kMax);
  out.paths.assign(n, {});
  std::vector<NetPlan> plans(n);
  for (size_t k = 0; k < n; ++k) {
    const mrf_net& nt = nets[idx[k]];
    plans[k].pads = sorted_unique(nt.pads, nt.n_pads);
    plans[k].vpads = sorted_unique(nt.via_pads, nt.n_via_pads);
    if (prices) {
      plans[k].ovr_cells = (*prices)[idx[k]].ovr_cells;
      plans[k].ovr_add = (*prices)[idx[k]].ovr_add;
      plans[k].own_group = (*prices)[idx[k]].own_group;
    }
  }
  size_t budget_words = std::max<size_t>(c.label_budget / 8, 1);
  size_t k0 = 0;
  while (k0 < n) {
    size_t words = 0, k1 = k0;
    while (k1 < n) {
      const mrf_window& w = wins[k1];
      size_t fw = mrf::field_label_words(c.shape, w.x1 - w.x0 + 1, w.y1 - w.y0 + 1, g.d.layers);
      if (k1 > k0 && words + fw > budget_words) break;
      words += fw;
      ++k1;
    }
    std::vector<mrf::FieldReq> reqs(k1 - k0);
    for (size_t k = k0; k < k1; ++k) {
      const mrf_net& nt = nets[idx[k]];
      mrf::FieldReq& q = reqs[k - k0];
      q.x0 = wins[k].x0;
      q.y0 = wins[k].y0;
      q.x1 = wins[k].x1;
      q.y1 = wins[k].y1;
      q.sources = {nt.src};
      q.dst = nt.dst;
      q.pads = plans[k].pads;
      q.via_pads = plans[k].vpads;
      q.own_id = net_ids ? net_ids[idx[k]] : -1;
      q.ovr_cells = plans[k].ovr_cells;
      q.ovr_add = plans[k].ovr_add;
      q.own_group = plans[k].own_group;
      if (!lee) {
        q.forced = {nt.src, nt.dst};
        q.flags = mrf::F_FORCE;
      }
    }
    mrf::SolveStats st;
    mrf::solve_fields(c, b, reqs, &st);
    std::vector<uint8_t> routed;
    std::vector<uint32_t> cost, cells;
    std::vector<uint64_t> off;
    mrf::reconstruct_paths(c, b, reqs.size(), routed, cost, off, cells, &st);
    if (c.inject == 1) throw Error(MRF_BACKEND_UNAVAILABLE, "injected backend failure");
    if (c.inject == 2 && !cells.empty()) cells[cells.size() / 2] ^= 1u;
    // Re-validate on up to 16 host threads (the shared workers); any failure
    // fails the whole call.
    const auto tv0 = std::chrono::steady_clock::now();
    std::vector<uint8_t> bad(k1 - k0, 0);
    auto check = [&](size_t a, size_t b) {
      for (size_t k = a; k < b; ++k) {
        size_t j = k - k0;
        out.routed[k] = routed[j];
        if (!routed[j]) continue;
        const uint32_t* p = cells.data() + off[j];
        size_t len = off[j + 1] - off[j];
        int32_t own = net_ids ? net_ids[idx[k]] : -1;
        if (!path_ok(g, nets[idx[k]], plans[k], wins[k], own, p, len, cost[j], lee)) {
          bad[j] = 1;
          continue;
        }
        out.cost[k] = cost[j];
        out.paths[k].assign(p, p + len);
      }
    };
    mrf::host::parallel_for(k1 - k0, 4, [&](size_t a, size_t b) { check(k0 + a, k0 + b); });
    c.last.validate_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tv0).count();
    for (uint8_t v : bad)
      if (v) throw Error(MRF_BACKEND_UNAVAILABLE, "device path failed host re-validation");
    k0 = k1;
  }
}



struct Lane {
  std::unique_lock<std::mutex> lk;
  // `solve`: the call runs relaxations; its statistics replace the last call's.
  explicit Lane(mrf::Ctx& c, bool solve = true) : lk(c.lane, std::try_to_lock) {
    if (!lk.owns_lock()) throw Error(MRF_BACKEND_UNAVAILABLE, "field engine lane is busy");
    if (hipSetDevice(c.device) != hipSuccess) throw Error(MRF_BACKEND_UNAVAILABLE, "hipSetDevice failed");
    if (solve) {
      c.last = mrf::SolveStats{};
      c.time_rounds = c.timing_every && (c.timing_calls++ % c.timing_every) == 0;
    }
  }
};

// Unit planar prices for the Lee / planar-field entries (no vias).
struct UnitEdges {
  std::vector<uint32_t> x, y, v;
  std::vector<uint8_t> a;
  mrf_edges e{};
  explicit UnitEdges(mrf_dims d)
      : x(std::max<uint32_t>(d.w, 1), 1u), y(std::max<uint32_t>(d.h, 1), 1u), v(std::max<uint32_t>(d.layers, 1), 0u),
        a(std::max<uint32_t>(d.layers, 1), 0u) {
    e.x = x.data();
    e.nx = d.w - 1;
    e.y = y.data();
    e.ny = d.h - 1;
    e.via = v.data();
    e.via_allowed = a.data();
    e.nvia = d.layers - 1;
  }
};

// transient: the board is freed before the calling entry point returns.
void upload(mrf::Ctx& c, mrf::DevBoard& b, mrf_dims d, const mrf_grid* g, const mrf_edges* e,
            const mrf_pricing* pr, bool transient) {
  const uint32_t* add = nullptr;
  const int32_t* halo = nullptr;
  uint32_t pen = 0;
  size_t n = dims_len(d);
  if (pr) {
    if (pr->n_additive) {
      if (pr->n_additive != n || !pr->additive) throw Error(MRF_MALFORMED_GRID, "additive plane length does not match dims");
      add = pr->additive;
    }
    if (pr->n_halo_owner) {
      if (pr->n_halo_owner != n || !pr->halo_owner) throw Error(MRF_MALFORMED_GRID, "halo plane length does not match dims");
      halo = pr->halo_owner;
      pen = pr->halo_penalty;
    }
  }
  mrf::board_upload(c, b, d.w, d.h, d.layers, g->cost, g->n_via_forbidden ? g->via_forbidden : nullptr,
                    g->n_board_constraint ? g->board_constraint : nullptr, e->x, e->y, e->via, e->via_allowed,
                    add, halo, pen, transient);
}

int isolated(mrf_ctx* ctx, mrf_dims dims, const mrf_grid* grid, const mrf_net* nets, size_t n_nets,
             const mrf_window* windows, const mrf_edges* edges, mrf_paths* out, mrf_error* err, bool retry) {
  if (out) memset(out, 0, sizeof(*out));
  if (!ctx || !out) {
    set_err(err, MRF_BACKEND_UNAVAILABLE, "null context or output");
    return MRF_BACKEND_UNAVAILABLE;
  }
  mrf::DevBoard b;
  try {
    Lane lane(ctx->c);
    check_grid(dims, grid);
    check_edges(dims, edges);
    if (n_nets && (!nets || !windows)) throw Error(MRF_BACKEND_UNAVAILABLE, "isolated request windows are not net-aligned");
    HostGrid g = host_grid(dims, grid, edges);
    for (size_t i = 0; i < n_nets; ++i) check_window(dims, windows[i], i);
    for (size_t i = 0; i < n_nets; ++i) check_net(g, nets[i], i);
    Batch all;
    all.routed.assign(n_nets, 0);
    all.cost.assign(n_nets, kMax);
    all.paths.assign(n_nets, {});
    if (n_nets) {
      upload(ctx->c, b, dims, grid, edges, nullptr, true);
      std::vector<size_t> idx(n_nets);
      for (size_t i = 0; i < n_nets; ++i) idx[i] = i;
      std::vector<mrf_window> wins(windows, windows + n_nets);
      Batch first;
      solve_nets(ctx->c, b, g, nets, idx, wins, false, nullptr, first);
      std::vector<size_t> retry_idx;
      for (size_t i = 0; i < n_nets; ++i) {
        if (first.routed[i]) {
          all.routed[i] = 1;
          all.cost[i] = first.cost[i];
          all.paths[i] = std::move(first.paths[i]);
        } else {
          retry_idx.push_back(i);
        }
      }
      if (retry && !retry_idx.empty()) {
        mrf_window full{0, 0, dims.w - 1, dims.h - 1};
        std::vector<mrf_window> fw(retry_idx.size(), full);
        Batch second;
        solve_nets(ctx->c, b, g, nets, retry_idx, fw, false, nullptr, second);
        for (size_t k = 0; k < retry_idx.size(); ++k) {
          size_t i = retry_idx[k];
          all.routed[i] = second.routed[k];
          all.cost[i] = second.cost[k];
          all.paths[i] = std::move(second.paths[k]);
        }
      }
    }
    mrf::board_free(b);
    export_paths(all, out);
    if (err) set_err(err, MRF_OK, "ok");
    return MRF_OK;
  } catch (const Error& e) {
    mrf::board_free(b);
    mrf_free_paths(out);
    return fail(err, e);
  } catch (const std::exception& e) {
    mrf::board_free(b);
    mrf_free_paths(out);
    set_err(err, MRF_BACKEND_UNAVAILABLE, e.what());
    return MRF_BACKEND_UNAVAILABLE;
  }
}

}  // namespace

namespace {

inline uint64_t sat_add64(uint64_t a, uint64_t b) { return a + b < a ? UINT64_MAX : a + b; }
inline uint64_t sat_mul64(uint64_t a, uint64_t b) { return (b && a > UINT64_MAX / b) ? UINT64_MAX : a * b; }
inline uint64_t sat_sub64(uint64_t a, uint64_t b) { return a > b ? a - b : 0; }
inline uint32_t clamp32(uint64_t v) { return v < 0xFFFFFFFFull ? (uint32_t)v : 0xFFFFFFFFu; }
constexpr uint64_t kNegScale = 16;  // SCALE == CLEARANCE_NEG_WEIGHT (negotiated.rs:71, :99)

// Own cells of net i from a CSR list, sorted and unique; `counts` gets the
// summed multiplicities (entry weights `weight`, 1 each when null).
void own_cells(const uint64_t* off, const uint32_t* cells, const uint32_t* weight, size_t i, size_t n_cells,
               std::vector<uint32_t>& out, std::vector<uint64_t>* counts) {
  out.clear();
  if (counts) counts->clear();
  if (!off) return;
  if (off[i + 1] < off[i]) throw Error(MRF_BACKEND_UNAVAILABLE, "own-footprint offsets are not ascending");
  std::vector<std::pair<uint32_t, uint64_t>> v;
  v.reserve(off[i + 1] - off[i]);
  for (uint64_t k = off[i]; k < off[i + 1]; ++k) {
    if (cells[k] >= n_cells) throw Error(MRF_BACKEND_UNAVAILABLE, "own-footprint cell outside the board");
    v.emplace_back(cells[k], weight ? weight[k] : 1u);
  }
  std::sort(v.begin(), v.end());
  for (size_t k = 0; k < v.size();) {
    uint64_t sum = 0;
    size_t e = k;
    for (; e < v.size() && v[e].first == v[k].first; ++e) sum += v[e].second;
    out.push_back(v[k].first);
    if (counts) counts->push_back(sum);
    k = e;
  }
}

using mrf::host::parallel_for;

}  // namespace

extern "C" {

int mrf_abi_version(void) { return MRF_ABI_VERSION; }

int mrf_open(int device, mrf_ctx** out, mrf_error* err) {
  if (!out) return MRF_BACKEND_UNAVAILABLE;
  *out = nullptr;
  mrf_ctx* c = new (std::nothrow) mrf_ctx();
  if (!c) {
    set_err(err, MRF_BACKEND_UNAVAILABLE, "out of host memory");
    return MRF_BACKEND_UNAVAILABLE;
  }
  try {
    mrf::ctx_init(c->c, device);
    const char* inj = getenv("MRF_INJECT_FAULT");
    if (inj) c->c.inject = atoi(inj);
    const char* spin = getenv("MRF_STAMP_SPIN_US");
    if (spin) c->c.stamp_spin_us = (uint32_t)strtoul(spin, nullptr, 10);
    const char* qwait = getenv("MRF_QUEUE_WAIT_MS");
    if (qwait) c->c.queue_wait_ms = (uint32_t)strtoul(qwait, nullptr, 10);
  } catch (const Error& e) {
    mrf::ctx_free(c->c);
    delete c;
    return fail(err, e);
  }
  *out = c;
  if (err) set_err(err, MRF_OK, "ok");
  return MRF_OK;
}

void mrf_close(mrf_ctx* ctx) {
  if (!ctx) return;
  mrf::ctx_free(ctx->c);
  delete ctx;
}

int mrf_set_tiles(mrf_ctx* ctx, uint32_t tw, uint32_t th, uint32_t threads, uint32_t max_local, mrf_error* err) {
  if (!ctx || !mrf::tile_shape_supported(tw, th) || threads < 64 || threads > 256 || threads % 64 ||
      max_local < 1) {
    set_err(err, MRF_BACKEND_UNAVAILABLE, "unsupported tile shape");
    return MRF_BACKEND_UNAVAILABLE;
  }
  std::lock_guard<std::mutex> g(ctx->c.lane);
  ctx->c.shape.tw = tw;
  ctx->c.shape.th = th;
  ctx->c.shape.threads = threads;
  ctx->c.shape.max_local = max_local;
  return MRF_OK;
}

int mrf_set_mode(mrf_ctx* ctx, int mode) {
  if (!ctx || (mode != MRF_MODE_PLAIN && mode != MRF_MODE_PERSISTENT && mode != MRF_MODE_GRAPH &&
               mode != MRF_MODE_QUEUE))
    return MRF_BACKEND_UNAVAILABLE;
  std::lock_guard<std::mutex> g(ctx->c.lane);
  ctx->c.mode = mode;
  return MRF_OK;
}

int mrf_set_delta(mrf_ctx* ctx, uint32_t delta) {
  if (!ctx) return MRF_BACKEND_UNAVAILABLE;
  std::lock_guard<std::mutex> g(ctx->c.lane);
  ctx->c.delta = delta == 0 ? 1u : delta;
  return MRF_OK;
}

int mrf_last_stats(mrf_ctx* ctx, mrf_stats* out) {
  if (!ctx || !out) return MRF_BACKEND_UNAVAILABLE;
  const mrf::SolveStats& s = ctx->c.last;
  out->relax_ms = s.relax_ms;
  out->path_ms = s.path_ms;
  out->total_ms = s.total_ms;
  out->tile_visits = s.tile_visits;
  out->local_rounds = s.local_rounds;
  out->cells_visited = s.cells_visited;
  out->field_cells = s.field_cells;
  out->host_rounds = s.rounds;
  out->persistent_ms = s.persistent_ms;
  out->relax_kernel_ms = s.kernel_ms;
  out->relax_launches = s.launches;
  out->changed_visits = s.changed_visits;
  out->validate_ms = s.validate_ms;
  out->wave_launches = s.wave_launches;
  out->label_bytes = s.label_bytes;
  out->timed_launches = s.timed_launches;
  return MRF_OK;
}

int mrf_set_round_timing(mrf_ctx* ctx, uint32_t every) {
  if (!ctx) return MRF_BACKEND_UNAVAILABLE;
  std::lock_guard<std::mutex> g(ctx->c.lane);
  ctx->c.timing_every = every;
  ctx->c.timing_calls = 0;
  return MRF_OK;
}

int mrf_route_isolated_batch(mrf_ctx* ctx, mrf_dims dims, const mrf_grid* grid, const mrf_net* nets, size_t n_nets,
                             const mrf_window* windows, const mrf_edges* edges, mrf_paths* out, mrf_error* err) {
  return isolated(ctx, dims, grid, nets, n_nets, windows, edges, out, err, true);
}

int mrf_route_isolated_pass(mrf_ctx* ctx, mrf_dims dims, const mrf_grid* grid, const mrf_net* nets, size_t n_nets,
                            const mrf_window* windows, const mrf_edges* edges, mrf_paths* out, mrf_error* err) {
  return isolated(ctx, dims, grid, nets, n_nets, windows, edges, out, err, false);
}

int mrf_router_route(mrf_ctx* ctx, mrf_dims dims, const mrf_grid* grid, const mrf_net* nets, size_t n_nets,
                     mrf_paths* out, mrf_error* err) {
  if (out) memset(out, 0, sizeof(*out));
  if (!ctx || !out) return MRF_BACKEND_UNAVAILABLE;
  mrf::DevBoard b;
  try {
    Lane lane(ctx->c);
    check_grid(dims, grid);
    UnitEdges ue(dims);
    HostGrid g = host_grid(dims, grid, &ue.e);
    if (n_nets && !nets) throw Error(MRF_BACKEND_UNAVAILABLE, "null nets");
    // LeeRouter::route validation order (lee.rs:73-97): pads, then endpoints.
    for (size_t i = 0; i < n_nets; ++i) check_net(g, nets[i], i);
    Batch res;
    if (n_nets) {
      upload(ctx->c, b, dims, grid, &ue.e, nullptr, true);
      std::vector<size_t> idx(n_nets);
      for (size_t i = 0; i < n_nets; ++i) idx[i] = i;
      std::vector<mrf_window> wins(n_nets, mrf_window{0, 0, dims.w - 1, dims.h - 1});
      solve_nets(ctx->c, b, g, nets, idx, wins, true, nullptr, res);
    }
    mrf::board_free(b);
    export_paths(res, out);
    if (err) set_err(err, MRF_OK, "ok");
    return MRF_OK;
  } catch (const Error& e) {
    mrf::board_free(b);
    mrf_free_paths(out);
    return fail(err, e);
  } catch (const std::exception& e) {
    mrf::board_free(b);
    mrf_free_paths(out);
    set_err(err, MRF_BACKEND_UNAVAILABLE, e.what());
    return MRF_BACKEND_UNAVAILABLE;
  }
}

static int fields_impl(mrf::Ctx& c, const mrf::DevBoard& b, mrf_dims dims, const uint32_t* const* sources,
                       const size_t* n_sources, const uint32_t* dsts, const int32_t* net_ids, size_t n_fields,
                       const mrf_window* windows, bool force, bool lee, uint32_t* dist_out, uint32_t* hops_out) {
  size_t n = dims_len(dims);
  size_t budget_words = std::max<size_t>(c.label_budget / 8, 1);
  size_t k0 = 0;
  while (k0 < n_fields) {
    size_t words = 0, k1 = k0;
    while (k1 < n_fields) {
      mrf_window w = windows ? windows[k1] : mrf_window{0, 0, dims.w - 1, dims.h - 1};
      size_t fw = mrf::field_label_words(c.shape, w.x1 - w.x0 + 1, w.y1 - w.y0 + 1, dims.layers);
      if (k1 > k0 && words + fw > budget_words) break;
      words += fw;
      ++k1;
    }
    std::vector<mrf::FieldReq> reqs(k1 - k0);
    for (size_t k = k0; k < k1; ++k) {
      mrf::FieldReq& q = reqs[k - k0];
      mrf_window w = windows ? windows[k] : mrf_window{0, 0, dims.w - 1, dims.h - 1};
      q.x0 = w.x0;
      q.y0 = w.y0;
      q.x1 = w.x1;
      q.y1 = w.y1;
      for (size_t j = 0; j < n_sources[k]; ++j)
        if (sources[k][j] < n) q.sources.push_back(sources[k][j]);
      q.dst = (dsts && dsts[k] < n) ? dsts[k] : kMax;
      q.own_id = net_ids ? net_ids[k] : -1;
      if (force && !lee) {
        q.forced = q.sources;
        if (q.dst != kMax) q.forced.push_back(q.dst);
        q.flags = mrf::F_FORCE;
      }
    }
    mrf::SolveStats st;
    mrf::solve_fields(c, b, reqs, &st);
    if (dist_out)
      for (size_t k = k0; k < k1; ++k)
        mrf::read_field(c, b, k - k0, dist_out + k * n, hops_out ? hops_out + k * n : nullptr);
    k0 = k1;
  }
  return MRF_OK;
}

int mrf_sweep_fields(mrf_ctx* ctx, mrf_dims dims, const mrf_grid* grid, const uint32_t* sources, size_t n_sources,
                     uint32_t* dist_out, mrf_error* err) {
  if (!ctx) return MRF_BACKEND_UNAVAILABLE;
  mrf::DevBoard b;
  try {
    Lane lane(ctx->c);
    check_grid(dims, grid);
    if (n_sources && (!sources || !dist_out)) throw Error(MRF_BACKEND_UNAVAILABLE, "null sources or output");
    UnitEdges ue(dims);
    if (n_sources) {
      upload(ctx->c, b, dims, grid, &ue.e, nullptr, true);
      std::vector<const uint32_t*> sp(n_sources);
      std::vector<size_t> ns(n_sources, 1);
      for (size_t i = 0; i < n_sources; ++i) sp[i] = sources + i;
      fields_impl(ctx->c, b, dims, sp.data(), ns.data(), nullptr, nullptr, n_sources, nullptr, false, true, dist_out,
                  nullptr);
    }
    mrf::board_free(b);
    if (err) set_err(err, MRF_OK, "ok");
    return MRF_OK;
  } catch (const Error& e) {
    mrf::board_free(b);
    return fail(err, e);
  } catch (const std::exception& e) {
    mrf::board_free(b);
    set_err(err, MRF_BACKEND_UNAVAILABLE, e.what());
    return MRF_BACKEND_UNAVAILABLE;
  }
}

int mrf_fields(mrf_ctx* ctx, mrf_dims dims, const mrf_grid* grid, const mrf_edges* edges, const mrf_pricing* pricing,
               const uint32_t* const* sources, const size_t* n_sources, const uint32_t* dsts, size_t n_fields,
               const mrf_window* windows, uint32_t* dist_out, uint32_t* hops_out, mrf_error* err) {
  if (!ctx) return MRF_BACKEND_UNAVAILABLE;
  mrf::DevBoard b;
  try {
    Lane lane(ctx->c);
    check_grid(dims, grid);
    check_edges(dims, edges);
    if (n_fields && (!sources || !n_sources || !dist_out)) throw Error(MRF_BACKEND_UNAVAILABLE, "null sources or output");
    for (size_t i = 0; windows && i < n_fields; ++i) check_window(dims, windows[i], i);
    if (n_fields) {
      upload(ctx->c, b, dims, grid, edges, pricing, true);
      fields_impl(ctx->c, b, dims, sources, n_sources, dsts, pricing ? pricing->net_ids : nullptr, n_fields, windows,
                  true, false, dist_out, hops_out);
    }
    mrf::board_free(b);
    if (err) set_err(err, MRF_OK, "ok");
    return MRF_OK;
  } catch (const Error& e) {
    mrf::board_free(b);
    return fail(err, e);
  } catch (const std::exception& e) {
    mrf::board_free(b);
    set_err(err, MRF_BACKEND_UNAVAILABLE, e.what());
    return MRF_BACKEND_UNAVAILABLE;
  }
}

void mrf_free_paths(mrf_paths* p) {
  if (!p) return;
  free(p->routed);
  free(p->search_cost);
  free(p->offsets);
  free(p->cells);
  memset(p, 0, sizeof(*p));
}

// ---- resident boards ----------------------------------------------------------

int mrf_board_create(mrf_ctx* ctx, mrf_dims dims, const mrf_grid* grid, const mrf_edges* edges,
                     const mrf_pricing* pricing, mrf_board** out, mrf_error* err) {
  if (!ctx || !out) return MRF_BACKEND_UNAVAILABLE;
  *out = nullptr;
  mrf_board* bd = new (std::nothrow) mrf_board();
  if (!bd) return MRF_BACKEND_UNAVAILABLE;
  try {
    Lane lane(ctx->c);
    check_grid(dims, grid);
    check_edges(dims, edges);
    size_t n = dims_len(dims);
    bd->ctx = ctx;
    bd->dims = dims;
    bd->cost.assign(grid->cost, grid->cost + n);
    if (grid->n_via_forbidden) bd->vf.assign(grid->via_forbidden, grid->via_forbidden + n);
    if (grid->n_board_constraint) bd->board.assign(grid->board_constraint, grid->board_constraint + n);
    bd->xe.assign(edges->x, edges->x + edges->nx);
    bd->ye.assign(edges->y, edges->y + edges->ny);
    bd->ve.assign(edges->via, edges->via + edges->nvia);
    bd->va.assign(edges->via_allowed, edges->via_allowed + edges->nvia);
    bd->xe.push_back(0);
    bd->ye.push_back(0);
    bd->ve.push_back(0);
    bd->va.push_back(0);
    if (pricing && pricing->n_additive) bd->add.assign(pricing->additive, pricing->additive + pricing->n_additive);
    if (pricing && pricing->n_halo_owner) {
      bd->halo.assign(pricing->halo_owner, pricing->halo_owner + pricing->n_halo_owner);
      bd->halo_pen = pricing->halo_penalty;
    }
    upload(ctx->c, bd->b, dims, grid, edges, pricing, false);
  } catch (const Error& e) {
    mrf::board_free(bd->b);
    delete bd;
    return fail(err, e);
  } catch (const std::exception& e) {
    mrf::board_free(bd->b);
    delete bd;
    set_err(err, MRF_BACKEND_UNAVAILABLE, e.what());
    return MRF_BACKEND_UNAVAILABLE;
  }
  *out = bd;
  if (err) set_err(err, MRF_OK, "ok");
  return MRF_OK;
}

void mrf_board_destroy(mrf_board* bd) {
  if (!bd) return;
  if (bd->ctx) {
    std::lock_guard<std::mutex> g(bd->ctx->c.lane);
    (void)hipSetDevice(bd->ctx->c.device);
    mrf::board_free(bd->b);
  }
  delete bd;
}

size_t mrf_board_bytes(const mrf_board* bd) { return bd ? bd->b.bytes : 0; }

static HostGrid board_host(const mrf_board* bd) {
  HostGrid g;
  g.d = bd->dims;
  g.cost = bd->cost.data();
  g.vf = bd->vf.empty() ? nullptr : bd->vf.data();
  g.board = bd->board.empty() ? nullptr : bd->board.data();
  g.xe = bd->xe.data();
  g.ye = bd->ye.data();
  g.ve = bd->ve.data();
  g.va = bd->va.data();
  g.add = bd->add.empty() ? nullptr : bd->add.data();
  g.halo = bd->halo.empty() ? nullptr : bd->halo.data();
  g.halo_pen = bd->halo_pen;
  g.owner = bd->owner.empty() ? nullptr : bd->owner.data();
  g.lhalo = bd->lhalo.empty() ? nullptr : bd->lhalo.data();
  return g;
}

// Nets on a resident board: windows (full board when NULL), then the
// unreachable nets on the full board when retry_full.  Caller holds the lane.
static void solve_board_batch(mrf_ctx* ctx, mrf_board* bd, const mrf_net* nets, size_t n_nets,
                              const mrf_window* windows, int retry_full, const std::vector<NetPlan>* prices,
                              Batch& all) {
  HostGrid g = board_host(bd);
  mrf_dims dims = bd->dims;
  if (n_nets && !nets) throw Error(MRF_BACKEND_UNAVAILABLE, "null nets");
  for (size_t i = 0; windows && i < n_nets; ++i) check_window(dims, windows[i], i);
  for (size_t i = 0; i < n_nets; ++i) check_net(g, nets[i], i);
  mrf_window full{0, 0, dims.w - 1, dims.h - 1};
  std::vector<size_t> idx(n_nets);
  for (size_t i = 0; i < n_nets; ++i) idx[i] = i;
  std::vector<mrf_window> wins(n_nets, full);
  if (windows) wins.assign(windows, windows + n_nets);
  solve_nets(ctx->c, bd->b, g, nets, idx, wins, false, nullptr, all, prices);
  if (retry_full && windows) {
    std::vector<size_t> retry_idx;
    for (size_t i = 0; i < n_nets; ++i)
      if (!all.routed[i]) retry_idx.push_back(i);
    if (!retry_idx.empty()) {
      std::vector<mrf_window> fw(retry_idx.size(), full);
      Batch second;
      solve_nets(ctx->c, bd->b, g, nets, retry_idx, fw, false, nullptr, second, prices);
      for (size_t k = 0; k < retry_idx.size(); ++k) {
        size_t i = retry_idx[k];
        all.routed[i] = second.routed[k];
        all.cost[i] = second.cost[k];
        all.paths[i] = std::move(second.paths[k]);
      }
    }
  }
}

static void solve_board(mrf_ctx* ctx, mrf_board* bd, const mrf_net* nets, size_t n_nets, const mrf_window* windows,
                        int retry_full, const std::vector<NetPlan>* prices, mrf_paths* out) {
  Batch all;
  solve_board_batch(ctx, bd, nets, n_nets, windows, retry_full, prices, all);
  export_paths(all, out);
}

// The allocation-free mode of a context for one call: solves carve their
// device buffers from the device's scratch array, run the plain form, and
// chunk labels within a budget that leaves the rest of the array for the
// other buffers of a solve and its paths.
struct ArenaScope {
  mrf::Ctx& c;
  int mode;
  size_t budget;
  explicit ArenaScope(mrf::Ctx& cx) : c(cx), mode(cx.mode), budget(cx.label_budget) {
    mrf::arena_begin(c);
    c.mode = mrf::MODE_PLAIN;
    c.label_budget = std::min(budget, mrf::scratch_bytes() / 8 * 5);
  }
  ~ArenaScope() {
    c.mode = mode;
    c.label_budget = budget;
    mrf::arena_end(c);
  }
};

int mrf_solve(mrf_ctx* ctx, mrf_board* bd, const mrf_net* nets, size_t n_nets, const mrf_window* windows,
              int retry_full, mrf_paths* out, mrf_error* err) {
  if (out) memset(out, 0, sizeof(*out));
  if (!ctx || !bd || !out || bd->ctx != ctx) return MRF_BACKEND_UNAVAILABLE;
  try {
    Lane lane(ctx->c);
    solve_board(ctx, bd, nets, n_nets, windows, retry_full, nullptr, out);
    if (err) set_err(err, MRF_OK, "ok");
    return MRF_OK;
  } catch (const Error& e) {
    mrf_free_paths(out);
    return fail(err, e);
  } catch (const std::exception& e) {
    mrf_free_paths(out);
    set_err(err, MRF_BACKEND_UNAVAILABLE, e.what());
    return MRF_BACKEND_UNAVAILABLE;
  }
}

int mrf_board_set_additive(mrf_board* bd, const uint32_t* additive, size_t n, mrf_error* err) {
  if (!bd || !bd->ctx) return MRF_BACKEND_UNAVAILABLE;
  try {
    Lane lane(bd->ctx->c, false);
    if (!additive && n == 0) {
      mrf::board_clear_additive(bd->ctx->c, bd->b);
      bd->add.clear();
    } else {
      if (n != dims_len(bd->dims) || !additive)
        throw Error(MRF_MALFORMED_GRID, "additive plane length does not match dims");
      mrf::board_set_additive(bd->ctx->c, bd->b, additive);
      bd->add.assign(additive, additive + n);
    }
    if (err) set_err(err, MRF_OK, "ok");
    return MRF_OK;
  } catch (const Error& e) {
    return fail(err, e);
  } catch (const std::exception& e) {
    set_err(err, MRF_BACKEND_UNAVAILABLE, e.what());
    return MRF_BACKEND_UNAVAILABLE;
  }
}

int mrf_route_negotiated_batch(mrf_ctx* ctx, mrf_board* bd, const mrf_net* nets, size_t n_nets,
                               const mrf_window* windows, int retry_full, const mrf_negotiation* neg, mrf_paths* out,
                               mrf_error* err) {
  if (out) memset(out, 0, sizeof(*out));
  if (!ctx || !bd || !out || !neg || bd->ctx != ctx) return MRF_BACKEND_UNAVAILABLE;
  try {
    Lane lane(ctx->c);
    const size_t n = dims_len(bd->dims);
    if (!bd->halo.empty()) throw Error(MRF_BACKEND_UNAVAILABLE, "negotiated batches take boards without a halo-owner plane");
    if (neg->n_cells != n || !neg->history || !neg->present)
      throw Error(MRF_MALFORMED_GRID, "negotiation planes do not match dims");
    if ((neg->own_path_off && !neg->own_path) || (neg->own_halo_off && !neg->own_halo))
      throw Error(MRF_BACKEND_UNAVAILABLE, "own-footprint offsets without cells");
    const uint64_t pf = (uint64_t)neg->pfac * kNegScale;
    auto halo_at = [&](size_t c) -> uint64_t { return neg->present_halo ? neg->present_halo[c] : 0u; };
    // Shared congestion snapshot (negotiated.rs:1654-1667), folded with
    // history into the board's additive plane: what every net pays on cells
    // outside its own footprint.
    auto base_price = [&](size_t c) {
      const uint64_t shared = sat_add64(sat_mul64(pf, neg->present[c]), sat_mul64(pf, halo_at(c)));
      return clamp32(sat_add64(neg->history[c], shared));
    };
    // Own footprint of each net: priced_with_base's self-subtraction
    // (negotiated.rs:2777-2797) as additive overrides.  Built first: a
    // malformed footprint fails the call before the board's plane changes.
    std::vector<NetPlan> prices(n_nets);
    std::vector<std::string> bad(n_nets);
    parallel_for(n_nets, 16, [&](size_t a, size_t b) {
      std::vector<uint32_t> path, halo;
      std::vector<uint64_t> hcount;
      for (size_t i = a; i < b; ++i) {
        try {
          own_cells(neg->own_path_off, neg->own_path, nullptr, i, n, path, nullptr);
          own_cells(neg->own_halo_off, neg->own_halo, neg->own_halo_count, i, n, halo, &hcount);
        } catch (const Error& e) {
          bad[i] = e.what();
          continue;
        }
        std::vector<uint32_t> u;
        std::set_union(path.begin(), path.end(), halo.begin(), halo.end(), std::back_inserter(u));
        NetPlan& p = prices[i];
        for (uint32_t c : u) {
          const uint64_t self_present = std::binary_search(path.begin(), path.end(), c) ? 1u : 0u;
          const auto h = std::lower_bound(halo.begin(), halo.end(), c);
          const uint64_t self_halo = (h != halo.end() && *h == c) ? hcount[h - halo.begin()] : 0u;
          uint64_t foreign;
          if (neg->fused) {
            const uint64_t shared = sat_add64(sat_mul64(pf, neg->present[c]), sat_mul64(pf, halo_at(c)));
            foreign = sat_sub64(sat_sub64(shared, sat_mul64(pf, self_present)), sat_mul64(pf, self_halo));
          } else {
            foreign = sat_add64(sat_mul64(pf, sat_sub64(neg->present[c], self_present)),
                                sat_mul64(pf, sat_sub64(halo_at(c), self_halo)));
          }
          const uint32_t price = clamp32(sat_add64(neg->history[c], foreign));
          if (price != base_price(c)) {
            p.ovr_cells.push_back(c);
            p.ovr_add.push_back(price);
          }
        }
      }
    });
    for (const std::string& m : bad)
      if (!m.empty()) throw Error(MRF_BACKEND_UNAVAILABLE, m.c_str());
    // The plane, built in place in the board's page-locked host copy (kept
    // for host re-validation), then uploaded.
    bd->add.resize(n);
    uint32_t* plane = bd->add.data();
    std::mutex max_mu;
    uint32_t plane_max = 0;
    parallel_for(n, 1 << 16, [&](size_t a, size_t b) {
      uint32_t m = 0;
      for (size_t c = a; c < b; ++c) {
        plane[c] = base_price(c);
        m = std::max(m, plane[c]);
      }
      std::lock_guard<std::mutex> lk(max_mu);
      plane_max = std::max(plane_max, m);
    });
    mrf::board_set_additive(ctx->c, bd->b, plane, plane_max);
    solve_board(ctx, bd, nets, n_nets, windows, retry_full, &prices, out);
    if (err) set_err(err, MRF_OK, "ok");
    return MRF_OK;
  } catch (const Error& e) {
    mrf_free_paths(out);
    return fail(err, e);
  } catch (const std::exception& e) {
    mrf_free_paths(out);
    set_err(err, MRF_BACKEND_UNAVAILABLE, e.what());
    return MRF_BACKEND_UNAVAILABLE;
  }
}

int mrf_negotiation_settle(size_t n_cells, size_t n_nets, const uint64_t* path_off, const uint32_t* path_cells,
                           const int64_t* groups, uint32_t scale, uint32_t* present, uint32_t* history,
                           uint8_t* touches, uint32_t* overused, size_t* n_overused, mrf_error* err) {
  try {
    if (!present || !history || !overused || !n_overused || (n_nets && (!path_off || !groups || !touches)))
      throw Error(MRF_BACKEND_UNAVAILABLE, "null negotiation settle argument");
    const uint64_t total = n_nets ? path_off[n_nets] : 0;
    if (total && !path_cells) throw Error(MRF_BACKEND_UNAVAILABLE, "null path cells");
    for (size_t i = 0; i < n_nets; ++i)
      if (path_off[i] > path_off[i + 1]) throw Error(MRF_BACKEND_UNAVAILABLE, "path offsets must not decrease");
    for (uint64_t k = 0; k < total; ++k)
      if (path_cells[k] >= n_cells) throw Error(MRF_BACKEND_UNAVAILABLE, "path cell outside the board");
    // present: paths per cell (merge, negotiated.rs:1779-1823)
    memset(present, 0, n_cells * sizeof(uint32_t));
    for (uint64_t k = 0; k < total; ++k) present[path_cells[k]]++;
    // Over-use across groups (:1924-1945): cells held by two or more paths
    // are the candidates; a candidate is over-used when two of its paths
    // belong to different groups.  Two passes over the paths: the first
    // records the first group seen at each candidate (first[c]), the second
    // marks a candidate another group also holds.  The scratch planes stay
    // zero between calls (only the candidates are reset).
    thread_local std::vector<int64_t> first;
    thread_local std::vector<uint8_t> over;  // 2 candidate, 1 over-used
    if (first.size() < n_cells) {
      first.assign(n_cells, 0);
      over.assign(n_cells, 0);
    }
    std::vector<uint32_t> cells;
    for (size_t i = 0; i < n_nets; ++i)
      for (uint64_t k = path_off[i]; k < path_off[i + 1]; ++k) {
        const uint32_t c = path_cells[k];
        if (present[c] >= 2 && !over[c]) {
          first[c] = groups[i];
          over[c] = 2;  // candidate, not yet over-used
          cells.push_back(c);
        }
      }
    for (size_t i = 0; i < n_nets; ++i)
      for (uint64_t k = path_off[i]; k < path_off[i + 1]; ++k) {
        const uint32_t c = path_cells[k];
        if (over[c] == 2 && first[c] != groups[i]) over[c] = 1;
      }
    size_t n_over = 0;
    std::sort(cells.begin(), cells.end());
    for (uint32_t c : cells) {
      if (over[c] == 1) {
        overused[n_over++] = c;
        const uint64_t h = (uint64_t)history[c] + scale;
        history[c] = h > UINT32_MAX ? UINT32_MAX : (uint32_t)h;
      }
    }
    // per net: holds an over-used cell (the next iteration's dirty set)
    for (size_t i = 0; i < n_nets; ++i) {
      uint8_t t = 0;
      for (uint64_t k = path_off[i]; k < path_off[i + 1] && !t; ++k) t = over[path_cells[k]] == 1;
      touches[i] = t;
    }
    for (uint32_t c : cells) {
      first[c] = 0;
      over[c] = 0;
    }
    *n_overused = n_over;
    if (err) set_err(err, MRF_OK, "ok");
    return MRF_OK;
  } catch (const Error& e) {
    return fail(err, e);
  } catch (const std::exception& e) {
    set_err(err, MRF_BACKEND_UNAVAILABLE, e.what());
    return MRF_BACKEND_UNAVAILABLE;
  }
}

int mrf_route_legal_batch(mrf_ctx* ctx, mrf_board* bd, const mrf_net* nets, size_t n_nets,
                          const mrf_window* windows, int retry_full, const int64_t* owner, const int64_t* halo,
                          size_t n_cells, const int64_t* own_groups, mrf_paths* out, mrf_error* err) {
  if (out) memset(out, 0, sizeof(*out));
  if (!ctx || !bd || !out || bd->ctx != ctx) return MRF_BACKEND_UNAVAILABLE;
  try {
    Lane lane(ctx->c);
    const size_t n = dims_len(bd->dims);
    if (!bd->add.empty() || !bd->halo.empty())
      throw Error(MRF_BACKEND_UNAVAILABLE, "legal batches take boards without additive or halo-owner pricing planes");
    if ((owner || halo) && n_cells != n) throw Error(MRF_MALFORMED_GRID, "owner / halo plane length does not match dims");
    if (n_nets && !own_groups) throw Error(MRF_BACKEND_UNAVAILABLE, "null own groups");
    auto narrow32 = [](int64_t v) {
      if (v < INT32_MIN || v > INT32_MAX) throw Error(MRF_BACKEND_UNAVAILABLE, "group id outside the 32-bit range");
      return (int32_t)v;
    };
    std::vector<int32_t> o32, h32;
    if (owner) {
      o32.resize(n);
      for (size_t c = 0; c < n; ++c) o32[c] = narrow32(owner[c]);
    }
    if (halo) {
      h32.resize(n);
      for (size_t c = 0; c < n; ++c) h32[c] = narrow32(halo[c]);
    }
    std::vector<NetPlan> plans(n_nets);
    for (size_t i = 0; i < n_nets; ++i) plans[i].own_group = narrow32(own_groups[i]);
    mrf::board_set_legal(ctx->c, bd->b, owner ? o32.data() : nullptr, halo ? h32.data() : nullptr);
    bd->owner = std::move(o32);
    bd->lhalo = std::move(h32);
    try {
      solve_board(ctx, bd, nets, n_nets, windows, retry_full, &plans, out);
    } catch (...) {
      mrf::board_clear_legal(bd->b);
      bd->owner.clear();
      bd->lhalo.clear();
      throw;
    }
    mrf::board_clear_legal(bd->b);
    bd->owner.clear();
    bd->lhalo.clear();
    if (err) set_err(err, MRF_OK, "ok");
    return MRF_OK;
  } catch (const Error& e) {
    mrf_free_paths(out);
    return fail(err, e);
  } catch (const std::exception& e) {
    mrf_free_paths(out);
    set_err(err, MRF_BACKEND_UNAVAILABLE, e.what());
    return MRF_BACKEND_UNAVAILABLE;
  }
}

int mrf_autotune(mrf_ctx* ctx, mrf_board* bd, const mrf_net* nets, size_t n_nets, const mrf_window* windows,
                 size_t sample, mrf_tuning* out, mrf_error* err) {
  if (!ctx || !bd || !out || bd->ctx != ctx) return MRF_BACKEND_UNAVAILABLE;
  memset(out, 0, sizeof(*out));
  try {
    Lane lane(ctx->c);
    HostGrid g = board_host(bd);
    const mrf_dims dims = bd->dims;
    if (n_nets == 0 || !nets) throw Error(MRF_BACKEND_UNAVAILABLE, "autotune needs nets");
    for (size_t i = 0; windows && i < n_nets; ++i) check_window(dims, windows[i], i);
    for (size_t i = 0; i < n_nets; ++i) check_net(g, nets[i], i);
    // every k-th net: the sample every candidate is timed on, and the larger
    // set the fastest few are confirmed on (a small sample favours forms that
    // save per-round host time, which a full batch may not)
    const mrf_window full{0, 0, dims.w - 1, dims.h - 1};
    auto every_kth = [&](size_t want, std::vector<size_t>& idx, std::vector<mrf_window>& wins) {
      want = std::max<size_t>(1, std::min(want, n_nets));
      const size_t step = std::max<size_t>(n_nets / want, 1);
      idx.clear();
      wins.clear();
      for (size_t i = 0; i < n_nets && idx.size() < want; i += step) idx.push_back(i);
      for (size_t i : idx) wins.push_back(windows ? windows[i] : full);
    };
    const size_t want = sample ? sample : 64;
    std::vector<size_t> idx, cidx;
    std::vector<mrf_window> wins, cwins;
    every_kth(want, idx, wins);
    every_kth(std::max<size_t>(8 * want, 1024), cidx, cwins);
    mrf::Ctx& c = ctx->c;
    auto run_on = [&](const std::vector<size_t>& idx, const std::vector<mrf_window>& wins) {
      auto t0 = std::chrono::steady_clock::now();
      Batch b;
      solve_nets(c, bd->b, g, nets, idx, wins, false, nullptr, b);
      if (windows) {
        std::vector<size_t> retry;
        for (size_t k = 0; k < idx.size(); ++k)
          if (!b.routed[k]) retry.push_back(idx[k]);
        if (!retry.empty()) {
          Batch b2;
          solve_nets(c, bd->b, g, nets, retry, std::vector<mrf_window>(retry.size(), full), false, nullptr, b2);
        }
      }
      return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    auto run = [&]() { return run_on(idx, wins); };
    struct Cand {
      uint32_t tw, th, threads, max_local;
      int mode;
      uint32_t graph_rounds;
      uint32_t delta;
    };
    const mrf::TileShape keep_shape = c.shape;
    const int keep_mode = c.mode == MRF_MODE_PERSISTENT ? MRF_MODE_PLAIN : c.mode;
    const uint32_t keep_rounds = c.graph_rounds;
    auto apply = [&](const Cand& k) {
      c.shape.tw = k.tw;
      c.shape.th = k.th;
      c.shape.threads = k.threads;
      c.shape.max_local = k.max_local;
      c.mode = k.mode;
      c.graph_rounds = k.graph_rounds;
      c.delta = k.delta;
    };
    // Typical step price of the board (mean planar edge price x mean enter
    // weight + mean additive price) sets the scale of the delta candidates:
    // a few tile crossings.
    double e_mean = 0, w_mean = 0, a_mean = 0;
    size_t ne = 0, nw = 0;
    for (uint32_t v : bd->xe) e_mean += v, ++ne;
    for (uint32_t v : bd->ye) e_mean += v, ++ne;
    e_mean = ne ? e_mean / ne : 1.0;
    for (uint32_t v : bd->cost)
      if (v != kMax) w_mean += v, ++nw;
    w_mean = nw ? w_mean / nw : 1.0;
    for (uint32_t v : bd->add) a_mean += v;
    a_mean = bd->add.empty() ? 0.0 : a_mean / bd->add.size();
    const double typical = std::max(1.0, e_mean * w_mean + a_mean);
    auto timed = [&](const Cand& k) {
      apply(k);
      run();  // warm-up: workspace growth, graph capture
      return std::min(run(), run());
    };
    Cand base{keep_shape.tw, keep_shape.th, keep_shape.threads, keep_shape.max_local, keep_mode, keep_rounds,
              c.delta};
    double best_ms = timed(base);
    out->default_ms = best_ms;
    Cand best = base;
    uint32_t tried = 1;
    std::vector<std::pair<double, Cand>> scored{{best_ms, base}};
    const uint32_t shapes[][2] = {{32, 16}, {16, 16}, {16, 8}, {8, 4}};
    for (auto& sh : shapes) {
      if (mrf::lds_bytes_for(sh[0], sh[1], dims.layers) > 160 * 1024 - 64) continue;
      // 64 threads on 16 x 16 tiles is the wave form (plain form, and the
      // queue form's wave visits; the graph form runs that shape at 256
      // threads)
      for (uint32_t threads : {64u, 128u, 256u})
        for (uint32_t ml : {4u, 32u})
          for (int mode : {(int)MRF_MODE_PLAIN, (int)MRF_MODE_GRAPH, (int)MRF_MODE_QUEUE}) {
            if (threads == 64 && (sh[0] != 16 || sh[1] != 16 || mode == MRF_MODE_GRAPH)) continue;
            std::vector<std::pair<uint32_t, uint32_t>> forms;  // (graph rounds, delta)
            if (mode == MRF_MODE_GRAPH) {
              forms = {{4, kMax}, {16, kMax}};
            } else if (mode == MRF_MODE_QUEUE) {
              forms = {{8, kMax}};  // one launch: no rounds to group, no delta
            } else {
              forms = {{8, kMax}};
              for (double crossings : {2.0, 8.0})
                forms.push_back({8, (uint32_t)std::min(4e9, crossings * sh[0] * typical)});
            }
            for (auto& fm : forms) {
              Cand k{sh[0], sh[1], threads, ml, mode, fm.first, fm.second};
              double ms = timed(k);
              ++tried;
              scored.push_back({ms, k});
            }
          }
    }
    // Confirmation: the three fastest on the sample and the context's own
    // configuration, timed on the larger set; the fastest there is kept.
    std::stable_sort(scored.begin(), scored.end(),
                     [](const std::pair<double, Cand>& a, const std::pair<double, Cand>& b) { return a.first < b.first; });
    std::vector<std::pair<double, Cand>> finalists(scored.begin(), scored.begin() + std::min<size_t>(3, scored.size()));
    bool has_base = false;
    for (auto& f : finalists)
      has_base = has_base || (f.second.tw == base.tw && f.second.th == base.th && f.second.threads == base.threads &&
                              f.second.max_local == base.max_local && f.second.mode == base.mode &&
                              f.second.graph_rounds == base.graph_rounds && f.second.delta == base.delta);
    if (!has_base) finalists.push_back({out->default_ms, base});
    double best_full = 0;
    for (size_t k = 0; k < finalists.size(); ++k) {
      apply(finalists[k].second);
      run_on(cidx, cwins);  // warm-up
      const double ms = std::min(run_on(cidx, cwins), run_on(cidx, cwins));
      if (k == 0 || ms < best_full) {
        best_full = ms;
        best = finalists[k].second;
        best_ms = finalists[k].first;
      }
    }
    apply(best);
    out->tile_w = best.tw;
    out->tile_h = best.th;
    out->threads = best.threads;
    out->max_local = best.max_local;
    out->mode = best.mode;
    out->graph_rounds = best.graph_rounds;
    out->delta = best.delta;
    out->sample_ms = best_ms;
    out->tried = tried;
    if (err) set_err(err, MRF_OK, "ok");
    return MRF_OK;
  } catch (const Error& e) {
    return fail(err, e);
  } catch (const std::exception& e) {
    set_err(err, MRF_BACKEND_UNAVAILABLE, e.what());
    return MRF_BACKEND_UNAVAILABLE;
  }
}

size_t mrf_noalloc_scratch_bytes(void) { return mrf::scratch_bytes(); }

int mrf_solve_noalloc(mrf_ctx* ctx, mrf_board* bd, const mrf_net* nets, size_t n_nets, const mrf_window* windows,
                      int retry_full, uint8_t* routed, uint32_t* search_cost, uint64_t* offsets, uint32_t* cells,
                      size_t cells_cap, size_t* n_cells, mrf_error* err) {
  if (!ctx || !bd || bd->ctx != ctx || !offsets || !n_cells || (n_nets && (!routed || !search_cost)) ||
      (cells_cap && !cells)) {
    set_err(err, MRF_BACKEND_UNAVAILABLE, "null context, board or output");
    return MRF_BACKEND_UNAVAILABLE;
  }
  try {
    Lane lane(ctx->c);
    Batch all;
    {
      ArenaScope arena(ctx->c);
      solve_board_batch(ctx, bd, nets, n_nets, windows, retry_full, nullptr, all);
    }
    size_t total = 0;
    for (const auto& p : all.paths) total += p.size();
    *n_cells = total;
    if (total > cells_cap)
      throw Error(MRF_BACKEND_UNAVAILABLE, "path cells (" + std::to_string(total) + ") exceed cells_cap (" +
                                               std::to_string(cells_cap) + ")");
    offsets[0] = 0;
    for (size_t i = 0; i < n_nets; ++i) {
      routed[i] = all.routed[i];
      search_cost[i] = all.routed[i] ? all.cost[i] : kMax;
      std::copy(all.paths[i].begin(), all.paths[i].end(), cells + offsets[i]);
      offsets[i + 1] = offsets[i] + all.paths[i].size();
    }
    if (err) set_err(err, MRF_OK, "ok");
    return MRF_OK;
  } catch (const Error& e) {
    return fail(err, e);
  } catch (const std::exception& e) {
    set_err(err, MRF_BACKEND_UNAVAILABLE, e.what());
    return MRF_BACKEND_UNAVAILABLE;
  }
}

int mrf_solve_fields(mrf_ctx* ctx, mrf_board* bd, const uint32_t* const* sources, const size_t* n_sources,
                     const uint32_t* dsts, const int32_t* net_ids, size_t n_fields, const mrf_window* windows,
                     mrf_error* err) {
  if (!ctx || !bd || bd->ctx != ctx) return MRF_BACKEND_UNAVAILABLE;
  try {
    Lane lane(ctx->c);
    mrf_dims dims = bd->dims;
    if (n_fields && (!sources || !n_sources)) throw Error(MRF_BACKEND_UNAVAILABLE, "null sources");
    for (size_t i = 0; windows && i < n_fields; ++i) check_window(dims, windows[i], i);
    size_t total = 0;
    for (size_t k = 0; k < n_fields; ++k) {
      mrf_window w = windows ? windows[k] : mrf_window{0, 0, dims.w - 1, dims.h - 1};
      total += mrf::field_label_words(ctx->c.shape, w.x1 - w.x0 + 1, w.y1 - w.y0 + 1, dims.layers);
    }
    if (total * 8 > ctx->c.label_budget)
      throw Error(MRF_BACKEND_UNAVAILABLE, "resident fields exceed the label budget; split the request");
    fields_impl(ctx->c, bd->b, dims, sources, n_sources, dsts, net_ids, n_fields, windows, true, false, nullptr,
                nullptr);
    if (err) set_err(err, MRF_OK, "ok");
    return MRF_OK;
  } catch (const Error& e) {
    return fail(err, e);
  } catch (const std::exception& e) {
    set_err(err, MRF_BACKEND_UNAVAILABLE, e.what());
    return MRF_BACKEND_UNAVAILABLE;
  }
}

int mrf_read_field(mrf_ctx* ctx, mrf_board* bd, size_t field, uint32_t* dist_out, uint32_t* hops_out,
                   mrf_error* err) {
  if (!ctx || !bd || !dist_out) return MRF_BACKEND_UNAVAILABLE;
  try {
    Lane lane(ctx->c, false);
    mrf::read_field(ctx->c, bd->b, field, dist_out, hops_out);
    return MRF_OK;
  } catch (const Error& e) {
    return fail(err, e);
  }
}

int mrf_set_label_budget(mrf_ctx* ctx, size_t bytes) {
  if (!ctx) return MRF_BACKEND_UNAVAILABLE;
  std::lock_guard<std::mutex> g(ctx->c.lane);
  ctx->c.label_budget = bytes ? std::min(bytes, ctx->c.label_budget_default) : ctx->c.label_budget_default;
  return MRF_OK;
}

size_t mrf_get_label_budget(mrf_ctx* ctx) { return ctx ? ctx->c.label_budget : 0; }

int mrf_congestion_from(mrf_dims dims, const mrf_paths* paths, uint32_t* out, mrf_error* err) {
  try {
    const size_t n = dims_len(dims);
    if (!paths || (n && !out)) throw Error(MRF_BACKEND_UNAVAILABLE, "null paths or output");
    std::fill(out, out + n, 0u);
    // Per path, a cell counts once however often the path visits it: the
    // stamp of the last path that counted the cell.
    std::vector<uint32_t> seen(n, 0);
    for (size_t i = 0; i < paths->n; ++i) {
      if (!paths->routed[i]) continue;
      const uint32_t stamp = (uint32_t)i + 1;
      for (uint64_t k = paths->offsets[i]; k < paths->offsets[i + 1]; ++k) {
        const uint32_t c = paths->cells[k];
        if (c >= n) throw Error(MRF_BACKEND_UNAVAILABLE, "path cell outside the board");
        if (seen[c] != stamp) {
          seen[c] = stamp;
          ++out[c];
        }
      }
    }
    if (err) set_err(err, MRF_OK, "ok");
    return MRF_OK;
  } catch (const Error& e) {
    return fail(err, e);
  } catch (const std::exception& e) {
    set_err(err, MRF_BACKEND_UNAVAILABLE, e.what());
    return MRF_BACKEND_UNAVAILABLE;
  }
}

}  // extern "C"
