// Host-side request validation and fail-closed path re-validation shared by
// the C ABI (capi.cpp) and the serving loop (serve.cpp).
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/mrfield.h"
#include "engine.h"

namespace mrf {
namespace host {

inline size_t dims_len(mrf_dims d) { return (size_t)d.w * d.h * d.layers; }

// Host view of a grid + prices, shared by validation and path checks.
struct HostGrid {
  mrf_dims d{};
  const uint32_t* cost = nullptr;
  const uint8_t* vf = nullptr;
  const uint8_t* board = nullptr;
  const uint32_t* xe = nullptr;
  const uint32_t* ye = nullptr;
  const uint32_t* ve = nullptr;
  const uint8_t* va = nullptr;
  const uint32_t* add = nullptr;
  const int32_t* halo = nullptr;
  uint32_t halo_pen = 0;
  // legal searches: committed copper owners and clearance halo (or null)
  const int32_t* owner = nullptr;
  const int32_t* lhalo = nullptr;
  bool obst(uint32_t c) const { return cost[c] == kMax; }
  bool trace(uint32_t c) const { return board && (board[c] & MRF_BOARD_TRACE_NODE); }
  bool vianode(uint32_t c) const { return board && (board[c] & MRF_BOARD_VIA_NODE); }
  // whether c is not the group's own copper (a foreign halo blocks it)
  bool owner_free_of(uint32_t c, int32_t grp) const { return !(owner && owner[c] == grp); }
  bool viaforb(uint32_t c) const { return vf && vf[c]; }
  // mr-core is_board_planar_step_forbidden (lib.rs:466-496) for neighbours.
  bool planar_forbidden(uint32_t u, uint32_t v) const {
    if (!board) return false;
    uint32_t plane = d.w * d.h;
    uint32_t ur = u % plane, vr = v % plane;
    uint32_t ux = ur % d.w, uy = ur / d.w, vx = vr % d.w, vy = vr / d.w;
    uint8_t sb, db;
    if (ux == vx && uy == vy + 1) {
      sb = MRF_BOARD_EDGE_NEG_Y;
      db = MRF_BOARD_EDGE_POS_Y;
    } else if (uy == vy && ux == vx + 1) {
      sb = MRF_BOARD_EDGE_NEG_X;
      db = MRF_BOARD_EDGE_POS_X;
    } else if (uy == vy && ux + 1 == vx) {
      sb = MRF_BOARD_EDGE_POS_X;
      db = MRF_BOARD_EDGE_NEG_X;
    } else {
      sb = MRF_BOARD_EDGE_POS_Y;
      db = MRF_BOARD_EDGE_NEG_Y;
    }
    return (board[u] & sb) || (board[v] & db);
  }
};

inline bool sorted_has(const std::vector<uint32_t>& a, uint32_t v) { return std::binary_search(a.begin(), a.end(), v); }

inline std::vector<uint32_t> sorted_unique(const uint32_t* p, size_t n) {
  std::vector<uint32_t> v(p, p + n);
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  return v;
}

inline void check_grid(mrf_dims d, const mrf_grid* g) {
  size_t n = dims_len(d);
  if (!g || !g->cost || g->n_cost != n || (g->n_via_forbidden != 0 && g->n_via_forbidden != n) ||
      (g->n_board_constraint != 0 && g->n_board_constraint != n))
    throw Error(MRF_MALFORMED_GRID, "grid is malformed: cost length does not match dims");
  if (d.w == 0 || d.h == 0 || d.layers == 0) throw Error(MRF_MALFORMED_GRID, "grid has no cells");
  if (n > 0xFFFFFFFEull) throw Error(MRF_BACKEND_UNAVAILABLE, "grid exceeds 32-bit cell indexing");
  if (d.layers > 16) throw Error(MRF_BACKEND_UNAVAILABLE, "at most 16 layers are supported");
}

inline void check_edges(mrf_dims d, const mrf_edges* e) {
  if (!e || e->nx != (size_t)(d.w - 1) || e->ny != (size_t)(d.h - 1) || e->nvia != (size_t)(d.layers - 1) ||
      (d.w > 1 && !e->x) || (d.h > 1 && !e->y) || (d.layers > 1 && (!e->via || !e->via_allowed)))
    throw Error(MRF_BACKEND_UNAVAILABLE, "isolated batch shape does not match grid/nets");
}

inline void check_window(mrf_dims d, const mrf_window& w, size_t i) {
  if (!(w.x0 <= w.x1 && w.y0 <= w.y1 && w.x1 < d.w && w.y1 < d.h))
    throw Error(MRF_BACKEND_UNAVAILABLE, "net " + std::to_string(i) + " has an invalid isolated-route window");
}

// Endpoint contract of provider_alone_paths / metal_route_isolated_batch
// (negotiated.rs:1396-1407, mr-metal lib.rs:446-466).
inline void check_net(const HostGrid& g, const mrf_net& n, size_t i) {
  size_t len = dims_len(g.d);
  auto bad = [&]() { throw Error(MRF_INVALID_ENDPOINT, "net " + std::to_string(i) + ": endpoint out of bounds or on an obstacle", (uint32_t)i); };
  if ((n.n_pads && !n.pads) || (n.n_via_pads && !n.via_pads)) bad();
  for (size_t k = 0; k < n.n_pads; ++k)
    if (n.pads[k] >= len) bad();
  std::vector<uint32_t> pads = sorted_unique(n.pads, n.n_pads);
  for (size_t k = 0; k < n.n_via_pads; ++k)
    if (n.via_pads[k] >= len || !sorted_has(pads, n.via_pads[k])) bad();
  for (uint32_t c : {n.src, n.dst}) {
    if (c >= len || g.trace(c) || (g.obst(c) && !sorted_has(pads, c))) bad();
  }
}

struct NetPlan {
  std::vector<uint32_t> pads, vpads;
  // Per-net additive prices replacing the board plane's (cells ascending).
  std::vector<uint32_t> ovr_cells, ovr_add;
  int32_t own_group = -1;  // legal searches: the net's group
};

inline uint32_t step_price(uint32_t e, uint32_t w, uint64_t a) {
  uint64_t p = (uint64_t)e * w + a;
  return p < 0xFFFFFFFEull ? (uint32_t)p : 0xFFFFFFFEu;
}

// Fail-closed re-validation of a device path: provider_path_is_valid
// (negotiated.rs:1284-1331) plus the ragged result checks of
// mr-metal lib.rs:902-974 (window, recomputed cost == search cost).
inline bool path_ok(const HostGrid& g, const mrf_net& n, const NetPlan& p, const mrf_window& w, int32_t own_id,
             const uint32_t* path, size_t len, uint32_t search_cost, bool lee) {
  size_t total = dims_len(g.d);
  if (len == 0 || len > total || path[0] != n.src || path[len - 1] != n.dst) return false;
  uint32_t plane = g.d.w * g.d.h;
  // simple path: no cell twice (open addressing over at least twice the
  // path's length; the empty slot value is no cell, as total < 2^32)
  thread_local std::vector<uint32_t> seen;
  size_t cap = 16;
  while (cap < 2 * len) cap <<= 1;
  seen.assign(cap, UINT32_MAX);
  uint64_t cost = 0;
  uint32_t al = 0, ar = 0, ax = 0, ay = 0;  // the previous cell's layer, in-plane index, x, y
  for (size_t i = 0; i < len; ++i) {
    uint32_t c = path[i];
    if (c >= total || g.trace(c)) return false;
    for (size_t h = (c * 0x9E3779B1u) & (cap - 1);; h = (h + 1) & (cap - 1)) {
      if (seen[h] == c) return false;
      if (seen[h] == UINT32_MAX) {
        seen[h] = c;
        break;
      }
    }
    const bool endpoint = c == n.src || c == n.dst;
    if (g.owner && g.owner[c] >= 0 && g.owner[c] != p.own_group) return false;  // foreign copper
    if (!endpoint && g.lhalo && g.owner_free_of(c, p.own_group) &&
        (g.lhalo[c] == -2 || (g.lhalo[c] >= 0 && g.lhalo[c] != p.own_group)))
      return false;  // foreign halo
    if (g.obst(c) && !sorted_has(p.pads, c)) return false;
    const uint32_t cl = c / plane, r = c - cl * plane, y = r / g.d.w, x = r - y * g.d.w;
    if (x < w.x0 || x > w.x1 || y < w.y0 || y > w.y1) return false;
    const uint32_t a = i ? path[i - 1] : 0u;
    const uint32_t pl = al, pr = ar, px = ax, py = ay;
    al = cl;
    ar = r;
    ax = x;
    ay = y;
    if (i == 0) continue;
    uint32_t e;
    if (pl == cl) {
      bool adj = (px == x && (py + 1 == y || y + 1 == py)) || (py == y && (px + 1 == x || x + 1 == px));
      if (!adj || g.planar_forbidden(a, c)) return false;
      e = lee ? 1u : (px != x ? g.xe[std::min(px, x)] : g.ye[std::min(py, y)]);
    } else {
      uint32_t lo = std::min(pl, cl), hi = std::max(pl, cl);
      if (lee || pr != r || hi != lo + 1 || !g.va[lo] || g.vianode(a) || g.vianode(c)) return false;
      if ((g.viaforb(a) && !sorted_has(p.vpads, a)) || (g.viaforb(c) && !sorted_has(p.vpads, c))) return false;
      e = g.ve[lo];
    }
    uint32_t wgt = (g.obst(c) && sorted_has(p.pads, c)) ? 1u : g.cost[c];
    uint64_t add = 0;
    if (!p.ovr_cells.empty() && std::binary_search(p.ovr_cells.begin(), p.ovr_cells.end(), c))
      add += p.ovr_add[std::lower_bound(p.ovr_cells.begin(), p.ovr_cells.end(), c) - p.ovr_cells.begin()];
    else if (g.add)
      add += g.add[c];
    if (g.halo && g.halo[c] != -1 && g.halo[c] != own_id) add += g.halo_pen;
    cost += step_price(e, wgt, add);
    if (cost >= kMax) return false;
  }
  return cost == search_cost;
}


inline HostGrid host_grid(mrf_dims d, const mrf_grid* g, const mrf_edges* e) {
  HostGrid h;
  h.d = d;
  h.cost = g->cost;
  h.vf = g->n_via_forbidden ? g->via_forbidden : nullptr;
  h.board = g->n_board_constraint ? g->board_constraint : nullptr;
  if (e) {
    h.xe = e->x;
    h.ye = e->y;
    h.ve = e->via;
    h.va = e->via_allowed;
  }
  return h;
}

struct Batch {
  std::vector<uint8_t> routed;
  std::vector<uint32_t> cost;
  std::vector<std::vector<uint32_t>> paths;
};

inline void export_paths(const Batch& b, mrf_paths* out) {
  size_t n = b.routed.size();
  size_t total = 0;
  for (auto& p : b.paths) total += p.size();
  out->n = n;
  out->routed = (uint8_t*)malloc(std::max<size_t>(n, 1));
  out->search_cost = (uint32_t*)malloc(std::max<size_t>(n, 1) * 4);
  out->offsets = (uint64_t*)malloc((n + 1) * 8);
  out->cells = (uint32_t*)malloc(std::max<size_t>(total, 1) * 4);
  if (!out->routed || !out->search_cost || !out->offsets || !out->cells) throw std::bad_alloc();
  out->offsets[0] = 0;
  for (size_t i = 0; i < n; ++i) {
    out->routed[i] = b.routed[i];
    out->search_cost[i] = b.cost[i];
    memcpy(out->cells + out->offsets[i], b.paths[i].data(), b.paths[i].size() * 4);
    out->offsets[i + 1] = out->offsets[i] + b.paths[i].size();
  }
}

}  // namespace host
}  // namespace mrf
