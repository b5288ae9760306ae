// Seeded synthetic workloads for the benchmark configurations.
//
// Random numbers come from xorshift64 (x ^= x << 13; x ^= x >> 7; x ^= x << 17),
// the generator of the reference's stress tests (mr-metal/src/lib.rs:1655-1661).
//
// config 0  reference bench_board (mr-metal/src/lib.rs:2761-2800): 128x128x1,
//           side/8 single-column walls leaving row 0 open, unit cost, 64 nets,
//           endpoints by xorshift from 0x9E3779B97F4A7C15 on free cells.
// config 1  128x128x2, 64 two-pin nets.  Rectangles with sides uniform in
//           [2,12] on a uniform layer until >= 15% of cells are blocked.  Planar
//           gap price 10, cell weight 1, via price 50.  Endpoints uniform on
//           free cells; unreachable pairs are kept.  Full-board windows.
// config 2  512x512x4, 256 two-pin nets.  Additive history per cell
//           round(logU[1,64]); gap price 10, weight 1, via 100.  Rectangles
//           (sides [2,12]) inflated by a 1-cell clearance halo until >= 20%
//           of cells are blocked.  Board outline: convex 12-gon (sorted uniform
//           angles, radii uniform in [0.35,0.5]*side about the centre).  Cells
//           whose centre lies outside or within 0.5 of the outline are
//           BOARD_TRACE_NODE; within 1.5 are BOARD_VIA_NODE; a planar edge
//           whose midpoint lies within 0.75 of the outline is forbidden through
//           one directed bit (source side on even x+y, destination side on odd),
//           exercising both halves of is_board_planar_step_forbidden.
//           Full-board windows (complete fields).
// config 3  2048x2048x2, 1024 two-pin nets, each in a square window whose side
//           is round(logU[16,512]), placed uniformly; endpoints uniform on free
//           cells of the window.  History uniform in [1,32], gap 10, via 80,
//           15% rectangle obstacles.
// config 4  1024x1024x6, 512 multi-pin nets with k ~ U[2,12] pins (sources).
//           History round(logU[1,128]); +40 on cells 4-adjacent (same layer) to
//           a pin of another net (halo-owner plane); gap 10, via 120, 25%
//           rectangle obstacles.  Fields only (no target).
// serving   mrf_synth_board: config-2 generator at a given side / layers /
//           net count (config 5 of the benchmark plan).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/mrfield.h"

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed ? seed : 0x9E3779B97F4A7C15ull) {}
  uint64_t next() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
  }
  uint32_t below(uint32_t n) { return (uint32_t)(next() % n); }
  double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  uint32_t range(uint32_t lo, uint32_t hi) { return lo + below(hi - lo + 1); }
  double logu(double lo, double hi) { return std::exp(std::log(lo) + unit() * (std::log(hi) - std::log(lo))); }
};

struct Gen {
  uint32_t w = 0, h = 0, L = 0;
  std::vector<uint32_t> cost;
  std::vector<uint8_t> board;
  std::vector<uint32_t> add;
  std::vector<int32_t> halo;
  uint32_t pen = 0;
  uint32_t planar = 10, via = 50;
  std::vector<uint32_t> src, dst;
  std::vector<uint64_t> pin_off;
  std::vector<uint32_t> pins;
  std::vector<mrf_window> windows;
  size_t n() const { return (size_t)w * h * L; }
  uint32_t idx(uint32_t x, uint32_t y, uint32_t l) const { return (l * h + y) * w + x; }
};

void rect_obstacles(Gen& g, Rng& r, double frac, uint32_t halo) {
  size_t target = (size_t)std::ceil(frac * (double)g.n());
  size_t blocked = 0;
  while (blocked < target) {
    uint32_t sw = r.range(2, 12), sh = r.range(2, 12);
    uint32_t x0 = r.below(g.w), y0 = r.below(g.h), l = r.below(g.L);
    uint32_t xa = x0 >= halo ? x0 - halo : 0, ya = y0 >= halo ? y0 - halo : 0;
    uint32_t xb = std::min(g.w - 1, x0 + sw - 1 + halo), yb = std::min(g.h - 1, y0 + sh - 1 + halo);
    for (uint32_t y = ya; y <= yb; ++y)
      for (uint32_t x = xa; x <= xb; ++x) {
        uint32_t i = g.idx(x, y, l);
        if (g.cost[i] != 0xFFFFFFFFu) {
          g.cost[i] = 0xFFFFFFFFu;
          ++blocked;
        }
      }
  }
}

bool free_cell(const Gen& g, uint32_t c) {
  return g.cost[c] != 0xFFFFFFFFu && (g.board.empty() || !(g.board[c] & MRF_BOARD_TRACE_NODE));
}

uint32_t pick_free(const Gen& g, Rng& r, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1) {
  for (int tries = 0; tries < 1 << 20; ++tries) {
    uint32_t x = x0 + r.below(x1 - x0 + 1), y = y0 + r.below(y1 - y0 + 1), l = r.below(g.L);
    uint32_t c = g.idx(x, y, l);
    if (free_cell(g, c)) return c;
  }
  return g.idx(x0, y0, 0);
}

// Distance from point p to segment ab.
double seg_dist(double px, double py, double ax, double ay, double bx, double by) {
  double vx = bx - ax, vy = by - ay, wx = px - ax, wy = py - ay;
  double t = (vx * wx + vy * wy) / (vx * vx + vy * vy);
  t = std::max(0.0, std::min(1.0, t));
  double dx = ax + t * vx - px, dy = ay + t * vy - py;
  return std::sqrt(dx * dx + dy * dy);
}

void outline(Gen& g, Rng& r) {
  const int K = 12;
  double cx = g.w / 2.0, cy = g.h / 2.0, side = std::min(g.w, g.h);
  std::vector<double> ang(K);
  for (int k = 0; k < K; ++k) ang[k] = r.unit() * 2 * M_PI;
  std::sort(ang.begin(), ang.end());
  std::vector<double> px(K), py(K);
  for (int k = 0; k < K; ++k) {
    double rad = (0.35 + 0.15 * r.unit()) * side;
    px[k] = cx + rad * std::cos(ang[k]);
    py[k] = cy + rad * std::sin(ang[k]);
  }
  // convex hull (gift wrapping is unnecessary: sorted angles around the centre
  // give a star polygon; take its hull so the outline is convex)
  std::vector<int> ids(K);
  for (int k = 0; k < K; ++k) ids[k] = k;
  std::sort(ids.begin(), ids.end(), [&](int a, int b) { return px[a] < px[b] || (px[a] == px[b] && py[a] < py[b]); });
  auto cross = [&](int o, int a, int b) { return (px[a] - px[o]) * (py[b] - py[o]) - (py[a] - py[o]) * (px[b] - px[o]); };
  std::vector<int> hull(2 * K);
  int m = 0;
  for (int i = 0; i < K; ++i) {
    while (m >= 2 && cross(hull[m - 2], hull[m - 1], ids[i]) <= 0) --m;
    hull[m++] = ids[i];
  }
  for (int i = K - 2, t = m + 1; i >= 0; --i) {
    while (m >= t && cross(hull[m - 2], hull[m - 1], ids[i]) <= 0) --m;
    hull[m++] = ids[i];
  }
  hull.resize(m - 1);
  int H = (int)hull.size();
  auto inside = [&](double x, double y) {
    for (int k = 0; k < H; ++k) {
      int a = hull[k], b = hull[(k + 1) % H];
      if ((px[b] - px[a]) * (y - py[a]) - (py[b] - py[a]) * (x - px[a]) < 0) return false;
    }
    return true;
  };
  auto bdist = [&](double x, double y) {
    double d = 1e30;
    for (int k = 0; k < H; ++k) {
      int a = hull[k], b = hull[(k + 1) % H];
      d = std::min(d, seg_dist(x, y, px[a], py[a], px[b], py[b]));
    }
    return d;
  };
  g.board.assign(g.n(), 0);
  std::vector<uint8_t> plane((size_t)g.w * g.h, 0);
  for (uint32_t y = 0; y < g.h; ++y)
    for (uint32_t x = 0; x < g.w; ++x) {
      double fx = x + 0.5, fy = y + 0.5;
      uint8_t b = 0;
      bool in = inside(fx, fy);
      double d = in ? bdist(fx, fy) : 0.0;
      if (!in || d < 0.5) b |= MRF_BOARD_TRACE_NODE;
      if (!in || d < 1.5) b |= MRF_BOARD_VIA_NODE;
      if (!(b & MRF_BOARD_TRACE_NODE)) {
        if (x + 1 < g.w) {
          double mx = fx + 0.5;
          if (inside(mx, fy) && bdist(mx, fy) < 0.75) {
            if (((x + y) & 1) == 0)
              b |= MRF_BOARD_EDGE_POS_X;
            else
              plane[(size_t)y * g.w + x + 1] |= MRF_BOARD_EDGE_NEG_X;
          }
        }
        if (y + 1 < g.h) {
          double my = fy + 0.5;
          if (inside(fx, my) && bdist(fx, my) < 0.75) {
            if (((x + y) & 1) == 0)
              b |= MRF_BOARD_EDGE_POS_Y;
            else
              plane[(size_t)(y + 1) * g.w + x] |= MRF_BOARD_EDGE_NEG_Y;
          }
        }
      }
      plane[(size_t)y * g.w + x] |= b;
    }
  for (uint32_t l = 0; l < g.L; ++l) memcpy(g.board.data() + (size_t)l * g.w * g.h, plane.data(), plane.size());
}

void base(Gen& g, uint32_t w, uint32_t h, uint32_t L) {
  g.w = w;
  g.h = h;
  g.L = L;
  g.cost.assign(g.n(), 1u);
}

void two_pin(Gen& g, Rng& r, size_t nets) {
  for (size_t i = 0; i < nets; ++i) {
    g.src.push_back(pick_free(g, r, 0, 0, g.w - 1, g.h - 1));
    g.dst.push_back(pick_free(g, r, 0, 0, g.w - 1, g.h - 1));
  }
}

void board_like_c2(Gen& g, Rng& r, uint32_t side, uint32_t L, size_t nets) {
  base(g, side, side, L);
  g.planar = 10;
  g.via = 100;
  g.add.resize(g.n());
  for (auto& a : g.add) a = (uint32_t)std::lround(r.logu(1.0, 64.0));
  rect_obstacles(g, r, 0.20, 1);
  outline(g, r);
  two_pin(g, r, nets);
}

void export_gen(Gen& g, mrf_synth* o) {
  memset(o, 0, sizeof(*o));
  size_t n = g.n();
  o->dims = {g.w, g.h, g.L};
  o->cost = (uint32_t*)malloc(n * 4);
  memcpy(o->cost, g.cost.data(), n * 4);
  if (!g.board.empty()) {
    o->board_constraint = (uint8_t*)malloc(n);
    memcpy(o->board_constraint, g.board.data(), n);
  }
  if (!g.add.empty()) {
    o->additive = (uint32_t*)malloc(n * 4);
    memcpy(o->additive, g.add.data(), n * 4);
  }
  if (!g.halo.empty()) {
    o->halo_owner = (int32_t*)malloc(n * 4);
    memcpy(o->halo_owner, g.halo.data(), n * 4);
  }
  o->halo_penalty = g.pen;
  o->x_edges = (uint32_t*)malloc(std::max<size_t>(g.w, 1) * 4);
  o->y_edges = (uint32_t*)malloc(std::max<size_t>(g.h, 1) * 4);
  o->via_edges = (uint32_t*)malloc(std::max<size_t>(g.L, 1) * 4);
  o->via_allowed = (uint8_t*)malloc(std::max<size_t>(g.L, 1));
  for (uint32_t k = 0; k < g.w; ++k) o->x_edges[k] = g.planar;
  for (uint32_t k = 0; k < g.h; ++k) o->y_edges[k] = g.planar;
  for (uint32_t k = 0; k < g.L; ++k) {
    o->via_edges[k] = g.via;
    o->via_allowed[k] = 1;
  }
  size_t nn = !g.pin_off.empty() ? g.pin_off.size() - 1 : g.src.size();
  o->n_nets = nn;
  o->src = (uint32_t*)malloc(std::max<size_t>(nn, 1) * 4);
  o->dst = (uint32_t*)malloc(std::max<size_t>(nn, 1) * 4);
  for (size_t i = 0; i < nn; ++i) {
    o->src[i] = g.src.size() > i ? g.src[i] : g.pins[g.pin_off[i]];
    o->dst[i] = g.dst.size() > i ? g.dst[i] : 0xFFFFFFFFu;
  }
  if (!g.pin_off.empty()) {
    o->pin_off = (uint64_t*)malloc(g.pin_off.size() * 8);
    memcpy(o->pin_off, g.pin_off.data(), g.pin_off.size() * 8);
    o->pins = (uint32_t*)malloc(std::max<size_t>(g.pins.size(), 1) * 4);
    memcpy(o->pins, g.pins.data(), g.pins.size() * 4);
  }
  if (!g.windows.empty()) {
    o->windows = (mrf_window*)malloc(g.windows.size() * sizeof(mrf_window));
    memcpy(o->windows, g.windows.data(), g.windows.size() * sizeof(mrf_window));
  }
}

}  // namespace

extern "C" {

int mrf_synth_config(int config, uint64_t seed, uint32_t scale_nets, mrf_synth* out, mrf_error* err) {
  if (!out) return MRF_BACKEND_UNAVAILABLE;
  Gen g;
  Rng r(seed);
  auto nets_or = [&](size_t d) { return scale_nets ? (size_t)scale_nets : d; };
  switch (config) {
    case 0: {
      // mr-metal/src/lib.rs:2761-2800
      uint32_t side = 128;
      base(g, side, side, 1);
      g.planar = 1;
      g.via = 0;
      uint32_t y_hi = std::max<uint32_t>(side / 2, 1) - 1;
      for (uint32_t k = 0; k < side / 8; ++k) {
        uint32_t x = std::min(2 + k * 5, side - 2);
        for (uint32_t y = 1; y <= y_hi; ++y) g.cost[g.idx(x, y, 0)] = 0xFFFFFFFFu;
      }
      Rng e(0x9E3779B97F4A7C15ull);
      size_t nets = nets_or(64);
      auto pick = [&]() {
        for (;;) {
          uint32_t i = (uint32_t)(e.next() % g.n());
          if (g.cost[i] != 0xFFFFFFFFu) return i;
        }
      };
      for (size_t k = 0; k < nets; ++k) {
        uint32_t s = pick();
        uint32_t d = pick();
        g.src.push_back(s);
        g.dst.push_back(d);
      }
      break;
    }
    case 1:
      base(g, 128, 128, 2);
      g.planar = 10;
      g.via = 50;
      rect_obstacles(g, r, 0.15, 0);
      two_pin(g, r, nets_or(64));
      break;
    case 2:
      board_like_c2(g, r, 512, 4, nets_or(256));
      break;
    case 3: {
      base(g, 2048, 2048, 2);
      g.planar = 10;
      g.via = 80;
      g.add.resize(g.n());
      for (auto& a : g.add) a = r.range(1, 32);
      rect_obstacles(g, r, 0.15, 0);
      size_t nets = nets_or(1024);
      for (size_t i = 0; i < nets; ++i) {
        uint32_t s = (uint32_t)std::lround(r.logu(16.0, 512.0));
        s = std::min(s, g.w);
        uint32_t x0 = r.below(g.w - s + 1), y0 = r.below(g.h - s + 1);
        mrf_window win{x0, y0, x0 + s - 1, y0 + s - 1};
        g.windows.push_back(win);
        g.src.push_back(pick_free(g, r, win.x0, win.y0, win.x1, win.y1));
        g.dst.push_back(pick_free(g, r, win.x0, win.y0, win.x1, win.y1));
      }
      break;
    }
    case 4: {
      base(g, 1024, 1024, 6);
      g.planar = 10;
      g.via = 120;
      g.add.resize(g.n());
      for (auto& a : g.add) a = (uint32_t)std::lround(r.logu(1.0, 128.0));
      rect_obstacles(g, r, 0.25, 0);
      size_t nets = nets_or(512);
      g.pin_off.push_back(0);
      g.halo.assign(g.n(), -1);
      g.pen = 40;
      for (size_t i = 0; i < nets; ++i) {
        uint32_t k = r.range(2, 12);
        for (uint32_t j = 0; j < k; ++j) g.pins.push_back(pick_free(g, r, 0, 0, g.w - 1, g.h - 1));
        g.pin_off.push_back(g.pins.size());
      }
      uint32_t plane = g.w * g.h;
      for (size_t i = 0; i < nets; ++i)
        for (uint64_t j = g.pin_off[i]; j < g.pin_off[i + 1]; ++j) {
          uint32_t p = g.pins[j], rr = p % plane, x = rr % g.w, y = rr / g.w;
          uint32_t nb[4];
          int m = 0;
          if (y > 0) nb[m++] = p - g.w;
          if (x > 0) nb[m++] = p - 1;
          if (x + 1 < g.w) nb[m++] = p + 1;
          if (y + 1 < g.h) nb[m++] = p + g.w;
          for (int q = 0; q < m; ++q) {
            int32_t& o = g.halo[nb[q]];
            if (o == -1)
              o = (int32_t)i;
            else if (o != (int32_t)i)
              o = -2;
          }
        }
      break;
    }
    default:
      if (err) {
        err->code = MRF_BACKEND_UNAVAILABLE;
        snprintf(err->message, sizeof(err->message), "unknown synthetic config %d", config);
      }
      return MRF_BACKEND_UNAVAILABLE;
  }
  export_gen(g, out);
  return MRF_OK;
}

// Config-2 generator at an arbitrary size (serving requests).
int mrf_synth_board(uint32_t side, uint32_t layers, uint32_t nets, uint64_t seed, mrf_synth* out) {
  if (!out || side < 8 || layers < 1 || layers > 16) return MRF_BACKEND_UNAVAILABLE;
  Gen g;
  Rng r(seed);
  board_like_c2(g, r, side, layers, nets);
  export_gen(g, out);
  return MRF_OK;
}

void mrf_free_synth(mrf_synth* s) {
  if (!s) return;
  free(s->cost);
  free(s->board_constraint);
  free(s->additive);
  free(s->halo_owner);
  free(s->x_edges);
  free(s->y_edges);
  free(s->via_edges);
  free(s->via_allowed);
  free(s->src);
  free(s->dst);
  free(s->pin_off);
  free(s->pins);
  free(s->windows);
  memset(s, 0, sizeof(*s));
}

}  // extern "C"
