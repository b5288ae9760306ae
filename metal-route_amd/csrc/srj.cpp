// SimpleRouteJson problems rasterised into the engine's inputs (host code).
//
// A tscircuit SimpleRouteJson problem is continuous geometry: bounds, rect
// obstacles (pads) on named layers, and connections of k points.  The engine
// routes on a grid, so a board is rasterised once:
//
//   1. Grid lines per axis (a Hanan grid): the bounds, every connection point
//      and every obstacle edge are feature lines; every gap between two feature
//      lines wide enough for a centred, clearance-legal track gets fill lanes at
//      about the track pitch; a planar cell budget thins the fill lanes (the
//      most crowded first) and, when the true clearance is finer than the one
//      the raster inflates by, sub-pitch escape lanes go into tight pad gaps
//      within what the budget leaves.  (crates/mr-srj/src/lib.rs:848-1273)
//   2. Cells: node (x, y) sits on lines x and y; a continuous point belongs to
//      the node whose half-open Voronoi interval holds it (ties to the lower
//      line).  (:1275-1301, :2098-2120)
//   3. Obstacles: each rect blocks the cells its box covers on its layers, then
//      every blocked cell blocks every cell within a geometric Chebyshev
//      distance (line positions, per layer) of the foreign-copper margin.
//      (:1438-1763; crates/mr-grid/src/lib.rs:92-229)
//   4. Via keepout: with the exact clearance model on a multi-layer board, a
//      node whose Euclidean distance to an obstacle is under the via-pad radius
//      plus the via clearance may not host a via.  (:1698-1737, :2033-2053)
//   5. Board outline: when the problem declares an outline or a board-edge
//      clearance, nodes outside the polygon or within the trace / via keepout
//      of its edges, and planar steps whose capsule comes too close, are masked.
//      (:1765-1796; crates/mr-srj/src/board_outline.rs:72-108, :223-461)
//   6. Nets: each k-point connection becomes k-1 chained two-point nets with
//      the cells of their own pads (plus the clearance halo around them, clipped
//      by foreign pads) and the own-pad cells where a via may override the via
//      keepout.  (:1842-2031, :2122-2216)
//
// The CLI's rule selection (the typed uniform physical rules when they are
// coherent, else the legacy clearance rounding) is in `route_rules`
// (crates/mr-cli/src/lib.rs:751-834; uniform_physical_rules, mr-srj lib.rs:286-528).
//
// Floating-point steps are evaluated in the same order as the reference (the
// midpoints 0.5 * (a + b), the fill step gap / intervals, the distances), so
// line positions and every inclusion test agree bit for bit.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <utility>
#include <vector>

#include "../../include/mrfield.h"

namespace {

using std::string;
using std::vector;

constexpr double kLineEps = 1e-6;      // lines closer than this are one line (lib.rs:661)
constexpr double kDrcEps = 1e-9;       // DRC comparison tolerance (lib.rs:664)
constexpr double kGeomEps = 1e-9;      // outline geometry tolerance (board_outline.rs:14)
constexpr size_t kBudgetFloor = 160000; // default planar cell budget (lib.rs:728)
constexpr size_t kMaxFillPerGap = 256;  // (lib.rs:1196)
constexpr double kPadBandK = 0.75;      // foreign-pad band factor (lib.rs:805)
constexpr double kDefaultEdgeClearance = 0.2;  // board_outline.rs:18
constexpr double kCliTraceWidth = 0.15;        // mr-cli DEFAULT_TRACE_WIDTH (mr-cli lib.rs:47)
constexpr double kCliClearance = 0.15;         // DEFAULT_CLEARANCE_MM (:52)
constexpr double kCliViaPad = 0.45;            // VIA_PAD_MM (:56)
constexpr double kMinAnnularRing = 0.05;       // lib.rs:499
constexpr uint32_t kObstacle = 0xFFFFFFFFu;

struct Fail : std::runtime_error {
  explicit Fail(const string& m) : std::runtime_error(m) {}
};

// ---------------------------------------------------------------------------
// JSON (RFC 8259) into a small tree.  Numbers go through strtod, which rounds
// correctly, as serde_json does.
// ---------------------------------------------------------------------------
struct Json {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  double num = 0;
  string str;
  vector<Json> arr;
  vector<std::pair<string, Json>> obj;

  const Json* get(const char* key) const {
    if (kind != OBJ) return nullptr;
    const Json* hit = nullptr;
    for (const auto& kv : obj)
      if (kv.first == key) hit = &kv.second;  // the last duplicate wins
    return hit;
  }
};

class Parser {
 public:
  Parser(const char* s, size_t n) : p_(s), e_(s + n) {}
  Json document() {
    Json v = value(0);
    ws();
    if (p_ != e_) fail("trailing characters after the JSON value");
    return v;
  }

 private:
  const char* p_;
  const char* e_;
  [[noreturn]] void fail(const char* what) { throw Fail(string("malformed JSON: ") + what); }
  void ws() {
    while (p_ < e_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
  }
  bool lit(const char* w) {
    const size_t n = strlen(w);
    if ((size_t)(e_ - p_) >= n && memcmp(p_, w, n) == 0) {
      p_ += n;
      return true;
    }
    return false;
  }
  Json value(int depth) {
    if (depth > 256) fail("nesting too deep");
    ws();
    if (p_ >= e_) fail("unexpected end");
    Json v;
    const char c = *p_;
    if (c == '{') {
      v.kind = Json::OBJ;
      ++p_;
      ws();
      if (p_ < e_ && *p_ == '}') {
        ++p_;
        return v;
      }
      for (;;) {
        ws();
        if (p_ >= e_ || *p_ != '"') fail("object key expected");
        string k = text();
        ws();
        if (p_ >= e_ || *p_ != ':') fail("':' expected");
        ++p_;
        Json x = value(depth + 1);
        v.obj.emplace_back(std::move(k), std::move(x));
        ws();
        if (p_ < e_ && *p_ == ',') {
          ++p_;
          continue;
        }
        if (p_ < e_ && *p_ == '}') {
          ++p_;
          return v;
        }
        fail("',' or '}' expected");
      }
    }
    if (c == '[') {
      v.kind = Json::ARR;
      ++p_;
      ws();
      if (p_ < e_ && *p_ == ']') {
        ++p_;
        return v;
      }
      for (;;) {
        v.arr.push_back(value(depth + 1));
        ws();
        if (p_ < e_ && *p_ == ',') {
          ++p_;
          continue;
        }
        if (p_ < e_ && *p_ == ']') {
          ++p_;
          return v;
        }
        fail("',' or ']' expected");
      }
    }
    if (c == '"') {
      v.kind = Json::STR;
      v.str = text();
      return v;
    }
    if (lit("true")) {
      v.kind = Json::BOOL;
      v.b = true;
      return v;
    }
    if (lit("false")) {
      v.kind = Json::BOOL;
      return v;
    }
    if (lit("null")) return v;
    return number();
  }
  Json number() {
    const char* s = p_;
    if (p_ < e_ && *p_ == '-') ++p_;
    if (p_ >= e_ || !isdigit((unsigned char)*p_)) fail("number expected");
    if (*p_ == '0')
      ++p_;
    else
      while (p_ < e_ && isdigit((unsigned char)*p_)) ++p_;
    if (p_ < e_ && *p_ == '.') {
      ++p_;
      if (p_ >= e_ || !isdigit((unsigned char)*p_)) fail("digit expected after '.'");
      while (p_ < e_ && isdigit((unsigned char)*p_)) ++p_;
    }
    if (p_ < e_ && (*p_ == 'e' || *p_ == 'E')) {
      ++p_;
      if (p_ < e_ && (*p_ == '+' || *p_ == '-')) ++p_;
      if (p_ >= e_ || !isdigit((unsigned char)*p_)) fail("exponent digits expected");
      while (p_ < e_ && isdigit((unsigned char)*p_)) ++p_;
    }
    Json v;
    v.kind = Json::NUM;
    v.num = strtod(string(s, p_).c_str(), nullptr);
    return v;
  }
  static void utf8(string& o, uint32_t cp) {
    if (cp < 0x80) {
      o += (char)cp;
    } else if (cp < 0x800) {
      o += (char)(0xC0 | (cp >> 6));
      o += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      o += (char)(0xE0 | (cp >> 12));
      o += (char)(0x80 | ((cp >> 6) & 0x3F));
      o += (char)(0x80 | (cp & 0x3F));
    } else {
      o += (char)(0xF0 | (cp >> 18));
      o += (char)(0x80 | ((cp >> 12) & 0x3F));
      o += (char)(0x80 | ((cp >> 6) & 0x3F));
      o += (char)(0x80 | (cp & 0x3F));
    }
  }
  uint32_t hex4() {
    if (e_ - p_ < 4) fail("short \\u escape");
    uint32_t v = 0;
    for (int i = 0; i < 4; ++i) {
      const char c = *p_++;
      v <<= 4;
      if (c >= '0' && c <= '9')
        v |= (uint32_t)(c - '0');
      else if (c >= 'a' && c <= 'f')
        v |= (uint32_t)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F')
        v |= (uint32_t)(c - 'A' + 10);
      else
        fail("bad \\u escape");
    }
    return v;
  }
  string text() {
    ++p_;  // opening quote
    string o;
    for (;;) {
      if (p_ >= e_) fail("unterminated string");
      const char c = *p_++;
      if (c == '"') return o;
      if ((unsigned char)c < 0x20) fail("control character in string");
      if (c != '\\') {
        o += c;
        continue;
      }
      if (p_ >= e_) fail("unterminated escape");
      const char x = *p_++;
      switch (x) {
        case '"': o += '"'; break;
        case '\\': o += '\\'; break;
        case '/': o += '/'; break;
        case 'b': o += '\b'; break;
        case 'f': o += '\f'; break;
        case 'n': o += '\n'; break;
        case 'r': o += '\r'; break;
        case 't': o += '\t'; break;
        case 'u': {
          uint32_t cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00) {
            if (e_ - p_ < 6 || p_[0] != '\\' || p_[1] != 'u') fail("lone surrogate");
            p_ += 2;
            const uint32_t lo = hex4();
            if (lo < 0xDC00 || lo >= 0xE000) fail("bad surrogate pair");
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          } else if (cp >= 0xDC00 && cp < 0xE000) {
            fail("lone surrogate");
          }
          utf8(o, cp);
          break;
        }
        default: fail("bad escape");
      }
    }
  }
};

// ---------------------------------------------------------------------------
// The SimpleRouteJson model (crates/mr-srj/src/lib.rs:66-250): the fields the
// raster and the typed rules read; unknown fields are ignored.
// ---------------------------------------------------------------------------
using Opt = std::pair<bool, double>;  // (present, value)
struct Pt {
  double x = 0, y = 0;
  bool has_layer = false;
  string layer;
};
struct Obs {
  string kind;
  Pt c;
  double w = 0, h = 0;
  bool has_shape = false;
  string shape;
  Opt rot{false, 0};
  vector<string> layers, connected;
};
struct Conn {
  string name;
  bool has_root = false;
  string root;
  Opt nominal{false, 0}, width{false, 0};
  vector<Pt> pts;
};
struct Srj {
  uint32_t layer_count = 1;
  Opt min_trace{false, 0}, min_clearance{false, 0};
  // typed physical rules (lib.rs:166-219)
  Opt nominal_trace, default_margin, trace_to_pad, via_to_pad, hole_to_hole, pad_to_pad, board_edge;
  Opt hole_dia, hole_dia_snake, pad_dia, pad_dia_snake, via_dia;
  vector<std::pair<double, double>> outline;
  vector<Obs> obstacles;
  vector<Conn> conns;
  double min_x = 0, max_x = 0, min_y = 0, max_y = 0;
};

double need_num(const Json& v, const char* what) {
  if (v.kind != Json::NUM) throw Fail(string("SimpleRouteJson: '") + what + "' must be a number");
  return v.num;
}
Opt opt_num(const Json& o, const char* key) {
  const Json* v = o.get(key);
  if (!v || v->kind == Json::NUL) return {false, 0};
  return {true, need_num(*v, key)};
}
string need_str(const Json& v, const char* what) {
  if (v.kind != Json::STR) throw Fail(string("SimpleRouteJson: '") + what + "' must be a string");
  return v.str;
}
vector<string> str_list(const Json& o, const char* key) {
  vector<string> out;
  const Json* v = o.get(key);
  if (!v || v->kind == Json::NUL) return out;
  if (v->kind != Json::ARR) throw Fail(string("SimpleRouteJson: '") + key + "' must be an array");
  for (const Json& s : v->arr) out.push_back(need_str(s, key));
  return out;
}
Pt point(const Json& v, const char* what) {
  if (v.kind != Json::OBJ) throw Fail(string("SimpleRouteJson: '") + what + "' must be an object");
  Pt p;
  const Json* x = v.get("x");
  const Json* y = v.get("y");
  if (!x || !y) throw Fail(string("SimpleRouteJson: '") + what + "' needs x and y");
  p.x = need_num(*x, "x");
  p.y = need_num(*y, "y");
  const Json* l = v.get("layer");
  if (l && l->kind != Json::NUL) {
    p.has_layer = true;
    p.layer = need_str(*l, "layer");
  }
  return p;
}

Srj parse_srj(const char* text, size_t len) {
  const Json doc = Parser(text, len).document();
  if (doc.kind != Json::OBJ) throw Fail("SimpleRouteJson: the document must be an object");
  Srj s;
  const Json* lc = doc.get("layerCount");
  if (!lc) throw Fail("SimpleRouteJson: missing 'layerCount'");
  const double l = need_num(*lc, "layerCount");
  if (!(l >= 0) || l != std::floor(l) || l > 4294967295.0) throw Fail("SimpleRouteJson: bad 'layerCount'");
  s.layer_count = (uint32_t)l;
  s.min_trace = opt_num(doc, "minTraceWidth");
  s.min_clearance = opt_num(doc, "minClearance");
  s.nominal_trace = opt_num(doc, "nominalTraceWidth");
  s.default_margin = opt_num(doc, "defaultObstacleMargin");
  s.trace_to_pad = opt_num(doc, "minTraceToPadEdgeClearance");
  s.via_to_pad = opt_num(doc, "minViaEdgeToPadEdgeClearance");
  s.hole_to_hole = opt_num(doc, "minViaHoleEdgeToViaHoleEdgeClearance");
  s.pad_to_pad = opt_num(doc, "minPadEdgeToPadEdgeClearance");
  s.board_edge = opt_num(doc, "minBoardEdgeClearance");
  s.hole_dia = opt_num(doc, "minViaHoleDiameter");
  s.hole_dia_snake = opt_num(doc, "min_via_hole_diameter");
  s.pad_dia = opt_num(doc, "minViaPadDiameter");
  s.pad_dia_snake = opt_num(doc, "min_via_pad_diameter");
  s.via_dia = opt_num(doc, "minViaDiameter");
  if (const Json* o = doc.get("outline"); o && o->kind != Json::NUL) {
    if (o->kind != Json::ARR) throw Fail("SimpleRouteJson: 'outline' must be an array");
    for (const Json& v : o->arr) {
      const Pt p = point(v, "outline");
      s.outline.emplace_back(p.x, p.y);
    }
  }
  const Json* b = doc.get("bounds");
  if (!b || b->kind != Json::OBJ) throw Fail("SimpleRouteJson: missing 'bounds'");
  const char* bk[4] = {"minX", "maxX", "minY", "maxY"};
  double* bv[4] = {&s.min_x, &s.max_x, &s.min_y, &s.max_y};
  for (int i = 0; i < 4; ++i) {
    const Json* v = b->get(bk[i]);
    if (!v) throw Fail(string("SimpleRouteJson: bounds need '") + bk[i] + "'");
    *bv[i] = need_num(*v, bk[i]);
  }
  if (const Json* os = doc.get("obstacles"); os && os->kind != Json::NUL) {
    if (os->kind != Json::ARR) throw Fail("SimpleRouteJson: 'obstacles' must be an array");
    for (const Json& o : os->arr) {
      if (o.kind != Json::OBJ) throw Fail("SimpleRouteJson: an obstacle must be an object");
      Obs ob;
      const Json* t = o.get("type");
      const Json* c = o.get("center");
      const Json* w = o.get("width");
      const Json* h = o.get("height");
      if (!t || !c || !w || !h) throw Fail("SimpleRouteJson: an obstacle needs type, center, width, height");
      ob.kind = need_str(*t, "type");
      ob.c = point(*c, "center");
      ob.w = need_num(*w, "width");
      ob.h = need_num(*h, "height");
      if (const Json* sh = o.get("shape"); sh && sh->kind != Json::NUL) {
        ob.has_shape = true;
        ob.shape = need_str(*sh, "shape");
      }
      ob.rot = opt_num(o, "ccwRotationDegrees");
      ob.layers = str_list(o, "layers");
      ob.connected = str_list(o, "connectedTo");
      s.obstacles.push_back(std::move(ob));
    }
  }
  if (const Json* cs = doc.get("connections"); cs && cs->kind != Json::NUL) {
    if (cs->kind != Json::ARR) throw Fail("SimpleRouteJson: 'connections' must be an array");
    for (const Json& c : cs->arr) {
      if (c.kind != Json::OBJ) throw Fail("SimpleRouteJson: a connection must be an object");
      Conn cn;
      const Json* n = c.get("name");
      const Json* pts = c.get("pointsToConnect");
      if (!n || !pts) throw Fail("SimpleRouteJson: a connection needs name and pointsToConnect");
      cn.name = need_str(*n, "name");
      if (const Json* r = c.get("rootConnectionName"); r && r->kind != Json::NUL) {
        cn.has_root = true;
        cn.root = need_str(*r, "rootConnectionName");
      }
      cn.nominal = opt_num(c, "nominalTraceWidth");
      cn.width = opt_num(c, "width");
      if (pts->kind != Json::ARR) throw Fail("SimpleRouteJson: 'pointsToConnect' must be an array");
      for (const Json& p : pts->arr) cn.pts.push_back(point(p, "pointsToConnect"));
      s.conns.push_back(std::move(cn));
    }
  }
  return s;
}

// ---------------------------------------------------------------------------
// Layers (LayerMap::standard, crates/mr-core/src/lib.rs:259-295)
// ---------------------------------------------------------------------------
struct Layers {
  vector<string> names;
  explicit Layers(uint32_t count) {
    count = std::max<uint32_t>(count, 1);
    names.push_back("top");
    if (count == 1) return;
    for (uint32_t i = 1; i + 1 < count; ++i) names.push_back("inner" + std::to_string(i));
    names.push_back("bottom");
  }
  uint32_t size() const { return (uint32_t)names.size(); }
  int index_of(const string& n) const {
    for (size_t i = 0; i < names.size(); ++i)
      if (names[i] == n) return (int)i;
    return -1;
  }
};

// A point's layer: its named layer, else 0 (lib.rs:1798-1803).
uint32_t point_layer(const Pt& p, const Layers& L) {
  if (!p.has_layer) return 0;
  const int i = L.index_of(p.layer);
  return i < 0 ? 0u : (uint32_t)i;
}

// An obstacle's layers, ascending: the known named ones, or all when none is
// known (lib.rs:1808-1822).
vector<uint32_t> obstacle_layers(const Obs& o, const Layers& L) {
  vector<uint32_t> out;
  for (const string& n : o.layers) {
    const int i = L.index_of(n);
    if (i >= 0) out.push_back((uint32_t)i);
  }
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  if (out.empty())
    for (uint32_t i = 0; i < L.size(); ++i) out.push_back(i);
  return out;
}

// ---------------------------------------------------------------------------
// Grid lines (lib.rs:848-1273)
// ---------------------------------------------------------------------------

// Drop non-finite values, sort, and fold each run of values within kLineEps of
// the last kept value into that value (lib.rs:1188-1192).
void sort_dedup(vector<double>& v) {
  v.erase(std::remove_if(v.begin(), v.end(), [](double x) { return !std::isfinite(x); }), v.end());
  std::stable_sort(v.begin(), v.end());
  size_t k = 0;
  for (size_t i = 0; i < v.size(); ++i)
    if (k == 0 || !(std::fabs(v[i] - v[k - 1]) <= kLineEps)) v[k++] = v[i];
  v.resize(k);
}

// Index of the first element of sorted `a` not satisfying `pred`.
template <class P>
size_t first_not(const vector<double>& a, P pred) {
  return (size_t)(std::partition_point(a.begin(), a.end(), pred) - a.begin());
}

// Fill lanes between adjacent feature lines (lib.rs:1218-1273): a gap of at
// least track_w + 2 * clearance is cut into ceil(gap / pitch) equal intervals
// (at least 2, at most 257) and a lane goes on every interior boundary.
vector<double> fill_lines(const vector<double>& f, double track_w, double clearance) {
  vector<double> out;
  if (f.size() < 2 || !(std::isfinite(track_w) && track_w > 0)) return out;
  if (!(std::isfinite(clearance) && clearance > 0)) clearance = 0;
  const double channel = track_w + 2.0 * clearance;
  const double pitch = std::max(track_w, kLineEps);
  for (size_t i = 0; i + 1 < f.size(); ++i) {
    const double lo = f[i], gap = f[i + 1] - f[i];
    if (gap < channel) continue;
    const double q = std::ceil(gap / pitch);
    size_t n = q >= 1e18 ? (size_t)-1 : (size_t)q;
    n = std::min<size_t>(std::max<size_t>(n, 2), kMaxFillPerGap + 1);
    const double step = gap / (double)n;
    for (size_t k = 1; k < n; ++k) out.push_back(lo + step * (double)k);
  }
  return out;
}

// Sub-pitch escape lanes (lib.rs:950-979): one midpoint lane in each gap too
// tight for a regular lane but wide enough for a lane at the true clearance.
vector<double> escape_lines(const vector<double>& f, double track_w, double clearance, double esc) {
  vector<double> out;
  if (f.size() < 2 || !(std::isfinite(track_w) && track_w > 0)) return out;
  const double coarse = track_w + 2.0 * clearance;
  const double fine = track_w + 2.0 * esc;
  if (!(fine < coarse)) return out;
  for (size_t i = 0; i + 1 < f.size(); ++i) {
    const double lo = f[i], gap = f[i + 1] - f[i];
    if (gap + kLineEps >= coarse || gap + kLineEps < fine) continue;
    out.push_back(lo + gap * 0.5);
  }
  sort_dedup(out);
  return out;
}

// Remove fill lanes within kLineEps of a feature line (after sort_dedup of the
// fill itself) (lib.rs:1130-1137).
void coalesce_fill(const vector<double>& features, vector<double>& fill) {
  sort_dedup(fill);
  size_t k = 0;
  for (size_t i = 0; i < fill.size(); ++i) {
    const double f = fill[i];
    const size_t j = first_not(features, [&](double x) { return x < f - kLineEps; });
    if (!(j < features.size() && std::fabs(features[j] - f) <= kLineEps)) fill[k++] = f;
  }
  fill.resize(k);
}

// Remove the fill lane whose nearest neighbour among features + fill is
// closest; the first such lane on ties (lib.rs:1144-1186).
void drop_least_important(const vector<double>& features, vector<double>& fill) {
  if (fill.empty()) return;
  vector<double> all(features);
  all.insert(all.end(), fill.begin(), fill.end());
  std::stable_sort(all.begin(), all.end());
  const double inf = std::numeric_limits<double>::infinity();
  size_t best = 0;
  double best_gap = inf;
  for (size_t i = 0; i < fill.size(); ++i) {
    const double f = fill[i];
    const size_t pos = first_not(all, [&](double x) { return x < f; });
    const double left = pos > 0 ? f - all[pos - 1] : inf;
    size_t up = pos + 1;
    while (up < all.size() && std::fabs(all[up] - f) <= kLineEps) ++up;
    const double right = up < all.size() ? all[up] - f : inf;
    const double nn = std::fmin(left, right);
    if (nn < best_gap) {
      best_gap = nn;
      best = i;
    }
  }
  fill.erase(fill.begin() + (ptrdiff_t)best);
}

size_t sat_mul(size_t a, size_t b) {
  if (a != 0 && b > std::numeric_limits<size_t>::max() / a) return std::numeric_limits<size_t>::max();
  return a * b;
}

// f64 -> u32 the way Rust's `as` converts: NaN to 0, saturating at both ends.
uint32_t sat_u32(double v) {
  if (!(v > 0)) return 0;
  if (v >= 4294967295.0) return 0xFFFFFFFFu;
  return (uint32_t)v;
}

// Thin the fill lanes until features x fill fits the budget, the axis with more
// fill first; when the features alone exceed it every fill lane goes
// (lib.rs:1047-1128).  Returns the number of lanes removed after coalescing.
size_t enforce_budget(const vector<double>& xf, const vector<double>& yf, vector<double>& xfill,
                      vector<double>& yfill, size_t budget) {
  coalesce_fill(xf, xfill);
  coalesce_fill(yf, yfill);
  auto cells = [&](size_t a, size_t b) { return sat_mul(xf.size() + a, yf.size() + b); };
  if (cells(xfill.size(), yfill.size()) <= budget) return 0;
  if (sat_mul(xf.size(), yf.size()) > budget) {
    const size_t n = xfill.size() + yfill.size();
    xfill.clear();
    yfill.clear();
    return n;
  }
  size_t dropped = 0;
  while (cells(xfill.size(), yfill.size()) > budget) {
    bool thin_x;
    if (xfill.empty())
      thin_x = false;
    else if (yfill.empty())
      thin_x = true;
    else
      thin_x = xfill.size() >= yfill.size();
    vector<double>& fill = thin_x ? xfill : yfill;
    if (fill.empty()) break;
    drop_least_important(thin_x ? xf : yf, fill);
    ++dropped;
  }
  return dropped;
}

// Escape lanes that fit the budget's headroom, taken in ascending order from
// the axis with more remaining, ties to x (lib.rs:981-1045).
void add_escape_within_budget(const vector<double>& xf, const vector<double>& yf, vector<double>& xfill,
                              vector<double>& yfill, vector<double> xe, vector<double> ye, size_t budget) {
  coalesce_fill(xf, xe);
  coalesce_fill(yf, ye);
  coalesce_fill(xfill, xe);
  coalesce_fill(yfill, ye);
  auto cells = [&](size_t a, size_t b) { return sat_mul(xf.size() + a, yf.size() + b); };
  size_t xi = 0, yi = 0;
  for (;;) {
    const bool tx = xi < xe.size(), ty = yi < ye.size();
    if (!tx && !ty) break;
    const bool take_x = (tx && ty) ? (xe.size() - xi) >= (ye.size() - yi) : tx;
    if (take_x) {
      if (cells(xfill.size() + 1, yfill.size()) <= budget) xfill.push_back(xe[xi]);
      ++xi;
    } else {
      if (cells(xfill.size(), yfill.size() + 1) <= budget) yfill.push_back(ye[yi]);
      ++yi;
    }
  }
}

// The planar cell budget (lib.rs:759-792): `explicit_budget` when the caller
// gives one, else MR_CELL_BUDGET: unset or unparseable = 160,000; "adaptive" =
// max(160,000, features x 3/2); an integer pins it (0 = no ceiling).
size_t cell_budget(size_t nx, size_t ny, uint64_t explicit_budget) {
  if (explicit_budget == UINT64_MAX) return std::numeric_limits<size_t>::max();
  if (explicit_budget) return (size_t)explicit_budget;
  const char* env = getenv("MR_CELL_BUDGET");
  if (!env) return kBudgetFloor;
  string t(env);
  const size_t a = t.find_first_not_of(" \t\n\r\f\v");
  if (a == string::npos) return kBudgetFloor;
  t = t.substr(a, t.find_last_not_of(" \t\n\r\f\v") - a + 1);
  string lower(t);
  for (char& c : lower) c = (char)tolower((unsigned char)c);
  if (lower == "adaptive") return std::max(sat_mul(sat_mul(nx, ny), 3) / 2, kBudgetFloor);
  // usize::from_str: ASCII digits, an optional leading '+', nothing else
  size_t i = (t[0] == '+') ? 1 : 0;
  if (i >= t.size()) return kBudgetFloor;
  unsigned long long v = 0;
  for (; i < t.size(); ++i) {
    if (!isdigit((unsigned char)t[i])) return kBudgetFloor;
    const unsigned d = (unsigned)(t[i] - '0');
    if (v > (std::numeric_limits<unsigned long long>::max() - d) / 10) return kBudgetFloor;
    v = v * 10 + d;
  }
  return v == 0 ? std::numeric_limits<size_t>::max() : (size_t)v;
}

struct LineSet {
  vector<double> x, y;
  size_t dropped = 0;
};

// build_grid_lines (lib.rs:848-948).
LineSet grid_lines(const Srj& s, double track_w, double clearance, double escape_clearance, uint64_t budget_arg) {
  LineSet L;
  vector<double>& xs = L.x;
  vector<double>& ys = L.y;
  xs = {s.min_x, s.max_x};
  ys = {s.min_y, s.max_y};
  for (const Conn& c : s.conns)
    for (const Pt& p : c.pts) {
      xs.push_back(p.x);
      ys.push_back(p.y);
    }
  for (const Obs& o : s.obstacles) {
    xs.push_back(o.c.x - o.w / 2.0);
    xs.push_back(o.c.x + o.w / 2.0);
    ys.push_back(o.c.y - o.h / 2.0);
    ys.push_back(o.c.y + o.h / 2.0);
  }
  sort_dedup(xs);
  sort_dedup(ys);
  vector<double> xfill = fill_lines(xs, track_w, clearance);
  vector<double> yfill = fill_lines(ys, track_w, clearance);
  const size_t budget = cell_budget(xs.size(), ys.size(), budget_arg);
  L.dropped = enforce_budget(xs, ys, xfill, yfill, budget);
  if (std::isfinite(escape_clearance) && escape_clearance > 0 && escape_clearance + kLineEps < clearance) {
    add_escape_within_budget(xs, ys, xfill, yfill, escape_lines(xs, track_w, clearance, escape_clearance),
                             escape_lines(ys, track_w, clearance, escape_clearance), budget);
  }
  xs.insert(xs.end(), xfill.begin(), xfill.end());
  ys.insert(ys.end(), yfill.begin(), yfill.end());
  sort_dedup(xs);
  sort_dedup(ys);
  return L;
}

// Node of a continuous coordinate: the number of interior Voronoi boundaries
// 0.5 * (lines[i-1] + lines[i]) at or below p (lib.rs:1275-1301).
uint32_t nearest_line(const vector<double>& lines, double p) {
  if (!std::isfinite(p)) return p > 0 ? (uint32_t)lines.size() - 1 : 0u;
  size_t lo = 0, hi = lines.size() - 1;
  while (lo < hi) {
    const size_t mid = (lo + hi + 1) / 2;
    if (0.5 * (lines[mid - 1] + lines[mid]) <= p)
      lo = mid;
    else
      hi = mid - 1;
  }
  return (uint32_t)lo;
}

// Last node whose lower Voronoi boundary lies strictly below hi (lib.rs:2098-2120).
uint32_t cell_upper(const vector<double>& lines, double hi) {
  if (!std::isfinite(hi)) return hi > 0 ? (uint32_t)lines.size() - 1 : 0u;
  size_t lo = 0, top = lines.size() - 1;
  while (lo < top) {
    const size_t mid = (lo + top + 1) / 2;
    if (0.5 * (lines[mid - 1] + lines[mid]) < hi)
      lo = mid;
    else
      top = mid - 1;
  }
  return (uint32_t)lo;
}

// [start, end) of the lines inside [lo, hi]; the whole axis when a bound is not
// finite (lib.rs:2055-2064).
std::pair<uint32_t, uint32_t> lines_within(const vector<double>& lines, double lo, double hi) {
  if (!std::isfinite(lo) || !std::isfinite(hi)) return {0u, (uint32_t)lines.size()};
  const double a = std::fmin(lo, hi), b = std::fmax(lo, hi);
  const size_t s = first_not(lines, [&](double v) { return v < a; });
  const size_t e = first_not(lines, [&](double v) { return v <= b; });
  return {(uint32_t)s, (uint32_t)e};
}

// Inclusive band of lines within `c` of line `seed`, walking outward until the
// first line beyond (lib.rs:2071-2096; mr-grid line_span, crates/mr-grid/src/lib.rs:209-229).
std::pair<uint32_t, uint32_t> line_band(const vector<double>& lines, uint32_t seed, double c) {
  const uint32_t n = (uint32_t)lines.size();
  seed = std::min(seed, n - 1);
  const double pos = lines[seed];
  uint32_t lo = seed, hi = seed;
  while (lo > 0 && std::fabs(pos - lines[lo - 1]) <= c) --lo;
  while (hi + 1 < n && std::fabs(lines[hi + 1] - pos) <= c) ++hi;
  return {lo, hi};
}

struct Mapping {
  vector<double> x, y;
  uint32_t w = 1, h = 1, layers = 1;
  size_t plane() const { return (size_t)w * h; }
  size_t cells() const { return plane() * layers; }
  uint32_t idx3(uint32_t cx, uint32_t cy, uint32_t l) const { return (uint32_t)(((size_t)l * h + cy) * w + cx); }
  std::pair<uint32_t, uint32_t> xy_of(double px, double py) const { return {nearest_line(x, px), nearest_line(y, py)}; }
  uint32_t cell(double px, double py, uint32_t l) const {
    const auto q = xy_of(px, py);
    return idx3(q.first, q.second, std::min(l, layers - 1));
  }
};

// ---------------------------------------------------------------------------
// Planar geometry (crates/mr-drc/src/lib.rs:1086-1150, board_outline.rs:698-761)
// ---------------------------------------------------------------------------
using P2 = std::pair<double, double>;
double dist(P2 p, P2 q) {
  const double dx = p.first - q.first, dy = p.second - q.second;
  return std::sqrt(dx * dx + dy * dy);
}
double point_seg_dist(P2 p, P2 a, P2 b) {
  const double abx = b.first - a.first, aby = b.second - a.second;
  const double len2 = abx * abx + aby * aby;
  if (len2 <= std::numeric_limits<double>::min()) return dist(p, a);
  double t = ((p.first - a.first) * abx + (p.second - a.second) * aby) / len2;
  t = std::min(std::max(t, 0.0), 1.0);
  return dist(p, {a.first + t * abx, a.second + t * aby});
}
double clamp01(double v) { return std::min(std::max(v, 0.0), 1.0); }
// Closest distance between segments [p0, p1] and [q0, q1] by the usual
// parametric minimisation with the point cases apart.
double seg_seg_dist(P2 p0, P2 p1, P2 q0, P2 q1) {
  const double mn = std::numeric_limits<double>::min();
  const P2 d1{p1.first - p0.first, p1.second - p0.second};
  const P2 d2{q1.first - q0.first, q1.second - q0.second};
  const P2 r{p0.first - q0.first, p0.second - q0.second};
  const double a = d1.first * d1.first + d1.second * d1.second;
  const double e = d2.first * d2.first + d2.second * d2.second;
  const double f = d2.first * r.first + d2.second * r.second;
  if (a <= mn && e <= mn) return dist(p0, q0);
  double s, t;
  if (a <= mn) {
    s = 0.0;
    t = clamp01(f / e);
  } else {
    const double c = d1.first * r.first + d1.second * r.second;
    if (e <= mn) {
      t = 0.0;
      s = clamp01(-c / a);
    } else {
      const double b = d1.first * d2.first + d1.second * d2.second;
      const double den = a * e - b * b;
      s = den > mn ? clamp01((b * f - c * e) / den) : 0.0;
      double tt = (b * s + f) / e;
      if (tt < 0.0) {
        tt = 0.0;
        s = clamp01(-c / a);
      } else if (tt > 1.0) {
        tt = 1.0;
        s = clamp01((b - c) / a);
      }
      t = tt;
    }
  }
  return dist({p0.first + s * d1.first, p0.second + s * d1.second}, {q0.first + t * d2.first, q0.second + t * d2.second});
}
bool finite2(P2 p) { return std::isfinite(p.first) && std::isfinite(p.second); }
bool same_point(P2 a, P2 b) { return std::fabs(a.first - b.first) <= kGeomEps && std::fabs(a.second - b.second) <= kGeomEps; }
double orient(P2 a, P2 b, P2 c) {
  return (b.first - a.first) * (c.second - a.second) - (b.second - a.second) * (c.first - a.first);
}
bool on_segment(P2 p, P2 a, P2 b) {
  return std::fabs(orient(a, b, p)) <= kGeomEps && p.first >= std::fmin(a.first, b.first) - kGeomEps &&
         p.first <= std::fmax(a.first, b.first) + kGeomEps && p.second >= std::fmin(a.second, b.second) - kGeomEps &&
         p.second <= std::fmax(a.second, b.second) + kGeomEps;
}
bool segments_cross(P2 a0, P2 a1, P2 b0, P2 b1) {
  const double o1 = orient(a0, a1, b0), o2 = orient(a0, a1, b1);
  const double o3 = orient(b0, b1, a0), o4 = orient(b0, b1, a1);
  auto opposite = [](double u, double v) { return (u > kGeomEps && v < -kGeomEps) || (u < -kGeomEps && v > kGeomEps); };
  if (opposite(o1, o2) && opposite(o3, o4)) return true;
  return (std::fabs(o1) <= kGeomEps && on_segment(b0, a0, a1)) || (std::fabs(o2) <= kGeomEps && on_segment(b1, a0, a1)) ||
         (std::fabs(o3) <= kGeomEps && on_segment(a0, b0, b1)) || (std::fabs(o4) <= kGeomEps && on_segment(a1, b0, b1));
}
double axis_gap(double v, double lo, double hi) { return v < lo ? lo - v : (v > hi ? v - hi : 0.0); }
double interval_gap(double a0, double a1, double b0, double b1) {
  return a1 < b0 ? b0 - a1 : (b1 < a0 ? a0 - b1 : 0.0);
}

// ---------------------------------------------------------------------------
// Board outline (board_outline.rs:72-108, :223-461, :585-696)
// ---------------------------------------------------------------------------
struct Outline {
  vector<P2> v;
  double edge_clearance = 0, trace_r = 0, via_r = 0;
  double trace_keepout() const { return edge_clearance + trace_r; }
  double via_keepout() const { return edge_clearance + via_r; }
};

// Remove zero-area hairpins a -> b -> c (collinear legs that reverse), one at a
// time from the lowest index, folding the duplicate a closed hairpin leaves
// (board_outline.rs:607-637).
void remove_spurs(vector<P2>& v) {
  for (;;) {
    const size_t n = v.size();
    if (n < 3) return;
    size_t at = n;
    for (size_t i = 0; i < n && at == n; ++i) {
      const P2 a = v[(i + n - 1) % n], b = v[i], c = v[(i + 1) % n];
      const double ix = b.first - a.first, iy = b.second - a.second;
      const double ox = c.first - b.first, oy = c.second - b.second;
      if (orient(a, b, c) == 0.0 && ix * ox + iy * oy < 0.0) at = i;
    }
    if (at == n) return;
    v.erase(v.begin() + (ptrdiff_t)at);
    if (v.size() >= 2) {
      const size_t prev = (at + v.size() - 1) % v.size(), next = at % v.size();
      if (prev != next && same_point(v[prev], v[next])) v.erase(v.begin() + (ptrdiff_t)next);
    }
  }
}

// 0 = no active outline; 1 = active (filled in `o`); -1 = malformed active
// outline (the caller fails closed).
int outline_from(const Srj& s, double trace_w, double via_pad, Outline& o) {
  const bool has_outline = !s.outline.empty(), has_clear = s.board_edge.first;
  if (!has_outline && !has_clear) return 0;
  if (!std::isfinite(trace_w) || trace_w <= 0) return -1;
  if (!std::isfinite(via_pad) || via_pad <= 0) return -1;
  const double ec = has_clear ? s.board_edge.second : kDefaultEdgeClearance;
  if (!std::isfinite(ec) || ec < 0) return -1;
  vector<P2> v;
  if (has_outline) {
    if (s.outline.size() < 3) return -1;
    v = s.outline;
    for (const P2& p : v)
      if (!finite2(p)) return -1;
    if (v.size() > 3 && same_point(v.front(), v.back())) v.pop_back();
    remove_spurs(v);
    if (v.size() < 3) return -1;
  } else {
    if (!std::isfinite(s.min_x) || !std::isfinite(s.max_x) || !std::isfinite(s.min_y) || !std::isfinite(s.max_y) ||
        s.min_x >= s.max_x || s.min_y >= s.max_y)
      return -1;
    v = {{s.min_x, s.min_y}, {s.max_x, s.min_y}, {s.max_x, s.max_y}, {s.min_x, s.max_y}};
  }
  // a simple polygon: no zero-length edge, non-zero area, no two non-adjacent
  // edges touching (board_outline.rs:657-688)
  const size_t n = v.size();
  double area2 = 0;
  for (size_t i = 0; i < n; ++i) {
    const P2 a = v[i], b = v[(i + 1) % n];
    if (same_point(a, b)) return -1;
  }
  for (size_t i = 0; i < n; ++i) {
    const P2 a = v[i], b = v[(i + 1) % n];
    area2 += a.first * b.second - b.first * a.second;
  }
  if (std::fabs(area2) <= kGeomEps) return -1;
  for (size_t i = 0; i < n; ++i)
    for (size_t j = i + 1; j < n; ++j) {
      if (j == (i + 1) % n || (i == 0 && j + 1 == n)) continue;
      if (segments_cross(v[i], v[(i + 1) % n], v[j], v[(j + 1) % n])) return -1;
    }
  o.v = std::move(v);
  o.edge_clearance = ec;
  o.trace_r = trace_w / 2.0;
  o.via_r = via_pad / 2.0;
  return 1;
}

struct Edge {
  P2 a, b;
  double x0, x1, y0, y1;
};

// The layer-invariant outline mask over the node plane.  Per row (column) a
// list of the edges whose y (x) extent lies within the largest keepout of the
// row's (column's) line; every test below only needs edges that close, so the
// lists give the same answers as a scan of every edge (board_outline.rs:223-461).
vector<uint8_t> outline_plane(const Outline& o, const vector<double>& xl, const vector<double>& yl) {
  const size_t n = o.v.size();
  vector<Edge> E(n);
  for (size_t i = 0; i < n; ++i) {
    const P2 a = o.v[i], b = o.v[(i + 1) % n];
    E[i] = {a, b, std::fmin(a.first, b.first), std::fmax(a.first, b.first), std::fmin(a.second, b.second),
            std::fmax(a.second, b.second)};
  }
  const double radius = std::fmax(std::fmax(o.trace_keepout(), o.via_keepout()), kGeomEps);
  const double reach = radius + kGeomEps;
  vector<vector<uint32_t>> rows(yl.size()), cols(xl.size());
  for (uint32_t e = 0; e < n; ++e) {
    for (size_t r = 0; r < yl.size(); ++r)
      if (axis_gap(yl[r], E[e].y0, E[e].y1) <= reach) rows[r].push_back(e);
    for (size_t c = 0; c < xl.size(); ++c)
      if (axis_gap(xl[c], E[e].x0, E[e].x1) <= reach) cols[c].push_back(e);
  }
  auto inside = [&](P2 p, const vector<uint32_t>& cand) {
    if (!finite2(p)) return false;
    for (uint32_t e : cand)
      if (point_seg_dist(p, E[e].a, E[e].b) <= kGeomEps) return true;
    bool in = false;
    for (uint32_t e : cand) {
      const P2 a = E[e].a, b = E[e].b;
      if ((a.second > p.second) != (b.second > p.second)) {
        const double xc = a.first + (p.second - a.second) * (b.first - a.first) / (b.second - a.second);
        if (xc > p.first) in = !in;
      }
    }
    return in;
  };
  auto step_ok = [&](P2 a, P2 b, const vector<uint32_t>& cand) {
    const double need = o.edge_clearance + o.trace_r;
    const double lim = need + kGeomEps;
    const double x0 = std::fmin(a.first, b.first), x1 = std::fmax(a.first, b.first);
    const double y0 = std::fmin(a.second, b.second), y1 = std::fmax(a.second, b.second);
    for (uint32_t e : cand) {
      if (interval_gap(x0, x1, E[e].x0, E[e].x1) > lim || interval_gap(y0, y1, E[e].y0, E[e].y1) > lim) continue;
      const double d = seg_seg_dist(a, b, E[e].a, E[e].b);
      if (a != b && d <= kGeomEps && segments_cross(a, b, E[e].a, E[e].b)) return false;
      if (!(d - o.trace_r + kGeomEps >= o.edge_clearance)) return false;
    }
    return true;
  };
  const size_t W = xl.size(), H = yl.size();
  vector<uint8_t> mask(W * H, 0);
  vector<char> in(W * H, 0);
  const double tk = o.trace_keepout(), vk = o.via_keepout();
  for (size_t y = 0; y < H; ++y) {
    for (size_t x = 0; x < W; ++x) {
      const size_t i = y * W + x;
      const P2 p{xl[x], yl[y]};
      const bool is_in = inside(p, rows[y]);
      bool trace_ok = is_in, via_ok = is_in;
      if (is_in) {
        for (uint32_t e : rows[y]) {
          if (axis_gap(p.first, E[e].x0, E[e].x1) > reach) continue;
          const double d = point_seg_dist(p, E[e].a, E[e].b);
          if (d + kGeomEps < tk) trace_ok = false;
          if (d + kGeomEps < vk) via_ok = false;
          if (!trace_ok && !via_ok) break;
        }
      }
      in[i] = is_in;
      if (!trace_ok) mask[i] |= MRF_BOARD_TRACE_NODE;
      if (!via_ok) mask[i] |= MRF_BOARD_VIA_NODE;
      if (x > 0 && (!in[i - 1] || !is_in || !step_ok({xl[x - 1], yl[y]}, p, rows[y]))) {
        mask[i - 1] |= MRF_BOARD_EDGE_POS_X;
        mask[i] |= MRF_BOARD_EDGE_NEG_X;
      }
      if (y > 0 && (!in[i - W] || !is_in || !step_ok({xl[x], yl[y - 1]}, p, cols[x]))) {
        mask[i - W] |= MRF_BOARD_EDGE_POS_Y;
        mask[i] |= MRF_BOARD_EDGE_NEG_Y;
      }
    }
  }
  return mask;
}

// ---------------------------------------------------------------------------
// Rules
// ---------------------------------------------------------------------------
struct Rules {
  double trace_w, margin, t2p, v2p;
  Opt p2p;
  double via_pad, via_hole;
  Opt h2h;
};

// SimpleRouteJson::uniform_physical_rules (lib.rs:286-528): the coherent typed
// profile, or false.
bool uniform_rules(const Srj& s, Rules& out) {
  auto finite_nonneg = [](const Opt& v, double& r) {
    if (!v.first || !std::isfinite(v.second) || v.second < 0) return false;
    r = v.second;
    return true;
  };
  auto finite_pos = [](const Opt& v, double& r) {
    if (!v.first || !std::isfinite(v.second) || v.second <= 0) return false;
    r = v.second;
    return true;
  };
  // every present alias positive, finite and within kLineEps of the first;
  // false when one is not, else `have` tells whether any was present
  auto aliases = [](std::initializer_list<Opt> vs, bool& have, double& r) {
    have = false;
    for (const Opt& v : vs) {
      if (!v.first) continue;
      if (!std::isfinite(v.second) || !(v.second > 0)) return false;
      if (have && std::fabs(v.second - r) > kLineEps) return false;
      have = true;
      r = v.second;
    }
    return true;
  };
  for (const Obs& o : s.obstacles) {
    if (o.connected.empty() || o.kind != "rect" || !o.has_shape || (o.shape != "rect" && o.shape != "circle") ||
        !std::isfinite(o.c.x) || !std::isfinite(o.c.y) || !std::isfinite(o.w) || !std::isfinite(o.h) || o.w <= 0 ||
        o.h <= 0 || (o.rot.first && (!std::isfinite(o.rot.second) || std::fabs(o.rot.second) > kLineEps)))
      return false;
  }
  const Layers L(s.layer_count);
  auto group_of = [](const Conn& c) { return (c.has_root && !c.root.empty()) ? c.root : c.name; };
  std::map<string, std::set<string>> alias;
  for (const Conn& c : s.conns) {
    const string g = group_of(c);
    if (g.empty()) return false;
    alias[c.name].insert(g);
    if (c.has_root && !c.root.empty()) alias[c.root].insert(g);
  }
  for (const Conn& c : s.conns) {
    const string g = group_of(c);
    for (const Pt& p : c.pts) {
      if (!std::isfinite(p.x) || !std::isfinite(p.y)) return false;
      uint32_t layer = 0;
      if (p.has_layer) {
        const int i = L.index_of(p.layer);
        if (i < 0) return false;
        layer = (uint32_t)i;
      }
      bool covered = false;
      for (const Obs& o : s.obstacles) {
        const vector<uint32_t> ol = obstacle_layers(o, L);
        if (std::find(ol.begin(), ol.end(), layer) == ol.end()) continue;
        if (!(p.x >= o.c.x - o.w / 2.0 && p.x <= o.c.x + o.w / 2.0 && p.y >= o.c.y - o.h / 2.0 &&
              p.y <= o.c.y + o.h / 2.0))
          continue;
        covered = true;
        bool in_shape = false;
        if (o.shape == "rect") {
          in_shape = true;
        } else if (o.shape == "circle") {
          const double dx = (p.x - o.c.x) / (o.w / 2.0), dy = (p.y - o.c.y) / (o.h / 2.0);
          in_shape = dx * dx + dy * dy <= 1.0 + kLineEps;
        }
        if (!in_shape) return false;
        std::set<string> groups;
        for (const string& id : o.connected) {
          auto it = alias.find(id);
          if (it != alias.end()) {
            groups.insert(it->second.begin(), it->second.end());
            continue;
          }
          // producer metadata "pcb_port_<digits>" names no electrical group
          const string pre = "pcb_port_";
          bool meta = id.size() > pre.size() && id.compare(0, pre.size(), pre) == 0;
          for (size_t k = pre.size(); meta && k < id.size(); ++k) meta = isdigit((unsigned char)id[k]) != 0;
          if (!meta) return false;
        }
        if (groups.size() != 1 || !groups.count(g)) return false;
      }
      if (!covered) return false;
    }
  }
  double min_w;
  if (!finite_pos(s.min_trace, min_w)) return false;
  bool has_board_w = false;
  double board_w = 0;
  if (s.nominal_trace.first) {
    if (!finite_pos(s.nominal_trace, board_w)) return false;
    has_board_w = true;
  }
  if (has_board_w && board_w + kLineEps < min_w) return false;
  bool has_common = has_board_w;
  double common = board_w;
  for (const Conn& c : s.conns) {
    bool have;
    double declared = 0;
    if (!aliases({c.nominal, c.width}, have, declared)) return false;
    const double w = have ? declared : (has_board_w ? board_w : min_w);
    if (w + kLineEps < min_w) return false;
    if (has_common) {
      if (std::fabs(w - common) > kLineEps) return false;
    } else {
      has_common = true;
      common = w;
    }
  }
  out.trace_w = has_common ? common : min_w;
  double declared_margin;
  if (!finite_nonneg(s.default_margin, declared_margin)) return false;
  if (s.min_clearance.first) {
    if (!finite_nonneg(s.min_clearance, out.margin)) return false;
  } else {
    out.margin = declared_margin;
  }
  if (!finite_nonneg(s.trace_to_pad, out.t2p)) return false;
  if (!finite_nonneg(s.via_to_pad, out.v2p)) return false;
  bool have;
  if (!aliases({s.pad_dia, s.pad_dia_snake, s.via_dia}, have, out.via_pad) || !have) return false;
  if (!aliases({s.hole_dia, s.hole_dia_snake}, have, out.via_hole) || !have) return false;
  if ((out.via_pad - out.via_hole) / 2.0 + kLineEps < kMinAnnularRing) return false;
  auto opt_nonneg = [](const Opt& v, Opt& r) {
    if (!v.first) {
      r = {false, 0};
      return true;
    }
    if (!std::isfinite(v.second) || v.second < 0) return false;
    r = v;
    return true;
  };
  if (!opt_nonneg(s.pad_to_pad, out.p2p)) return false;
  if (!opt_nonneg(s.hole_to_hole, out.h2h)) return false;
  if (out.t2p + kLineEps < out.margin || out.v2p + kLineEps < out.margin || out.v2p + kLineEps < out.t2p ||
      (out.p2p.first && out.p2p.second + kLineEps < out.margin))
    return false;
  if (out.h2h.first && out.via_pad / 2.0 + out.v2p + kLineEps < out.via_hole / 2.0 + out.h2h.second) return false;
  return true;
}

// The raster's rule inputs (RasterRuleProjection, lib.rs:1428-1436).
struct Projection {
  uint32_t clearance_cells = 0;
  double generic = 0, t2p = 0, v2p = 0, via_pad = 0;
  Opt trace_w{false, 0};
  bool exact = false;
};

// select_clearance_raster_model (lib.rs:689-726).
bool exact_model(double res, uint32_t cells, double min_clear, double via_pad) {
  if (!std::isfinite(res) || res <= 0 || !std::isfinite(min_clear) || min_clear < 0 || !std::isfinite(via_pad) ||
      via_pad < 0)
    return false;
  if (min_clear == 0.0) return via_pad > 0.0;
  return (double)cells * res + kDrcEps >= min_clear + via_pad;
}

// ---------------------------------------------------------------------------
// The raster
// ---------------------------------------------------------------------------
struct NetOut {
  string name;
  uint32_t src, dst;
  vector<uint32_t> pads, vpads;
};
struct Problem {
  Mapping m;
  vector<uint32_t> cost;
  vector<uint8_t> via_forbidden, board;
  vector<NetOut> nets;
  std::map<uint32_t, P2> pins;
  size_t dropped = 0;
};

bool in_rect(double px, double py, const Obs& o) {
  return px >= o.c.x - o.w / 2.0 && px <= o.c.x + o.w / 2.0 && py >= o.c.y - o.h / 2.0 && py <= o.c.y + o.h / 2.0;
}

// Whether a via centre at (x, y) lies closer than `margin` (less the DRC
// tolerance) to the obstacle's rectangle, by Euclidean distance; malformed
// geometry counts as inside (lib.rs:2033-2053).
bool in_via_halo(const Obs& o, double x, double y, double margin) {
  if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(margin) || margin < 0 || !std::isfinite(o.c.x) ||
      !std::isfinite(o.c.y) || !std::isfinite(o.w) || !std::isfinite(o.h) || o.w < 0 || o.h < 0)
    return true;
  const double dx = std::fmax(std::fabs(x - o.c.x) - o.w / 2.0, 0.0);
  const double dy = std::fmax(std::fabs(y - o.c.y) - o.h / 2.0, 0.0);
  return std::hypot(dx, dy) < std::fmax(margin - kDrcEps, 0.0);
}

// Own-pad cells of an endpoint on its layer (lib.rs:1842-1965): the cells of
// every same-layer obstacle whose rect holds the point, plus the point's own
// cell; then every line within `clear` of one of them on both axes, except
// halo cells that fall within `foreign` of a same-layer obstacle not holding
// the point.  Ascending.
vector<uint32_t> pad_cells(double px, double py, uint32_t layer, const Mapping& m, const vector<const Obs*>& lay,
                           double clear, double foreign) {
  vector<std::pair<uint32_t, uint32_t>> own;
  for (const Obs* o : lay) {
    if (px < o->c.x - o->w / 2.0 || px > o->c.x + o->w / 2.0 || py < o->c.y - o->h / 2.0 || py > o->c.y + o->h / 2.0)
      continue;
    const auto lo = m.xy_of(o->c.x - o->w / 2.0, o->c.y - o->h / 2.0);
    const uint32_t x1 = cell_upper(m.x, o->c.x + o->w / 2.0), y1 = cell_upper(m.y, o->c.y + o->h / 2.0);
    if (x1 < lo.first || y1 < lo.second) continue;
    for (uint32_t y = lo.second; y <= y1; ++y)
      for (uint32_t x = lo.first; x <= x1; ++x) own.emplace_back(x, y);
  }
  own.push_back(m.xy_of(px, py));
  std::sort(own.begin(), own.end());
  own.erase(std::unique(own.begin(), own.end()), own.end());
  const double margin = std::fmax(foreign, 0.0);
  auto foreign_at = [&](double lx, double ly) {
    for (const Obs* o : lay) {
      const bool near = lx >= o->c.x - o->w / 2.0 - margin && lx <= o->c.x + o->w / 2.0 + margin &&
                        ly >= o->c.y - o->h / 2.0 - margin && ly <= o->c.y + o->h / 2.0 + margin;
      if (near && !in_rect(px, py, *o)) return true;
    }
    return false;
  };
  const double c = std::fmax(clear, 0.0);
  std::unordered_set<uint32_t> cand;
  for (const auto& q : own) {
    const auto bx = line_band(m.x, q.first, c), by = line_band(m.y, q.second, c);
    for (uint32_t y = by.first; y <= by.second; ++y)
      for (uint32_t x = bx.first; x <= bx.second; ++x) cand.insert(m.idx3(x, y, layer));
  }
  vector<uint32_t> out;
  out.reserve(cand.size());
  for (uint32_t cell : cand) {
    const uint32_t x = cell % m.w, y = (cell / m.w) % m.h;
    if (std::binary_search(own.begin(), own.end(), std::make_pair(x, y)) || !foreign_at(m.x[x], m.y[y]))
      out.push_back(cell);
  }
  std::sort(out.begin(), out.end());
  return out;
}

// Own-pad core cells where a via may override the via keepout (lib.rs:1967-2031):
// the line nodes inside each obstacle that holds the point on the endpoint's
// layer, less those within the via halo of any other obstacle on their layer.
vector<uint32_t> via_pad_cells(double px, double py, uint32_t layer, const Mapping& m, const vector<Obs>& obs,
                               const vector<vector<uint32_t>>& obs_layers, double via_margin) {
  vector<uint32_t> out;
  if (via_margin <= 0) return out;
  vector<char> owned(obs.size(), 0);
  for (size_t i = 0; i < obs.size(); ++i) {
    const Obs& o = obs[i];
    const bool on_layer = std::find(obs_layers[i].begin(), obs_layers[i].end(), layer) != obs_layers[i].end();
    owned[i] = on_layer && std::fabs(px - o.c.x) <= o.w / 2.0 && std::fabs(py - o.c.y) <= o.h / 2.0;
  }
  for (size_t i = 0; i < obs.size(); ++i) {
    if (!owned[i]) continue;
    const Obs& o = obs[i];
    const auto xr = lines_within(m.x, o.c.x - o.w / 2.0, o.c.x + o.w / 2.0);
    const auto yr = lines_within(m.y, o.c.y - o.h / 2.0, o.c.y + o.h / 2.0);
    for (uint32_t y = yr.first; y < yr.second; ++y)
      for (uint32_t x = xr.first; x < xr.second; ++x) out.push_back(m.idx3(x, y, layer));
  }
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  vector<uint32_t> kept;
  for (uint32_t cell : out) {
    const uint32_t x = cell % m.w, y = (cell / m.w) % m.h, l = cell / (uint32_t)m.plane();
    bool blocked = false;
    for (size_t i = 0; i < obs.size() && !blocked; ++i) {
      if (owned[i]) continue;
      if (std::find(obs_layers[i].begin(), obs_layers[i].end(), l) == obs_layers[i].end()) continue;
      blocked = in_via_halo(obs[i], m.x[x], m.y[y], via_margin);
    }
    if (!blocked) kept.push_back(cell);
  }
  return kept;
}

// rasterize_with_layers_model (lib.rs:1438-1763).
Problem rasterize(const Srj& s, double res, uint32_t layer_count, const Projection& pr, uint64_t budget) {
  const Layers L(layer_count);
  Problem P;
  // the declared trace width: the projection's, else minTraceWidth, if positive
  Opt declared = pr.trace_w.first ? pr.trace_w : s.min_trace;
  if (declared.first && !(declared.second > 0)) declared.first = false;
  const double track_w = declared.first ? declared.second : res;
  const double rounded = std::fmax((double)pr.clearance_cells * res, 0.0);
  const double clearance = (pr.exact && std::isfinite(pr.generic) && pr.generic >= 0) ? pr.generic : rounded;
  const double pad_clearance = (pr.exact && std::isfinite(pr.t2p) && pr.t2p >= 0) ? pr.t2p : rounded;
  const double track_block = declared.first ? declared.second / 2.0 : 0.0;
  double pad_band = track_block;
  if (pad_clearance > 0 && declared.first) pad_band = kPadBandK * declared.second;
  // the via-class reservation fraction is 0 (lib.rs:816), so the band stays
  const double block_margin = pad_clearance + pad_band;
  auto positive_min_clearance = [&](double fallback) {
    return (s.min_clearance.first && s.min_clearance.second > 0) ? s.min_clearance.second : fallback;
  };
  double foreign_margin = 0;
  if (block_margin > 0) {
    const double true_clear = pr.t2p > 0 ? pr.t2p : positive_min_clearance(pad_clearance);
    foreign_margin = std::fmin(true_clear, pad_clearance) + pad_band;
  }
  const double escape = pr.t2p > 0 ? pr.t2p : positive_min_clearance(0.0);
  LineSet ls = grid_lines(s, track_w, clearance, escape, budget);
  P.dropped = ls.dropped;
  Mapping& m = P.m;
  m.x = std::move(ls.x);
  m.y = std::move(ls.y);
  if (m.x.empty()) m.x.push_back(0.0);
  if (m.y.empty()) m.y.push_back(0.0);
  m.w = (uint32_t)m.x.size();
  m.h = (uint32_t)m.y.size();
  m.layers = L.size();
  if ((double)m.w * m.h * m.layers >= 4294967295.0) throw Fail("rasterised grid exceeds 2^32 cells");
  const double via_margin = (pr.exact && L.size() > 1 && std::isfinite(pr.via_pad) && pr.via_pad > 0)
                                ? std::fmax(pr.v2p, 0.0) + pr.via_pad / 2.0
                                : 0.0;
  vector<vector<uint32_t>> obs_layers;
  for (const Obs& o : s.obstacles) obs_layers.push_back(obstacle_layers(o, L));
  // obstacle cells (mark_rect_layer, crates/mr-grid/src/lib.rs:92-107)
  vector<uint8_t> blocked(m.cells(), 0);
  for (size_t i = 0; i < s.obstacles.size(); ++i) {
    const Obs& o = s.obstacles[i];
    const auto lo = m.xy_of(o.c.x - o.w / 2.0, o.c.y - o.h / 2.0);
    const uint32_t x1 = cell_upper(m.x, o.c.x + o.w / 2.0), y1 = cell_upper(m.y, o.c.y + o.h / 2.0);
    if (x1 < lo.first || y1 < lo.second) continue;
    for (uint32_t l : obs_layers[i]) {
      if (l >= m.layers) continue;
      for (uint32_t y = lo.second; y <= std::min(y1, m.h - 1); ++y)
        for (uint32_t x = lo.first; x <= std::min(x1, m.w - 1); ++x) blocked[m.idx3(x, y, l)] = 1;
    }
  }
  for (const Conn& c : s.conns)
    for (const Pt& p : c.pts) P.pins[m.cell(p.x, p.y, point_layer(p, L))] = {p.x, p.y};
  // geometric inflation (inflate_clearance, crates/mr-grid/src/lib.rs:131-184):
  // every blocked cell blocks the rectangle of lines within block_margin of its
  // own lines, on its layer; done as a row pass then a column pass over the
  // per-line bands, which yields the same set as one rectangle per seed.
  if (block_margin > 0) {
    vector<uint32_t> bx0(m.w), bx1(m.w), by0(m.h), by1(m.h);
    for (uint32_t x = 0; x < m.w; ++x) std::tie(bx0[x], bx1[x]) = line_band(m.x, x, block_margin);
    for (uint32_t y = 0; y < m.h; ++y) std::tie(by0[y], by1[y]) = line_band(m.y, y, block_margin);
    vector<uint8_t> rowd(m.plane());
    vector<int32_t> diff(std::max(m.w, m.h) + 1);
    for (uint32_t l = 0; l < m.layers; ++l) {
      const uint8_t* B = blocked.data() + (size_t)l * m.plane();
      for (uint32_t y = 0; y < m.h; ++y) {
        std::fill(diff.begin(), diff.begin() + m.w + 1, 0);
        for (uint32_t x = 0; x < m.w; ++x)
          if (B[(size_t)y * m.w + x]) {
            diff[bx0[x]]++;
            diff[bx1[x] + 1]--;
          }
        int32_t run = 0;
        for (uint32_t x = 0; x < m.w; ++x) {
          run += diff[x];
          rowd[(size_t)y * m.w + x] = run > 0;
        }
      }
      uint8_t* O = blocked.data() + (size_t)l * m.plane();
      for (uint32_t x = 0; x < m.w; ++x) {
        std::fill(diff.begin(), diff.begin() + m.h + 1, 0);
        for (uint32_t y = 0; y < m.h; ++y)
          if (rowd[(size_t)y * m.w + x]) {
            diff[by0[y]]++;
            diff[by1[y] + 1]--;
          }
        int32_t run = 0;
        for (uint32_t y = 0; y < m.h; ++y) {
          run += diff[y];
          if (run > 0) O[(size_t)y * m.w + x] = 1;
        }
      }
    }
  }
  P.cost.resize(m.cells());
  for (size_t i = 0; i < m.cells(); ++i) P.cost[i] = blocked[i] ? kObstacle : 1u;
  // static via keepout (lib.rs:1698-1737)
  if (via_margin > 0) {
    vector<uint8_t> vf(m.cells(), 0);
    bool any = false;
    for (size_t i = 0; i < s.obstacles.size(); ++i) {
      const Obs& o = s.obstacles[i];
      const auto xr = lines_within(m.x, o.c.x - o.w / 2.0 - via_margin, o.c.x + o.w / 2.0 + via_margin);
      const auto yr = lines_within(m.y, o.c.y - o.h / 2.0 - via_margin, o.c.y + o.h / 2.0 + via_margin);
      for (uint32_t l : obs_layers[i])
        for (uint32_t y = yr.first; y < yr.second; ++y)
          for (uint32_t x = xr.first; x < xr.second; ++x)
            if (in_via_halo(o, m.x[x], m.y[y], via_margin)) {
              vf[m.idx3(x, y, l)] = 1;
              any = true;
            }
    }
    if (any) P.via_forbidden = std::move(vf);
  }
  // board outline (lib.rs:1744-1796)
  {
    const double board_via = (std::isfinite(pr.via_pad) && pr.via_pad > 0) ? pr.via_pad : 0.45;
    Outline o;
    const int st = outline_from(s, track_w, board_via, o);
    if (st < 0) {
      P.board.assign(m.cells(), MRF_BOARD_TRACE_NODE | MRF_BOARD_VIA_NODE | MRF_BOARD_EDGE_NEG_Y |
                                    MRF_BOARD_EDGE_NEG_X | MRF_BOARD_EDGE_POS_X | MRF_BOARD_EDGE_POS_Y);
    } else if (st > 0) {
      const vector<uint8_t> plane = outline_plane(o, m.x, m.y);
      P.board.reserve(m.cells());
      for (uint32_t l = 0; l < m.layers; ++l) P.board.insert(P.board.end(), plane.begin(), plane.end());
    }
  }
  // nets (decompose_connections, lib.rs:2122-2216)
  vector<vector<const Obs*>> by_layer(m.layers);
  for (size_t i = 0; i < s.obstacles.size(); ++i)
    for (uint32_t l : obs_layers[i])
      if (l < m.layers) by_layer[l].push_back(&s.obstacles[i]);
  for (const Conn& c : s.conns) {
    if (c.pts.size() < 2) continue;
    const size_t segs = c.pts.size() - 1;
    for (size_t k = 0; k < segs; ++k) {
      const Pt& a = c.pts[k];
      const Pt& b = c.pts[k + 1];
      const uint32_t la = point_layer(a, L), lb = point_layer(b, L);
      NetOut n;
      string label = segs == 1 ? c.name : c.name + "#" + std::to_string(k);
      n.name = (c.has_root && !c.root.empty()) ? c.root + "#" + label : label;
      n.src = m.cell(a.x, a.y, la);
      n.dst = m.cell(b.x, b.y, lb);
      n.pads = pad_cells(a.x, a.y, la, m, by_layer[std::min(la, m.layers - 1)], block_margin, foreign_margin);
      const vector<uint32_t> pb = pad_cells(b.x, b.y, lb, m, by_layer[std::min(lb, m.layers - 1)], block_margin,
                                            foreign_margin);
      n.pads.insert(n.pads.end(), pb.begin(), pb.end());
      std::sort(n.pads.begin(), n.pads.end());
      n.pads.erase(std::unique(n.pads.begin(), n.pads.end()), n.pads.end());
      n.vpads = via_pad_cells(a.x, a.y, la, m, s.obstacles, obs_layers, via_margin);
      const vector<uint32_t> vb = via_pad_cells(b.x, b.y, lb, m, s.obstacles, obs_layers, via_margin);
      n.vpads.insert(n.vpads.end(), vb.begin(), vb.end());
      std::sort(n.vpads.begin(), n.vpads.end());
      n.vpads.erase(std::unique(n.vpads.begin(), n.vpads.end()), n.vpads.end());
      vector<uint32_t> keep;
      for (uint32_t v : n.vpads)
        if (std::binary_search(n.pads.begin(), n.pads.end(), v)) keep.push_back(v);
      n.vpads = std::move(keep);
      P.nets.push_back(std::move(n));
    }
  }
  return P;
}

// Library-owned result: the C struct plus the storage its pointers refer to.
struct Holder {
  mrf_srj_problem pub{};
  Problem prob;
  vector<mrf_srj_net> nets;
  vector<uint32_t> pin_cells;
  vector<double> pin_xy;
};

void set_error(mrf_error* err, int code, const string& msg) {
  if (!err) return;
  err->code = code;
  err->net = 0;
  snprintf(err->message, sizeof(err->message), "%s", msg.c_str());
}

void fill_rules(const Rules& r, mrf_srj_rules& o) {
  o.trace_width_mm = r.trace_w;
  o.obstacle_margin_mm = r.margin;
  o.trace_to_pad_clearance_mm = r.t2p;
  o.via_to_pad_clearance_mm = r.v2p;
  o.pad_to_pad_clearance_mm = r.p2p.first ? r.p2p.second : -1.0;
  o.via_pad_diameter_mm = r.via_pad;
  o.via_hole_diameter_mm = r.via_hole;
  o.via_hole_to_hole_clearance_mm = r.h2h.first ? r.h2h.second : -1.0;
}

}  // namespace

extern "C" int mrf_srj_rasterize(const char* json, size_t len, const mrf_srj_params* params, mrf_srj_problem** out,
                                 mrf_error* err) {
  if (out) *out = nullptr;
  if (!json || !out) {
    set_error(err, MRF_BACKEND_UNAVAILABLE, "null input or output");
    return MRF_BACKEND_UNAVAILABLE;
  }
  mrf_srj_params p{};
  if (params) p = *params;
  try {
    Srj s = parse_srj(json, len);
    double res = p.resolution;
    if (!(res > 0)) {
      // default_resolution (crates/mr-cli/src/lib.rs:462-473)
      const double span = std::fmax(s.max_x - s.min_x, s.max_y - s.min_y);
      res = (std::isfinite(span) && span > 0) ? span / 64.0 : 1.0;
    }
    if (!std::isfinite(res) || res <= 0) throw Fail("resolution must be a finite positive number");
    const uint32_t layer_count = std::max<uint32_t>(p.layers ? p.layers : s.layer_count, 1);
    std::unique_ptr<Holder> H(new Holder());
    mrf_srj_problem& o = H->pub;
    Projection pr;
    Rules rules{};
    bool physical = false;
    double trace_w = kCliTraceWidth, min_clear = 0, via_pad = 0;
    if (p.mode == MRF_SRJ_PHYSICAL && p.rules) {
      const mrf_srj_rules& r = *p.rules;
      rules = {r.trace_width_mm, r.obstacle_margin_mm, r.trace_to_pad_clearance_mm, r.via_to_pad_clearance_mm,
               {r.pad_to_pad_clearance_mm >= 0, std::fmax(r.pad_to_pad_clearance_mm, 0.0)},
               r.via_pad_diameter_mm, r.via_hole_diameter_mm,
               {r.via_hole_to_hole_clearance_mm >= 0, std::fmax(r.via_hole_to_hole_clearance_mm, 0.0)}};
      physical = true;
    } else if (p.mode == MRF_SRJ_ROUTE || p.mode == MRF_SRJ_PHYSICAL) {
      physical = uniform_rules(s, rules);
      if (p.mode == MRF_SRJ_PHYSICAL && !physical)
        throw Fail("the problem's typed physical rules do not resolve to a uniform profile");
    }
    if (physical) {
      // rasterize_with_uniform_physical_rules (lib.rs:1400-1426)
      pr.clearance_cells = (rules.margin > 0 && res > 0) ? sat_u32(std::ceil(rules.margin / res)) : 0u;
      pr.generic = rules.margin;
      pr.t2p = rules.t2p;
      pr.v2p = rules.v2p;
      pr.via_pad = rules.via_pad;
      pr.trace_w = {true, rules.trace_w};
      pr.exact = true;
      trace_w = rules.trace_w;
      min_clear = s.min_clearance.first ? std::fmax(s.min_clearance.second, 0.0) : std::fmax(rules.margin, 0.0);
      via_pad = rules.via_pad;
      fill_rules(rules, o.rules);
    } else if (p.mode == MRF_SRJ_ROUTE) {
      // route_problem_impl's legacy branch (crates/mr-cli/src/lib.rs:772-834):
      // minClearance (default 0.15 mm), ceil(minClearance / resolution) cells,
      // a 0.45 mm via pad; with an active board edge the raster takes the
      // CLI's 0.15 mm trace width (legacy_board_edge_raster_input, :688-704)
      min_clear = std::fmax(s.min_clearance.first ? s.min_clearance.second : kCliClearance, 0.0);
      pr.clearance_cells = (min_clear > 0 && res > 0) ? sat_u32(std::ceil(min_clear / res)) : 0u;
      pr.generic = pr.t2p = pr.v2p = min_clear;
      pr.via_pad = via_pad = kCliViaPad;
      pr.exact = exact_model(res, pr.clearance_cells, min_clear, via_pad);
      if ((!s.outline.empty() || s.board_edge.first) && !(s.min_trace.first && s.min_trace.second == kCliTraceWidth))
        s.min_trace = {true, kCliTraceWidth};
      trace_w = kCliTraceWidth;
    } else if (p.mode == MRF_SRJ_LAYERS) {
      pr.clearance_cells = p.clearance_cells;
      min_clear = p.min_clearance_mm;
      via_pad = p.via_pad_mm;
      pr.generic = pr.t2p = pr.v2p = min_clear;
      pr.via_pad = via_pad;
      if (p.model < 0 || p.model > 2) throw Fail("unknown clearance model");
      pr.exact = p.model == 0 ? exact_model(res, pr.clearance_cells, min_clear, via_pad) : p.model == 2;
      trace_w = s.min_trace.first && s.min_trace.second > 0 ? s.min_trace.second : res;
    } else {
      throw Fail("unknown rasterisation mode");
    }
    H->prob = rasterize(s, res, layer_count, pr, p.cell_budget);
    Problem& P = H->prob;
    o.dims = {P.m.w, P.m.h, P.m.layers};
    o.x_lines = P.m.x.data();
    o.y_lines = P.m.y.data();
    o.cost = P.cost.data();
    o.via_forbidden = P.via_forbidden.empty() ? nullptr : P.via_forbidden.data();
    o.board_constraint = P.board.empty() ? nullptr : P.board.data();
    H->nets.resize(P.nets.size());
    for (size_t i = 0; i < P.nets.size(); ++i) {
      NetOut& n = P.nets[i];
      H->nets[i] = {n.name.c_str(), n.src, n.dst, n.pads.data(), n.pads.size(), n.vpads.data(), n.vpads.size()};
    }
    o.n_nets = H->nets.size();
    o.nets = H->nets.data();
    for (const auto& kv : P.pins) {
      H->pin_cells.push_back(kv.first);
      H->pin_xy.push_back(kv.second.first);
      H->pin_xy.push_back(kv.second.second);
    }
    o.n_pins = H->pin_cells.size();
    o.pin_cells = H->pin_cells.data();
    o.pin_xy = H->pin_xy.data();
    o.resolution = res;
    o.trace_width_mm = trace_w;
    o.min_clearance_mm = min_clear;
    o.via_pad_mm = via_pad;
    o.clearance_cells = pr.clearance_cells;
    o.exact_model = pr.exact ? 1 : 0;
    o.physical = physical ? 1 : 0;
    o.fill_dropped = (uint32_t)std::min<size_t>(P.dropped, 0xFFFFFFFFu);
    *out = &H.release()->pub;
    set_error(err, MRF_OK, "ok");
    return MRF_OK;
  } catch (const std::bad_alloc&) {
    set_error(err, MRF_BACKEND_UNAVAILABLE, "out of host memory while rasterising");
  } catch (const std::exception& e) {
    set_error(err, MRF_BACKEND_UNAVAILABLE, e.what());
  }
  return MRF_BACKEND_UNAVAILABLE;
}

extern "C" void mrf_srj_free(mrf_srj_problem* p) {
  // `pub` is the first member of Holder, so the problem's address is the holder's
  if (p) delete reinterpret_cast<Holder*>(p);
}

extern "C" int mrf_srj_uniform_rules(const char* json, size_t len, mrf_srj_rules* out, int* ok, mrf_error* err) {
  if (!json || !out || !ok) {
    set_error(err, MRF_BACKEND_UNAVAILABLE, "null input or output");
    return MRF_BACKEND_UNAVAILABLE;
  }
  try {
    const Srj s = parse_srj(json, len);
    Rules r{};
    *ok = uniform_rules(s, r) ? 1 : 0;
    if (*ok) fill_rules(r, *out);
    set_error(err, MRF_OK, "ok");
    return MRF_OK;
  } catch (const std::exception& e) {
    set_error(err, MRF_BACKEND_UNAVAILABLE, e.what());
  }
  return MRF_BACKEND_UNAVAILABLE;
}

extern "C" int mrf_srj_enforce_budget(const double* x_features, size_t nxf, const double* y_features, size_t nyf,
                                      double* x_fill, size_t* n_x_fill, double* y_fill, size_t* n_y_fill,
                                      uint64_t budget) {
  if ((!x_features && nxf) || (!y_features && nyf) || !n_x_fill || !n_y_fill || (!x_fill && *n_x_fill) ||
      (!y_fill && *n_y_fill))
    return MRF_BACKEND_UNAVAILABLE;
  try {
    const vector<double> xf(x_features, x_features + nxf), yf(y_features, y_features + nyf);
    vector<double> xs(x_fill, x_fill + *n_x_fill), ys(y_fill, y_fill + *n_y_fill);
    enforce_budget(xf, yf, xs, ys, budget == UINT64_MAX ? std::numeric_limits<size_t>::max() : (size_t)budget);
    std::copy(xs.begin(), xs.end(), x_fill);
    std::copy(ys.begin(), ys.end(), y_fill);
    *n_x_fill = xs.size();
    *n_y_fill = ys.size();
    return MRF_OK;
  } catch (const std::exception&) {
    return MRF_BACKEND_UNAVAILABLE;
  }
}

extern "C" int mrf_srj_outline_plane(const char* json, size_t len, double trace_width_mm, double via_pad_mm,
                                     const double* x_lines, size_t nx, const double* y_lines, size_t ny, uint8_t* out,
                                     int* state, mrf_error* err) {
  if (!json || !state || (nx && !x_lines) || (ny && !y_lines) || (nx && ny && !out)) {
    set_error(err, MRF_BACKEND_UNAVAILABLE, "null input or output");
    return MRF_BACKEND_UNAVAILABLE;
  }
  try {
    const Srj s = parse_srj(json, len);
    Outline o;
    *state = outline_from(s, trace_width_mm, via_pad_mm, o);
    if (*state < 0) {
      std::fill(out, out + nx * ny, (uint8_t)(MRF_BOARD_TRACE_NODE | MRF_BOARD_VIA_NODE | MRF_BOARD_EDGE_NEG_Y |
                                              MRF_BOARD_EDGE_NEG_X | MRF_BOARD_EDGE_POS_X | MRF_BOARD_EDGE_POS_Y));
    } else if (*state > 0) {
      const vector<uint8_t> plane = outline_plane(o, vector<double>(x_lines, x_lines + nx),
                                                  vector<double>(y_lines, y_lines + ny));
      std::copy(plane.begin(), plane.end(), out);
    }
    set_error(err, MRF_OK, "ok");
    return MRF_OK;
  } catch (const std::exception& e) {
    set_error(err, MRF_BACKEND_UNAVAILABLE, e.what());
  }
  return MRF_BACKEND_UNAVAILABLE;
}
