"""CPU restatement of metalroute's SimpleRouteJson rasterisation (mr-srj) --
TEST INFRASTRUCTURE ONLY.  Only tests/ may import this module, as the checker
of the library's host producer (metal-route_amd/csrc/srj.cpp, mrf_srj_rasterize).

It follows the reference's steps with brute-force scans where the product
indexes or walks (each function cites the reference file:line it restates;
paths are relative to the reference repository root):

* nearest node / upper cell: a count of the Voronoi boundaries at or below
  (strictly below) the coordinate over every boundary, not a binary search;
* clearance inflation: every seed marks every line within the clearance
  (a scan of all lines, the "slow reference" of crates/mr-grid/src/lib.rs:413-446),
  not the product's per-line bands and separable row / column passes;
* via keepout: every node tested against every obstacle;
* board outline: the exact continuous predicates over every polygon edge
  (the naive mask of crates/mr-srj/src/lib.rs:5523-5560), not the product's
  per-row / per-column edge index;
* own-pad halos: every line tested, not a walk from the seed.

Pinned by (tests/test_srj.py): the asserted values of the reference's own
unit tests for this path (crates/mr-srj/src/lib.rs:2480-5856 and
crates/mr-grid/src/lib.rs:231-490), restated with their inputs; and, for
what those leave open, brute-force invariants (every free node clears every
foreign pad by the centreline margin, the outline mask equals the continuous
predicates node by node, inflation is the union of per-seed rectangles).
No function here is "parity unpinned".
"""
from __future__ import annotations

import bisect
import json
import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

LINE_EPS = 1e-6  # crates/mr-srj/src/lib.rs:661
DRC_EPS = 1e-9  # :664
GEOM_EPS = 1e-9  # crates/mr-srj/src/board_outline.rs:14
BUDGET_FLOOR = 160_000  # lib.rs:728
MAX_FILL_PER_GAP = 256  # lib.rs:1196
PAD_BAND_K = 0.75  # lib.rs:805
EDGE_CLEARANCE_DEFAULT = 0.2  # board_outline.rs:18
OBSTACLE = 0xFFFFFFFF
TRACE_NODE, VIA_NODE, EDGE_NEG_Y, EDGE_NEG_X, EDGE_POS_X, EDGE_POS_Y = 1, 2, 4, 8, 16, 32


# ---- model -------------------------------------------------------------------
def layer_names(count: int) -> List[str]:
    """LayerMap::standard (crates/mr-core/src/lib.rs:259-280)."""
    count = max(count, 1)
    if count == 1:
        return ["top"]
    return ["top"] + [f"inner{i}" for i in range(1, count - 1)] + ["bottom"]


def point_layer(p: dict, names: List[str]) -> int:
    """lib.rs:1798-1803."""
    n = p.get("layer")
    return names.index(n) if n in names else 0


def obstacle_layers(o: dict, names: List[str]) -> List[int]:
    """lib.rs:1808-1822."""
    out = sorted({names.index(n) for n in (o.get("layers") or []) if n in names})
    return out if out else list(range(len(names)))


def rect(o: dict) -> Tuple[float, float, float, float]:
    cx, cy = o["center"]["x"], o["center"]["y"]
    return cx - o["width"] / 2.0, cx + o["width"] / 2.0, cy - o["height"] / 2.0, cy + o["height"] / 2.0


# ---- grid lines (lib.rs:848-1273) ---------------------------------------------
def sort_dedup(v: List[float]) -> List[float]:
    """lib.rs:1188-1192."""
    out: List[float] = []
    for x in sorted(x for x in v if math.isfinite(x)):
        if not out or not abs(x - out[-1]) <= LINE_EPS:
            out.append(x)
    return out


def fill_lines(f: List[float], track_w: float, clearance: float) -> List[float]:
    """lib.rs:1218-1273."""
    if len(f) < 2 or not (math.isfinite(track_w) and track_w > 0):
        return []
    if not (math.isfinite(clearance) and clearance > 0):
        clearance = 0.0
    channel = track_w + 2.0 * clearance
    pitch = max(track_w, LINE_EPS)
    out = []
    for lo, hi in zip(f, f[1:]):
        gap = hi - lo
        if gap < channel:
            continue
        n = min(max(int(math.ceil(gap / pitch)), 2), MAX_FILL_PER_GAP + 1)
        step = gap / n
        out.extend(lo + step * k for k in range(1, n))
    return out


def escape_lines(f, track_w, clearance, esc) -> List[float]:
    """lib.rs:950-979."""
    if len(f) < 2 or not (math.isfinite(track_w) and track_w > 0):
        return []
    coarse, fine = track_w + 2.0 * clearance, track_w + 2.0 * esc
    if not fine < coarse:
        return []
    out = []
    for lo, hi in zip(f, f[1:]):
        gap = hi - lo
        if gap + LINE_EPS >= coarse or gap + LINE_EPS < fine:
            continue
        out.append(lo + gap * 0.5)
    return sort_dedup(out)


def coalesce_fill(features, fill) -> List[float]:
    """lib.rs:1130-1137: drop fill within LINE_EPS of a feature (a scan)."""
    return [f for f in sort_dedup(fill) if not any(abs(x - f) <= LINE_EPS for x in features)]


def drop_least_important(features, fill) -> List[float]:
    """lib.rs:1144-1186."""
    if not fill:
        return fill
    allv = sorted(list(features) + list(fill))
    best, best_gap = 0, math.inf
    for i, f in enumerate(fill):
        pos = bisect.bisect_left(allv, f)
        left = f - allv[pos - 1] if pos > 0 else math.inf
        up = pos + 1
        while up < len(allv) and abs(allv[up] - f) <= LINE_EPS:
            up += 1
        right = allv[up] - f if up < len(allv) else math.inf
        nn = min(left, right)
        if nn < best_gap:
            best, best_gap = i, nn
    return fill[:best] + fill[best + 1:]


def enforce_budget(xf, yf, xfill, yfill, budget):
    """lib.rs:1047-1128."""
    xfill, yfill = coalesce_fill(xf, xfill), coalesce_fill(yf, yfill)

    def cells(a, b):
        return (len(xf) + a) * (len(yf) + b)

    if cells(len(xfill), len(yfill)) <= budget:
        return xfill, yfill
    if len(xf) * len(yf) > budget:
        return [], []
    while cells(len(xfill), len(yfill)) > budget:
        thin_x = bool(xfill) and (not yfill or len(xfill) >= len(yfill))
        if thin_x:
            xfill = drop_least_important(xf, xfill)
        elif yfill:
            yfill = drop_least_important(yf, yfill)
        else:
            break
    return xfill, yfill


def add_escape_within_budget(xf, yf, xfill, yfill, xe, ye, budget):
    """lib.rs:981-1045."""
    xe = coalesce_fill(xfill, coalesce_fill(xf, xe))
    ye = coalesce_fill(yfill, coalesce_fill(yf, ye))
    xi = yi = 0
    while xi < len(xe) or yi < len(ye):
        tx, ty = xi < len(xe), yi < len(ye)
        take_x = (len(xe) - xi) >= (len(ye) - yi) if (tx and ty) else tx
        if take_x:
            if (len(xf) + len(xfill) + 1) * (len(yf) + len(yfill)) <= budget:
                xfill = xfill + [xe[xi]]
            xi += 1
        else:
            if (len(xf) + len(xfill)) * (len(yf) + len(yfill) + 1) <= budget:
                yfill = yfill + [ye[yi]]
            yi += 1
    return xfill, yfill


def cell_budget(nx: int, ny: int, explicit: int = 0) -> float:
    """lib.rs:759-792 (env MR_CELL_BUDGET), or an explicit ceiling."""
    if explicit == 2 ** 64 - 1:
        return math.inf
    if explicit:
        return explicit
    raw = os.environ.get("MR_CELL_BUDGET")
    if raw is None:
        return BUDGET_FLOOR
    t = raw.strip()
    if t.lower() == "adaptive":
        return max(nx * ny * 3 // 2, BUDGET_FLOOR)
    body = t[1:] if t.startswith("+") else t
    if body and body.isascii() and body.isdigit():
        v = int(body)
        return math.inf if v == 0 else v
    return BUDGET_FLOOR


def grid_lines(s: dict, track_w, clearance, escape, budget_arg=0):
    """build_grid_lines (lib.rs:848-948)."""
    b = s["bounds"]
    xs, ys = [b["minX"], b["maxX"]], [b["minY"], b["maxY"]]
    for c in s.get("connections") or []:
        for p in c["pointsToConnect"]:
            xs.append(p["x"])
            ys.append(p["y"])
    for o in s.get("obstacles") or []:
        x0, x1, y0, y1 = rect(o)
        xs += [x0, x1]
        ys += [y0, y1]
    xs, ys = sort_dedup(xs), sort_dedup(ys)
    budget = cell_budget(len(xs), len(ys), budget_arg)
    xfill, yfill = enforce_budget(xs, ys, fill_lines(xs, track_w, clearance), fill_lines(ys, track_w, clearance),
                                  budget)
    if math.isfinite(escape) and escape > 0 and escape + LINE_EPS < clearance:
        xfill, yfill = add_escape_within_budget(xs, ys, xfill, yfill, escape_lines(xs, track_w, clearance, escape),
                                                escape_lines(ys, track_w, clearance, escape), budget)
    return sort_dedup(xs + xfill) or [0.0], sort_dedup(ys + yfill) or [0.0]


def nearest_line(lines, p) -> int:
    """lib.rs:1275-1301: the number of boundaries 0.5*(l[i-1]+l[i]) <= p."""
    if not math.isfinite(p):
        return len(lines) - 1 if p > 0 else 0
    return sum(1 for i in range(1, len(lines)) if 0.5 * (lines[i - 1] + lines[i]) <= p)


def cell_upper(lines, hi) -> int:
    """lib.rs:2098-2120: the number of boundaries strictly below hi."""
    if not math.isfinite(hi):
        return len(lines) - 1 if hi > 0 else 0
    return sum(1 for i in range(1, len(lines)) if 0.5 * (lines[i - 1] + lines[i]) < hi)


# ---- planar geometry (crates/mr-drc/src/lib.rs:1086-1150; board_outline.rs:698-761)
def dist(p, q):
    return math.sqrt((p[0] - q[0]) ** 2 + (p[1] - q[1]) ** 2)


def clamp01(v):
    return v if v != v else min(max(v, 0.0), 1.0)


def point_seg_dist(p, a, b):
    abx, aby = b[0] - a[0], b[1] - a[1]
    len2 = abx * abx + aby * aby
    if len2 <= 2.2250738585072014e-308:
        return dist(p, a)
    t = clamp01(((p[0] - a[0]) * abx + (p[1] - a[1]) * aby) / len2)
    return dist(p, (a[0] + t * abx, a[1] + t * aby))


def seg_seg_dist(p0, p1, q0, q1):
    mn = 2.2250738585072014e-308
    d1 = (p1[0] - p0[0], p1[1] - p0[1])
    d2 = (q1[0] - q0[0], q1[1] - q0[1])
    r = (p0[0] - q0[0], p0[1] - q0[1])
    a = d1[0] * d1[0] + d1[1] * d1[1]
    e = d2[0] * d2[0] + d2[1] * d2[1]
    f = d2[0] * r[0] + d2[1] * r[1]
    if a <= mn and e <= mn:
        return dist(p0, q0)
    if a <= mn:
        s, t = 0.0, clamp01(f / e)
    else:
        c = d1[0] * r[0] + d1[1] * r[1]
        if e <= mn:
            t, s = 0.0, clamp01(-c / a)
        else:
            bb = d1[0] * d2[0] + d1[1] * d2[1]
            den = a * e - bb * bb
            s = clamp01((bb * f - c * e) / den) if den > mn else 0.0
            tt = (bb * s + f) / e
            if tt < 0.0:
                tt, s = 0.0, clamp01(-c / a)
            elif tt > 1.0:
                tt, s = 1.0, clamp01((bb - c) / a)
            t = tt
    return dist((p0[0] + s * d1[0], p0[1] + s * d1[1]), (q0[0] + t * d2[0], q0[1] + t * d2[1]))


def orient(a, b, c):
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])


def on_segment(p, a, b):
    return (abs(orient(a, b, p)) <= GEOM_EPS and min(a[0], b[0]) - GEOM_EPS <= p[0] <= max(a[0], b[0]) + GEOM_EPS
            and min(a[1], b[1]) - GEOM_EPS <= p[1] <= max(a[1], b[1]) + GEOM_EPS)


def segments_cross(a0, a1, b0, b1):
    o1, o2, o3, o4 = orient(a0, a1, b0), orient(a0, a1, b1), orient(b0, b1, a0), orient(b0, b1, a1)

    def opp(u, v):
        return (u > GEOM_EPS and v < -GEOM_EPS) or (u < -GEOM_EPS and v > GEOM_EPS)

    if opp(o1, o2) and opp(o3, o4):
        return True
    return ((abs(o1) <= GEOM_EPS and on_segment(b0, a0, a1)) or (abs(o2) <= GEOM_EPS and on_segment(b1, a0, a1))
            or (abs(o3) <= GEOM_EPS and on_segment(a0, b0, b1)) or (abs(o4) <= GEOM_EPS and on_segment(a1, b0, b1)))


def same_point(a, b):
    return abs(a[0] - b[0]) <= GEOM_EPS and abs(a[1] - b[1]) <= GEOM_EPS


# ---- board outline (board_outline.rs:72-206, :585-696) ---------------------------
@dataclass
class Outline:
    v: List[Tuple[float, float]]
    edge_clearance: float
    trace_r: float
    via_r: float

    def edges(self):
        n = len(self.v)
        return [(self.v[i], self.v[(i + 1) % n]) for i in range(n)]

    def inside(self, p):
        """point_in_polygon (board_outline.rs:712-727) over every edge."""
        if any(point_seg_dist(p, a, b) <= GEOM_EPS for a, b in self.edges()):
            return True
        c = False
        for a, b in self.edges():
            if (a[1] > p[1]) != (b[1] > p[1]):
                if a[0] + (p[1] - a[1]) * (b[0] - a[0]) / (b[1] - a[1]) > p[0]:
                    c = not c
        return c

    def point_legal(self, p, required):
        """point_with_keepout_is_legal (board_outline.rs:206-210)."""
        return (math.isfinite(p[0]) and math.isfinite(p[1]) and self.inside(p)
                and min(point_seg_dist(p, a, b) for a, b in self.edges()) + GEOM_EPS >= required)

    def segment_legal(self, a, b):
        """trace_segment_is_legal (board_outline.rs:146-189)."""
        if not (self.inside(a) and self.inside(b)):
            return False
        d = min(seg_seg_dist(a, b, p, q) for p, q in self.edges())
        if a != b and d <= GEOM_EPS and any(segments_cross(a, b, p, q) for p, q in self.edges()):
            return False
        return (d - self.trace_r) + GEOM_EPS >= self.edge_clearance

    def mask_plane(self, xl, yl) -> np.ndarray:
        """The naive mask of crates/mr-srj/src/lib.rs:5523-5560 (one plane)."""
        W, H = len(xl), len(yl)
        m = np.zeros(W * H, dtype=np.uint8)
        tk, vk = self.edge_clearance + self.trace_r, self.edge_clearance + self.via_r
        for y in range(H):
            for x in range(W):
                i, p = y * W + x, (xl[x], yl[y])
                if not self.point_legal(p, tk):
                    m[i] |= TRACE_NODE
                if not self.point_legal(p, vk):
                    m[i] |= VIA_NODE
                if x + 1 < W and not self.segment_legal(p, (xl[x + 1], yl[y])):
                    m[i] |= EDGE_POS_X
                    m[i + 1] |= EDGE_NEG_X
                if y + 1 < H and not self.segment_legal(p, (xl[x], yl[y + 1])):
                    m[i] |= EDGE_POS_Y
                    m[i + W] |= EDGE_NEG_Y
        return m


def outline_from(s: dict, trace_w: float, via_pad: float):
    """BoardOutlineConstraint::from_srj (board_outline.rs:72-108): None when
    inactive, "malformed" when active but invalid, else an Outline."""
    pts = s.get("outline") or []
    has_clear = s.get("minBoardEdgeClearance") is not None
    if not pts and not has_clear:
        return None
    if not (math.isfinite(trace_w) and trace_w > 0) or not (math.isfinite(via_pad) and via_pad > 0):
        return "malformed"
    ec = s["minBoardEdgeClearance"] if has_clear else EDGE_CLEARANCE_DEFAULT
    if not (math.isfinite(ec) and ec >= 0):
        return "malformed"
    if pts:
        if len(pts) < 3:
            return "malformed"
        v = [(p["x"], p["y"]) for p in pts]
        if not all(math.isfinite(a) and math.isfinite(b) for a, b in v):
            return "malformed"
        if len(v) > 3 and same_point(v[0], v[-1]):
            v.pop()
        while len(v) >= 3:  # collinear reversing spurs (board_outline.rs:607-637)
            n = len(v)
            at = next((i for i in range(n) if orient(v[(i - 1) % n], v[i], v[(i + 1) % n]) == 0.0 and
                       (v[i][0] - v[(i - 1) % n][0]) * (v[(i + 1) % n][0] - v[i][0]) +
                       (v[i][1] - v[(i - 1) % n][1]) * (v[(i + 1) % n][1] - v[i][1]) < 0.0), None)
            if at is None:
                break
            del v[at]
            if len(v) >= 2:
                prev, nxt = (at - 1) % len(v), at % len(v)
                if prev != nxt and same_point(v[prev], v[nxt]):
                    del v[nxt]
        if len(v) < 3:
            return "malformed"
    else:
        b = s["bounds"]
        if not (b["minX"] < b["maxX"] and b["minY"] < b["maxY"]):
            return "malformed"
        v = [(b["minX"], b["minY"]), (b["maxX"], b["minY"]), (b["maxX"], b["maxY"]), (b["minX"], b["maxY"])]
    n = len(v)
    if any(same_point(v[i], v[(i + 1) % n]) for i in range(n)):
        return "malformed"
    if abs(sum(v[i][0] * v[(i + 1) % n][1] - v[(i + 1) % n][0] * v[i][1] for i in range(n))) <= GEOM_EPS:
        return "malformed"
    for i in range(n):
        for j in range(i + 1, n):
            if j == (i + 1) % n or (i == 0 and j + 1 == n):
                continue
            if segments_cross(v[i], v[(i + 1) % n], v[j], v[(j + 1) % n]):
                return "malformed"
    return Outline(v, ec, trace_w / 2.0, via_pad / 2.0)


# ---- the raster (lib.rs:1438-1763, :1842-2216) ------------------------------------
def in_via_halo(r, x, y, margin):
    """lib.rs:2033-2053 (finite geometry)."""
    cx, cy = (r[0] + r[1]) / 2.0, (r[2] + r[3]) / 2.0
    return math.hypot(max(abs(x - cx) - (r[1] - r[0]) / 2.0, 0.0),
                      max(abs(y - cy) - (r[3] - r[2]) / 2.0, 0.0)) < max(margin - DRC_EPS, 0.0)


def halo_plane(o: dict, X: np.ndarray, Y: np.ndarray, margin: float) -> np.ndarray:
    """in_via_halo at every node of the plane (an H x W mask)."""
    cx, cy, hw, hh = o["center"]["x"], o["center"]["y"], o["width"] / 2.0, o["height"] / 2.0
    dx = np.maximum(np.abs(X - cx) - hw, 0.0)[None, :]
    dy = np.maximum(np.abs(Y - cy) - hh, 0.0)[:, None]
    return np.hypot(dx, dy) < max(margin - DRC_EPS, 0.0)


@dataclass
class Raster:
    w: int
    h: int
    layers: int
    x: List[float]
    y: List[float]
    cost: np.ndarray
    via_forbidden: Optional[np.ndarray]
    board: Optional[np.ndarray]
    nets: List[dict] = field(default_factory=list)
    pins: Dict[int, Tuple[float, float]] = field(default_factory=dict)

    def idx(self, x, y, l=0):
        return (l * self.h + y) * self.w + x

    def cell(self, px, py, l=0):
        return self.idx(nearest_line(self.x, px), nearest_line(self.y, py), min(l, self.layers - 1))


def rasterize(s: dict, res: float, layers: int, cells: int = 0, generic: float = 0.0, t2p: float = 0.0,
              v2p: float = 0.0, via_pad: float = 0.0, trace_w: Optional[float] = None, exact: bool = False,
              budget: int = 0) -> Raster:
    """rasterize_with_layers_model (lib.rs:1438-1763) over the projection
    (RasterRuleProjection, :1428-1436)."""
    names = layer_names(layers)
    L = len(names)
    obs = s.get("obstacles") or []
    mt = s.get("minTraceWidth")
    declared = trace_w if trace_w is not None else mt
    if declared is not None and not declared > 0:
        declared = None
    track_w = declared if declared is not None else res
    rounded = max(cells * res, 0.0)
    clearance = generic if exact and math.isfinite(generic) and generic >= 0 else rounded
    pad_clear = t2p if exact and math.isfinite(t2p) and t2p >= 0 else rounded
    track_block = declared / 2.0 if declared is not None else 0.0
    pad_band = PAD_BAND_K * declared if pad_clear > 0 and declared is not None else track_block
    block = pad_clear + pad_band
    mc = s.get("minClearance")
    pos_mc = mc if mc is not None and mc > 0 else None
    foreign = 0.0
    if block > 0:
        true_c = t2p if t2p > 0 else (pos_mc if pos_mc is not None else pad_clear)
        foreign = min(true_c, pad_clear) + pad_band
    escape = t2p if t2p > 0 else (pos_mc if pos_mc is not None else 0.0)
    xl, yl = grid_lines(s, track_w, clearance, escape, budget)
    W, H = len(xl), len(yl)
    X, Y = np.array(xl), np.array(yl)
    via_margin = max(v2p, 0.0) + via_pad / 2.0 if exact and L > 1 and math.isfinite(via_pad) and via_pad > 0 \
        else 0.0
    olay = [obstacle_layers(o, names) for o in obs]
    blocked = np.zeros((L, H, W), dtype=bool)
    for o, ls in zip(obs, olay):
        x0, x1, y0, y1 = rect(o)
        cx0, cy0, cx1, cy1 = nearest_line(xl, x0), nearest_line(yl, y0), cell_upper(xl, x1), cell_upper(yl, y1)
        if cx1 < cx0 or cy1 < cy0:
            continue
        for l in ls:
            blocked[l, cy0:cy1 + 1, cx0:cx1 + 1] = True
    if block > 0:
        out = blocked.copy()
        for l, y, x in zip(*np.nonzero(blocked)):
            xs_ok = np.abs(X - X[x]) <= block
            ys_ok = np.abs(Y - Y[y]) <= block
            out[l][np.ix_(ys_ok, xs_ok)] = True
        blocked = out
    cost = np.where(blocked.reshape(-1), OBSTACLE, 1).astype(np.uint32)
    vf = None
    if via_margin > 0:
        vfa = np.zeros((L, H, W), dtype=bool)
        for o, ls in zip(obs, olay):
            hit = halo_plane(o, X, Y, via_margin)
            for l in ls:
                vfa[l] |= hit
        if vfa.any():
            vf = vfa.reshape(-1).astype(np.uint8)
    board = None
    o = outline_from(s, track_w, via_pad if math.isfinite(via_pad) and via_pad > 0 else 0.45)
    if o == "malformed":
        board = np.full(L * W * H, 63, dtype=np.uint8)
    elif o is not None:
        board = np.tile(o.mask_plane(xl, yl), L)
    R = Raster(W, H, L, xl, yl, cost, vf, board)
    for c in s.get("connections") or []:
        for p in c["pointsToConnect"]:
            R.pins[R.cell(p["x"], p["y"], point_layer(p, names))] = (p["x"], p["y"])

    def pad_cells(px, py, layer):
        """lib.rs:1842-1965, every line tested."""
        lay = [rect(o) for o, ls in zip(obs, olay) if layer in ls]
        own = np.zeros((H, W), dtype=bool)
        for r in lay:
            if px < r[0] or px > r[1] or py < r[2] or py > r[3]:
                continue
            cx0, cy0, cx1, cy1 = nearest_line(xl, r[0]), nearest_line(yl, r[2]), cell_upper(xl, r[1]), \
                cell_upper(yl, r[3])
            own[cy0:cy1 + 1, cx0:cx1 + 1] = True
        own[nearest_line(yl, py), nearest_line(xl, px)] = True
        c, m = max(block, 0.0), max(foreign, 0.0)
        cand = np.zeros((H, W), dtype=bool)
        for y, x in zip(*np.nonzero(own)):
            cand[np.ix_(np.abs(Y - Y[y]) <= c, np.abs(X - X[x]) <= c)] = True
        near_foreign = np.zeros((H, W), dtype=bool)
        for r in lay:
            if r[0] <= px <= r[1] and r[2] <= py <= r[3]:
                continue
            near_foreign |= (((Y >= r[2] - m) & (Y <= r[3] + m))[:, None] & ((X >= r[0] - m) & (X <= r[1] + m))[None, :])
        keep = cand & (own | ~near_foreign)
        return {R.idx(int(x), int(y), layer) for y, x in zip(*np.nonzero(keep))}

    def via_cells(px, py, layer):
        """lib.rs:1967-2031, every line tested."""
        if via_margin <= 0:
            return set()
        owned = [layer in ls and abs(px - o["center"]["x"]) <= o["width"] / 2.0 and
                 abs(py - o["center"]["y"]) <= o["height"] / 2.0 for o, ls in zip(obs, olay)]
        core = np.zeros((H, W), dtype=bool)
        for o, own in zip(obs, owned):
            if own:
                r = rect(o)
                core |= ((Y >= r[2]) & (Y <= r[3]))[:, None] & ((X >= r[0]) & (X <= r[1]))[None, :]
        for o, ls, own in zip(obs, olay, owned):
            if not own and layer in ls:
                core &= ~halo_plane(o, X, Y, via_margin)
        return {R.idx(int(x), int(y), layer) for y, x in zip(*np.nonzero(core))}

    for c in s.get("connections") or []:
        pts = c["pointsToConnect"]
        if len(pts) < 2:
            continue
        segs = len(pts) - 1
        root = c.get("rootConnectionName")
        for k in range(segs):
            a, b = pts[k], pts[k + 1]
            la, lb = point_layer(a, names), point_layer(b, names)
            label = c["name"] if segs == 1 else f"{c['name']}#{k}"
            pads = pad_cells(a["x"], a["y"], la) | pad_cells(b["x"], b["y"], lb)
            vp = (via_cells(a["x"], a["y"], la) | via_cells(b["x"], b["y"], lb)) & pads
            R.nets.append({"name": f"{root}#{label}" if root else label, "src": R.cell(a["x"], a["y"], la),
                           "dst": R.cell(b["x"], b["y"], lb), "pads": sorted(pads), "via_pads": sorted(vp)})
    return R


def load(path_or_text) -> dict:
    if isinstance(path_or_text, dict):
        return path_or_text
    if os.path.exists(str(path_or_text)):
        with open(path_or_text) as f:
            return json.load(f)
    return json.loads(path_or_text)
