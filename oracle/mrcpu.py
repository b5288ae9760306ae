"""ctypes bindings for the CPU ORACLE (oracle/build/libmrcpu.so).

TEST INFRASTRUCTURE ONLY: the product engine never imports this module.  Only
tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg use it,
as the reference results and as the timed CPU baseline.

The C library restates the reference CPU router (crate mr-cpu); see the header
of oracle/mrcpu.c for the file:line of every function and what pins it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libmrcpu.so")

MAX = 0xFFFFFFFF
MODE_ALONE = 0
MODE_LEE = 1

BOARD_TRACE_NODE = 0x01
BOARD_VIA_NODE = 0x02
BOARD_EDGE_NEG_Y = 0x04
BOARD_EDGE_NEG_X = 0x08
BOARD_EDGE_POS_X = 0x10
BOARD_EDGE_POS_Y = 0x20

U32P = C.POINTER(C.c_uint32)
U8P = C.POINTER(C.c_uint8)
I32P = C.POINTER(C.c_int32)


class Dims(C.Structure):
    _fields_ = [("w", C.c_uint32), ("h", C.c_uint32), ("layers", C.c_uint32)]


class GridS(C.Structure):
    _fields_ = [("dims", Dims), ("cost", U32P), ("via_forbidden", U8P), ("board", U8P)]


class WindowS(C.Structure):
    _fields_ = [("x0", C.c_uint32), ("y0", C.c_uint32), ("x1", C.c_uint32), ("y1", C.c_uint32)]


class EdgesS(C.Structure):
    _fields_ = [("x", U32P), ("y", U32P), ("via", U32P), ("via_allowed", U8P)]


class SpecS(C.Structure):
    _fields_ = [
        ("grid", C.POINTER(GridS)),
        ("edges", EdgesS),
        ("additive", U32P),
        ("halo_owner", I32P),
        ("halo_penalty", C.c_uint32),
        ("own_id", C.c_int32),
        ("pads", U32P),
        ("n_pads", C.c_size_t),
        ("via_pads", U32P),
        ("n_via_pads", C.c_size_t),
        ("window", WindowS),
        ("mode", C.c_int),
        ("use_heuristic", C.c_int),
        ("min_via_cost", C.c_uint32),
        ("neg_history", U32P),
        ("neg_present", U32P),
        ("neg_halo", U32P),
        ("neg_shared", C.POINTER(C.c_uint64)),
        ("neg_pfac", C.c_uint32),
        ("has_own_path", C.c_int),
        ("own_path", U32P),
        ("n_own_path", C.c_size_t),
        ("own_halo", U32P),
        ("own_halo_count", U32P),
        ("n_own_halo", C.c_size_t),
        ("legal_owner", C.POINTER(C.c_int64)),
        ("legal_halo", C.POINTER(C.c_int64)),
        ("own_group", C.c_int64),
    ]


class NetS(C.Structure):
    _fields_ = [
        ("src", C.c_uint32),
        ("dst", C.c_uint32),
        ("pads", U32P),
        ("n_pads", C.c_size_t),
        ("via_pads", U32P),
        ("n_via_pads", C.c_size_t),
        ("own_id", C.c_int32),
        ("has_own_path", C.c_int),
        ("own_path", U32P),
        ("n_own_path", C.c_size_t),
        ("own_halo", U32P),
        ("own_halo_count", U32P),
        ("n_own_halo", C.c_size_t),
        ("own_group", C.c_int64),
    ]


def build() -> None:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.mrc_edge_cost.restype = C.c_uint32
        L.mrc_edge_cost.argtypes = [C.c_double]
        L.mrc_axis_gap_cost.restype = C.c_uint32
        L.mrc_axis_gap_cost.argtypes = [C.POINTER(C.c_double), C.c_size_t, C.c_uint32]
        L.mrc_axis_edge_costs.restype = None
        L.mrc_axis_edge_costs.argtypes = [C.POINTER(C.c_double), C.c_size_t, C.c_uint32, U32P]
        L.mrc_net_window.restype = WindowS
        L.mrc_net_window.argtypes = [Dims, C.c_uint32, C.c_uint32, U32P, C.c_size_t]
        L.mrc_bfs_distance_field.argtypes = [C.POINTER(GridS), C.c_uint32, U32P]
        L.mrc_sweep_distance_field.argtypes = [C.POINTER(GridS), C.c_uint32, U32P]
        L.mrc_path_from_field.restype = C.c_size_t
        L.mrc_path_from_field.argtypes = [C.POINTER(GridS), U32P, C.c_uint32, C.c_uint32, U32P]
        L.mrc_searchbuf_new.restype = C.c_void_p
        L.mrc_searchbuf_new.argtypes = [C.c_size_t]
        L.mrc_searchbuf_free.argtypes = [C.c_void_p]
        L.mrc_astar.restype = C.c_size_t
        L.mrc_astar.argtypes = [C.c_void_p, C.POINTER(SpecS), C.c_uint32, C.c_uint32, U32P, U32P]
        L.mrc_field_lex.argtypes = [C.POINTER(SpecS), U32P, C.c_size_t, U32P, C.c_size_t, U32P, U32P]
        L.mrc_reconstruct.restype = C.c_size_t
        L.mrc_reconstruct.argtypes = [C.POINTER(SpecS), U32P, U32P, C.c_uint32, U32P]
        L.mrc_provider_path_is_valid.restype = C.c_int
        L.mrc_provider_path_is_valid.argtypes = [
            C.POINTER(GridS), C.c_uint32, C.c_uint32, U32P, C.c_size_t, U32P, C.c_size_t,
            C.POINTER(EdgesS), U32P, C.c_size_t]
        L.mrc_grid_path_cost.restype = C.c_uint32
        L.mrc_grid_path_cost.argtypes = [C.POINTER(GridS), U32P, C.c_size_t, U32P, C.c_size_t]
        L.mrc_step_cost.restype = C.c_uint32
        L.mrc_step_cost.argtypes = [C.POINTER(SpecS), C.c_uint32, C.c_uint32]
        L.mrc_alone_batch.argtypes = [
            C.POINTER(SpecS), C.POINTER(NetS), C.c_size_t, C.POINTER(WindowS), C.c_int, C.c_int,
            U32P, C.POINTER(C.c_uint64), U32P, U32P]
        L.mrc_fields_batch.argtypes = [
            C.POINTER(SpecS), C.POINTER(NetS), C.c_size_t, C.POINTER(U32P), C.POINTER(C.c_size_t),
            C.POINTER(WindowS), C.c_int, U32P, U32P]
        L.mrc_congestion_from.argtypes = [Dims, C.c_size_t, C.POINTER(C.c_uint64), U32P, U32P]
        _lib = L
    return _lib


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _counted(own_halo):
    """CountedCellSet from a {cell: count} dict or a list of cells repeated by multiplicity."""
    if isinstance(own_halo, dict):
        keys = sorted(int(c) for c, k in own_halo.items() if k)
        return keys, [int(own_halo[c]) for c in keys]
    hc, hn = np.unique(np.asarray(list(own_halo), dtype=np.int64), return_counts=True)
    return [int(c) for c in hc], [int(k) for k in hn]


def _p32(a: Optional[np.ndarray]):
    return a.ctypes.data_as(U32P) if a is not None else None


def _p8(a: Optional[np.ndarray]):
    return a.ctypes.data_as(U8P) if a is not None else None


@dataclass
class Grid:
    """mr-core Grid (mr-core/src/lib.rs:394-408); masks may be None (empty Vec)."""

    w: int
    h: int
    layers: int
    cost: np.ndarray
    via_forbidden: Optional[np.ndarray] = None
    board: Optional[np.ndarray] = None

    def __post_init__(self):
        self.cost = _u32(self.cost)
        if self.via_forbidden is not None:
            self.via_forbidden = np.ascontiguousarray(np.asarray(self.via_forbidden, dtype=np.uint8))
        if self.board is not None:
            self.board = np.ascontiguousarray(np.asarray(self.board, dtype=np.uint8))
        self._s = GridS(Dims(self.w, self.h, self.layers), _p32(self.cost), _p8(self.via_forbidden),
                        _p8(self.board))

    @property
    def n(self) -> int:
        return self.w * self.h * self.layers

    def idx3(self, x, y, l=0):
        return (l * self.h + y) * self.w + x

    @staticmethod
    def filled(w, h, layers, fill):
        return Grid(w, h, layers, np.full(w * h * layers, fill, dtype=np.uint32))


@dataclass
class Edges:
    """Rounded request prices (negotiated.rs:530-537); via entries None = forbidden."""

    x: Sequence[int]
    y: Sequence[int]
    via: Sequence[Optional[int]]

    def __post_init__(self):
        self._x = _u32(list(self.x) + [0])
        self._y = _u32(list(self.y) + [0])
        self._v = _u32([0 if v is None else v for v in self.via] + [0])
        self._a = np.ascontiguousarray(np.array([0 if v is None else 1 for v in self.via] + [0], dtype=np.uint8))
        self._s = EdgesS(_p32(self._x), _p32(self._y), _p32(self._v), _p8(self._a))

    @staticmethod
    def uniform(w, h, layers, planar=16, via=160):
        return Edges([planar] * (w - 1), [planar] * (h - 1), [via] * (layers - 1))


@dataclass
class Net:
    src: int
    dst: int
    pads: List[int] = field(default_factory=list)
    via_pads: List[int] = field(default_factory=list)
    own_id: int = -1
    # negotiated searches: own copper (None = not passed) and own halo
    # ({cell: count}, or cells repeated by multiplicity)
    own_path: Optional[List[int]] = None
    own_halo: object = field(default_factory=list)
    own_group: int = -1      # legal searches: the net's group


@dataclass
class Negotiation:
    """Snapshot of route_negotiated's pricing inputs (negotiated.rs:2746-2797)."""
    history: np.ndarray
    present: np.ndarray
    present_halo: Optional[np.ndarray]
    pfac: int
    fused: bool

    def shared(self) -> np.ndarray:
        """shared_congestion of the Jacobi batch (negotiated.rs:1654-1667)."""
        pf = np.uint64(self.pfac * 16)
        p = np.asarray(self.present, dtype=np.uint64) * pf
        h = (np.asarray(self.present_halo, dtype=np.uint64) * pf if self.present_halo is not None
             else np.zeros_like(p))
        s = p + h
        s[s < p] = np.uint64(0xFFFFFFFFFFFFFFFF)
        return s


class Spec:
    """One search specification (see mrc_spec in oracle/mrcpu.h)."""

    def __init__(self, grid: Grid, edges: Optional[Edges] = None, *, mode=MODE_ALONE,
                 pads=(), via_pads=(), window=None, additive=None, halo_owner=None,
                 halo_penalty=0, own_id=-1, use_heuristic=False, min_via_cost=0,
                 negotiation: Optional[Negotiation] = None, own_path=None, own_halo=(),
                 legal_owner=None, legal_halo=None, own_group=-1):
        self.grid = grid
        self.edges = edges if edges is not None else Edges.uniform(grid.w, grid.h, grid.layers, 1, 0)
        self.pads = _u32(sorted(set(int(p) for p in pads)) + [0])
        self.n_pads = len(set(pads))
        self.via_pads = _u32(sorted(set(int(p) for p in via_pads)) + [0])
        self.n_via_pads = len(set(via_pads))
        self.additive = _u32(additive) if additive is not None else None
        self.halo_owner = (np.ascontiguousarray(np.asarray(halo_owner, dtype=np.int32))
                           if halo_owner is not None else None)
        if window is None:
            window = (0, 0, grid.w - 1, grid.h - 1)
        self.s = SpecS(
            C.pointer(grid._s), self.edges._s, _p32(self.additive),
            self.halo_owner.ctypes.data_as(I32P) if self.halo_owner is not None else None,
            halo_penalty, own_id, _p32(self.pads), self.n_pads, _p32(self.via_pads), self.n_via_pads,
            WindowS(*window), mode, 1 if use_heuristic else 0, min_via_cost)
        self.negotiation = negotiation
        if negotiation is not None:
            self._neg = [_u32(negotiation.history), _u32(negotiation.present)]
            self._neg_halo = _u32(negotiation.present_halo) if negotiation.present_halo is not None else None
            self._neg_shared = (np.ascontiguousarray(negotiation.shared(), dtype=np.uint64)
                                if negotiation.fused else None)
            self.s.neg_history = _p32(self._neg[0])
            self.s.neg_present = _p32(self._neg[1])
            self.s.neg_halo = _p32(self._neg_halo)
            self.s.neg_shared = (self._neg_shared.ctypes.data_as(C.POINTER(C.c_uint64))
                                 if self._neg_shared is not None else None)
            self.s.neg_pfac = int(negotiation.pfac)
        if own_path is not None:
            self._own = _u32(sorted(set(int(c) for c in own_path)) + [0])
            self.s.has_own_path = 1
            self.s.own_path = _p32(self._own)
            self.s.n_own_path = len(self._own) - 1
        hc, hn = _counted(own_halo)
        self._own_halo = (_u32(list(hc) + [0]), _u32(list(hn) + [0]))
        self.s.own_halo = _p32(self._own_halo[0])
        self.s.own_halo_count = _p32(self._own_halo[1])
        self.s.n_own_halo = len(hc)
        # route_legal_once hard blocks (copper owners, clearance halo)
        self._legal = [np.ascontiguousarray(np.asarray(a, dtype=np.int64)) if a is not None else None
                       for a in (legal_owner, legal_halo)]
        P64 = C.POINTER(C.c_int64)
        self.s.legal_owner = self._legal[0].ctypes.data_as(P64) if self._legal[0] is not None else None
        self.s.legal_halo = self._legal[1].ctypes.data_as(P64) if self._legal[1] is not None else None
        self.s.own_group = int(own_group)


# ---- thin wrappers ---------------------------------------------------------

def edge_cost(length: float) -> int:
    return lib().mrc_edge_cost(float(length))


def axis_edge_costs(lines: Sequence[float], count: int) -> List[int]:
    arr = (C.c_double * max(len(lines), 1))(*lines)
    out = np.zeros(max(count - 1, 0) + 1, dtype=np.uint32)
    lib().mrc_axis_edge_costs(arr, len(lines), count, _p32(out))
    return [int(v) for v in out[: max(count - 1, 0)]]


def net_window(grid: Grid, src: int, dst: int, pads=()) -> tuple:
    p = _u32(list(pads) + [0])
    w = lib().mrc_net_window(Dims(grid.w, grid.h, grid.layers), src, dst, _p32(p), len(pads))
    return (w.x0, w.y0, w.x1, w.y1)


def bfs_distance_field(grid: Grid, src: int) -> np.ndarray:
    out = np.zeros(grid.n, dtype=np.uint32)
    lib().mrc_bfs_distance_field(C.byref(grid._s), src, _p32(out))
    return out


def sweep_distance_field(grid: Grid, src: int) -> np.ndarray:
    out = np.zeros(grid.n, dtype=np.uint32)
    lib().mrc_sweep_distance_field(C.byref(grid._s), src, _p32(out))
    return out


def path_from_field(grid: Grid, dist: np.ndarray, src: int, dst: int) -> Optional[List[int]]:
    d = _u32(dist)
    out = np.zeros(grid.n + 1, dtype=np.uint32)
    n = lib().mrc_path_from_field(C.byref(grid._s), _p32(d), src, dst, _p32(out))
    return [int(v) for v in out[:n]] if n else None


class SearchBuf:
    def __init__(self, n: int):
        self.p = lib().mrc_searchbuf_new(n)
        self.n = n

    def __del__(self):
        try:
            lib().mrc_searchbuf_free(self.p)
        except Exception:
            pass


def astar(spec: Spec, src: int, dst: int, buf: Optional[SearchBuf] = None):
    """astar_buf restatement: returns (path, search_cost) or None."""
    n = spec.grid.n
    b = buf or SearchBuf(n)
    out = np.zeros(n + 1, dtype=np.uint32)
    cost = C.c_uint32(0)
    k = lib().mrc_astar(b.p, C.byref(spec.s), src, dst, _p32(out), C.byref(cost))
    if k == 0:
        return None
    return [int(v) for v in out[:k]], int(cost.value)


def field_lex(spec: Spec, sources: Sequence[int], forced: Sequence[int] = ()):
    """Complete lexicographic (dist, hops) field; returns (dist, hops)."""
    n = spec.grid.n
    s = _u32(list(sources) + [0])
    f = _u32(list(forced) + [0])
    dist = np.zeros(n, dtype=np.uint32)
    hops = np.zeros(n, dtype=np.uint32)
    lib().mrc_field_lex(C.byref(spec.s), _p32(s), len(sources), _p32(f), len(forced), _p32(dist), _p32(hops))
    return dist, hops


def reconstruct(spec: Spec, dist: np.ndarray, hops: np.ndarray, dst: int) -> Optional[List[int]]:
    n = spec.grid.n
    out = np.zeros(n + 1, dtype=np.uint32)
    d = _u32(dist)
    h = _u32(hops)
    k = lib().mrc_reconstruct(C.byref(spec.s), _p32(d), _p32(h), dst, _p32(out))
    return [int(v) for v in out[:k]] if k else None


def step_cost(spec: Spec, u: int, v: int) -> int:
    return int(lib().mrc_step_cost(C.byref(spec.s), u, v))


def provider_path_is_valid(grid: Grid, net: Net, edges: Edges, path: Sequence[int]) -> bool:
    pads = _u32(sorted(set(net.pads)) + [0])
    vp = _u32(sorted(set(net.via_pads)) + [0])
    p = _u32(list(path) + [0])
    return bool(lib().mrc_provider_path_is_valid(
        C.byref(grid._s), net.src, net.dst, _p32(pads), len(set(net.pads)), _p32(vp),
        len(set(net.via_pads)), C.byref(edges._s), _p32(p), len(path)))


def grid_path_cost(grid: Grid, pads: Sequence[int], path: Sequence[int]) -> int:
    pp = _u32(sorted(set(pads)) + [0])
    p = _u32(list(path) + [0])
    return int(lib().mrc_grid_path_cost(C.byref(grid._s), _p32(pp), len(set(pads)), _p32(p), len(path)))


def _nets_array(nets: Sequence[Net]):
    keep = []
    arr = (NetS * max(len(nets), 1))()
    for i, n in enumerate(nets):
        p = _u32(sorted(set(n.pads)) + [0])
        vp = _u32(sorted(set(n.via_pads)) + [0])
        own = sorted(set(n.own_path)) if n.own_path is not None else []
        op = _u32(own + [0])
        hc, hn = _counted(n.own_halo)
        hcell = _u32(list(hc) + [0])
        hcnt = _u32(list(hn) + [0])
        keep += [p, vp, op, hcell, hcnt]
        arr[i] = NetS(n.src, n.dst, _p32(p), len(set(n.pads)), _p32(vp), len(set(n.via_pads)), n.own_id,
                      1 if n.own_path is not None else 0, _p32(op), len(own), _p32(hcell), _p32(hcnt), len(hc),
                      int(n.own_group))
    return arr, keep


def alone_batch(tmpl: Spec, nets: Sequence[Net], windows: Optional[Sequence[tuple]], *,
                retry_full=True, threads=1):
    """CPU alone batch (negotiated.rs:2115-2184): window first, full-board retry.

    Returns (paths, costs) with None for unreachable nets.
    """
    n = tmpl.grid.n
    nn = len(nets)
    arr, keep = _nets_array(nets)
    wins = None
    if windows is not None:
        wins = (WindowS * max(nn, 1))(*[WindowS(*w) for w in windows])
    # Path capacity per net is the field size (a simple path cannot be longer).
    cells = np.zeros(max(nn, 1) * n + 1, dtype=np.uint32)
    off = np.ascontiguousarray(np.arange(max(nn, 1), dtype=np.uint64) * n)
    plen = np.zeros(max(nn, 1), dtype=np.uint32)
    cost = np.zeros(max(nn, 1), dtype=np.uint32)
    lib().mrc_alone_batch(C.byref(tmpl.s), arr, nn, wins, 1 if retry_full else 0, threads,
                          _p32(cells), off.ctypes.data_as(C.POINTER(C.c_uint64)), _p32(plen), _p32(cost))
    paths, costs = [], []
    for i in range(nn):
        if plen[i]:
            start = i * n
            paths.append([int(v) for v in cells[start: start + int(plen[i])]])
            costs.append(int(cost[i]))
        else:
            paths.append(None)
            costs.append(None)
    return paths, costs


def fields_batch(tmpl: Spec, nets: Sequence[Net], sources: Sequence[Sequence[int]],
                 windows: Optional[Sequence[tuple]] = None, *, threads=1, want_hops=True):
    """Complete multi-source canonical fields, one per net: (dist[F,n], hops[F,n])."""
    n = tmpl.grid.n
    nn = len(nets)
    arr, keep = _nets_array(nets)
    srcs = [_u32(list(s) + [0]) for s in sources]
    sp = (U32P * max(nn, 1))(*[_p32(s) for s in srcs])
    ns = (C.c_size_t * max(nn, 1))(*[len(s) for s in sources])
    wins = None
    if windows is not None:
        wins = (WindowS * max(nn, 1))(*[WindowS(*w) for w in windows])
    dist = np.zeros((max(nn, 1), n), dtype=np.uint32)
    hops = np.zeros((max(nn, 1), n), dtype=np.uint32) if want_hops else None
    lib().mrc_fields_batch(C.byref(tmpl.s), arr, nn, sp, ns, wins, threads, _p32(dist),
                           _p32(hops) if hops is not None else None)
    return dist[:nn], (hops[:nn] if hops is not None else None)


def congestion_from(dims, paths) -> np.ndarray:
    """BoardRoute::congestion_from (mr-core/src/lib.rs:575-593) over a list of
    paths (lists of cells; None = unrouted)."""
    w, h, L = dims
    cells_l = [list(p) if p is not None else [] for p in paths]
    off = np.zeros(len(cells_l) + 1, dtype=np.uint64)
    for i, c in enumerate(cells_l):
        off[i + 1] = off[i] + len(c)
    flat = _u32([v for c in cells_l for v in c] + [0])
    out = np.zeros(max(w * h * L, 1), dtype=np.uint32)
    lib().mrc_congestion_from(Dims(w, h, L), len(cells_l), off.ctypes.data_as(C.POINTER(C.c_uint64)), _p32(flat),
                              _p32(out))
    return out[: w * h * L]
